// ibg_mlp.h — learned SA closure: features, BF16 MFMA network (gfx950).
//
// The network stands in for the point-wise part of the SA source that
// ComptMutSource evaluates per cell (NS_Solver.cpp:700-767): the production
// minus destruction Q = cb1 (1-ft2) S~ nu - (cw1 fw - cb1 ft2/kappa^2)
// (nu/d)^2 / Re.  It predicts the bounded ratio z = Q / (nu S^) with the
// production scale S^ = |Omega| + nu/(kappa^2 d^2 Re), from 8 per-cell
// features (raw, then standardised with the training statistics):
//   0 ln(|Omega| + eps)      vorticity magnitude (the reference's sqrt2|curl/2|)
//   1 a = |Omega| / S^       vorticity share of the production scale, in [0, 1]
//   2 ln(d)                  wall distance
//   3 ln(chi + eps)          chi = nu / (mu/roe), eddy-viscosity ratio
//   4 ln(|u| d Re/nu_l + eps) local Reynolds number
//   5 roe
//   6 p / p_inf
//   7 ln(mu/roe)             laminar kinematic viscosity
// The derivative terms of the source (advection, the cb2 gradient term and
// the frozen face flux) stay in their discrete form.  A cell whose raw
// features leave the training range [lo, hi] falls back to the plain
// closure.
//
// Network: 8 -> 64 -> 64 -> 1, ReLU, BF16 weights and activations, FP32
// accumulation:  X = bf16(x),  H1 = bf16(relu(X W1 + b1)),
// H2 = bf16(relu(H1 W2 + b2)),  y = H2 w3 + b3.  The two inner layers run
// on v_mfma_f32_16x16x32_bf16 (16x16 output tiles, K = 32 per instruction):
//   A fragment  lane l: row l%16, k = 8*(l/16) .. +7
//   B fragment  lane l: col l%16, k = 8*(l/16) .. +7
//   D           lane l: col l%16, rows 4*(l/16) .. +3
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ibg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 8;
constexpr int kMlpHid = 64;
constexpr int kX = 40;     // LDS row stride (bf16) of the padded 32-wide input rows
constexpr int kH = 72;     // LDS row stride (bf16) of the 64-wide hidden rows
// fp32 parameter vector: W1[8][64] b1[64] W2[64][64] b2[64] w3[64] b3
constexpr int kOffW1 = 0, kOffB1 = 512, kOffW2 = 576, kOffB2 = 4672, kOffW3 = 4736, kOffB3 = 4800;
constexpr int kNParam = 4801;

// Device copy of the network in the layouts the kernels read.
struct MlpDev {
  const __bf16 *W1t;   // [64][kX]  W1 transposed, inputs padded to 32
  const __bf16 *W2t;   // [64][kH]  W2 transposed (forward B operand)
  const __bf16 *W2;    // [64][kH]  W2 (backward B operand)
  const __bf16 *W2p;   // [64][64]  W2 permuted and swizzled for the register forward (mlp_fwd_reg)
  const float *w3;     // [64] bf16-rounded values
  const float *b1, *b2, *b3;
  const float *fmean, *fstd;   // [8] feature standardisation
  const float *finv;           // [8] 1 / fstd (x_s = (x - fmean) * finv)
  const float *flo, *fhi;      // [8] raw feature range of the training data
  const float *zst;            // [2] target mean, std
};

// Hidden units: ReLU, one v_max per unit (and 0/1 as the derivative).  A
// smooth gate (SiLU: one v_exp_f32 and one v_rcp_f32 per unit, both
// quarter rate) made the learned-closure stage kernel about 20 % slower for
// the same closure error.
__device__ __forceinline__ float mlp_act(float z) { return fmaxf(z, 0.0f); }
__device__ __forceinline__ float mlp_dact(float z) { return z > 0.0f ? 1.0f : 0.0f; }
// Two units to a bf16 pair with the ReLU applied after the rounding, as a
// signed 16-bit max against 0 (v_cvt_pk_bf16_f32 + v_pk_max_i16): a
// negative value rounds to a bf16 with the sign bit set (or -0), a negative
// integer, so for finite z, bf16(relu(z)) = max_i16(bf16(z), 0) bit for
// bit.  Non-finite z differ: a NaN with the sign bit clear stays NaN here
// while fmaxf(NaN, 0) = 0 in the training path (mlp_act); both only arise
// from diverged weights.  fmaxf on the FP32 values costs two v_max_f32 each
// (IEEE-mode quieting of the MFMA result, then the max).
typedef short s16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x2_t mlp_act_pk(float a, float b) {
  typedef float f32x2_t __attribute__((ext_vector_type(2)));
  const bf16x2_t h = __builtin_convertvector((f32x2_t{a, b}), bf16x2_t);
  const s16x2_t m = __builtin_elementwise_max(__builtin_bit_cast(s16x2_t, h), s16x2_t{0, 0});
  return __builtin_bit_cast(bf16x2_t, m);
}
__device__ __forceinline__ float bf16r(float x) { return (float)(__bf16)x; }

__device__ __forceinline__ floatx4_t mfma16(bf16x8_t a, bf16x8_t b, floatx4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Weights into LDS (cooperative over the workgroup).
struct MlpLds {
  __bf16 *W1t;   // [64][kX]
  __bf16 *W2t;   // [64][kH]
  float *w3, *b1, *b2;   // [64] each
};

__device__ __forceinline__ void mlp_load_lds(const MlpDev &M, const MlpLds &S) {
  for (int i = threadIdx.x; i < 64 * kX; i += blockDim.x) S.W1t[i] = M.W1t[i];
  for (int i = threadIdx.x; i < 64 * kH; i += blockDim.x) S.W2t[i] = M.W2t[i];
  for (int i = threadIdx.x; i < 64; i += blockDim.x) {
    S.w3[i] = M.w3[i];
    S.b1[i] = M.b1[i];
    S.b2[i] = M.b2[i];
  }
}

// Forward of 16 rows by one wave.  Xs: [16][kX] bf16 rows (inputs in columns
// 0..7, zeros in 8..31).  Hs: [16][kH] bf16 scratch of the wave.  The
// pre-activations are optionally kept (Z1s/Z2s [16][kH] as relu' in bf16)
// and the activations H1/H2 (H1s, H2s) for the backward pass.  Returns
// y for rows 4*(l/16)+v in yv[v] of the lanes with l%16 == 0 (summed over
// the 16 lanes of each group), b3 not included.  All 64 lanes must call it
// together; the barriers are workgroup barriers, so every wave of the
// workgroup calls it the same number of times.
__device__ __forceinline__ void mlp_fwd16(const __bf16 *Xs, __bf16 *H1s, __bf16 *H2s,
                                          __bf16 *S1s, __bf16 *S2s, const MlpLds &W,
                                          float *yv) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  // layer 1: [16 x 32] x [32 x 64]
  floatx4_t acc[4];
  const bf16x8_t a1 = *reinterpret_cast<const bf16x8_t *>(Xs + r * kX + 8 * q);
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float bb = W.b1[16 * nt + r];
    acc[nt] = floatx4_t{bb, bb, bb, bb};
    const bf16x8_t b = *reinterpret_cast<const bf16x8_t *>(W.W1t + (16 * nt + r) * kX + 8 * q);
    acc[nt] = mfma16(a1, b, acc[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 4 * q + v, col = 16 * nt + r;
      H1s[row * kH + col] = (__bf16)mlp_act(acc[nt][v]);
      if (S1s) S1s[row * kH + col] = (__bf16)mlp_dact(acc[nt][v]);
    }
  __syncthreads();
  // layer 2: [16 x 64] x [64 x 64]
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float bb = W.b2[16 * nt + r];
    acc[nt] = floatx4_t{bb, bb, bb, bb};
  }
#pragma unroll
  for (int kt = 0; kt < 2; ++kt) {
    const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(H1s + r * kH + 32 * kt + 8 * q);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const bf16x8_t b = *reinterpret_cast<const bf16x8_t *>(W.W2t + (16 * nt + r) * kH + 32 * kt + 8 * q);
      acc[nt] = mfma16(a, b, acc[nt]);
    }
  }
  // layer 3: y = bf16(relu(Z2)) . w3, reduced over the 16 lanes of a group
#pragma unroll
  for (int v = 0; v < 4; ++v) yv[v] = 0.0f;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 4 * q + v, col = 16 * nt + r;
      const __bf16 h = (__bf16)mlp_act(acc[nt][v]);
      if (H2s) H2s[row * kH + col] = h;
      if (S2s) S2s[row * kH + col] = (__bf16)mlp_dact(acc[nt][v]);
      yv[v] += (float)h * W.w3[col];
    }
#pragma unroll
  for (int v = 0; v < 4; ++v)
    for (int off = 1; off < 16; off <<= 1) yv[v] += __shfl_xor(yv[v], off, 64);
  __syncthreads();
}

// Transposed forward with the activations in registers (the stage kernel),
// on v_mfma_f32_32x32x16_bf16 (32x32 output tiles, K = 16):
//   A fragment  lane l: row l%32, k = 8*(l/32) .. +7
//   B fragment  lane l: col l%32, k = 8*(l/32) .. +7
//   D           lane l: col l%32, rows 8*g + 4*(l/32) + v (acc[4g+v], g,v < 4)
// The wave's 64 cells are two blocks of 32: block sb is the cells of the
// lanes with l/32 == sb, each lane holding its own cell's 8 standardised
// features (bf16x8 xf).  Lane (n, h) = lane n + 32 h.
// Layer 1, Z1^T = W1^T X^T per 32 hidden units t: both K halves of the A
// operand hold W1 and only the K half sb of the B operand is nonzero, so
// column n is cell 32 sb + n.  Lane (n, h) receives hidden units
// 32 t + 8 g + 4 h + v of that cell; ReLU and bf16 pairs turn them straight
// into the layer-2 B fragments: K chunk kt (16 inputs) holds, in lane
// (n, h), the units pi(kt, h, e) = 32 (kt/2) + 8 (2 (kt%2) + e/4) + 4 h + e%4,
// e < 8, the order in which W2p stores each row: logical column
// 16 kt + 8 h + e of row j holds W2[pi(kt, h, e)][j], and the 16-byte chunk
// of logical index q of row j sits at chunk q ^ (j % 8) (rows of 128 bytes:
// the eight rows a group of eight lanes reads then fall on distinct banks).
// Layer 2 runs two output tiles of 32 units one after the other, each followed by its ReLU and the output layer as bf16 v_dot2
// with FP32 accumulation; the two halves of a cell are summed across lanes
// n and n + 32.  Returns y (b3 not included) of cell 32 sb + n in lanes
// (n, 0) and (n, 1).  Per block: 10 MFMA, weights from LDS (W1t rows of
// stride W1S of which the first 8 columns are read, W2p as above),
// biases and the bf16 w3 from LDS; no LDS for activations, no barriers.
typedef float floatx16_t __attribute__((ext_vector_type(16)));
__device__ __forceinline__ floatx16_t mfma32(bf16x8_t a, bf16x8_t b, floatx16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// 16 units of a lane: acc[4g + v] of 32-unit tile `base` start at unit
// base + 8 g + 4 h + v (bias and w3 reads, four 4-vectors)
__device__ __forceinline__ floatx16_t mlp_bias16(const float *b, int base, int h) {
  floatx16_t c;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 bb = *reinterpret_cast<const float4 *>(b + base + 8 * g + 4 * h);
    c[4 * g] = bb.x; c[4 * g + 1] = bb.y; c[4 * g + 2] = bb.z; c[4 * g + 3] = bb.w;
  }
  return c;
}

template <int W1S = kX>
__device__ __forceinline__ float mlp_fwd_reg(bf16x8_t xf, int sb, const __bf16 *W1t,
                                             const __bf16 *W2p, const float *b1, const float *b2,
                                             const __bf16 *w3b) {
  const int lane = threadIdx.x & 63;
  const int n = lane & 31, h = lane >> 5;
  const bf16x8_t zero8 = {};
  const bf16x8_t xm = (h == sb) ? xf : zero8;
  // layer 1 -> the four layer-2 B fragments
  bf16x8_t b2f[4];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    // compiler barriers between the tiles hold the LDS reads of one tile
    // back until the previous tile is done (hoisted, they exceed the
    // register budget of eight waves per SIMD)
    asm volatile("" ::: "memory");
    const bf16x8_t aw = *reinterpret_cast<const bf16x8_t *>(W1t + (32 * t + n) * W1S);
    const floatx16_t z = mfma32(aw, xm, mlp_bias16(b1, 32 * t, h));
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2) {
      const int o = 8 * c2;
      const bf16x2_t p0 = mlp_act_pk(z[o], z[o + 1]), p1 = mlp_act_pk(z[o + 2], z[o + 3]);
      const bf16x2_t p2 = mlp_act_pk(z[o + 4], z[o + 5]), p3 = mlp_act_pk(z[o + 6], z[o + 7]);
      b2f[2 * t + c2] = __builtin_shufflevector(__builtin_shufflevector(p0, p1, 0, 1, 2, 3),
                                                __builtin_shufflevector(p2, p3, 0, 1, 2, 3), 0, 1,
                                                2, 3, 4, 5, 6, 7);
    }
  }
  // layer 2 and the output layer, one 32-unit tile at a time
  float y = 0.0f;
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    asm volatile("" ::: "memory");
    floatx16_t acc = mlp_bias16(b2, 32 * nt, h);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      const bf16x8_t aw =
          *reinterpret_cast<const bf16x8_t *>(W2p + (32 * nt + n) * 64 + 8 * ((2 * kt + h) ^ (n & 7)));
      acc = mfma32(aw, b2f[kt], acc);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const bf16x2_t *ww = reinterpret_cast<const bf16x2_t *>(w3b + 32 * nt + 8 * g + 4 * h);
      const bf16x2_t h01 = mlp_act_pk(acc[4 * g], acc[4 * g + 1]);
      const bf16x2_t h23 = mlp_act_pk(acc[4 * g + 2], acc[4 * g + 3]);
      y = __builtin_amdgcn_fdot2_f32_bf16(h01, ww[0], y, false);
      y = __builtin_amdgcn_fdot2_f32_bf16(h23, ww[1], y, false);
    }
  }
  y += __shfl_xor(y, 32, 64);
  return y;
}

// natural log: libm in FP64, the hardware log2 (v_log_f32) in FP32
template <typename R> __device__ __forceinline__ R r_log(R x);
template <> __device__ __forceinline__ double r_log<double>(double x) { return log(x); }
template <> __device__ __forceinline__ float r_log<float>(float x) {
  return __builtin_amdgcn_logf(x) * 0.6931471805599453f;
}

// Velocity gradient of one cell from central differences (physical units).
template <typename R>
struct VelGrad {
  R dux, duy, duz, dvx, dvy, dvz, dwx, dwy, dwz;
};

// |Omega| as the reference forms it: sqrt(2) |curl(u)/2| (NS_Solver.cpp:688-731)
template <typename R>
__device__ __forceinline__ R sa_olen(const VelGrad<R> &g) {
  const R ox = R(0.5) * (g.dwy - g.dvz), oy = R(0.5) * (g.duz - g.dwx), oz = R(0.5) * (g.dvx - g.duy);
  return v_sqrt(R(2.0)) * v_sqrt(ox * ox + oy * oy + oz * oz);
}

// Point-wise production - destruction of the SA source for one cell, the
// terms of NS_Solver.cpp:700-767 (production cb1(1-ft2) S~ nu, destruction
// (cw1 fw - cb1 ft2/kappa^2)(nu/d)^2/Re), from the working variable, the
// laminar viscosity mu, roe, the wall distance and |Omega|.
template <typename R>
__device__ __forceinline__ R sa_pointwise(R nu, R mu0, R roe, R sd, R olen, R Re) {
  const R vis0 = v_div(mu0, roe);
  const R bigx = v_div(nu, vis0);
  const R b3 = bigx * bigx * bigx;
  const R fv1 = v_div(b3, b3 + R(SA::cv1three));
  const R fv2 = R(1.0) - v_div(bigx, R(1.0) + bigx * fv1);
  R segdis, vis2d;
  if (r_abs(sd) < R(0.00000001)) {
    segdis = R(1.0);
    vis2d = R(1.0);
  } else {
    segdis = R(SA::kappa) * sd;
    vis2d = v_div(nu, sd);
  }
  segdis *= segdis;
  R bigs = olen + v_div(v_div(nu * fv2, segdis), Re);
  bigs = (R(0.3) * olen < bigs) ? bigs : R(0.3) * olen;
  const R ft2 = R(SA::ct3) * r_exp(-R(SA::ct4) * bigx * bigx);
  R rr = v_div(v_div(nu, bigs * segdis), Re);
  rr = (R(10.0) < rr) ? R(10.0) : rr;
  const R rsix = r_pow6(rr);
  const R g = rr + R(SA::cw2) * (rsix - rr);
  const R gsix = r_pow6(g);
  const R fw = g * r_root6(v_div(R(65.0), gsix + R(64.0)));
  const R prod = R(SA::cb1) * (R(1.0) - ft2) * bigs * nu;
  const R dest = v_div((R(SA::cw1) * fw - v_div(R(SA::cb1) * ft2, R(SA::kappa * SA::kappa))) * (vis2d * vis2d), Re);
  return prod - dest;
}

// Raw features (see the header) and the production scale nu S^.
// Returns false when the cell is outside the network's domain by
// construction (no wall distance or no working variable).
template <typename R>
__device__ __forceinline__ bool sa_features(R u, R v, R w, R roe, R p, R mu0, R nu, R d,
                                            const VelGrad<R> &g, R olen, R Re, R pinf, R *x,
                                            R &scale) {
  const R eps = R(1e-8);
  const R nul = v_div(mu0, roe);
  const R speed = v_sqrt(u * u + v * v + w * w);
  const R shat = olen + v_div(nu, R(SA::kappa * SA::kappa) * d * d * Re);
  x[0] = r_log(olen + eps);
  x[1] = shat > R(0) ? v_div(olen, shat) : R(0);
  x[2] = r_log(d > eps ? d : eps);
  x[3] = r_log(v_div(nu, nul) + eps);
  x[4] = r_log(v_div(speed * d * Re, nul) + eps);
  x[5] = roe;
  x[6] = v_div(p, pinf);
  x[7] = r_log(nul);
  scale = nu * shat;
  (void)g;
  return d > eps && nu > eps && scale > R(0);
}

}  // namespace ibg
