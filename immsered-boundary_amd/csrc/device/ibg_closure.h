// ibg_closure.h — training-data sampler, training and stand-alone inference
// kernels of the learned SA closure (network in ibg_mlp.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ibg_kernels.h"

namespace ibg {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  // splitmix64 finaliser
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Central-difference velocity gradient and |Omega| of cell c from the
// Box::neib neighbours (ComptMutSource NS_Solver.cpp:662-703), FP64.
template <int D>
__device__ __forceinline__ VelGrad<double> cell_grad(const Pack<double, D> &W, const int *nb27,
                                                     int c, const BlkInfo &bi) {
  const int *nb = nb27 + (size_t)c * 27;
  const Prim<double, D> xl = pload<double, D>(W, nb[0 * 9 + 1 * 3 + 1]);
  const Prim<double, D> xr = pload<double, D>(W, nb[2 * 9 + 1 * 3 + 1]);
  const Prim<double, D> yl = pload<double, D>(W, nb[1 * 9 + 0 * 3 + 1]);
  const Prim<double, D> yr = pload<double, D>(W, nb[1 * 9 + 2 * 3 + 1]);
  VelGrad<double> g;
  g.dux = (xr.u - xl.u) * 0.5 * bi.invh[0];
  g.dvx = (xr.v - xl.v) * 0.5 * bi.invh[0];
  g.dwx = (xr.w - xl.w) * 0.5 * bi.invh[0];
  g.duy = (yr.u - yl.u) * 0.5 * bi.invh[1];
  g.dvy = (yr.v - yl.v) * 0.5 * bi.invh[1];
  g.dwy = (yr.w - yl.w) * 0.5 * bi.invh[1];
  g.duz = g.dvz = g.dwz = 0.0;
  if (D == 3) {
    const Prim<double, D> zl = pload<double, D>(W, nb[1 * 9 + 1 * 3 + 0]);
    const Prim<double, D> zr = pload<double, D>(W, nb[1 * 9 + 1 * 3 + 2]);
    g.duz = (zr.u - zl.u) * 0.5 * bi.invh[2];
    g.dvz = (zr.v - zl.v) * 0.5 * bi.invh[2];
    g.dwz = (zr.w - zl.w) * 0.5 * bi.invh[2];
  }
  return g;
}

// Training data from the solver's own state: every normal fluid cell that
// the update advances (F_UPDATE, signdis > 0) is kept with probability frac
// (seeded hash of the cell), with its raw features and the target
// z = Q / (nu S^) of the plain closure in FP64.  One workgroup per tile, in
// two passes so that the rows land in a fixed order (tile, then cell):
// kWrite = false stores the tile's count of kept cells, k_closure_offsets
// turns the counts into row offsets after the rows already stored, and
// kWrite = true writes each kept cell at its offset plus its rank in the
// tile.  Rows past the capacity are dropped.
template <int D, bool kWrite>
__global__ void __launch_bounds__(kTile) k_closure_sample(const TileInfo *tiles, const BlkInfo *blk,
                                                         const DevPhys *phys, const uint8_t *flags,
                                                         Pack<double, D> W, const double *nu,
                                                         const double *signdis, const int *nb27,
                                                         double frac, uint64_t seed, float *X,
                                                         float *Z, int *Cell, int *tcount,
                                                         const long long *toff, long long cap) {
  __shared__ int wsum[kTile / 64];
  const TileInfo ti = tiles[blockIdx.x];
  const int c = ti.cell0 + threadIdx.x;
  bool keep = (int)threadIdx.x < ti.ncell && (flags[c] & F_UPDATE) && signdis[c] > 0.0;
  if (keep) {
    const double u01 = (double)(mix64(seed ^ ((uint64_t)c * 0x632be59bd9b4e019ull)) >> 11) *
                       (1.0 / 9007199254740992.0);
    keep = u01 < frac;
  }
  double x[kMlpIn], z = 0.0;
  if (keep) {
    const BlkInfo &bi = blk[ti.blk];
    const DevPhys &P = phys[bi.case_id];
    const Prim<double, D> me = pload<double, D>(W, c);
    const double T = me.p / (me.r * P.Rg);
    const double mu0 = sqrt(T * T * T) * ((1.0 + P.SoT) / (T + P.SoT));
    const VelGrad<double> g = cell_grad<D>(W, nb27, c, bi);
    const double olen = sa_olen<double>(g);
    double scale;
    keep = sa_features<double>(me.u, me.v, me.w, me.r, me.p, mu0, nu[c], signdis[c], g, olen,
                               P.Re, P.init[4], x, scale);
    if (keep) z = sa_pointwise<double>(nu[c], mu0, me.r, signdis[c], olen, P.Re) / scale;
  }
  // rank of the kept cell within the tile: wave prefix + earlier waves
  const unsigned long long bal = __ballot(keep);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int below = __popcll(bal & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wid] = __popcll(bal);
  __syncthreads();
  int before = 0, total = 0;
  for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) {
    if (w < wid) before += wsum[w];
    total += wsum[w];
  }
  if constexpr (!kWrite) {
    if (threadIdx.x == 0) tcount[blockIdx.x] = total;
  } else {
    if (keep) {
      const long long k = toff[blockIdx.x] + before + below;
      if (k < cap) {
        for (int j = 0; j < kMlpIn; ++j) X[(size_t)k * kMlpIn + j] = (float)x[j];
        Z[k] = (float)z;
        Cell[k] = c;
      }
    }
  }
}

// Exclusive prefix sum of the per-tile counts after the count of rows
// already stored (count[0]), then count[0] += the total.  One workgroup.
__global__ void __launch_bounds__(1024) k_closure_offsets(const int *tcount, int ntiles,
                                                          long long *toff,
                                                          unsigned long long *count) {
  __shared__ long long part[1024];
  const int t = threadIdx.x, nt = blockDim.x;
  const int per = (ntiles + nt - 1) / nt, a = t * per, b = min(ntiles, a + per);
  long long s = 0;
  for (int i = a; i < b; ++i) s += tcount[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long long run = (long long)count[0];
    for (int k = 0; k < nt; ++k) { const long long v = part[k]; part[k] = run; run += v; }
    count[1] = (unsigned long long)run;   // new total, published below
  }
  __syncthreads();
  long long run = part[t];
  for (int i = a; i < b; ++i) { toff[i] = run; run += tcount[i]; }
  __syncthreads();
  if (t == 0) count[0] = count[1];
}

// Per-block partial moments of the dataset: for each of the 8 features and
// the target: sum, sum of squares, min, max (double).
__global__ void k_closure_moments(const float *X, const float *Z, long long n, double *part) {
  __shared__ double red[256][4];
  double acc[kMlpIn + 1][4];
  for (int j = 0; j <= kMlpIn; ++j) { acc[j][0] = 0; acc[j][1] = 0; acc[j][2] = 1e300; acc[j][3] = -1e300; }
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    for (int j = 0; j <= kMlpIn; ++j) {
      const double v = (j < kMlpIn) ? (double)X[i * kMlpIn + j] : (double)Z[i];
      acc[j][0] += v; acc[j][1] += v * v;
      acc[j][2] = fmin(acc[j][2], v); acc[j][3] = fmax(acc[j][3], v);
    }
  }
  for (int j = 0; j <= kMlpIn; ++j) {
    red[threadIdx.x][0] = acc[j][0]; red[threadIdx.x][1] = acc[j][1];
    red[threadIdx.x][2] = acc[j][2]; red[threadIdx.x][3] = acc[j][3];
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0, s2 = 0, mn = 1e300, mx = -1e300;
      for (int t = 0; t < (int)blockDim.x; ++t) {
        s += red[t][0]; s2 += red[t][1]; mn = fmin(mn, red[t][2]); mx = fmax(mx, red[t][3]);
      }
      double *o = part + ((size_t)blockIdx.x * (kMlpIn + 1) + j) * 4;
      o[0] = s; o[1] = s2; o[2] = mn; o[3] = mx;
    }
    __syncthreads();
  }
}

// fp32 master parameters -> the bf16 layouts of the kernels.
__global__ void k_closure_pack(const float *p, __bf16 *W1t, __bf16 *W2t, __bf16 *W2, __bf16 *W2p,
                               float *w3, float *b1, float *b2, float *b3) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 64 * kX) {
    const int j = t / kX, k = t % kX;
    W1t[t] = (__bf16)(k < kMlpIn ? p[kOffW1 + k * 64 + j] : 0.0f);
  }
  if (t < 64 * kH) {
    const int a = t / kH, b = t % kH;
    W2t[t] = (__bf16)(b < 64 ? p[kOffW2 + b * 64 + a] : 0.0f);   // W2t[j][i] = W2[i][j]
    W2[t] = (__bf16)(b < 64 ? p[kOffW2 + a * 64 + b] : 0.0f);    // W2[i][j]
  }
  if (t < 64 * 64) {
    // W2p row j, physical 16-byte chunk pc holds logical chunk pc ^ (j % 8);
    // logical column 16 kt + 8 h + e = W2[pi(kt, h, e)][j] with
    // pi(kt, h, e) = 32 (kt/2) + 8 (2 (kt%2) + e/4) + 4 h + e%4 (mlp_fwd_reg)
    const int j = t / 64, pc = (t % 64) / 8, e = t % 8;
    const int lc = pc ^ (j & 7), kt = lc / 2, hh = lc % 2;
    const int pk = 32 * (kt / 2) + 8 * (2 * (kt % 2) + e / 4) + 4 * hh + e % 4;
    W2p[t] = (__bf16)p[kOffW2 + pk * 64 + j];
  }
  if (t < 64) {
    w3[t] = (float)(__bf16)p[kOffW3 + t];
    b1[t] = p[kOffB1 + t];
    b2[t] = p[kOffB2 + t];
  }
  if (t == 0) b3[0] = p[kOffB3];
}

// Stand-alone inference over raw feature rows: z = zmean + zstd * y, with
// the register forward of the stage kernel (mlp_fwd_reg): 64 rows per wave,
// one row per lane.
__global__ void __launch_bounds__(256) k_closure_infer(MlpDev M, const float *X, float *Zout, long long n) {
  __shared__ __align__(16) __bf16 sW1t[64 * kX];
  __shared__ __align__(16) __bf16 sW2p[64 * 64];
  __shared__ __align__(16) float sb1[64], sb2[64];
  __shared__ __align__(16) __bf16 sw3[64];
  for (int i = threadIdx.x; i < 64 * kX; i += blockDim.x) sW1t[i] = M.W1t[i];
  for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) sW2p[i] = M.W2p[i];
  for (int i = threadIdx.x; i < 64; i += blockDim.x) { sw3[i] = (__bf16)M.w3[i]; sb1[i] = M.b1[i]; sb2[i] = M.b2[i]; }
  __syncthreads();
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bf16x8_t xf;
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k)
    xf[k] = (__bf16)(row < n ? (X[row * kMlpIn + k] - M.fmean[k]) * M.finv[k] : 0.0f);
  const int lane = threadIdx.x & 63;
  float yown = 0.0f;
#pragma unroll 1
  for (int sb = 0; sb < 2; ++sb) {
    const float y = mlp_fwd_reg(xf, sb, sW1t, sW2p, sb1, sb2, sw3);
    if ((lane >> 5) == sb) yown = y;
  }
  if (row < n) Zout[row] = M.zst[0] + M.zst[1] * (yown + M.b3[0]);
}

// Forward + backward of 64 rows per workgroup (4 waves x 16 rows), loss
// L = (1/B) sum (y - t)^2 on the standardised target.  Rows come from the
// dataset through a seeded hash (sampling with replacement) or, when
// explicit is set, rows blockIdx.x*64 .. +63 in order (the gradient check).
// Each workgroup writes its full gradient into grad[blockIdx.x][kNParam]
// and its loss sum into lossp[blockIdx.x].
struct TrainArgs {
  MlpDev M;
  const float *X, *Z;
  long long n;         // dataset rows
  int batch;           // B (loss normalisation)
  int explicit_rows;
  uint64_t seed;
  float *grad, *lossp;
};

__global__ void __launch_bounds__(256) k_closure_train(TrainArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __bf16 *sW1t = reinterpret_cast<__bf16 *>(smem_raw);
  __bf16 *sW2t = sW1t + 64 * kX;
  __bf16 *sX = sW2t + 64 * kH;
  __bf16 *sH1 = sX + 64 * kX;
  __bf16 *sS1 = sH1 + 64 * kH;
  __bf16 *sH2 = sS1 + 64 * kH;
  __bf16 *sS2 = sH2 + 64 * kH;
  float *sw3 = reinterpret_cast<float *>(sS2 + 64 * kH);
  float *sb1 = sw3 + 64, *sb2 = sb1 + 64, *sy = sb2 + 64, *st = sy + 64, *red = st + 64;   // red [4][64]
  const MlpDev &M = a.M;
  MlpLds Wl{sW1t, sW2t, sw3, sb1, sb2};
  mlp_load_lds(M, Wl);
  // batch rows
  if (threadIdx.x < 64) {
    long long row;
    if (a.explicit_rows) row = (long long)blockIdx.x * 64 + threadIdx.x;
    else row = (long long)(mix64(a.seed + (uint64_t)blockIdx.x * 64 + threadIdx.x) % (uint64_t)a.n);
    float t = 0.0f;
    if (row < a.n) t = (a.Z[row] - M.zst[0]) / M.zst[1];
    st[threadIdx.x] = t;
    for (int k = 0; k < kX; ++k) {
      float v = 0.0f;
      if (k < kMlpIn && row < a.n) v = (a.X[row * kMlpIn + k] - M.fmean[k]) * M.finv[k];
      sX[threadIdx.x * kX + k] = (__bf16)v;
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  float yv[4];
  mlp_fwd16(sX + 16 * w * kX, sH1 + 16 * w * kH, sH2 + 16 * w * kH, sS1 + 16 * w * kH,
            sS2 + 16 * w * kH, Wl, yv);
  if (r == 0)
    for (int v = 0; v < 4; ++v) sy[16 * w + 4 * q + v] = yv[v] + M.b3[0];
  __syncthreads();
  // dy, loss (rows past n in the explicit mode contribute nothing)
  const float invB = 1.0f / (float)a.batch;
  if (threadIdx.x < 64) {
    const bool valid = !a.explicit_rows || (long long)blockIdx.x * 64 + threadIdx.x < a.n;
    const float e = valid ? sy[threadIdx.x] - st[threadIdx.x] : 0.0f;
    sy[threadIdx.x] = 2.0f * e * invB;    // sy now holds dy
    red[threadIdx.x] = e * e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < 64; ++i) s += red[i];
    a.lossp[blockIdx.x] = s;
  }
  // dZ2 = bf16(dy w3 relu'(Z2)), in place of S2
  for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
    const int row = i >> 6, j = i & 63;
    sS2[row * kH + j] = (__bf16)(sy[row] * sw3[j] * (float)sS2[row * kH + j]);
  }
  __syncthreads();
  float *g = a.grad + (size_t)blockIdx.x * kNParam;
  // column sums: dw3[j] = sum H2 dy, db2[j] = sum dZ2
  {
    const int j = threadIdx.x & 63, gr = threadIdx.x >> 6;
    float s3 = 0.0f, s2 = 0.0f;
    for (int row = 16 * gr; row < 16 * gr + 16; ++row) {
      s3 += (float)sH2[row * kH + j] * sy[row];
      s2 += (float)sS2[row * kH + j];
    }
    red[gr * 64 + j] = s3;
    __syncthreads();
    if (threadIdx.x < 64) g[kOffW3 + j] = red[j] + red[64 + j] + red[128 + j] + red[192 + j];
    __syncthreads();
    red[gr * 64 + j] = s2;
    __syncthreads();
    if (threadIdx.x < 64) g[kOffB2 + j] = red[j] + red[64 + j] + red[128 + j] + red[192 + j];
    if (threadIdx.x == 0) {
      float s = 0.0f;
      for (int i = 0; i < 64; ++i) s += sy[i];
      g[kOffB3] = s;
    }
    __syncthreads();
  }
  // dW2[i][j] = sum_r H1[r][i] dZ2[r][j]: wave w -> rows i in 16w..16w+15
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    floatx4_t acc = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
      bf16x8_t av, bv;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int row = 32 * kt + 8 * q + e;
        av[e] = sH1[row * kH + 16 * w + r];
        bv[e] = sS2[row * kH + 16 * nt + r];
      }
      acc = mfma16(av, bv, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) g[kOffW2 + (16 * w + 4 * q + v) * 64 + 16 * nt + r] = acc[v];
  }
  // dZ1 = bf16((dZ2 W2^T) relu'(Z1)) for this wave's rows, in place of S1
  {
    floatx4_t acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
      const bf16x8_t av = *reinterpret_cast<const bf16x8_t *>(sS2 + (16 * w + r) * kH + 32 * kt + 8 * q);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const bf16x8_t bv = *reinterpret_cast<const bf16x8_t *>(M.W2 + (16 * nt + r) * kH + 32 * kt + 8 * q);
        acc[nt] = mfma16(av, bv, acc[nt]);
      }
    }
    __syncthreads();   // every wave has finished reading dZ2/H1 above
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * w + 4 * q + v, col = 16 * nt + r;
        sS1[row * kH + col] = (__bf16)(acc[nt][v] * (float)sS1[row * kH + col]);
      }
  }
  __syncthreads();
  // db1[i] = sum dZ1; dW1[k][i] = sum_r X[r][k] dZ1[r][i]: wave w -> i in 16w..
  {
    const int j = threadIdx.x & 63, gr = threadIdx.x >> 6;
    float s1 = 0.0f;
    for (int row = 16 * gr; row < 16 * gr + 16; ++row) s1 += (float)sS1[row * kH + j];
    red[gr * 64 + j] = s1;
    __syncthreads();
    if (threadIdx.x < 64) g[kOffB1 + j] = red[j] + red[64 + j] + red[128 + j] + red[192 + j];
  }
  {
    floatx4_t acc = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
      bf16x8_t av, bv;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int row = 32 * kt + 8 * q + e;
        av[e] = sX[row * kX + r];
        bv[e] = sS1[row * kH + 16 * w + r];
      }
      acc = mfma16(av, bv, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = 4 * q + v;
      if (k < kMlpIn) g[kOffW1 + k * 64 + 16 * w + r] = acc[v];
    }
  }
}

// Adam over the summed workgroup gradients (beta1 0.9, beta2 0.999, eps 1e-8).
// With gsum_out set, only the summed gradient is written (no update).
__global__ void k_closure_adam(float *p, float *m, float *v, const float *grad, int nwg, float lr,
                               int step, float *gsum_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNParam) return;
  float g = 0.0f;
  for (int k = 0; k < nwg; ++k) g += grad[(size_t)k * kNParam + i];
  if (gsum_out) {
    gsum_out[i] = g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f;
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)step));
  const float vh = vi / (1.0f - powf(b2, (float)step));
  p[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

}  // namespace ibg
