// ibg_kernels.h — HIP kernels of the residual engine (gfx950).
//
// Conventions: R is the arithmetic type of the flux kernels (double in the
// plain mode, float in the mixed mode); D is 2 or 3; NV = D+2 values per
// cell.  Every ghost fill, the CFL time step, the wall force and the
// conserved state are FP64 in both modes.
#pragma once
#include <hip/hip_runtime.h>

#include "ibg_common.h"
#include "ibg_flux.h"
#include "ibg_mlp.h"

#include <type_traits>

namespace ibg {

constexpr int kTile = 256;   // cells per workgroup of the tiled kernels (max)

// ------------------------------------------------------------ pack access
template <typename R, int D>
__device__ __forceinline__ Prim<R, D> pload(const Pack<R, D> &W, int c) {
  const typename V4<R>::T q = W.a[c];
  Prim<R, D> o;
  o.r = q.x; o.u = q.y; o.v = q.z;
  if (D == 2) { o.w = R(0); o.p = q.w; }
  else { o.w = q.w; o.p = W.e[c]; }
  return o;
}

// velocity of cell c (the 4-vector only; w = 0 in 2D)
template <typename R, int D>
__device__ __forceinline__ void uvw_load(const Pack<R, D> &W, int c, R &u, R &v, R &w) {
  const typename V4<R>::T q = W.a[c];
  u = q.y; v = q.z;
  w = (D == 3) ? q.w : R(0);
}

template <typename R, int D>
__device__ __forceinline__ void pstore(const Pack<R, D> &W, int c, R r, R u, R v, R w, R p) {
  typename V4<R>::T q;
  q.x = r; q.y = u; q.z = v;
  if (D == 2) { q.w = p; }
  else { q.w = w; W.e[c] = p; }
  W.a[c] = q;
}

// NV-value records (conserved state, residuals): values 0..3 in a, 4 in e
template <typename R, int D>
__device__ __forceinline__ void vload(const Pack<R, D> &P, int c, R *o) {
  const typename V4<R>::T q = P.a[c];
  o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
  if (D == 3) o[4] = P.e[c];
}

template <typename R, int D>
__device__ __forceinline__ void vstore(const Pack<R, D> &P, int c, const R *v) {
  typename V4<R>::T q;
  q.x = v[0]; q.y = v[1]; q.z = v[2]; q.w = v[3];
  P.a[c] = q;
  if (D == 3) P.e[c] = v[4];
}

// Ghost-fill write: the FP64 master and every stencil copy (see StateOut).
template <int D>
__device__ __forceinline__ void put_all(const StateOut<D> &o, int c, double r, double u, double v,
                                        double w, double p) {
  pstore<double, D>(o.m, c, r, u, v, w, p);
  if (o.nd > 0) pstore<double, D>(o.d1, c, r, u, v, w, p);
  if (o.nd > 1) pstore<double, D>(o.d2, c, r, u, v, w, p);
  if (o.nf > 0) pstore<float, D>(o.f0, c, (float)r, (float)u, (float)v, (float)w, (float)p);
  if (o.nf > 1) pstore<float, D>(o.f1, c, (float)r, (float)u, (float)v, (float)w, (float)p);
  if (o.nf > 2) pstore<float, D>(o.f2, c, (float)r, (float)u, (float)v, (float)w, (float)p);
}

template <int D>
__device__ __forceinline__ void put_nu(const StateOut<D> &o, int c, double nu) {
  if (!o.nm) return;
  o.nm[c] = nu;
  if (o.nd > 0) o.nd1[c] = nu;
  if (o.nd > 1) o.nd2[c] = nu;
  if (o.nf > 0) o.nf0[c] = (float)nu;
  if (o.nf > 1) o.nf1[c] = (float)nu;
  if (o.nf > 2) o.nf2[c] = (float)nu;
}

// error word: ((cell << 16) | code), the smallest cell wins
__device__ __forceinline__ void report_error(unsigned long long *err, int cell, int code) {
  unsigned long long v = ((unsigned long long)(unsigned)cell << 16) | (unsigned)code;
  atomicMin(err, v);
}

// ------------------------------------------------------------ stage kernel
// ----------------------------------------- viscous flux + step preparation
template <typename R, int D>
struct ViscArgs {
  const TileInfo *tiles;
  const int4 *tst;           // L, R from .y, .z
  const int4 *tvs;           // [nface][D-1] tangential stencil
  const int8_t *tdir;
  const uint16_t *slots;
  const BlkInfo *blk;
  const DevPhys *phys;
  Pack<R, D> W;              // step-start stencil state
  Pack<R, D> Rv;
  const R *nu;               // SA working variable (turb)
  R *mutf;                   // [Nn] frozen face-flux term (turb)
  int viscous, turb;
  const unsigned long long *err;
  // optional coarse/fine balance: tile face -> entry of rbalv (-1: none);
  // rbalv NULL when the balance is off
  const int *tmod;
  const R *rbalv;
};

template <typename R, int D>
__device__ __forceinline__ ViscPhys<R, D> visc_phys(const BlkInfo &bi, const DevPhys &P) {
  ViscPhys<R, D> vp;
  vp.Rg = (R)P.Rg; vp.Re = (R)P.Re; vp.qk = (R)P.qk; vp.SoT = (R)P.SoT;
  vp.ih0 = (R)bi.invh[0]; vp.ih1 = (R)bi.invh[1]; vp.ih2 = (R)bi.invh[2];
  vp.ar0 = (R)bi.area[0]; vp.ar1 = (R)bi.area[1]; vp.ar2 = (R)bi.area[2];
  vp.qkt = (R)P.qkt_den;
  return vp;
}

// The viscous flux (and, with the SA closure, the working variable's face
// flux mf) of tile face k (an index into tst / tvs / tdir).
template <typename R, int D>
__device__ __forceinline__ void visc_face(const ViscArgs<R, D> &a, const ViscPhys<R, D> &vp, int k,
                                          R *out, R &mf) {
  const int4 st = a.tst[k];
  const int dir = a.tdir[k];
  const Prim<R, D> L = pload<R, D>(a.W, st.y);
  const Prim<R, D> Rr = pload<R, D>(a.W, st.z);
  Prim<R, D> t[(D == 3) ? 8 : 4];
  const int4 v0 = a.tvs[(size_t)k * (D - 1)];
  t[0] = pload<R, D>(a.W, v0.x); t[1] = pload<R, D>(a.W, v0.y);
  t[2] = pload<R, D>(a.W, v0.z); t[3] = pload<R, D>(a.W, v0.w);
  if (D == 3) {
    const int4 v1 = a.tvs[(size_t)k * (D - 1) + 1];
    t[4 % ((D == 3) ? 8 : 4)] = pload<R, D>(a.W, v1.x);
    t[5 % ((D == 3) ? 8 : 4)] = pload<R, D>(a.W, v1.y);
    t[6 % ((D == 3) ? 8 : 4)] = pload<R, D>(a.W, v1.z);
    t[7 % ((D == 3) ? 8 : 4)] = pload<R, D>(a.W, v1.w);
  }
  mf = R(0);
  if (a.turb)
    visc_flux<R, D>(L, Rr, t, dir, vp, out, true, a.nu[st.y], a.nu[st.z], &mf);
  else
    visc_flux<R, D>(L, Rr, t, dir, vp, out);
}

// Once per step (ComptViscousStress NS_Solver.cpp:365-629 is frozen across
// the stages, NS_Solver.H:669-699): viscous fluxes of the tile faces into
// LDS, then per cell Rv = sum_d (Fv_hi - Fv_lo).  With the SA closure the
// face flux of the working variable (mutflux, :534-542) is reduced the same
// way into sum_d (mutflux_hi - mutflux_lo) / V / Re (:791-800).
// halt: an LDS flag that thread 0 sets from herr (the error word, loaded at
// kernel start) just before the face phase's barrier and that every thread
// tests after it: a halted workgroup writes nothing.
template <typename R, int D>
__device__ __forceinline__ void visc_prep_tile(const ViscArgs<R, D> &a, const TileInfo &ti,
                                               unsigned char *smem_raw, int *halt, bool herr) {
  constexpr int NV = D + 2;
  using V = typename V4<R>::T;
  V *fl4 = reinterpret_cast<V *>(smem_raw);
  R *fl1 = reinterpret_cast<R *>(smem_raw + sizeof(V) * (size_t)ti.nface);
  R *mfl = reinterpret_cast<R *>(smem_raw + (sizeof(V) + (D == 3 ? sizeof(R) : 0)) * (size_t)ti.nface);
  const BlkInfo &bi = a.blk[ti.blk];
  const DevPhys &P = a.phys[bi.case_id];
  if (a.viscous) {
    const ViscPhys<R, D> vp = visc_phys<R, D>(bi, P);
    for (int f = threadIdx.x; f < ti.nface; f += blockDim.x) {
      R out[NV];
      R mf;
      visc_face<R, D>(a, vp, ti.face0 + f, out, mf);
      if (a.rbalv) {
        // a coarse/fine face of the optional balance: its balanced flux
        const int q = a.tmod[ti.face0 + f];
        if (q >= 0)
          for (int v = 0; v < NV; ++v) out[v] = a.rbalv[(size_t)q * NV + v];
      }
      V o;
      o.x = out[0]; o.y = out[1]; o.z = out[2]; o.w = out[3];
      fl4[f] = o;
      if (D == 3) fl1[f] = out[4];
      if (a.turb) mfl[f] = mf;
    }
  }
  if (threadIdx.x == 0) *halt = herr ? 1 : 0;
  __syncthreads();
  if (*halt) return;
  const int i = threadIdx.x;
  if (i >= ti.ncell) return;   // no barrier follows in this function
  const int c = ti.cell0 + i;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  double tot = 0.0;
  if (a.viscous) {
    int sl[2 * D];
    if (D == 2) {
      const ushort4 s4 = *reinterpret_cast<const ushort4 *>(a.slots + (size_t)c * 4);
      sl[0] = s4.x; sl[1] = s4.y; sl[2] = s4.z; sl[3] = s4.w;
    } else {
#pragma unroll
      for (int q = 0; q < 2 * D; ++q) sl[q] = a.slots[(size_t)c * 2 * D + q];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int lo = sl[2 * d], hi = sl[2 * d + 1];
      const V flo = fl4[lo], fhi = fl4[hi];
      s[0] += (double)fhi.x - (double)flo.x;
      s[1] += (double)fhi.y - (double)flo.y;
      s[2] += (double)fhi.z - (double)flo.z;
      s[3] += (double)fhi.w - (double)flo.w;
      if (D == 3) s[4] += (double)fl1[hi] - (double)fl1[lo];
      if (a.turb) {
        // total_face_flux += dirsign[f0] * mutflux, dirsign = {-1, +1}
        tot += -(double)mfl[lo];
        tot += (double)mfl[hi];
      }
    }
  }
  R rv[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) rv[k] = (R)s[k];
  vstore<R, D>(a.Rv, c, rv);
  if (a.turb) a.mutf[c] = (R)(tot / bi.vol / P.Re);
}

template <typename R, int D>
__global__ void __launch_bounds__(kTile) k_visc_prep(ViscArgs<R, D> a) {
  extern __shared__ __align__(32) unsigned char smem_raw[];
  __shared__ int halt;
  const bool herr = halted(a.err);
  const TileInfo ti = a.tiles[dispatch_tile()];
  visc_prep_tile<R, D>(a, ti, smem_raw, &halt, herr);
}

template <typename R, int D>
struct StageArgs {
  const TileInfo *tiles;
  const int4 *tst;           // tile face stencil LL,L,R,RR
  const int8_t *tdir;        // tile face direction
  const uint16_t *slots;     // [Nn][2D] tile-local face slot of each cell face
  const uint8_t *flags;      // [Nn]
  const BlkInfo *blk;
  const DevPhys *phys;
  const double *dt;          // [ncases]
  Pack<R, D> Win;            // stencil state this stage reads
  Pack<R, D> Wout;           // stencil state this stage writes (another buffer)
  Pack<double, D> W64;       // FP64 master: the state at the start of the step
                             // until the last stage writes it (mixed mode)
  Pack<R, D> Rv;             // frozen viscous residual
  // SA closure (turb != 0): working variable in the same buffer roles
  const R *nuin;
  R *nuout;
  double *nu64;
  const R *mutf;             // [Nn] frozen face-flux term of the step
  const double *signdis;
  const float *sd32;         // signdis in FP32 (mixed mode)
  int turb;
  MlpDev mlp;                // learned closure (the ML instantiation only)
  unsigned long long *mlcount;   // [2] cells on the network / on the fallback
  double alpha;
  int limiter;
  int last;                  // last stage: residual partials (+ W64 in mixed)
  double *rpart;             // [ntiles][kRpart]
  double *src;               // optional probe [N][5] written at the last stage
  double *nusrc;             // optional probe [N] (SA source, last stage)
  unsigned long long *err;
  // first stage of a step: the frozen viscous residual of the tile first
  // (the work of k_visc_prep, into the same Rv / mutf arrays)
  int fuse_visc;
  ViscArgs<R, D> va;
  // persistent schedule (the P instantiations): gridDim.x workgroups loop
  // over the ntile tiles of the launch (XCD-contiguous runs), with the
  // learned closure's weights staged once per workgroup at byte mlw_off of
  // the dynamic LDS
  int ntile;
  unsigned mlw_off;
  // optional coarse/fine balance (last stage): tile face -> entry of rbal
  // (-1: none); rbal NULL when the balance is off or not this stage's
  const int *tmod;
  const R *rbal;
};

// The fused first stage for the 2D FP64 variant only.  There the viscous
// part needs fewer registers than the stage and the fused launch saves most
// of the separate pass (1.14 -> 1.10 ms per step on the 2.6M-cell NACA0015
// SA case).  In the FP32 variants the separate k_visc_prep runs at a higher
// occupancy than the stage kernel, and fusing measured even (plain closure)
// or slower (learned closure: +100 us on the first stage); in 3D it would
// lower the occupancy of every stage launch.
template <typename R, int D>
constexpr bool kFuseVisc = (D == 2 && std::is_same<R, double>::value);

constexpr int kRpart = 12;   // per tile: max[5], sumsq[5], SA max, SA sumsq
// learned-closure cell counters: kMlSlots pairs (network, fallback), one
// 128-byte line apart, each workgroup adding into slot blockIdx % kMlSlots;
// the reader sums the slots
constexpr int kMlSlots = 64, kMlSlotStride = 16;

// LDS of the learned-closure part of k_stage, after the face fluxes (it
// overwrites the face-cell indices, dead by then; the fluxes are read after
// the network): W1t as 64 rows of its 8 used columns, W2p (64 rows of 64,
// 16-byte chunks XOR-swizzled by the row, see mlp_fwd_reg), b1, b2 and the
// bf16 w3.
constexpr size_t kStageMlSmem = (size_t)64 * 8 * 2 + (size_t)64 * 64 * 2 + 2 * 64 * 4 + 64 * 2;
// byte offset of that region for a tile of nface faces (16-byte aligned)
template <typename R, int D>
__host__ __device__ constexpr size_t ml_smem_offset(int nface) {
  return ((sizeof(typename V4<R>::T) + (D == 3 ? sizeof(R) : 0)) * (size_t)nface + 15) & ~(size_t)15;
}

// One workgroup per tile of Morton-consecutive normal cells of one
// (case, level).  Phase 1: every face touching the tile — MUSCL + all-speed
// Roe (NS_Solver.H:704-832) — into LDS (with the SA closure, also the two
// face cells' indices; phase 2 gathers their velocity and working variable
// from cache, phase 1 having just read them).  Phase 2: ComptSource
// (NS_Solver.cpp:282-363), ComptMutSource (:632-840), UpdateConsVars
// (:844-985) and ComptNewPrimVar (:3-194) for the tile's cells.  The ML
// instantiation (mixed mode, launched as k_stage_ml) evaluates the
// point-wise SA term with the network for the cells inside its training
// range, between the two halves of phase 2, on MFMA.  The 2D FP64 first
// stage of a time-march step also does the viscous pass of the tile first
// (fuse_visc).
//
// Win and Wout are different buffers, so no workgroup reads a state another
// one has already advanced within the launch.  U_old is cons(W64): the
// master holds the step-start state until the last stage rewrites it, and
// each cell's master entry is read and written by the same thread.
// The network's weights to LDS at dst (W1t rows' 8 used columns, then W2p
// as stored, b1, b2, the bf16 w3), in 16-byte chunks with all loads issued
// before the stores.  Cooperative over the workgroup; the caller syncs.
__device__ __forceinline__ void ml_fill_weights(const MlpDev &M, unsigned char *dst_raw) {
  __bf16 *mW1t = reinterpret_cast<__bf16 *>(dst_raw);   // [64][8]
  __bf16 *mW2p = mW1t + 64 * 8;                          // [64][64]
  float *mb1 = reinterpret_cast<float *>(mW2p + 64 * 64), *mb2 = mb1 + 64;
  __bf16 *mw3 = reinterpret_cast<__bf16 *>(mb2 + 64);
  constexpr int kW1Rows = 64, kWChunks = kW1Rows + 64 * 64 * 2 / 16;
  static_assert(kWChunks == 576, "weight chunk count");
  uint4 *dst = reinterpret_cast<uint4 *>(mW1t);   // W2p follows W1t in LDS
  const uint4 *s2 = reinterpret_cast<const uint4 *>(M.W2p);   // chunks [64, 576)
  // chunk id of W1t or W2p; both indices stay in range whichever is read
  auto chunk = [&](int id) {
    const uint4 x1 = *reinterpret_cast<const uint4 *>(M.W1t + (size_t)(id < kW1Rows ? id : 0) * kX);
    const uint4 x2 = s2[id < kW1Rows ? 0 : id - kW1Rows];
    return id < kW1Rows ? x1 : x2;
  };
  if (blockDim.x == 256) {
    // chunks t, t+256 always exist; t+512 for t < 64
    const int t = threadIdx.x;
    const uint4 c0 = chunk(t);
    const uint4 c1 = s2[t + 256 - kW1Rows];
    const uint4 c2 = s2[(t < kWChunks - 512 ? t + 512 : 512) - kW1Rows];
    dst[t] = c0;
    dst[t + 256] = c1;
    if (t < kWChunks - 512) dst[t + 512] = c2;
  } else {
    for (int id = threadIdx.x; id < kWChunks; id += blockDim.x) dst[id] = chunk(id);
  }
  for (int i = threadIdx.x; i < 64; i += blockDim.x) {
    mw3[i] = (__bf16)M.w3[i]; mb1[i] = M.b1[i]; mb2[i] = M.b2[i];
  }
}

// Reduction of 16 per-lane slots over a wave: slots 0..7 by maximum, 8..15
// by sum.  A butterfly that halves the slots a lane keeps at each exchange
// (offsets 32, 16, 8, 4: 8 + 4 + 2 + 1 exchanges; the lane with the offset's
// bit set keeps the upper half, so after the first exchange lanes 0..31
// hold maxima and lanes 32..63 sums), then two exchanges of the one slot
// left: lane l returns slot l / 4 of the whole wave.  17 exchanges instead
// of 6 per slot.  The partner of an exchange has the same slot set (their
// lanes differ only in bits not yet used), so each slot combines the same
// 64 values as a per-slot reduction, in another order.
__device__ __forceinline__ double wave_slots16(double *v) {
  const int lane = threadIdx.x & 63;
  const bool sum = (lane & 32) != 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const double keep = sum ? v[j + 8] : v[j], send = sum ? v[j] : v[j + 8];
    const double o = __shfl_xor(send, 32, 64);
    v[j] = sum ? keep + o : fmax(keep, o);
  }
#pragma unroll
  for (int m = 4, off = 16; m >= 1; m >>= 1, off >>= 1) {
    const bool up = (lane & off) != 0;
#pragma unroll
    for (int j = 0; j < m; ++j) {
      const double keep = up ? v[j + m] : v[j], send = up ? v[j] : v[j + m];
      const double o = __shfl_xor(send, off, 64);
      v[j] = sum ? keep + o : fmax(keep, o);
    }
  }
  double x = v[0];
#pragma unroll
  for (int off = 2; off >= 1; off >>= 1) {
    const double o = __shfl_xor(x, off, 64);
    x = sum ? x + o : fmax(x, o);
  }
  return x;
}

// One tile.  WRES: the network's weights are resident at a.mlw_off
// (persistent schedule); otherwise the ML instantiation stages them after
// the face fluxes, over the face-cell indices (dead by then).
template <typename R, int D, bool ML, bool WRES>
__device__ __forceinline__ void stage_tile(const StageArgs<R, D> &a, int tile) {
  constexpr bool wres = WRES;
  constexpr int NV = D + 2;
  constexpr bool kMixed = std::is_same<R, float>::value;
  using V = typename V4<R>::T;
  extern __shared__ __align__(32) unsigned char smem_raw[];
  // the error word: loaded here, its load in flight while the face phase
  // runs, published by thread 0 just before the face phase's barrier and
  // tested after it (the face phase writes only LDS)
  __shared__ int halt;
  __shared__ unsigned mlc[2];   // learned closure: network / fallback cells of this tile
  const bool herr = halted(a.err);
  if (ML && threadIdx.x == 0) { mlc[0] = 0u; mlc[1] = 0u; }
  const TileInfo ti = a.tiles[tile];
  if constexpr (kFuseVisc<R, D>) {
    if (a.fuse_visc) {
      // Rv and mutf of this tile's cells from the step-start state (the
      // buffer this stage reads); each thread reads back only its own
      // cell's entries in phase 2, so no other workgroup is involved
      visc_prep_tile<R, D>(a.va, ti, smem_raw, &halt, herr);
      if (halt) return;  // uniform: read after the barrier in visc_prep_tile
      __syncthreads();   // the face arrays are reused by phase 1
    }
  }
  V *fl4 = reinterpret_cast<V *>(smem_raw);
  R *fl1 = reinterpret_cast<R *>(smem_raw + sizeof(V) * (size_t)ti.nface);   // 3D 5th value
  // SA: the low / high cell of each face; the cell phase gathers their
  // velocity and working variable (cache hits: phase 1 has just read them)
  int *fL = reinterpret_cast<int *>(smem_raw + (sizeof(V) + (D == 3 ? sizeof(R) : 0)) * (size_t)ti.nface);
  int *fR = fL + ti.nface;
  __shared__ double red[kRpart][kTile / 64];

  const BlkInfo &bi = a.blk[ti.blk];
  const DevPhys &P = a.phys[bi.case_id];
  PhysR<R> ph;
  ph.gama = (R)P.gama; ph.Rg = (R)P.Rg; ph.Cv = (R)P.Cv; ph.Ma = (R)P.Ma;
  const R areas[3] = {(R)bi.area[0], (R)bi.area[1], (R)bi.area[2]};

  for (int f = threadIdx.x; f < ti.nface; f += blockDim.x) {
    const int4 st = a.tst[ti.face0 + f];
    const int dir = a.tdir[ti.face0 + f];
    const Prim<R, D> q0 = pload<R, D>(a.Win, st.x);
    const Prim<R, D> q1 = pload<R, D>(a.Win, st.y);
    const Prim<R, D> q2 = pload<R, D>(a.Win, st.z);
    const Prim<R, D> q3 = pload<R, D>(a.Win, st.w);
    R flux[NV];
    if constexpr (kMixed)
      roe_flux_fast<D>(q0, q1, q2, q3, dir, areas[dir], ph, a.limiter, flux);
    else
      roe_flux<R, D>(q0, q1, q2, q3, dir, areas[dir], ph, a.limiter, flux);
    if (a.rbal) {
      // a coarse/fine face of the optional balance: its balanced flux
      const int q = a.tmod[ti.face0 + f];
      if (q >= 0)
        for (int v = 0; v < NV; ++v) flux[v] = a.rbal[(size_t)q * NV + v];
    }
    V o;
    o.x = flux[0]; o.y = flux[1]; o.z = flux[2]; o.w = flux[3];
    fl4[f] = o;
    if (D == 3) fl1[f] = flux[4];
    if (a.turb) {
      fL[f] = st.y;
      fR[f] = st.z;
    }
  }
  if (threadIdx.x == 0) halt = herr ? 1 : 0;
  __syncthreads();
  if (halt) return;   // uniform: the flag was written before the barrier

  const double level_ts = a.alpha * a.dt[bi.case_id] * bi.step_num;
  const double inv_vol = 1.0 / bi.vol;
  double mx[NV], sq[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) { mx[k] = 0.0; sq[k] = 0.0; }
  double nmx = 0.0, nsq = 0.0;
  const int i = threadIdx.x;
  const bool live = i < ti.ncell;
  const int c = ti.cell0 + (live ? i : 0);
  const uint8_t fg = live ? a.flags[c] : 0;
  // ---- phase 2a: residual of the cell, the SA terms
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  double snu = 0.0;
  bool sa_on = false;
  bool use_ml = false;
  R ml_x[kMlpIn];
  R ml_scale = R(0), sa_rest = R(0), sa_nu = R(0);
  // the cell's face slots in the tile's face list
  auto load_slots = [&](int *sl) {
    if (D == 2) {
      const ushort4 s4 = *reinterpret_cast<const ushort4 *>(a.slots + (size_t)c * 4);
      sl[0] = s4.x; sl[1] = s4.y; sl[2] = s4.z; sl[3] = s4.w;
    } else {
#pragma unroll
      for (int q = 0; q < 2 * D; ++q) sl[q] = a.slots[(size_t)c * 2 * D + q];
    }
  };
  int sl[2 * D];
  if (live) load_slots(sl);
  // ComptSource: per direction (F_lo - F_hi), plus the frozen viscous
  // residual sum_d (Fv_hi - Fv_lo).  The learned-closure instantiation
  // forms it after the network (slots read again), so that neither the sum
  // nor the slots are held in registers across the network (they spill).
  auto residual = [&]() {
    if constexpr (ML) load_slots(sl);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const V lo = fl4[sl[2 * d]], hi = fl4[sl[2 * d + 1]];
      s[0] += (double)lo.x - (double)hi.x;
      s[1] += (double)lo.y - (double)hi.y;
      s[2] += (double)lo.z - (double)hi.z;
      s[3] += (double)lo.w - (double)hi.w;
      if (D == 3) s[4] += (double)fl1[sl[2 * d]] - (double)fl1[sl[2 * d + 1]];
    }
    R rv[NV];
    vload<R, D>(a.Rv, c, rv);
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += (double)rv[k];
  };
  if (live) {
    if constexpr (!ML) residual();
    // ComptMutSource for a non-solid updated cell (zero otherwise: the
    // reference never writes mutsource of solid cells)
    if (a.turb && (fg & F_UPDATE) && (fg & F_SA)) {
      sa_on = true;
      const Prim<R, D> me = pload<R, D>(a.Win, c);
      R nl[3], nr[3], ul[3], ur[3], vl[3], vr[3], wl[3], wr[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (d < D) {
          const int lo = fL[sl[2 * d]], hi = fR[sl[2 * d + 1]];
          uvw_load<R, D>(a.Win, lo, ul[d], vl[d], wl[d]);
          uvw_load<R, D>(a.Win, hi, ur[d], vr[d], wr[d]);
          nl[d] = a.nuin[lo];
          nr[d] = a.nuin[hi];
        } else {
          ul[d] = ur[d] = vl[d] = vr[d] = wl[d] = wr[d] = nl[d] = nr[d] = R(0);
        }
      }
      const R T = v_div(me.p, me.r * (R)P.Rg);
      const R nu = a.nuin[c];
      const R sd = kMixed ? (R)a.sd32[c] : (R)a.signdis[c];
      if constexpr (!ML) {
        snu = (double)sa_source<R, D>(me.u, me.v, me.w, T, me.r, nu, sd, nl, nr, ul, ur, vl, vr,
                                      wl, wr, (R)bi.invh[0], (R)bi.invh[1], (R)bi.invh[2],
                                      (R)P.SoT, (R)P.Re, a.mutf[c]);
      } else {
        const R h = R(0.5);
        const R i0 = (R)bi.invh[0], i1 = (R)bi.invh[1], i2 = (R)bi.invh[2];
        VelGrad<R> g;
        g.dux = ((ur[0] - ul[0]) * h) * i0; g.dvx = ((vr[0] - vl[0]) * h) * i0; g.dwx = ((wr[0] - wl[0]) * h) * i0;
        g.duy = ((ur[1] - ul[1]) * h) * i1; g.dvy = ((vr[1] - vl[1]) * h) * i1; g.dwy = ((wr[1] - wl[1]) * h) * i1;
        if (D == 3) {
          g.duz = ((ur[2] - ul[2]) * h) * i2; g.dvz = ((vr[2] - vl[2]) * h) * i2; g.dwz = ((wr[2] - wl[2]) * h) * i2;
        } else {
          g.duz = g.dvz = g.dwz = R(0);
        }
        const R dvisx = ((nr[0] - nl[0]) * h) * i0, dvisy = ((nr[1] - nl[1]) * h) * i1;
        const R dvisz = (D == 3) ? ((nr[2] - nl[2]) * h) * i2 : R(0);
        const R Re = (R)P.Re;
        const R olen = sa_olen<R>(g);
        const R mu0 = v_sqrt(T * T * T) * v_div(R(1.0) + (R)P.SoT, T + (R)P.SoT);
        // derivative terms in their discrete form
        sa_rest = -(me.u * dvisx + me.v * dvisy + me.w * dvisz) +
                  v_div(v_div(R(SA::cb2) * (dvisx * dvisx + dvisy * dvisy + dvisz * dvisz), R(SA::segma)), Re) +
                  a.mutf[c];
        sa_nu = nu;
        bool ok = sa_features<R>(me.u, me.v, me.w, me.r, me.p, mu0, nu, sd, g, olen, Re,
                                 (R)P.init[4], ml_x, ml_scale);
        // the range test without short-circuit branches (a chain of
        // exec-mask branches, one per bound, otherwise)
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k)
          ok = ok & (ml_x[k] >= a.mlp.flo[k]) & (ml_x[k] <= a.mlp.fhi[k]);
        use_ml = ok;
        if (!ok) snu = (double)(sa_rest + sa_pointwise<R>(nu, mu0, me.r, sd, olen, Re));
      }
    }
  }
  // ---- learned closure: the network for the cells in its range
  if constexpr (ML) {
    unsigned char *wbase = smem_raw + (wres ? (size_t)a.mlw_off : ml_smem_offset<R, D>(ti.nface));
    if (!wres) {
      __syncthreads();   // every phase-2a read of the face-cell indices is done
      ml_fill_weights(a.mlp, wbase);
      __syncthreads();
    }
    __bf16 *mW1t = reinterpret_cast<__bf16 *>(wbase);   // [64][8]
    __bf16 *mW2p = mW1t + 64 * 8;                        // [64][64]
    float *mb1 = reinterpret_cast<float *>(mW2p + 64 * 64), *mb2 = mb1 + 64;
    __bf16 *mw3 = reinterpret_cast<__bf16 *>(mb2 + 64);
    const int lane = threadIdx.x & 63;
    bf16x8_t xf;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k)
      xf[k] = (__bf16)(use_ml ? ((float)ml_x[k] - a.mlp.fmean[k]) * a.mlp.finv[k] : 0.0f);
    float yown = 0.0f;
    // one 32-cell block per pass, skipped (wave-uniformly) when none of its
    // cells is on the network; the compiler barrier keeps the weight
    // fragments from being hoisted out of the loop
    const unsigned long long on = __ballot(use_ml);
#pragma unroll 1
    for (int sb = 0; sb < 2; ++sb) {
      asm volatile("" ::: "memory");
      if (((on >> (32 * sb)) & 0xffffffffull) == 0ull) continue;
      const float y = mlp_fwd_reg<8>(xf, sb, mW1t, mW2p, mb1, mb2, mw3);
      if ((lane >> 5) == sb) yown = y;
    }
    if (use_ml) {
      const float z = a.mlp.zst[0] + a.mlp.zst[1] * (yown + a.mlp.b3[0]);
      snu = (double)(sa_rest + (R)z * ml_scale);
    }
    if (a.last && a.mlcount) {
      // cells on the network / on the fallback in this tile (LDS sums,
      // added to the device counters once per workgroup below)
      unsigned long long nm = __ballot(use_ml), nf = __ballot(sa_on && !use_ml);
      if (lane == 0) {
        atomicAdd(&mlc[0], (unsigned)__popcll(nm));
        atomicAdd(&mlc[1], (unsigned)__popcll(nf));
      }
    }
    if (live) residual();   // the face fluxes are still in LDS
  }
  // ---- phase 2b: update
  if (live) {
    bool write = false;
    double wr = 0, wu = 0, wv = 0, ww = 0, wp = 0;
    double nu_new = 0.0;
    if (fg & F_UPDATE) {
      // U_old = cons(W64): Get_Cons_Vars FlowVariables.H:43-48,
      // e = Cv T + |u|^2/2, T = p/(roe Rg)
      double U[NV];
      {
        const Prim<double, D> w0 = pload<double, D>(a.W64, c);
        const double T0 = w0.p / (w0.r * P.Rg);
        const double ke0 = (D == 3) ? 0.5 * (w0.u * w0.u + w0.v * w0.v + w0.w * w0.w)
                                    : 0.5 * (w0.u * w0.u + w0.v * w0.v);
        const double e0 = P.Cv * T0 + ke0;
        U[0] = w0.r; U[1] = w0.r * w0.u; U[2] = w0.r * w0.v;
        if (D == 3) { U[3] = w0.r * w0.w; U[4] = w0.r * e0; } else U[3] = w0.r * e0;
      }
      if constexpr (kMixed) {
        const double sc = level_ts * inv_vol;
#pragma unroll
        for (int k = 0; k < NV; ++k) U[k] = U[k] + s[k] * sc;
      } else {
        // cv_old + source * level_ts / v  (NS_Solver.cpp:870)
#pragma unroll
        for (int k = 0; k < NV; ++k) U[k] = U[k] + s[k] * level_ts / bi.vol;
      }
      if (a.turb) {
        // viseddy = max(c_vis_old + mutsource * level_ts, 0) (:874-876, :955-957)
        nu_new = a.nu64[c] + snu * level_ts;
        nu_new = (nu_new < 0.0) ? 0.0 : nu_new;
      }
      if (U[0] < 0.0) report_error(a.err, c, (fg & F_FINEST) ? 611 : 523);
      if (fg & F_PRIM) {
        // Get_Prim_Vars FlowVariables.H:50-57
        wr = U[0];
        double e;
        if constexpr (kMixed) {
          const double ir = 1.0 / U[0];
          wu = U[1] * ir;
          wv = U[2] * ir;
          if (D == 3) { ww = U[3] * ir; e = U[4] * ir; } else e = U[3] * ir;
        } else {
          wu = U[1] / U[0];
          wv = U[2] / U[0];
          if (D == 3) { ww = U[3] / U[0]; e = U[4] / U[0]; } else e = U[3] / U[0];
        }
        const double ke = (D == 3) ? 0.5 * (wu * wu + wv * wv + ww * ww) : 0.5 * (wu * wu + wv * wv);
        wp = (e - ke) * (P.gama - 1.0) * wr;
        // T = p/(roe Rg) < 0  (NS_Solver.cpp:55-66), Rg > 0
        if ((wp < 0.0) != (wr < 0.0)) report_error(a.err, c, 37);
        write = true;
      }
    }
    if (!(fg & F_PRIM) && (fg & F_FREE)) {
      // m_data = initvar, the working variable included
      wr = P.init[0]; wu = P.init[1]; wv = P.init[2]; ww = P.init[3]; wp = P.init[4];
      nu_new = P.nut_inf;
      write = true;
    }
    if (write) {
      pstore<R, D>(a.Wout, c, (R)wr, (R)wu, (R)wv, (R)ww, (R)wp);
      if (kMixed && a.last) pstore<double, D>(a.W64, c, wr, wu, wv, ww, wp);
    }
    if (a.turb && ((fg & F_UPDATE) || (fg & F_FREE))) {
      a.nuout[c] = (R)nu_new;
      if (kMixed && a.last) a.nu64[c] = nu_new;
    }
    if (a.last) {
      if (a.src) {
        double *o = a.src + (size_t)c * 5;
        o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
        if (D == 3) { o[3] = s[3]; o[4] = s[4]; } else { o[3] = 0.0; o[4] = s[3]; }
      }
      if (a.nusrc) a.nusrc[c] = snu;
      if (fg & F_RESID) {
        bool nan = (snu != snu);
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          mx[k] = fabs(s[k]);
          sq[k] = s[k] * s[k];
          nan = nan || (s[k] != s[k]);
        }
        nmx = fabs(snu);
        nsq = snu * snu;
        // a NaN residual aborts the march (ResinDomain, NS_Solver.H:355-375)
        if (nan) report_error(a.err, c, 268);
      }
    }
  }
  if (a.last) {
    // the wave's partials as 16 slots: maxima (|s_k|, then |snu|) in 0..7,
    // sums of squares in the same order in 8..15, zeros in the rest
    double v[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      v[j] = j < NV ? mx[j < NV ? j : 0] : (j == NV ? nmx : 0.0);
      v[8 + j] = j < NV ? sq[j < NV ? j : 0] : (j == NV ? nsq : 0.0);
    }
    const double r = wave_slots16(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if ((lane & 3) == 0) {
      // lane 4 s holds slot s of the whole wave
      const int sl = lane >> 2, j = sl & 7;
      if (j < NV) red[sl < 8 ? j : 5 + j][wid] = r;
      else if (j == NV) red[sl < 8 ? 10 : 11][wid] = r;
    }
    __syncthreads();
    // the tile's partials, one entry per thread: entries 0..4 the maxima of
    // (roe, u, v, w, e) (in 2D w is 0 and e the fourth variable), 5..9 the
    // sums of squares, 10 and 11 the working variable's
    if (threadIdx.x < kRpart) {
      const int nw = (blockDim.x + 63) >> 6;
      const int j = threadIdx.x, jv = j < 10 ? j % 5 : 0;
      const int k = (j >= 10) ? j : ((D == 3) ? jv : (jv == 4 ? 3 : (jv == 3 ? -1 : jv)));
      double r = 0.0;
      if (k >= 0) {
        const int row = (j >= 10) ? j : (j < 5 ? k : 5 + k);
        const bool is_max = (j < 5) || (j == 10);
        for (int w = 0; w < nw; ++w) r = is_max ? fmax(r, red[row][w]) : r + red[row][w];
      }
      a.rpart[(size_t)tile * kRpart + j] = r;
    }
    if (threadIdx.x == kRpart) {
      if constexpr (ML) {
        if (a.mlcount) {
          unsigned long long *slot = a.mlcount + (size_t)(blockIdx.x % kMlSlots) * kMlSlotStride;
          if (mlc[0]) atomicAdd(&slot[0], (unsigned long long)mlc[0]);
          if (mlc[1]) atomicAdd(&slot[1], (unsigned long long)mlc[1]);
        }
      }
    }
  }
}

// Schedule P (a separate instantiation each, so that the loop costs the
// one-workgroup-per-tile kernels no registers).  Persistent: XCD x (blocks
// b = x mod 8) owns the x-th contiguous run of the launch's tiles and its
// workgroups take them in turn, so concurrently processed tiles are
// Morton neighbours sharing that XCD's L2.
template <typename R, int D, bool ML, bool P>
__device__ __forceinline__ void stage_launch(const StageArgs<R, D> &a) {
  if constexpr (!P) {
    stage_tile<R, D, ML, false>(a, dispatch_tile());
    return;
  }
  extern __shared__ __align__(32) unsigned char smem_raw[];
  if constexpr (ML) {
    ml_fill_weights(a.mlp, smem_raw + a.mlw_off);
    __syncthreads();
  }
  constexpr int kXcd = 8;
  const int b = (int)blockIdx.x, x = b % kXcd, G = (int)gridDim.x;
  const int per = a.ntile / kXcd, rem = a.ntile % kXcd;
  const int t0 = x * per + (x < rem ? x : rem), t1 = t0 + per + (x < rem ? 1 : 0);
  const int nwk = (G - x + kXcd - 1) / kXcd;   // workgroups of XCD x
  for (int t = t0 + b / kXcd; t < t1; t += nwk) {
    // nothing is hoisted out of the loop: values kept live across the
    // tiles would cost the loop body registers (and occupancy)
    asm volatile("" ::: "memory");
    stage_tile<R, D, ML, true>(a, t);
    __syncthreads();   // the next tile rewrites the face arrays
  }
}

template <typename R, int D, bool ML, bool P = false>
__global__ void __launch_bounds__(kTile) k_stage(StageArgs<R, D> a) {
  stage_launch<R, D, ML, P>(a);
}

// The learned-closure stage at seven waves per SIMD: its network needs more
// than the 64 registers eight waves allow (held to 64, the compiler spills
// to scratch and the launch takes 112 instead of 93 us on the 2.6M-cell
// NACA0015 case).
template <int D, bool P = false>
__global__ void __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_stage_ml(StageArgs<float, D> a) {
  stage_launch<float, D, true, P>(a);
}

// Stages with a single substep write into buffer 1; this copies the marched
// cells back into the home buffer afterwards.
template <typename R, int D>
__global__ void k_copy_back(int n, Pack<R, D> src, Pack<R, D> dst, const R *nusrc, R *nudst,
                            const unsigned long long *err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || halted(err)) return;
  dst.a[c] = src.a[c];
  if (D == 3) dst.e[c] = src.e[c];
  if (nudst) nudst[c] = nusrc[c];
}

// ---------------------------------------------------------- ghost fills
// BCBoxTreat BCValues.H:51-188: -13 pressure inlet (:191-206), -12 pressure
// outlet (:208-214), -50 outflow (:216-220), -81 wall of the TURBULENCE
// build (:89-125); entries (cell, refcell, type, case).  SA working
// variable: copied with the state for -12/-50, the inlet value for -13
// (:203-205), kappa y+ at a wall.
template <int D>
__global__ void k_bc(int n, const int4 *dmg, const DmgWall *wall, StateOut<D> o,
                     const DevPhys *phys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || halted(o.err)) return;
  const int4 e = dmg[i];
  const int g = e.x, ty = e.z;
  const DevPhys &P = phys[e.w];
  const Prim<double, D> q = pload<double, D>(o.m, e.y);
  if (ty == -81) {
    // the wall of the TURBULENCE build (BCValues.H:89-125); nothing in the
    // laminar build, where the branch does not exist
    if (!o.nm || !wall) return;
    const DmgWall W = wall[i];
    if (W.dnbr > 1.0) {
      put_all<D>(o, g, q.r, q.u, q.v, q.w, q.p);
      put_nu<D>(o, g, o.nm[e.y]);
      return;
    }
    // ImageViseddy_Domain Turbulence.cpp:140-195 with the reference cell's
    // neighbour as the reference point and a wall at rest
    const Prim<double, D> s = pload<double, D>(o.m, W.nbr);
    const double *nm = W.nmv;
    const double Tr = s.p / (s.r * P.Rg);
    const double ib_rdis = W.dref / W.dnbr;
    const double hvn = s.u * nm[0] + s.v * nm[1] + s.w * nm[2];
    const double hn[3] = {hvn * nm[0], hvn * nm[1], hvn * nm[2]};
    const double ht[3] = {s.u - hn[0], s.v - hn[1], s.w - hn[2]};
    const double df_velt = sqrt(ht[0] * ht[0] + ht[1] * ht[1] + ht[2] * ht[2]);
    const double hg_mu = sqrt(Tr * Tr * Tr) * ((1.0 + P.SoT) / (Tr + P.SoT));
    const double hg_vis = hg_mu / s.r;
    const double hg_re = df_velt * W.dnbr / hg_vis * P.Re;
    const double hg_yplus = sa_yplus(hg_re);
    const double ut = hg_yplus * hg_vis / W.dnbr / P.Re;
    const double yplus = ib_rdis * hg_yplus;
    const double uplus = sa_uplus(yplus);
    const double ib_velt = uplus * ut;
    const double ib_r0 = (fabs(df_velt) < 0.00000001) ? 0.0 : ib_velt / df_velt;
    double ou[3];
    for (int k = 0; k < 3; ++k) ou[k] = hn[k] * ib_rdis + ht[k] * ib_r0;
    put_all<D>(o, g, s.p / (P.Rg * Tr), ou[0], ou[1], (D == 3) ? ou[2] : 0.0, s.p);
    put_nu<D>(o, g, SA::kappa * yplus);
    return;
  }
  if (ty == -50) {
    put_all<D>(o, g, q.r, q.u, q.v, q.w, q.p);
    if (o.nm) put_nu<D>(o, g, o.nm[e.y]);
  } else if (ty == -12) {
    const double T = q.p / (q.r * P.Rg);
    const double p = P.outlet_p;
    put_all<D>(o, g, p / (P.Rg * T), q.u, q.v, q.w, p);
    if (o.nm) put_nu<D>(o, g, o.nm[e.y]);
  } else if (ty == -13) {
    const double speed0 = sqrt(q.u * q.u + q.v * q.v + q.w * q.w);
    const double u = speed0 * P.inlet_cos, v = speed0 * P.inlet_sin;
    const double T = P.inlet_T0 - (0.5 * (u * u + v * v)) / P.Cp;
    const double p = P.inlet_p0 * pow((T / P.inlet_T0), 3.5);
    put_all<D>(o, g, p / (P.Rg * T), u, v, 0.0, p);
    put_nu<D>(o, g, P.nut_inlet);
  }
}

// PairRestriction[_2d] TwoLevelOp.H:217-231 with AverageCell[_2d] :45-87
// (the SA variant :129-179 averages the working variable without weights)
template <int D>
__global__ void k_restrict(int n, const int *ci, const int *fi, StateOut<D> o, int twod,
                           const DevPhys *phys, const int *cellcase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || halted(o.err)) return;
  const int c = ci[i];
  const DevPhys &P = phys[cellcase[c]];
  double r = 0, ru = 0, rv = 0, rw = 0, re = 0, nu = 0;
  for (int q = 0; q < 8; ++q) {
    if (twod && (q & 1)) continue;
    const int a = fi[(size_t)i * 8 + q];
    const Prim<double, D> s = pload<double, D>(o.m, a);
    const double T = s.p / (s.r * P.Rg);
    const double e = P.Cv * T + ((D == 3) ? 0.5 * (s.u * s.u + s.v * s.v + s.w * s.w) : 0.5 * (s.u * s.u + s.v * s.v));
    r += s.r;
    ru += s.u * s.r;
    rv += s.v * s.r;
    rw += s.w * s.r;
    re += e * s.r;
    if (o.nm) nu += o.nm[a];
  }
  const double u = ru / r, v = rv / r, w = rw / r, e = re / r;
  const double rho = r * (twod ? 0.25 : 0.125);
  const double ke = (D == 3) ? 0.5 * (u * u + v * v + w * w) : 0.5 * (u * u + v * v);
  const double p = (e - ke) * (P.gama - 1.0) * rho;
  put_all<D>(o, c, rho, u, v, w, p);
  if (o.nm) put_nu<D>(o, c, nu * (twod ? 0.25 : 0.125));
}

// PairProlongation[_2d] TwoLevelOp.H:183-215: central gradients of the
// conserved variables over the 6 face neighbours (cneib, :352-358), sons at
// +-1/4 (fincrem); the SA variant (:91-127) does the same for the working
// variable.
template <int D>
__global__ void k_prolong(int n, const int *ci, const int *fi, const int *nb27, StateOut<D> o,
                          int twod, const DevPhys *phys, const int *cellcase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || halted(o.err)) return;
  const int c = ci[i];
  const DevPhys &P = phys[cellcase[c]];
  const int *nb = nb27 + (size_t)c * 27;
  auto cons = [&](int a, double *U) {
    const Prim<double, D> q = pload<double, D>(o.m, a);
    const double T = q.p / (q.r * P.Rg);
    const double e = P.Cv * T + ((D == 3) ? 0.5 * (q.u * q.u + q.v * q.v + q.w * q.w) : 0.5 * (q.u * q.u + q.v * q.v));
    U[0] = q.r; U[1] = q.r * q.u; U[2] = q.r * q.v; U[3] = q.r * q.w; U[4] = q.r * e;
  };
  const int ax0 = nb[0 * 9 + 1 * 3 + 1], ax1 = nb[2 * 9 + 1 * 3 + 1];
  const int ay0 = nb[1 * 9 + 0 * 3 + 1], ay1 = nb[1 * 9 + 2 * 3 + 1];
  const int az0 = nb[1 * 9 + 1 * 3 + 0], az1 = nb[1 * 9 + 1 * 3 + 2];
  double Uc[5], Ux0[5], Ux1[5], Uy0[5], Uy1[5], Uz0[5], Uz1[5];
  cons(c, Uc);
  cons(ax0, Ux0); cons(ax1, Ux1);
  cons(ay0, Uy0); cons(ay1, Uy1);
  cons(az0, Uz0); cons(az1, Uz1);
  double gx[5], gy[5], gz[5];
  for (int t = 0; t < 5; ++t) {
    gx[t] = (Ux1[t] - Ux0[t]) * 0.5;
    gy[t] = (Uy1[t] - Uy0[t]) * 0.5;
    gz[t] = (Uz1[t] - Uz0[t]) * 0.5;
  }
  double nc = 0, ngx = 0, ngy = 0, ngz = 0;
  if (o.nm) {
    nc = o.nm[c];
    ngx = (o.nm[ax1] - o.nm[ax0]) * 0.5;
    ngy = (o.nm[ay1] - o.nm[ay0]) * 0.5;
    ngz = (o.nm[az1] - o.nm[az0]) * 0.5;
  }
  const double fincrem[2] = {-0.25, 0.25};
  for (int si = 0; si < 2; ++si)
    for (int sj = 0; sj < 2; ++sj)
      for (int sk = 0; sk < (twod ? 1 : 2); ++sk) {
        const int f = fi[(size_t)i * 8 + si * 4 + sj * 2 + sk];
        double U[5];
        for (int t = 0; t < 5; ++t) {
          if (twod) U[t] = Uc[t] + fincrem[si] * gx[t] + fincrem[sj] * gy[t];
          else U[t] = Uc[t] + fincrem[si] * gx[t] + fincrem[sj] * gy[t] + fincrem[sk] * gz[t];
        }
        const double r = U[0], u = U[1] / U[0], v = U[2] / U[0], w = U[3] / U[0], e = U[4] / U[0];
        const double ke = (D == 3) ? 0.5 * (u * u + v * v + w * w) : 0.5 * (u * u + v * v);
        const double p = (e - ke) * (P.gama - 1.0) * r;
        put_all<D>(o, f, r, u, v, w, p);
        if (o.nm) {
          const double nf = twod ? nc + fincrem[si] * ngx + fincrem[sj] * ngy
                                 : nc + fincrem[si] * ngx + fincrem[sj] * ngy + fincrem[sk] * ngz;
          put_nu<D>(o, f, nf);
        }
      }
}

// Body::IBConditions_Image Body.cpp:1366-1496: image state from the
// bracketing cells (Mesh::IntpHGgcellVar Mesh.cpp:1532-1547).  Laminar:
// LinearDist (Body.cpp:183-204) for signdis >= 0, the mirror otherwise.  SA
// closure: ImageViseddy (Turbulence.cpp:57-138, a wall function on the
// tangential velocity, Crocco-Busemann wall temperature, nu = kappa y+) for
// signdis >= 0, and initvar with the wall velocity otherwise
// (Body.cpp:1403-1407).  The image cells of an IB cell are never IB cells
// (checked at upload), so one pass equals the reference's compute-all-then-
// write-all.
template <int D>
__global__ void k_ib(int n, const int *ndev, const int *cell, const int *intp, const double *coef,
                     const double *hgdis, const double *attach, const double *nvec, const int *body,
                     const int *ibcase, const double *signdis, const Motion *mot, int have_motion,
                     StateOut<D> o, const DevPhys *phys, unsigned long long *err) {
  // n: list length, or its capacity when the length lives on the device
  // (moving bodies rebuild the list every step)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (ndev && i >= ndev[0]) || halted(err)) return;
  const DevPhys &P = phys[ibcase[i]];
  double T = 0, p = 0, u = 0, v = 0, w = 0;
  for (int q = 0; q < 8; ++q) {
    const int a = intp[(size_t)i * 8 + q];
    if (a < 0) continue;
    const double cq = coef[(size_t)i * 8 + q];
    const Prim<double, D> s = pload<double, D>(o.m, a);
    T += cq * (s.p / (s.r * P.Rg));
    p += cq * s.p;
    u += cq * s.u;
    v += cq * s.v;
    w += cq * s.w;
  }
  double pv[3] = {0.0, 0.0, 0.0};
  if (have_motion) {
    const Motion M = mot[body[i]];
    const double *at = attach + (size_t)i * 3;
    const double r1[3] = {at[0] - M.center[0], at[1] - M.center[1], at[2] - M.center[2]};
    pv[0] = M.tra[0] + (M.rot[1] * r1[2] - M.rot[2] * r1[1]);
    pv[1] = M.tra[1] + (M.rot[2] * r1[0] - M.rot[0] * r1[2]);
    pv[2] = M.tra[2] + (M.rot[0] * r1[1] - M.rot[1] * r1[0]);
  }
  const int c = cell[i];
  const double sd = signdis[c];
  if (T < 0.0) report_error(err, c, 1050);
  if (!o.nm) {
    double ou, ov, ow;
    if (sd < 0.0) {
      ou = 2.0 * pv[0] - u; ov = 2.0 * pv[1] - v; ow = 2.0 * pv[2] - w;
    } else {
      const double rt = sd + hgdis[i];
      const double r0 = sd / rt;
      ou = pv[0] + (u - pv[0]) * r0;
      ov = pv[1] + (v - pv[1]) * r0;
      ow = pv[2] + (w - pv[2]) * r0;
    }
    put_all<D>(o, c, p / (P.Rg * T), ou, ov, (D == 3) ? ow : 0.0, p);
    return;
  }
  if (sd < 0.0) {
    put_all<D>(o, c, P.init[0], pv[0], pv[1], (D == 3) ? pv[2] : 0.0, P.init[4]);
    put_nu<D>(o, c, P.nut_inf);
    return;
  }
  // ImageViseddy
  const double *nm = nvec + (size_t)i * 3;
  const double rho_hg = p / (P.Rg * T);
  const double rt0 = hgdis[i] + sd;
  const double ib_rdis = sd / rt0;
  const double hvn = u * nm[0] + v * nm[1] + w * nm[2];
  const double hn[3] = {hvn * nm[0], hvn * nm[1], hvn * nm[2]};
  const double ht[3] = {u - hn[0], v - hn[1], w - hn[2]};
  const double wvn = pv[0] * nm[0] + pv[1] * nm[1] + pv[2] * nm[2];
  const double wn[3] = {nm[0] * wvn, nm[1] * wvn, nm[2] * wvn};
  const double wt[3] = {pv[0] - wn[0], pv[1] - wn[1], pv[2] - wn[2]};
  const double df[3] = {ht[0] - wt[0], ht[1] - wt[1], ht[2] - wt[2]};
  const double df_velt = sqrt(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
  const double hg_mu = sqrt(T * T * T) * ((1.0 + P.SoT) / (T + P.SoT));
  const double hg_vis = hg_mu / rho_hg;
  const double hg_re = df_velt * rt0 / hg_vis * P.Re;
  const double hg_yplus = sa_yplus(hg_re);
  const double ut = hg_yplus * hg_vis / rt0 / P.Re;
  const double yplus = ib_rdis * hg_yplus;
  const double uplus = sa_uplus(yplus);
  const double ib_velt = uplus * ut;
  const double ib_r0 = (fabs(df_velt) < 0.00000001) ? 0.0 : ib_velt / df_velt;
  double ou[3];
  for (int k = 0; k < 3; ++k) {
    const double ibn = wn[k] + (hn[k] - wn[k]) * ib_rdis;
    const double ibt = wt[k] + df[k] * ib_r0;
    ou[k] = ibn + ibt;
  }
  const double Tib = T + 0.5 * P.pr_cuberoot / P.Cp * (df_velt * df_velt - ib_velt * ib_velt);
  put_all<D>(o, c, p / (P.Rg * Tib), ou[0], ou[1], (D == 3) ? ou[2] : 0.0, p);
  put_nu<D>(o, c, SA::kappa * yplus);
}

// Comptdt NS_Solver.cpp:229-262 per tile: min over the tile's cells of
// CFL V / sum_d (|u_d| + c) A_d (all three directions, also in 2D).
template <int D>
__global__ void __launch_bounds__(kTile) k_cfl(const TileInfo *tiles, Pack<double, D> W64,
                                               const BlkInfo *blk, const DevPhys *phys,
                                               double *cflpart, unsigned long long *err) {
  __shared__ double red[kTile / 64];
  if (wg_halted(err)) return;
  const TileInfo ti = tiles[blockIdx.x];
  const BlkInfo &bi = blk[ti.blk];
  const DevPhys &P = phys[bi.case_id];
  double d0 = 1.0e300;
  if ((int)threadIdx.x < ti.ncell) {
    const int c = ti.cell0 + threadIdx.x;
    const Prim<double, D> q = pload<double, D>(W64, c);
    const double a = sqrt(P.gama * P.Rg * (q.p / (q.r * P.Rg)));
    const double c0 = (fabs(q.u) + a) * bi.area[0];
    const double c1 = (fabs(q.v) + a) * bi.area[1];
    const double c2 = (fabs(q.w) + a) * bi.area[2];
    d0 = P.cfl * bi.vol / (c0 + c1 + c2);
    if (d0 < 0.0) report_error(err, c, 424);
  }
  for (int off = 32; off > 0; off >>= 1) d0 = fmin(d0, __shfl_xor(d0, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d0;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) m = fmin(m, red[w]);
    cflpart[blockIdx.x] = m;
  }
}

// dt = min(CFL_dt, dt*dt_ratio) per case, CFL_dt capped at 10000
// (NS_Solver.cpp:198, 263).  One workgroup per case over the case's tile
// range of each level (ct[case][level] = {first, end}).
__global__ void k_dt_final(int L, const int *ct, const double *cflpart, double *dt,
                           const DevPhys *phys, int ncases, const unsigned long long *err) {
  const int cs = blockIdx.x;
  if (cs >= ncases) return;
  __shared__ double red[256];
  if (wg_halted(err)) return;
  double m = 10000.0;
  for (int l = 0; l < L; ++l) {
    const int a = ct[(cs * L + l) * 2], b = ct[(cs * L + l) * 2 + 1];
    for (int t = a + (int)threadIdx.x; t < b; t += blockDim.x) m = fmin(m, cflpart[t]);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = red[0];
    for (int k = 1; k < (int)blockDim.x; ++k) v = fmin(v, red[k]);
    const double d = dt[cs] * phys[cs].dt_ratio;
    dt[cs] = (v < d) ? v : d;
  }
}

// Stencil copies of the FP64 master after ibg_set_state / ibg_set_nut (one
// upload, the conversions on the device).
template <typename R, int D>
__global__ void k_fill_from_master(int n, Pack<double, D> src, Pack<R, D> dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const D4 q = src.a[i];
  typename V4<R>::T o;
  o.x = (R)q.x; o.y = (R)q.y; o.z = (R)q.z; o.w = (R)q.w;
  dst.a[i] = o;
  if (D == 3) dst.e[i] = (R)src.e[i];
}
template <typename R>
__global__ void k_fill_nu_from_master(int n, const double *src, R *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (R)src[i];
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Body::ComptWallForce Body.cpp:460-555, one workgroup per case, summed in
// a fixed order (deterministic).
template <int D>
__global__ void k_forces(int npatch, const int *pcase, const int *intp, const double *coef,
                         const double *hgdis, const double *nv, const double *pc,
                         const double *area, const double *ccenter, const Motion *mot,
                         int have_motion, Pack<double, D> W64, const DevPhys *phys, int turb,
                         double *out) {
  const int c = blockIdx.x;
  const DevPhys &P = phys[c];
  __shared__ double acc[256][12];
  double s[12];
  for (int k = 0; k < 12; ++k) s[k] = 0.0;
  const double *bc = ccenter + (size_t)c * 3;
  for (int i = threadIdx.x; i < npatch; i += blockDim.x) {
    if (pcase[i] != c) continue;
    double T = 0, p = 0, u = 0, v = 0, w = 0;
    for (int q = 0; q < 8; ++q) {
      const int a = intp[(size_t)i * 8 + q];
      if (a < 0) continue;
      const double cq = coef[(size_t)i * 8 + q];
      const Prim<double, D> sp = pload<double, D>(W64, a);
      T += cq * (sp.p / (sp.r * P.Rg));
      p += cq * sp.p;
      u += cq * sp.u;
      v += cq * sp.v;
      w += cq * sp.w;
    }
    const double *x = pc + (size_t)i * 3;
    const double *n = nv + (size_t)i * 3;
    double pv[3] = {0, 0, 0};
    if (have_motion) {
      const Motion M = mot[c];
      const double r1[3] = {x[0] - M.center[0], x[1] - M.center[1], x[2] - M.center[2]};
      pv[0] = M.tra[0] + (M.rot[1] * r1[2] - M.rot[2] * r1[1]);
      pv[1] = M.tra[1] + (M.rot[2] * r1[0] - M.rot[0] * r1[2]);
      pv[2] = M.tra[2] + (M.rot[0] * r1[1] - M.rot[1] * r1[0]);
    }
    const double A = area[i];
    double df0[3], df1[3];
    if (!turb) {
      const double du[3] = {u - pv[0], v - pv[1], w - pv[2]};
      const double dn = du[0] * n[0] + du[1] * n[1] + du[2] * n[2];
      for (int k = 0; k < 3; ++k) {
        df0[k] = ((du[k] - n[k] * dn) / hgdis[i]) * P.mu * A;
        df1[k] = n[k] * (-p * A);
      }
    } else {
      // Patchut (Turbulence.cpp:197-226): wall-function friction velocity
      // and wall temperature; shear = roe_w ut^2 along the slip direction.
      // FiltSinglePatchPressure (Filt.cpp:44-67) has coefficients b = {1,0..},
      // a = {0..} (Filt.cpp:5-6): the identity, so p is used as is.
      const double roe = p / (P.Rg * T);
      const double hv = u * n[0] + v * n[1] + w * n[2];
      const double pn = pv[0] * n[0] + pv[1] * n[1] + pv[2] * n[2];
      double df[3];
      df[0] = (u - n[0] * hv) - (pv[0] - n[0] * pn);
      df[1] = (v - n[1] * hv) - (pv[1] - n[1] * pn);
      df[2] = (w - n[2] * hv) - (pv[2] - n[2] * pn);
      const double dv = sqrt(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
      double vt[3] = {0.0, 0.0, 0.0};
      if (!(dv < 1e-8)) { vt[0] = df[0] / dv; vt[1] = df[1] / dv; vt[2] = df[2] / dv; }
      const double hg_mu = sqrt(T * T * T) * ((1.0 + P.SoT) / (T + P.SoT));
      const double hg_re = roe * dv * hgdis[i] / hg_mu * P.Re;
      const double yplus = sa_yplus(hg_re);
      const double ut = yplus * (hg_mu / roe) / hgdis[i] / P.Re;
      const double wall_T = T + 0.5 * P.pr_cuberoot / P.Cp * (dv * dv);
      const double wall_roe = p / (P.Rg * wall_T);
      const double Cf = wall_roe * (ut * ut);
      for (int k = 0; k < 3; ++k) {
        df0[k] = vt[k] * Cf * A;
        df1[k] = n[k] * (-p * A);
      }
    }
    const double dx[3] = {x[0] - bc[0], x[1] - bc[1], x[2] - bc[2]};
    const double dm0[3] = {dx[1] * df0[2] - dx[2] * df0[1], dx[2] * df0[0] - dx[0] * df0[2], dx[0] * df0[1] - dx[1] * df0[0]};
    const double dm1[3] = {dx[1] * df1[2] - dx[2] * df1[1], dx[2] * df1[0] - dx[0] * df1[2], dx[0] * df1[1] - dx[1] * df1[0]};
    for (int k = 0; k < 3; ++k) { s[k] += df1[k]; s[3 + k] += df0[k]; s[6 + k] += dm1[k]; s[9 + k] += dm0[k]; }
  }
  for (int k = 0; k < 12; ++k) acc[threadIdx.x][k] = s[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[12];
    for (int k = 0; k < 12; ++k) t[k] = 0.0;
    for (int j = 0; j < (int)blockDim.x; ++j)
      for (int k = 0; k < 12; ++k) t[k] += acc[j][k];
    for (int k = 0; k < 3; ++k) {
      // the SA build scales the wall shear by p_ref*area_ref (Body.cpp:547-553)
      const double sr = turb ? P.p_ref_area : P.shear_ref_area;
      out[c * 6 + k] = t[k] * P.p_ref_area + t[3 + k] * sr;
      out[c * 6 + 3 + k] = t[6 + k] * P.p_ref_area + t[9 + k] * sr;
    }
  }
}

// residual finalize: per case, over the tiles of the last march
// out[case*12 + {0..4}] = L-inf, {5..9} = L2 of the flow residual, {10, 11}
// = L-inf, L2 of the SA residual (res_vis_sum of NS_Solver.H:297, 321).
__global__ void k_resid_final(int ntiles, int L, const int *ct, const double *rpart, double *out,
                              int ncases) {
  // one workgroup per case over its tiles of the last march (tiles < ntiles)
  const int cs = blockIdx.x;
  if (cs >= ncases) return;
  __shared__ double red[kRpart][256];
  double m[6] = {0, 0, 0, 0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0};
  for (int l = 0; l < L; ++l) {
    const int a = ct[(cs * L + l) * 2], b = min(ct[(cs * L + l) * 2 + 1], ntiles);
    for (int t = a + (int)threadIdx.x; t < b; t += blockDim.x) {
      for (int k = 0; k < 5; ++k) {
        m[k] = fmax(m[k], rpart[(size_t)t * kRpart + k]);
        q[k] += rpart[(size_t)t * kRpart + 5 + k];
      }
      m[5] = fmax(m[5], rpart[(size_t)t * kRpart + 10]);
      q[5] += rpart[(size_t)t * kRpart + 11];
    }
  }
  for (int k = 0; k < 6; ++k) { red[k][threadIdx.x] = m[k]; red[6 + k][threadIdx.x] = q[k]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 6; ++k) {
      double mm = 0.0, qq = 0.0;
      for (int w = 0; w < (int)blockDim.x; ++w) { mm = fmax(mm, red[k][w]); qq += red[6 + k][w]; }
      m[k] = mm;
      q[k] = qq;
    }
    for (int k = 0; k < 5; ++k) { out[cs * kRpart + k] = m[k]; out[cs * kRpart + 5 + k] = sqrt(q[k]); }
    out[cs * kRpart + 10] = m[5];
    out[cs * kRpart + 11] = sqrt(q[5]);
  }
}

// parity probes: inviscid / viscous flux of every face, [face][5]
template <typename R, int D>
__global__ void k_face_probe(int nface, const int4 *st, const int8_t *dir, const int16_t *fblk,
                             Pack<R, D> W, const BlkInfo *blk, const DevPhys *phys, int lim,
                             double *out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface) return;
  constexpr int NV = D + 2;
  const BlkInfo &bi = blk[fblk[f]];
  const DevPhys &P = phys[bi.case_id];
  PhysR<R> ph;
  ph.gama = (R)P.gama; ph.Rg = (R)P.Rg; ph.Cv = (R)P.Cv; ph.Ma = (R)P.Ma;
  const int4 s = st[f];
  R flux[NV];
  roe_flux<R, D>(pload<R, D>(W, s.x), pload<R, D>(W, s.y), pload<R, D>(W, s.z),
                 pload<R, D>(W, s.w), dir[f], (R)bi.area[dir[f]], ph, lim, flux);
  double *o = out + (size_t)f * 5;
  if (D == 3) { for (int k = 0; k < 5; ++k) o[k] = flux[k]; }
  else { o[0] = flux[0]; o[1] = flux[1]; o[2] = flux[2]; o[3] = 0.0; o[4] = flux[3]; }
}

template <typename R, int D>
__global__ void k_visc_probe(int nface, const int4 *st, const int4 *vst, const int8_t *dir,
                             const int16_t *fblk, Pack<R, D> W, const BlkInfo *blk,
                             const DevPhys *phys, const R *nu, double *out, double *mout) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface) return;
  constexpr int NV = D + 2;
  const BlkInfo &bi = blk[fblk[f]];
  const DevPhys &P = phys[bi.case_id];
  ViscPhys<R, D> vp;
  vp.Rg = (R)P.Rg; vp.Re = (R)P.Re; vp.qk = (R)P.qk; vp.SoT = (R)P.SoT;
  vp.ih0 = (R)bi.invh[0]; vp.ih1 = (R)bi.invh[1]; vp.ih2 = (R)bi.invh[2];
  vp.ar0 = (R)bi.area[0]; vp.ar1 = (R)bi.area[1]; vp.ar2 = (R)bi.area[2];
  vp.qkt = (R)P.qkt_den;
  const int4 s = st[f];
  Prim<R, D> t[(D == 3) ? 8 : 4];
  const int4 v0 = vst[(size_t)f * (D - 1)];
  t[0] = pload<R, D>(W, v0.x); t[1] = pload<R, D>(W, v0.y); t[2] = pload<R, D>(W, v0.z); t[3] = pload<R, D>(W, v0.w);
  if (D == 3) {
    const int4 v1 = vst[(size_t)f * (D - 1) + 1];
    t[4 % ((D == 3) ? 8 : 4)] = pload<R, D>(W, v1.x);
    t[5 % ((D == 3) ? 8 : 4)] = pload<R, D>(W, v1.y);
    t[6 % ((D == 3) ? 8 : 4)] = pload<R, D>(W, v1.z);
    t[7 % ((D == 3) ? 8 : 4)] = pload<R, D>(W, v1.w);
  }
  R o5[NV];
  R mf = R(0);
  if (nu)
    visc_flux<R, D>(pload<R, D>(W, s.y), pload<R, D>(W, s.z), t, dir[f], vp, o5, true, nu[s.y],
                    nu[s.z], &mf);
  else
    visc_flux<R, D>(pload<R, D>(W, s.y), pload<R, D>(W, s.z), t, dir[f], vp, o5);
  if (mout) mout[f] = mf;
  double *o = out + (size_t)f * 5;
  if (D == 3) { for (int k = 0; k < 5; ++k) o[k] = o5[k]; }
  else { o[0] = o5[0]; o[1] = o5[1]; o[2] = o5[2]; o[3] = 0.0; o[4] = o5[3]; }
}

// plain copy (bandwidth / counter calibration probe)
__global__ void k_copy(const F4 *__restrict__ a, F4 *__restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

}  // namespace ibg
