// ibg_engine.hip — MI355X (gfx950) residual engine behind include/ibgpu.h.
//
// Device layout.  All cells of all cases live in one index space reordered
// for the hot path: normal cells first, sorted by level from finest to
// coarsest (then case, then Morton order), ghost cells after.  Under the
// reference's temporal sub-cycling (mpiparams.cpp:186-212) the levels
// marched in a step are always the finest ones, so the marched cells, faces
// and tiles of every step are a prefix of their arrays and every kernel is a
// single launch over a prefix.
//
// State is SoA: W64 (roe,u,v,[w],p) in FP64 is the master copy; in mixed mode
// a W32 FP32 shadow feeds the flux stencils.  T and e are derived
// (T = p/(roe Rg), e = Cv T + |u|^2/2, FlowVariables.H:23,33).
//
// Hot kernel: k_stage, one workgroup per tile of TILE Morton-consecutive
// normal cells.  Phase 1 computes every face touching the tile (MUSCL + Roe)
// into LDS; phase 2 sums each cell's faces, adds the frozen viscous residual,
// applies the stage update in FP64 and recovers the primitive state.  Face
// fluxes never touch HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "ibgpu.h"
#include "ibg_flux.h"

using namespace ibg;

namespace {

constexpr int kDefaultTile = 256;
constexpr int kMaxTile = 1024;

enum CellFlag : uint8_t {
  F_UPDATE = 1,      // UpdateConsVars touches the cell
  F_PRIM = 2,        // ComptNewPrimVar recovers W from U
  F_FREE = 4,        // ComptNewPrimVar resets W to initvar
  F_RESID = 8,       // counted by ResinDomain
  F_FINEST = 16      // finest level with bodies (abort code 611 vs 523)
};

struct DevPhys {
  double gama, Ma, Re, Rg, Cv, Cp, mu, qk, SoT;
  double init[IBG_NFV];
  double inlet_T0, inlet_p0, inlet_cos, inlet_sin, outlet_p;
  double cfl, dt_ratio;
  double force_ref, p_ref_area, shear_ref_area;
};

struct BlkInfo {
  double h[3], invh[3], area[3], vol, step_num;
  int case_id, level, finest, pad;
};

struct TileInfo {
  int cell0, ncell, face0, nface, blk, pad0, pad1, pad2;
};

// error word: ((cell << 16) | code), minimum wins
__device__ __forceinline__ void report_error(unsigned long long *err, int cell, int code) {
  unsigned long long v = ((unsigned long long)(unsigned)cell << 16) | (unsigned)code;
  atomicMin(err, v);
}

template <typename R>
struct WPtr {
  R *r, *u, *v, *w, *p;
};

struct Planes64 {
  double *r, *u, *v, *w, *p;
};

// ------------------------------------------------------------ stage kernel
template <typename R, int D, bool MIXED>
struct StageArgs {
  const TileInfo *tiles;
  const int4 *tst;          // tile face stencil LL,L,R,RR
  const int8_t *tdir;       // tile face direction
  const uint16_t *slots;    // [Nn][2D] tile-local face slot per cell face
  const uint8_t *flags;     // [Nn]
  const BlkInfo *blk;
  const DevPhys *phys;
  const double *dt;         // [ncases]
  WPtr<R> W;                // stencil state (W64 or W32)
  Planes64 W64;             // master (written in FP64 mode every stage,
                            // in mixed mode at the last stage)
  const double *U0;         // [NV][Nn]
  const R *Rv;              // [NV][Nn]
  int Nn;
  double alpha;
  int limiter;
  int last;                 // last stage: residual partials, W64 write
  double *rpart;            // [ntiles][10]
  double *src;              // optional probe [N][5] (last stage)
  unsigned long long *err;
};

template <typename R, int D>
__device__ __forceinline__ Prim<R, D> load_prim(const WPtr<R> &W, int c) {
  Prim<R, D> q;
  q.r = W.r[c];
  q.u = W.u[c];
  q.v = W.v[c];
  q.w = (D == 3) ? W.w[c] : R(0);
  q.p = W.p[c];
  return q;
}

template <typename R, int D, bool MIXED>
__global__ void __launch_bounds__(kDefaultTile)
k_stage(StageArgs<R, D, MIXED> a) {
  constexpr int NV = D + 2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R *fl = reinterpret_cast<R *>(smem_raw);
  __shared__ double red[2 * NV][kDefaultTile / 64];

  const TileInfo ti = a.tiles[blockIdx.x];
  const BlkInfo bi = a.blk[ti.blk];
  const DevPhys &P = a.phys[bi.case_id];
  PhysR<R> ph;
  ph.gama = (R)P.gama;
  ph.Rg = (R)P.Rg;
  ph.Cv = (R)P.Cv;
  ph.Ma = (R)P.Ma;

  // phase 1: faces of the tile
  for (int f = threadIdx.x; f < ti.nface; f += blockDim.x) {
    const int4 st = a.tst[ti.face0 + f];
    const int dir = a.tdir[ti.face0 + f];
    Prim<R, D> q0 = load_prim<R, D>(a.W, st.x);
    Prim<R, D> q1 = load_prim<R, D>(a.W, st.y);
    Prim<R, D> q2 = load_prim<R, D>(a.W, st.z);
    Prim<R, D> q3 = load_prim<R, D>(a.W, st.w);
    R flux[NV];
    roe_flux<R, D>(q0, q1, q2, q3, dir, (R)bi.area[dir], ph, a.limiter, flux);
#pragma unroll
    for (int k = 0; k < NV; ++k) fl[f * NV + k] = flux[k];
  }
  __syncthreads();

  // phase 2: cells of the tile
  const double level_ts = a.alpha * a.dt[bi.case_id] * bi.step_num;
  double mx[NV], sq[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) { mx[k] = 0.0; sq[k] = 0.0; }
  const int i = threadIdx.x;
  if (i < ti.ncell) {
    const int c = ti.cell0 + i;
    const uint8_t fg = a.flags[c];
    // ComptSource: sum over directions of (F_lo - F_hi), plus the frozen
    // viscous part sum of (Fv_hi - Fv_lo)
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int lo = a.slots[(size_t)c * (2 * D) + 2 * d];
      const int hi = a.slots[(size_t)c * (2 * D) + 2 * d + 1];
#pragma unroll
      for (int k = 0; k < NV; ++k) s[k] += (double)fl[lo * NV + k] - (double)fl[hi * NV + k];
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += (double)a.Rv[(size_t)k * a.Nn + c];
    if (fg & F_UPDATE) {
      double U[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) U[k] = a.U0[(size_t)k * a.Nn + c] + s[k] * level_ts / bi.vol;
      if (U[0] < 0.0) report_error(a.err, c, (fg & F_FINEST) ? 611 : 523);
      double wr, wu, wv, ww = 0.0, wp;
      bool write = false;
      if (fg & F_PRIM) {
        // Get_Prim_Vars FlowVariables.H:50-57
        wr = U[0];
        wu = U[1] / U[0];
        wv = U[2] / U[0];
        double e;
        if (D == 3) { ww = U[3] / U[0]; e = U[4] / U[0]; }
        else e = U[NV - 1] / U[0];
        double ke = (D == 3) ? 0.5 * (wu * wu + wv * wv + ww * ww) : 0.5 * (wu * wu + wv * wv);
        wp = (e - ke) * (P.gama - 1.0) * wr;
        double T = wp / (wr * P.Rg);
        if (T < 0.0) report_error(a.err, c, 37);
        write = true;
      } else if (fg & F_FREE) {
        wr = P.init[0]; wu = P.init[1]; wv = P.init[2]; ww = P.init[3]; wp = P.init[4];
        write = true;
      }
      if (write) {
        if (!MIXED || a.last) {
          a.W64.r[c] = wr; a.W64.u[c] = wu; a.W64.v[c] = wv; a.W64.p[c] = wp;
          if (D == 3) a.W64.w[c] = ww;
        }
        if (MIXED) {
          a.W.r[c] = (R)wr; a.W.u[c] = (R)wu; a.W.v[c] = (R)wv; a.W.p[c] = (R)wp;
          if (D == 3) a.W.w[c] = (R)ww;
        }
      }
    } else if (fg & F_FREE) {
      // finest-level solid cells: W <- initvar (NS_Solver.cpp:170-176)
      if (!MIXED || a.last) {
        a.W64.r[c] = P.init[0]; a.W64.u[c] = P.init[1]; a.W64.v[c] = P.init[2]; a.W64.p[c] = P.init[4];
        if (D == 3) a.W64.w[c] = P.init[3];
      }
      if (MIXED) {
        a.W.r[c] = (R)P.init[0]; a.W.u[c] = (R)P.init[1]; a.W.v[c] = (R)P.init[2]; a.W.p[c] = (R)P.init[4];
        if (D == 3) a.W.w[c] = (R)P.init[3];
      }
    }
    if (a.last) {
      if (a.src) {
        // probe layout [cell][5] = (mass, x, y, z, energy)
        double *o = a.src + (size_t)c * 5;
        o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
        if (D == 3) { o[3] = s[3]; o[4] = s[4]; } else { o[3] = 0.0; o[4] = s[3]; }
      }
      if (fg & F_RESID) {
#pragma unroll
        for (int k = 0; k < NV; ++k) { mx[k] = fabs(s[k]); sq[k] = s[k] * s[k]; }
      }
    }
  }
  if (a.last) {
    // wave64 reductions, then across the waves of the workgroup
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      for (int off = 32; off > 0; off >>= 1) {
        mx[k] = fmax(mx[k], __shfl_xor(mx[k], off, 64));
        sq[k] += __shfl_xor(sq[k], off, 64);
      }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) { red[k][wid] = mx[k]; red[NV + k][wid] = sq[k]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = (blockDim.x + 63) >> 6;
      double *o = a.rpart + (size_t)blockIdx.x * 10;
      for (int k = 0; k < 5; ++k) { o[k] = 0.0; o[5 + k] = 0.0; }
      for (int k = 0; k < NV; ++k) {
        double m = 0.0, q = 0.0;
        for (int w = 0; w < nw; ++w) { m = fmax(m, red[k][w]); q += red[NV + k][w]; }
        int kk = (D == 3) ? k : (k == 3 ? 4 : k);
        o[kk] = m;
        o[5 + kk] = q;
      }
    }
  }
}

// ------------------------------------------------- viscous face flux kernel
// ComptViscousStress NS_Solver.cpp:385-621, laminar.  One thread per face of
// the marched prefix; writes vs[k][F] in the kernel's arithmetic type.
template <typename R, int D>
__global__ void k_viscous_face(int nface, const int2 *face_lohi, const int8_t *face_dir,
                               const int16_t *face_blk, const int *nb27, WPtr<R> W,
                               const BlkInfo *blk, const DevPhys *phys, R *vs, int F) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface) return;
  constexpr int NV = D + 2;
  const int2 lh = face_lohi[f];
  const int d = face_dir[f];
  const BlkInfo bi = blk[face_blk[f]];
  const DevPhys &P = phys[bi.case_id];
  const R Rg = (R)P.Rg;
  const int *nb = nb27 + (size_t)lh.x * 27;
  auto T_of = [&](int c) -> R { return W.p[c] / (W.r[c] * Rg); };
  // neighbour (i,j,k) in 0..2 of abox
#define NB(i, j, k) nb[(i) * 9 + (j) * 3 + (k)]
  R g[4][3];
  const int kc = 1;
  for (int q = 0; q < 4; ++q) {
    if (D == 2 && q == 2) { g[q][0] = g[q][1] = g[q][2] = R(0); continue; }
    auto val = [&](int c) -> R {
      if (q == 0) return W.u[c];
      if (q == 1) return W.v[c];
      if (q == 2) return W.w[c];
      return T_of(c);
    };
    if (d == 0) {
      g[q][0] = val(NB(2, 1, kc)) - val(NB(1, 1, kc));
      g[q][1] = R(0.25) * (val(NB(1, 2, kc)) + val(NB(2, 2, kc)) - val(NB(1, 0, kc)) - val(NB(2, 0, kc)));
      g[q][2] = (D == 3) ? R(0.25) * (val(NB(1, 1, 2)) + val(NB(2, 1, 2)) - val(NB(1, 1, 0)) - val(NB(2, 1, 0))) : R(0);
    } else if (d == 1) {
      g[q][1] = val(NB(1, 2, kc)) - val(NB(1, 1, kc));
      g[q][0] = R(0.25) * (val(NB(2, 1, kc)) + val(NB(2, 2, kc)) - val(NB(0, 1, kc)) - val(NB(0, 2, kc)));
      g[q][2] = (D == 3) ? R(0.25) * (val(NB(1, 1, 2)) + val(NB(1, 2, 2)) - val(NB(1, 1, 0)) - val(NB(1, 2, 0))) : R(0);
    } else {
      g[q][2] = val(NB(1, 1, 2)) - val(NB(1, 1, 1));
      g[q][0] = R(0.25) * (val(NB(2, 1, 1)) + val(NB(2, 1, 2)) - val(NB(0, 1, 1)) - val(NB(0, 1, 2)));
      g[q][1] = R(0.25) * (val(NB(1, 2, 1)) + val(NB(1, 2, 2)) - val(NB(1, 0, 1)) - val(NB(1, 0, 2)));
    }
  }
#undef NB
  const R k0 = (R)bi.invh[0], k1 = (R)bi.invh[1], k2 = (R)bi.invh[2];
  R dux = g[0][0] * k0, duy = g[0][1] * k1, duz = g[0][2] * k2;
  R dvx = g[1][0] * k0, dvy = g[1][1] * k1, dvz = g[1][2] * k2;
  R dwx = g[2][0] * k0, dwy = g[2][1] * k1, dwz = g[2][2] * k2;
  R dTx = g[3][0] * k0, dTy = g[3][1] * k1, dTz = g[3][2] * k2;
  const int L = lh.x, Rr = lh.y;
  R ftemp = (T_of(L) + T_of(Rr)) * R(0.5);
  const R SoT = (R)P.SoT;
  // Sutherland: sqrt(pow(T,3)) (1+S)/(T+S)  (NS_Solver.cpp:532)
  R mu0 = r_sqrt(ftemp * ftemp * ftemp) * ((R(1.0) + SoT) / (ftemp + SoT));
  R qk0 = (R)P.qk * mu0;
  const R Re = (R)P.Re;
  R divu = (D == 3) ? (dux + dvy + dwz) : (dux + dvy);
  R tauxx = R(2.0) * mu0 * (dux - divu / R(3.0)) / Re;
  R tauyy = R(2.0) * mu0 * (dvy - divu / R(3.0)) / Re;
  R tauzz = R(2.0) * mu0 * (dwz - divu / R(3.0)) / Re;
  R tauxy = mu0 / Re * (duy + dvx);
  R tauxz = mu0 / Re * (duz + dwx);
  R tauyz = mu0 / Re * (dvz + dwy);
  R fu = (W.u[L] + W.u[Rr]) * R(0.5), fv = (W.v[L] + W.v[Rr]) * R(0.5);
  R fw = (D == 3) ? (W.w[L] + W.w[Rr]) * R(0.5) : R(0);
  R bx, by, bz;
  if (D == 3) {
    bx = fu * tauxx + fv * tauxy + fw * tauxz + qk0 * dTx;
    by = fu * tauxy + fv * tauyy + fw * tauyz + qk0 * dTy;
    bz = fu * tauxz + fv * tauyz + fw * tauzz + qk0 * dTz;
  } else {
    bx = fu * tauxx + fv * tauxy + qk0 * dTx;
    by = fu * tauxy + fv * tauyy + qk0 * dTy;
    bz = R(0);
  }
  const R area = (R)bi.area[d];
  R o[5];
  o[0] = R(0);
  if (d == 0) { o[1] = tauxx * area; o[2] = tauxy * area; o[3] = tauxz * area; o[4] = bx * area; }
  else if (d == 1) { o[1] = tauxy * area; o[2] = tauyy * area; o[3] = tauyz * area; o[4] = by * area; }
  else { o[1] = tauxz * area; o[2] = tauyz * area; o[3] = tauzz * area; o[4] = bz * area; }
  if (D == 3) {
    for (int k = 0; k < 5; ++k) vs[(size_t)k * F + f] = o[k];
  } else {
    vs[(size_t)0 * F + f] = o[0];
    vs[(size_t)1 * F + f] = o[1];
    vs[(size_t)2 * F + f] = o[2];
    vs[(size_t)3 * F + f] = o[4];
  }
  (void)NV;
}

// U_old = cons(W64) (ComptOldConsVar NS_Solver.H:392-424) and the frozen
// viscous residual Rv = sum_d (Fv_hi - Fv_lo) for the marched normal cells.
template <typename R, int D>
__global__ void k_prep_cells(int ncell, Planes64 W64, const DevPhys *phys, const BlkInfo *blk,
                             const int16_t *cellblk, const int *cfaces, const R *vs, int F,
                             double *U0, R *Rv, int Nn, int viscous) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  constexpr int NV = D + 2;
  const DevPhys &P = phys[blk[cellblk[c]].case_id];
  const double r = W64.r[c], u = W64.u[c], v = W64.v[c], p = W64.p[c];
  const double w = (D == 3) ? W64.w[c] : 0.0;
  const double T = p / (r * P.Rg);
  const double ke = (D == 3) ? 0.5 * (u * u + v * v + w * w) : 0.5 * (u * u + v * v);
  const double e = P.Cv * T + ke;
  U0[(size_t)0 * Nn + c] = r;
  U0[(size_t)1 * Nn + c] = r * u;
  U0[(size_t)2 * Nn + c] = r * v;
  if (D == 3) {
    U0[(size_t)3 * Nn + c] = r * w;
    U0[(size_t)4 * Nn + c] = r * e;
  } else {
    U0[(size_t)3 * Nn + c] = r * e;
  }
  double s[NV];
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  if (viscous) {
    for (int d = 0; d < D; ++d) {
      const int flo = cfaces[(size_t)c * 2 * D + 2 * d];
      const int fhi = cfaces[(size_t)c * 2 * D + 2 * d + 1];
      for (int k = 0; k < NV; ++k)
        s[k] += (double)vs[(size_t)k * F + fhi] - (double)vs[(size_t)k * F + flo];
    }
  }
  for (int k = 0; k < NV; ++k) Rv[(size_t)k * Nn + c] = (R)s[k];
}

// ------------------------------------------------ ghost fills (FP64 master)
__device__ __forceinline__ void store_w(Planes64 W64, WPtr<float> W32, bool mixed, int D, int c,
                                        double r, double u, double v, double w, double p) {
  W64.r[c] = r; W64.u[c] = u; W64.v[c] = v; W64.p[c] = p;
  if (D == 3) W64.w[c] = w;
  if (mixed) {
    W32.r[c] = (float)r; W32.u[c] = (float)u; W32.v[c] = (float)v; W32.p[c] = (float)p;
    if (D == 3) W32.w[c] = (float)w;
  }
}

// BCBoxTreat BCValues.H:51-188: -13 pressure inlet, -12 pressure outlet,
// -50 outflow; the entries are (cell, refcell, type, case).
__global__ void k_bc(int n, const int4 *dmg, Planes64 W64, WPtr<float> W32, int mixed, int D,
                     const DevPhys *phys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int4 e = dmg[i];
  const int g = e.x, rf = e.y, ty = e.z;
  const DevPhys &P = phys[e.w];
  const double rr = W64.r[rf], ru = W64.u[rf], rv = W64.v[rf], rp = W64.p[rf];
  const double rw = (D == 3) ? W64.w[rf] : 0.0;
  if (ty == -50) {
    store_w(W64, W32, mixed, D, g, rr, ru, rv, rw, rp);
  } else if (ty == -12) {
    const double T = rp / (rr * P.Rg);
    const double p = P.outlet_p;
    const double r = p / (P.Rg * T);
    store_w(W64, W32, mixed, D, g, r, ru, rv, rw, p);
  } else if (ty == -13) {
    const double speed0 = sqrt(ru * ru + rv * rv + rw * rw);
    const double u = speed0 * P.inlet_cos, v = speed0 * P.inlet_sin;
    const double T = P.inlet_T0 - (0.5 * (u * u + v * v)) / P.Cp;
    const double p = P.inlet_p0 * pow((T / P.inlet_T0), 3.5);
    const double r = p / (P.Rg * T);
    store_w(W64, W32, mixed, D, g, r, u, v, 0.0, p);
  }
}

// PairRestriction[_2d] TwoLevelOp.H:217-231 with AverageCell[_2d] :45-87
__global__ void k_restrict(int n, const int *ci, const int *fi, Planes64 W64, WPtr<float> W32,
                           int mixed, int D, int twod, const DevPhys *phys, const BlkInfo *blk,
                           const int16_t *cellblk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = ci[i];
  const DevPhys &P = phys[blk[cellblk[c]].case_id];
  double r = 0, ru = 0, rv = 0, rw = 0, re = 0;
  for (int q = 0; q < 8; ++q) {
    if (twod && (q & 1)) continue;
    const int s = fi[(size_t)i * 8 + q];
    const double sr = W64.r[s], su = W64.u[s], sv = W64.v[s], sp = W64.p[s];
    const double sw = (D == 3) ? W64.w[s] : 0.0;
    const double T = sp / (sr * P.Rg);
    const double e = P.Cv * T + ((D == 3) ? 0.5 * (su * su + sv * sv + sw * sw) : 0.5 * (su * su + sv * sv));
    r += sr;
    ru += su * sr;
    rv += sv * sr;
    rw += sw * sr;
    re += e * sr;
  }
  const double u = ru / r, v = rv / r, w = rw / r, e = re / r;
  const double rho = r * (twod ? 0.25 : 0.125);
  const double ke = (D == 3) ? 0.5 * (u * u + v * v + w * w) : 0.5 * (u * u + v * v);
  const double p = (e - ke) * (P.gama - 1.0) * rho;
  store_w(W64, W32, mixed, D, c, rho, u, v, w, p);
}

// PairProlongation[_2d] TwoLevelOp.H:183-215: central gradients of the
// conserved variables over the 6 face neighbours, sons at +-1/4.
__global__ void k_prolong(int n, const int *ci, const int *fi, const int *nb27, Planes64 W64,
                          WPtr<float> W32, int mixed, int D, int twod, const DevPhys *phys,
                          const BlkInfo *blk, const int16_t *cellblk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = ci[i];
  const DevPhys &P = phys[blk[cellblk[c]].case_id];
  const int *nb = nb27 + (size_t)c * 27;
  auto cons = [&](int a, double *U) {
    const double r = W64.r[a], u = W64.u[a], v = W64.v[a], p = W64.p[a];
    const double w = (D == 3) ? W64.w[a] : 0.0;
    const double T = p / (r * P.Rg);
    const double e = P.Cv * T + ((D == 3) ? 0.5 * (u * u + v * v + w * w) : 0.5 * (u * u + v * v));
    U[0] = r; U[1] = r * u; U[2] = r * v; U[3] = r * w; U[4] = r * e;
  };
  double Uc[5], Ux0[5], Ux1[5], Uy0[5], Uy1[5], Uz0[5], Uz1[5];
  cons(c, Uc);
  cons(nb[0 * 9 + 1 * 3 + 1], Ux0); cons(nb[2 * 9 + 1 * 3 + 1], Ux1);
  cons(nb[1 * 9 + 0 * 3 + 1], Uy0); cons(nb[1 * 9 + 2 * 3 + 1], Uy1);
  cons(nb[1 * 9 + 1 * 3 + 0], Uz0); cons(nb[1 * 9 + 1 * 3 + 2], Uz1);
  double gx[5], gy[5], gz[5];
  for (int t = 0; t < 5; ++t) {
    gx[t] = (Ux1[t] - Ux0[t]) * 0.5;
    gy[t] = (Uy1[t] - Uy0[t]) * 0.5;
    gz[t] = (Uz1[t] - Uz0[t]) * 0.5;
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
        store_w(W64, W32, mixed, D, f, r, u, v, w, p);
      }
}

struct Motion {
  double tra[3], rot[3], center[3];
};

// Body::IBConditions_Image Body.cpp:1366-1496, laminar: image state from
// the 2x2(x2) interpolation cells (Mesh.cpp:1532-1547), then LinearDist
// (Body.cpp:183-204) or the mirror.  The interpolation cells of an image
// point are never IB cells (checked at upload), so one pass is the same as
// the reference's compute-all-then-write-all.
__global__ void k_ib(int n, const int *cell, const int *intp, const double *coef,
                     const double *hgdis, const double *attach, const int *body, const int *ibcase,
                     const double *signdis, const Motion *mot, int have_motion, Planes64 W64,
                     WPtr<float> W32, int mixed, int D, const DevPhys *phys,
                     unsigned long long *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevPhys &P = phys[ibcase[i]];
  double T = 0, p = 0, u = 0, v = 0, w = 0;
  for (int q = 0; q < 8; ++q) {
    const int a = intp[(size_t)i * 8 + q];
    if (a < 0) continue;
    const double cq = coef[(size_t)i * 8 + q];
    const double ar = W64.r[a], ap = W64.p[a];
    const double aT = ap / (ar * P.Rg);
    T += cq * aT;
    p += cq * ap;
    u += cq * W64.u[a];
    v += cq * W64.v[a];
    if (D == 3) w += cq * W64.w[a];
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
  const double rho = p / (P.Rg * T);
  if (T < 0.0) report_error(err, c, 1050);
  store_w(W64, W32, mixed, D, c, rho, ou, ov, (D == 3) ? ow : 0.0, p);
}

// Comptdt NS_Solver.cpp:229-262: per-case minimum of CFL V / sum_d
// (|u_d| + c) A_d over every normal cell; a wave64 minimum, then one
// atomicMin per wave on the bit pattern of the (positive) double.
__global__ void k_cfl(int n, Planes64 W64, int D, const int16_t *cellblk, const BlkInfo *blk,
                      const DevPhys *phys, unsigned long long *dtred, unsigned long long *err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double d0 = 1.0e300;
  int cs = -1;
  if (c < n) {
    const BlkInfo bi = blk[cellblk[c]];
    const DevPhys &P = phys[bi.case_id];
    cs = bi.case_id;
    const double r = W64.r[c], p = W64.p[c];
    const double T = p / (r * P.Rg);
    const double a = sqrt(P.gama * P.Rg * T);
    const double wz = (D == 3) ? W64.w[c] : 0.0;
    const double c0 = (fabs(W64.u[c]) + a) * bi.area[0];
    const double c1 = (fabs(W64.v[c]) + a) * bi.area[1];
    const double c2 = (fabs(wz) + a) * bi.area[2];
    d0 = P.cfl * bi.vol / (c0 + c1 + c2);
    if (d0 < 0.0) report_error(err, c, 424);
  }
  const int cs0 = __shfl(cs, 0, 64);
  const bool uniform = __all(cs == cs0 || cs < 0);
  if (uniform) {
    double m = d0;
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0 && cs0 >= 0 && m > 0.0)
      atomicMin(&dtred[cs0], (unsigned long long)__double_as_longlong(m));
  } else if (cs >= 0 && d0 > 0.0) {
    atomicMin(&dtred[cs], (unsigned long long)__double_as_longlong(d0));
  }
}

__global__ void k_dt_final(int ncases, const unsigned long long *dtred, double *dt,
                           const DevPhys *phys) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncases) return;
  double cfl_dt = __longlong_as_double((long long)dtred[c]);
  double d = dt[c] * phys[c].dt_ratio;
  dt[c] = (cfl_dt < d) ? cfl_dt : d;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Body::ComptWallForce Body.cpp:460-555, one workgroup per case, summed over
// that case's patches in a fixed order (deterministic).
__global__ void k_forces(int npatch, const int *pcase, const int *intp, const double *coef,
                         const double *hgdis, const double *nv, const double *pc,
                         const double *area, const double *bcenter, const Motion *mot,
                         int have_motion, Planes64 W64, int D, const DevPhys *phys,
                         double *out) {
  const int c = blockIdx.x;
  const DevPhys &P = phys[c];
  __shared__ double acc[256][12];
  double s[12];
  for (int k = 0; k < 12; ++k) s[k] = 0.0;
  const double *bc = bcenter + (size_t)c * 3;
  for (int i = threadIdx.x; i < npatch; i += blockDim.x) {
    if (pcase[i] != c) continue;
    double T = 0, p = 0, u = 0, v = 0, w = 0;
    for (int q = 0; q < 8; ++q) {
      const int a = intp[(size_t)i * 8 + q];
      if (a < 0) continue;
      const double cq = coef[(size_t)i * 8 + q];
      T += cq * (W64.p[a] / (W64.r[a] * P.Rg));
      p += cq * W64.p[a];
      u += cq * W64.u[a];
      v += cq * W64.v[a];
      if (D == 3) w += cq * W64.w[a];
    }
    (void)T;
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
    const double du[3] = {u - pv[0], v - pv[1], w - pv[2]};
    const double dn = du[0] * n[0] + du[1] * n[1] + du[2] * n[2];
    const double A = area[i];
    double df0[3], df1[3];
    for (int k = 0; k < 3; ++k) {
      df0[k] = ((du[k] - n[k] * dn) / hgdis[i]) * P.mu * A;
      df1[k] = n[k] * (-p * A);
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
      out[c * 6 + k] = t[k] * P.p_ref_area + t[3 + k] * P.shear_ref_area;
      out[c * 6 + 3 + k] = t[6 + k] * P.p_ref_area + t[9 + k] * P.shear_ref_area;
    }
  }
}

// residual finalize: per case, over the tiles of the last march
__global__ void k_resid_final(int ntiles, const TileInfo *tiles, const BlkInfo *blk,
                              const double *rpart, double *out /* [ncases][10] */, int ncases) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncases) return;
  double m[5] = {0, 0, 0, 0, 0}, q[5] = {0, 0, 0, 0, 0};
  for (int t = 0; t < ntiles; ++t) {
    if (blk[tiles[t].blk].case_id != c) continue;
    for (int k = 0; k < 5; ++k) {
      m[k] = fmax(m[k], rpart[(size_t)t * 10 + k]);
      q[k] += rpart[(size_t)t * 10 + 5 + k];
    }
  }
  for (int k = 0; k < 5; ++k) { out[c * 10 + k] = m[k]; out[c * 10 + 5 + k] = sqrt(q[k]); }
}

// parity probe: inviscid flux of every face of the marched prefix
template <typename R, int D>
__global__ void k_face_probe(int nface, const int4 *st, const int8_t *dir, const int16_t *fblk,
                             WPtr<R> W, const BlkInfo *blk, const DevPhys *phys, int lim,
                             double *out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nface) return;
  constexpr int NV = D + 2;
  const BlkInfo bi = blk[fblk[f]];
  const DevPhys &P = phys[bi.case_id];
  PhysR<R> ph;
  ph.gama = (R)P.gama; ph.Rg = (R)P.Rg; ph.Cv = (R)P.Cv; ph.Ma = (R)P.Ma;
  const int4 s = st[f];
  R flux[NV];
  roe_flux<R, D>(load_prim<R, D>(W, s.x), load_prim<R, D>(W, s.y), load_prim<R, D>(W, s.z),
                 load_prim<R, D>(W, s.w), dir[f], (R)bi.area[dir[f]], ph, lim, flux);
  double *o = out + (size_t)f * 5;
  if (D == 3) { for (int k = 0; k < 5; ++k) o[k] = flux[k]; }
  else { o[0] = flux[0]; o[1] = flux[1]; o[2] = flux[2]; o[3] = 0.0; o[4] = flux[3]; }
}

}  // namespace

// ===================================================================== host
struct ibg_ctx {
  ibg_config cfg{};
  hipStream_t s = nullptr;
  bool mixed = false;
  int D = 2, NV = 4, ncases = 0, L = 0, nblk = 0;
  int N = 0, Nn = 0, F = 0, ntiles = 0, tile = kDefaultTile, max_tile_faces = 0;
  int nbody = 0, npatch = 0, nib = 0;
  bool have_mesh = false;
  std::vector<int> h2d, d2h;                      // cell order maps
  std::vector<int> fh2d;                          // face order map (host face -> device face)
  std::vector<int> cell_end, face_end, tile_end;  // prefix ends per lmin
  std::vector<int> res_off, pro_off;              // per coarse level
  std::vector<int> blk_twod;
  std::vector<int16_t> hcellblk;
  std::vector<DevPhys> hphys;
  std::vector<double> body_center;
  std::vector<int> hbody_case;
  int ndmg_edge = 0, ndmg_corner = 0;
  // device buffers
  Planes64 W64{};
  WPtr<float> W32{};
  double *U0 = nullptr;
  void *Rv = nullptr, *vs = nullptr;
  uint8_t *flags = nullptr;
  int16_t *cellblk = nullptr, *face_blk = nullptr, *tface_blk = nullptr;
  int *nb27 = nullptr, *cfaces = nullptr;
  int2 *face_lohi = nullptr;
  int8_t *face_dir = nullptr;
  int4 *face_st = nullptr;
  TileInfo *tiles = nullptr;
  int4 *tst = nullptr;
  int8_t *tdir = nullptr;
  uint16_t *slots = nullptr;
  BlkInfo *blk = nullptr;
  DevPhys *phys = nullptr;
  int4 *dmg = nullptr;
  int *res_ci = nullptr, *res_fi = nullptr, *pro_ci = nullptr, *pro_fi = nullptr;
  int *ib_cell = nullptr, *ib_intp = nullptr, *ib_body = nullptr, *ib_case = nullptr;
  double *ib_coef = nullptr, *ib_hgdis = nullptr, *ib_attach = nullptr, *signdis = nullptr;
  int *pat_intp = nullptr, *pat_case = nullptr;
  double *pat_coef = nullptr, *pat_hgdis = nullptr, *pat_nv = nullptr, *pat_pc = nullptr,
         *pat_area = nullptr, *case_center = nullptr;
  Motion *motion = nullptr;
  int have_motion = 0;
  double *dt = nullptr;
  unsigned long long *dtred = nullptr, *err = nullptr;
  double *rpart = nullptr, *resid = nullptr;
  double *probe = nullptr;     // N*5 or F*5 scratch for probes
  int last_ntiles = 0;
  // graph cache
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int gts_phase_len = 0;
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_total_ms = 0.0, last_stage_ms = 0.0;
  int64_t last_stage_launches = 0, last_updates = 0;
  std::vector<int64_t> upd_per_lmin;   // cell-stage count per step for each lmin
  std::vector<void *> allocs;
  std::string errmsg;
};

namespace {

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "ibgpu: %s failed: %s (%s:%d)\n", #x,              \
              hipGetErrorString(e_), __FILE__, __LINE__);                \
      return IBG_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

template <typename T>
ibg_status dalloc(ibg_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  size_t bytes = std::max<size_t>(n * sizeof(T), 16);
  HIPCHK(hipMalloc(&q, bytes));
  c->allocs.push_back(q);
  *p = reinterpret_cast<T *>(q);
  return IBG_OK;
}

template <typename T>
ibg_status dupload(ibg_ctx *c, T **p, const std::vector<T> &v) {
  ibg_status st = dalloc(c, p, v.size());
  if (st) return st;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return IBG_OK;
}

void free_all(ibg_ctx *c) {
  for (int k = 0; k < 2; ++k)
    if (c->gexec[k]) { (void)hipGraphExecDestroy(c->gexec[k]); c->gexec[k] = nullptr; }
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  c->have_mesh = false;
}

inline int marched(const ibg_ctx *c, int level, int ts) {
  if (!c->cfg.temporal_refine) return 1;
  return (ts % (1 << (c->L - 1 - level))) == 0;
}
inline int lmin_of(const ibg_ctx *c, int ts) {
  int l = c->L - 1;
  while (l > 0 && marched(c, l - 1, ts)) --l;
  return l;
}
inline int transfer_due(const ibg_ctx *c, int cl, int ts) {
  if (!c->cfg.temporal_refine) return 1;
  const int i = cl + 1;
  return (ts % (1 << (c->L - i))) == (1 << (c->L - i - 1));
}
inline unsigned grid(int n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------- kernels run
namespace {

template <typename R, int D, bool MIXED>
ibg_status launch_stage(ibg_ctx *c, int ntile, int subt, bool last, double *src) {
  StageArgs<R, D, MIXED> a;
  a.tiles = c->tiles; a.tst = c->tst; a.tdir = c->tdir; a.slots = c->slots;
  a.flags = c->flags; a.blk = c->blk; a.phys = c->phys; a.dt = c->dt;
  if (MIXED) {
    a.W.r = (R *)c->W32.r; a.W.u = (R *)c->W32.u; a.W.v = (R *)c->W32.v; a.W.w = (R *)c->W32.w; a.W.p = (R *)c->W32.p;
  } else {
    a.W.r = (R *)c->W64.r; a.W.u = (R *)c->W64.u; a.W.v = (R *)c->W64.v; a.W.w = (R *)c->W64.w; a.W.p = (R *)c->W64.p;
  }
  a.W64 = c->W64;
  a.U0 = c->U0;
  a.Rv = (const R *)c->Rv;
  a.Nn = c->Nn;
  a.alpha = c->cfg.alpha[subt];
  a.limiter = c->cfg.limiter;
  a.last = last ? 1 : 0;
  a.rpart = c->rpart;
  a.src = src;
  a.err = c->err;
  if (ntile <= 0) return IBG_OK;
  const size_t shm = (size_t)c->max_tile_faces * (D + 2) * sizeof(R);
  hipLaunchKernelGGL((k_stage<R, D, MIXED>), dim3(ntile), dim3(c->tile), shm, c->s, a);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

ibg_status run_stage(ibg_ctx *c, int ts, int subt, double *src) {
  const int lmin = lmin_of(c, ts);
  const int nt = c->tile_end[lmin];
  const bool last = (subt == c->cfg.nstage - 1);
  if (last) c->last_ntiles = nt;
  if (c->D == 2) {
    return c->mixed ? launch_stage<float, 2, true>(c, nt, subt, last, src)
                    : launch_stage<double, 2, false>(c, nt, subt, last, src);
  }
  return c->mixed ? launch_stage<float, 3, true>(c, nt, subt, last, src)
                  : launch_stage<double, 3, false>(c, nt, subt, last, src);
}

template <typename R, int D>
ibg_status launch_viscous(ibg_ctx *c, int nf) {
  WPtr<R> W;
  if (std::is_same<R, float>::value) {
    W.r = (R *)c->W32.r; W.u = (R *)c->W32.u; W.v = (R *)c->W32.v; W.w = (R *)c->W32.w; W.p = (R *)c->W32.p;
  } else {
    W.r = (R *)c->W64.r; W.u = (R *)c->W64.u; W.v = (R *)c->W64.v; W.w = (R *)c->W64.w; W.p = (R *)c->W64.p;
  }
  if (nf > 0) {
    hipLaunchKernelGGL((k_viscous_face<R, D>), dim3(grid(nf, 256)), dim3(256), 0, c->s, nf,
                       c->face_lohi, c->face_dir, c->face_blk, c->nb27, W, c->blk, c->phys,
                       (R *)c->vs, c->F);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

template <typename R, int D>
ibg_status launch_prep(ibg_ctx *c, int nc) {
  if (nc > 0) {
    hipLaunchKernelGGL((k_prep_cells<R, D>), dim3(grid(nc, 256)), dim3(256), 0, c->s, nc, c->W64,
                       c->phys, c->blk, c->cellblk, c->cfaces, (const R *)c->vs, c->F, c->U0,
                       (R *)c->Rv, c->Nn, c->cfg.viscous);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

ibg_status run_viscous_and_prep(ibg_ctx *c, int ts) {
  const int lmin = lmin_of(c, ts);
  const int nf = c->face_end[lmin], nc = c->cell_end[lmin];
  ibg_status st = IBG_OK;
  if (c->cfg.viscous) {
    if (c->D == 2) st = c->mixed ? launch_viscous<float, 2>(c, nf) : launch_viscous<double, 2>(c, nf);
    else st = c->mixed ? launch_viscous<float, 3>(c, nf) : launch_viscous<double, 3>(c, nf);
    if (st) return st;
  }
  if (c->D == 2) return c->mixed ? launch_prep<float, 2>(c, nc) : launch_prep<double, 2>(c, nc);
  return c->mixed ? launch_prep<float, 3>(c, nc) : launch_prep<double, 3>(c, nc);
}

ibg_status run_bc(ibg_ctx *c, int ts) {
  // domain ghosts live on level 0 (BCBoxTreat runs on marched levels only)
  if (!marched(c, 0, ts)) return IBG_OK;
  if (c->ndmg_edge > 0) {
    hipLaunchKernelGGL(k_bc, dim3(grid(c->ndmg_edge, 256)), dim3(256), 0, c->s, c->ndmg_edge,
                       c->dmg, c->W64, c->W32, (int)c->mixed, c->D, c->phys);
    HIPCHK(hipGetLastError());
  }
  if (c->ndmg_corner > 0) {
    hipLaunchKernelGGL(k_bc, dim3(grid(c->ndmg_corner, 256)), dim3(256), 0, c->s, c->ndmg_corner,
                       c->dmg + c->ndmg_edge, c->W64, c->W32, (int)c->mixed, c->D, c->phys);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

ibg_status run_exchange(ibg_ctx *c, int ts) {
  for (int l = 0; l + 1 < c->L; ++l) {
    if (!transfer_due(c, l, ts)) continue;
    const int n = c->res_off[l + 1] - c->res_off[l];
    if (n > 0) {
      hipLaunchKernelGGL(k_restrict, dim3(grid(n, 256)), dim3(256), 0, c->s, n,
                         c->res_ci + c->res_off[l], c->res_fi + (size_t)8 * c->res_off[l], c->W64,
                         c->W32, (int)c->mixed, c->D, c->blk_twod[l + 1], c->phys, c->blk,
                         c->cellblk);
      HIPCHK(hipGetLastError());
    }
  }
  for (int l = 0; l + 1 < c->L; ++l) {
    if (!transfer_due(c, l, ts)) continue;
    const int n = c->pro_off[l + 1] - c->pro_off[l];
    if (n > 0) {
      hipLaunchKernelGGL(k_prolong, dim3(grid(n, 256)), dim3(256), 0, c->s, n,
                         c->pro_ci + c->pro_off[l], c->pro_fi + (size_t)8 * c->pro_off[l], c->nb27,
                         c->W64, c->W32, (int)c->mixed, c->D, c->blk_twod[l + 1], c->phys, c->blk,
                         c->cellblk);
      HIPCHK(hipGetLastError());
    }
  }
  return IBG_OK;
}

ibg_status run_ib(ibg_ctx *c) {
  if (c->nib > 0) {
    hipLaunchKernelGGL(k_ib, dim3(grid(c->nib, 256)), dim3(256), 0, c->s, c->nib, c->ib_cell,
                       c->ib_intp, c->ib_coef, c->ib_hgdis, c->ib_attach, c->ib_body, c->ib_case,
                       c->signdis, c->motion, c->have_motion, c->W64, c->W32, (int)c->mixed, c->D,
                       c->phys, c->err);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

ibg_status run_dt(ibg_ctx *c) {
  const unsigned long long cap = (unsigned long long)0x40C3880000000000ull;   // 10000.0
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->s, c->dtred, c->ncases, cap);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_cfl, dim3(grid(c->Nn, 256)), dim3(256), 0, c->s, c->Nn, c->W64, c->D,
                     c->cellblk, c->blk, c->phys, c->dtred, c->err);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(64), 0, c->s, c->ncases, c->dtred, c->dt, c->phys);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

// one full step: main.cpp:87-97
ibg_status run_step(ibg_ctx *c, int ts) {
  ibg_status st;
  if ((st = run_exchange(c, ts))) return st;
  if ((st = run_ib(c))) return st;
  if (c->cfg.dt_var_step > 0 && ts % c->cfg.dt_var_step == 0)
    if ((st = run_dt(c))) return st;
  if ((st = run_bc(c, ts))) return st;
  if ((st = run_viscous_and_prep(c, ts))) return st;
  for (int k = 0; k < c->cfg.nstage; ++k)
    if ((st = run_stage(c, ts, k, nullptr))) return st;
  return IBG_OK;
}

ibg_status reset_err(ibg_ctx *c) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->s, c->err, 1, ~0ull);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

ibg_status read_err(ibg_ctx *c, int64_t *cell, int32_t *code) {
  unsigned long long v = 0;
  HIPCHK(hipMemcpyAsync(&v, c->err, sizeof(v), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (v == ~0ull) { *cell = -1; *code = 0; return IBG_OK; }
  int dcell = (int)(v >> 16);
  *code = (int32_t)(v & 0xffff);
  *cell = (dcell >= 0 && dcell < c->N) ? c->d2h[dcell] : -1;
  return IBG_OK;
}

}  // namespace

// Host-side validation of every index a kernel will dereference, so that a
// malformed mesh is rejected before any launch (the reference asserts the
// same facts under -D DEBUG: NS_Solver.H:741, NS_Solver.cpp:421,
// TwoLevelOp.H:188, Mesh.cpp:1539).
namespace {
std::string g_check_msg;

ibg_status validate_mesh(const ibg_mesh_soa *m) {
  auto fail = [](const std::string &s) { g_check_msg = s; return IBG_ERR_ARG; };
  if (!m) return fail("null mesh");
  if (m->dim != 2 && m->dim != 3) return fail("dim must be 2 or 3");
  const int L = m->nlevels, nblk = m->ncases * L, N = m->ncells, D = m->dim;
  if (m->ncases < 1 || L < 1) return fail("empty mesh");
  if (m->blk_start[0] != 0 || m->blk_start[nblk] != N) return fail("blk_start does not span the cells");
  for (int b = 0; b < nblk; ++b) {
    if (m->blk_start[b + 1] < m->blk_start[b]) return fail("blk_start not monotone");
    if (m->blk_nnormal[b] < 0 || m->blk_nnormal[b] > m->blk_start[b + 1] - m->blk_start[b])
      return fail("blk_nnormal out of range");
    for (int g = m->blk_start[b]; g < m->blk_start[b] + m->blk_nnormal[b]; ++g)
      if (m->kind[g] != IBG_NORMAL) return fail("a ghost inside the normal range of a block");
  }
  auto in = [N](int a) { return a >= 0 && a < N; };
  auto nb = [m](int c, int i, int j, int k) { return m->nb[(size_t)c * 27 + i * 9 + j * 3 + k]; };
  for (int b = 0; b < nblk; ++b)
    for (int g = m->blk_start[b]; g < m->blk_start[b] + m->blk_nnormal[b]; ++g)
      for (int d = 0; d < D; ++d)
        for (int s2 = 0; s2 < 2; ++s2) {
          const int f = m->cell_faces[(size_t)g * 6 + d * 2 + s2];
          if (f < 0 || f >= m->nfaces) return fail("normal cell without face");
          if (m->face_cells[2 * f + (1 - s2)] != g) return fail("cell_faces and face_cells disagree");
        }
  for (int f = 0; f < m->nfaces; ++f) {
    const int lo = m->face_cells[2 * f], hi = m->face_cells[2 * f + 1], d = m->face_dir[f];
    if (!in(lo) || !in(hi) || d < 0 || d >= D) return fail("face out of range");
    int o[3] = {1, 1, 1};
    o[d] = 0;
    int LL = nb(lo, o[0], o[1], o[2]);
    o[d] = 2;
    int RR = nb(hi, o[0], o[1], o[2]);
    int LLf = LL, RRf = RR;
    if (LL < 0) LLf = RR;
    if (RR < 0) RRf = LL;
    if (!in(LLf) || !in(RRf)) return fail("face stencil without LL/RR");
    // viscous stencil of the low cell (NS_Solver.cpp:413-425)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          int idx[3] = {i, j, k};
          if (idx[d] == 0) continue;
          if (D == 2 && k != 1) continue;
          if (!in(nb(lo, i, j, k))) return fail("viscous stencil neighbour missing");
        }
  }
  for (int i = 0; i < m->nres; ++i) {
    if (!in(m->res_ci[i])) return fail("restriction coarse cell");
    const int b = m->res_blk[i];
    if (b < 0 || b >= nblk || b % L == L - 1) return fail("restriction block");
    const int twod = m->blk_twod[b + 1];
    for (int q = 0; q < 8; ++q) {
      if (twod && (q & 1)) continue;
      if (!in(m->res_fi[(size_t)i * 8 + q])) return fail("restriction son missing");
    }
  }
  for (int i = 0; i < m->npro; ++i) {
    const int c = m->pro_ci[i];
    if (!in(c)) return fail("prolongation coarse cell");
    const int b = m->pro_blk[i];
    if (b < 0 || b >= nblk || b % L == L - 1) return fail("prolongation block");
    const int cn[7][3] = {{1, 1, 0}, {1, 1, 2}, {1, 0, 1}, {1, 2, 1}, {0, 1, 1}, {2, 1, 1}, {1, 1, 1}};
    for (int k = 0; k < 7; ++k)
      if (!in(nb(c, cn[k][0], cn[k][1], cn[k][2]))) return fail("prolongation gradient neighbour missing");
    const int twod = m->blk_twod[b + 1];
    for (int q = 0; q < 8; ++q) {
      if (twod && (q & 1)) continue;
      if (!in(m->pro_fi[(size_t)i * 8 + q])) return fail("prolongation son missing");
    }
  }
  for (int i = 0; i < m->ndmg; ++i)
    if (!in(m->dmg_cell[i]) || !in(m->dmg_ref[i])) return fail("domain ghost out of range");
  for (int i = 0; i < m->nib; ++i) {
    if (!in(m->ib_cell[i])) return fail("IB cell out of range");
    if (m->ib_body[i] < 0 || m->ib_body[i] >= m->nbody) return fail("IB body out of range");
    for (int q = 0; q < 8; ++q) {
      if (D == 2 && (q & 1) == 0) continue;
      const int a = m->ib_intp[(size_t)i * 8 + q];
      if (!in(a)) return fail("IB interpolation cell missing");
      if (m->ibidx[a] >= 0) return fail("an image point interpolates from an IB cell");
    }
  }
  for (int i = 0; i < m->npatch; ++i) {
    if (m->pat_case[i] < 0 || m->pat_case[i] >= m->ncases) return fail("patch case");
    for (int q = 0; q < 8; ++q) {
      if (D == 2 && (q & 1) == 0) continue;
      if (!in(m->pat_intp[(size_t)i * 8 + q])) return fail("patch interpolation cell missing");
    }
  }
  return IBG_OK;
}
}  // namespace

// ================================================================== C ABI
extern "C" {

const char *ibg_version(void) { return "ibgpu 0.1 gfx950"; }

ibg_status ibg_check_mesh(const ibg_mesh_soa *m, const char **msg) {
  ibg_status st = validate_mesh(m);
  if (msg) *msg = st ? g_check_msg.c_str() : "";
  return st;
}

ibg_status ibg_create(const ibg_config *cfg, ibg_ctx **out) {
  if (!cfg || !out) return IBG_ERR_ARG;
  if (!(cfg->nstage == 1 || cfg->nstage == 3 || cfg->nstage == 4)) return IBG_ERR_ARG;
  if (cfg->mode != IBG_MODE_FP64 && cfg->mode != IBG_MODE_MIXED) return IBG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    fprintf(stderr, "ibgpu: no HIP device\n");
    return IBG_ERR_HIP;
  }
  ibg_ctx *c = new ibg_ctx();
  c->cfg = *cfg;
  c->mixed = (cfg->mode == IBG_MODE_MIXED);
  c->tile = cfg->tile > 0 ? cfg->tile : kDefaultTile;
  if (c->tile % 64 != 0 || c->tile > kDefaultTile) { delete c; return IBG_ERR_ARG; }
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) { delete c; return IBG_ERR_HIP; }
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  *out = c;
  return IBG_OK;
}

ibg_status ibg_destroy(ibg_ctx *c) {
  if (!c) return IBG_ERR_ARG;
  (void)hipStreamSynchronize(c->s);
  free_all(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return IBG_OK;
}

ibg_status ibg_upload_mesh(ibg_ctx *c, const ibg_mesh_soa *m, const ibg_physics *ph) {
  if (!c || !m || !ph) return IBG_ERR_ARG;
  if (m->dim != 2 && m->dim != 3) return IBG_ERR_ARG;
  if (m->ncases < 1 || m->ncases > IBG_MAX_CASES || m->nlevels < 1 || m->nlevels > IBG_MAX_LEVELS)
    return IBG_ERR_ARG;
  {
    ibg_status vs = validate_mesh(m);
    if (vs) { c->errmsg = g_check_msg; fprintf(stderr, "ibgpu: mesh rejected: %s\n", g_check_msg.c_str()); return vs; }
  }
  (void)hipStreamSynchronize(c->s);
  free_all(c);
  const int D = m->dim, NV = D + 2, L = m->nlevels, nc = m->ncases, nblk = nc * L;
  c->D = D; c->NV = NV; c->L = L; c->ncases = nc; c->nblk = nblk;
  const int N = m->ncells;
  c->N = N;
  // ---- block / level of every host cell
  std::vector<int> hblk(N);
  for (int b = 0; b < nblk; ++b)
    for (int g = m->blk_start[b]; g < m->blk_start[b + 1]; ++g) hblk[g] = b;
  // ---- device cell order: normal cells by (level desc, case, host order),
  // then ghosts
  std::vector<int> order;
  order.reserve(N);
  c->cell_end.assign(L, 0);
  for (int l = L - 1; l >= 0; --l) {
    for (int cs = 0; cs < nc; ++cs) {
      int b = cs * L + l;
      for (int g = m->blk_start[b]; g < m->blk_start[b] + m->blk_nnormal[b]; ++g) order.push_back(g);
    }
    c->cell_end[l] = (int)order.size();
  }
  const int Nn = (int)order.size();
  c->Nn = Nn;
  for (int b = 0; b < nblk; ++b)
    for (int g = m->blk_start[b] + m->blk_nnormal[b]; g < m->blk_start[b + 1]; ++g) order.push_back(g);
  if ((int)order.size() != N) return IBG_ERR_ARG;
  c->d2h = order;
  c->h2d.assign(N, -1);
  for (int i = 0; i < N; ++i) c->h2d[order[i]] = i;
  const std::vector<int> &h2d = c->h2d;
  // ---- faces: device order by (level desc, case)
  const int F = m->nfaces;
  c->F = F;
  std::vector<int> forder;
  forder.reserve(F);
  c->face_end.assign(L, 0);
  for (int l = L - 1; l >= 0; --l) {
    for (int cs = 0; cs < nc; ++cs) {
      int b = cs * L + l;
      for (int f = m->blk_face_start[b]; f < m->blk_face_start[b + 1]; ++f) forder.push_back(f);
    }
    c->face_end[l] = (int)forder.size();
  }
  c->fh2d.assign(F, -1);
  for (int i = 0; i < F; ++i) c->fh2d[forder[i]] = i;
  // ---- block info and physics
  std::vector<BlkInfo> hb(nblk);
  c->blk_twod.assign(L, 0);
  for (int b = 0; b < nblk; ++b) {
    BlkInfo &B = hb[b];
    const double *h = &m->blk_h[3 * b];
    for (int d = 0; d < 3; ++d) { B.h[d] = h[d]; B.invh[d] = 1.0 / h[d]; }
    B.area[0] = h[1] * h[2]; B.area[1] = h[2] * h[0]; B.area[2] = h[0] * h[1];
    B.vol = h[0] * h[1] * h[2];
    const int l = b % L;
    B.step_num = c->cfg.temporal_refine ? (double)(1 << (L - 1 - l)) : 1.0;
    B.case_id = b / L;
    B.level = l;
    B.finest = (l == L - 1) && (m->nbody > 0);
    if (b < L) c->blk_twod[l] = m->blk_twod[b];
  }
  c->hphys.resize(nc);
  for (int cs = 0; cs < nc; ++cs) {
    DevPhys &P = c->hphys[cs];
    const ibg_physics &p = ph[cs];
    P.gama = p.gama; P.Ma = p.Ma; P.Re = p.Re; P.Rg = p.Rg; P.Cv = p.Cv; P.Cp = p.Cp; P.mu = p.mu;
    P.qk = p.qk; P.SoT = p.S_over_T_ref;
    for (int k = 0; k < IBG_NFV; ++k) P.init[k] = p.initvar[k];
    P.inlet_T0 = p.inlet_T0; P.inlet_p0 = p.inlet_p0;
    P.inlet_cos = cos(p.inlet_angle); P.inlet_sin = sin(p.inlet_angle);
    P.outlet_p = p.outlet_p; P.cfl = p.cfl; P.dt_ratio = p.dt_ratio;
    P.force_ref = p.force_ref; P.p_ref_area = p.p_ref_area; P.shear_ref_area = p.shear_ref_area;
  }
  // ---- per-cell arrays in device order
  std::vector<int16_t> cellblk(N);
  std::vector<int> nb27((size_t)N * 27);
  std::vector<double> sd(N);
  for (int i = 0; i < N; ++i) {
    const int g = order[i];
    cellblk[i] = (int16_t)hblk[g];
    sd[i] = m->signdis[g];
    for (int q = 0; q < 27; ++q) {
      int a = m->nb[(size_t)g * 27 + q];
      nb27[(size_t)i * 27 + q] = a < 0 ? -1 : h2d[a];
    }
  }
  std::vector<uint8_t> flags(Nn, 0);
  for (int i = 0; i < Nn; ++i) {
    const int g = order[i];
    const BlkInfo &B = hb[hblk[g]];
    const bool ib = m->ibidx[g] >= 0;
    const bool solid = m->signdis[g] < 0.0;
    uint8_t f = 0;
    if (!B.finest) {
      f = F_UPDATE | F_RESID | ((m->signdis[g] > 0.0) ? F_PRIM : F_FREE);
    } else {
      f = F_FINEST;
      if (!solid && !ib) f |= F_UPDATE;
      if (m->signdis[g] > 0.0 && !ib) f |= F_PRIM | F_RESID;
      else if (!ib) f |= F_FREE;
    }
    flags[i] = f;
  }
  std::vector<int> cfaces((size_t)Nn * 2 * D);
  for (int i = 0; i < Nn; ++i) {
    const int g = order[i];
    for (int d = 0; d < D; ++d)
      for (int s2 = 0; s2 < 2; ++s2) {
        int f = m->cell_faces[(size_t)g * 6 + d * 2 + s2];
        if (f < 0) return IBG_ERR_ARG;
        cfaces[(size_t)i * 2 * D + 2 * d + s2] = c->fh2d[f];
      }
  }
  // ---- faces
  std::vector<int2> face_lohi(F);
  std::vector<int8_t> face_dir(F);
  std::vector<int16_t> face_blk(F);
  std::vector<int4> face_st(F);
  auto nbh = [&](int dcell, int d, int sgn) {
    int off[3] = {1, 1, 1};
    off[d] += sgn;
    return nb27[(size_t)dcell * 27 + off[0] * 9 + off[1] * 3 + off[2]];
  };
  for (int i = 0; i < F; ++i) {
    const int f = forder[i];
    const int lo = h2d[m->face_cells[2 * f]], hi = h2d[m->face_cells[2 * f + 1]];
    const int d = m->face_dir[f];
    face_lohi[i] = make_int2(lo, hi);
    face_dir[i] = (int8_t)d;
    face_blk[i] = (int16_t)hblk[m->face_cells[2 * f + 1]];
    // ComputeFaceFlux stencil with fallback NS_Solver.H:735-755
    int LL = nbh(lo, d, -1);
    if (LL < 0) LL = nbh(hi, d, +1);
    int RR = nbh(hi, d, +1);
    if (RR < 0) RR = nbh(lo, d, -1);
    if (LL < 0 || RR < 0) return IBG_ERR_ARG;
    face_st[i] = make_int4(LL, lo, hi, RR);
  }
  // ---- tiles over each (level, case) run of normal cells
  std::vector<TileInfo> tiles;
  std::vector<int4> tst;
  std::vector<int8_t> tdir;
  std::vector<uint16_t> slots((size_t)Nn * 2 * D, 0);
  c->tile_end.assign(L, 0);
  std::vector<int> slot_of(F, -1);
  int maxf = 0;
  {
    int pos = 0;
    for (int l = L - 1; l >= 0; --l) {
      for (int cs = 0; cs < nc; ++cs) {
        const int b = cs * L + l;
        const int n = m->blk_nnormal[b];
        for (int t0 = 0; t0 < n; t0 += c->tile) {
          TileInfo ti{};
          ti.cell0 = pos + t0;
          ti.ncell = std::min(c->tile, n - t0);
          ti.face0 = (int)tst.size();
          ti.blk = b;
          std::vector<int> used;
          for (int k = 0; k < ti.ncell; ++k) {
            const int dc = ti.cell0 + k;
            for (int q = 0; q < 2 * D; ++q) {
              const int f = cfaces[(size_t)dc * 2 * D + q];
              if (slot_of[f] < 0) {
                slot_of[f] = (int)used.size();
                used.push_back(f);
                tst.push_back(face_st[f]);
                tdir.push_back(face_dir[f]);
              }
              slots[(size_t)dc * 2 * D + q] = (uint16_t)slot_of[f];
            }
          }
          for (int f : used) slot_of[f] = -1;
          ti.nface = (int)used.size();
          if (ti.nface > 65535) return IBG_ERR_ARG;
          maxf = std::max(maxf, ti.nface);
          tiles.push_back(ti);
        }
        pos += n;
      }
      c->tile_end[l] = (int)tiles.size();
    }
  }
  c->ntiles = (int)tiles.size();
  c->max_tile_faces = std::max(maxf, 1);
  // ---- domain ghosts: edge entries first, then corners (their refcell is
  // an edge ghost); the mesh lists them in that order
  std::vector<int4> dmg;
  {
    std::vector<int4> edge, corner;
    for (int i = 0; i < m->ndmg; ++i) {
      const int g = m->dmg_cell[i], r = m->dmg_ref[i];
      int4 e = make_int4(h2d[g], h2d[r], m->dmg_type[i], hblk[g] / L);
      if (m->kind[r] == IBG_DMGHOST) corner.push_back(e);
      else edge.push_back(e);
    }
    dmg = edge;
    dmg.insert(dmg.end(), corner.begin(), corner.end());
    c->ndmg_edge = (int)edge.size();
    c->ndmg_corner = (int)corner.size();
  }
  // ---- pairs grouped by coarse level
  std::vector<int> res_ci, res_fi, pro_ci, pro_fi;
  c->res_off.assign(L, 0);
  c->pro_off.assign(L, 0);
  for (int l = 0; l < L; ++l) {
    c->res_off[l] = (int)res_ci.size();
    for (int i = 0; i < m->nres; ++i) {
      if (m->res_blk[i] % L != l) continue;
      res_ci.push_back(h2d[m->res_ci[i]]);
      for (int q = 0; q < 8; ++q) {
        int a = m->res_fi[(size_t)i * 8 + q];
        res_fi.push_back(a < 0 ? -1 : h2d[a]);
      }
    }
    c->pro_off[l] = (int)pro_ci.size();
    for (int i = 0; i < m->npro; ++i) {
      if (m->pro_blk[i] % L != l) continue;
      pro_ci.push_back(h2d[m->pro_ci[i]]);
      for (int q = 0; q < 8; ++q) {
        int a = m->pro_fi[(size_t)i * 8 + q];
        pro_fi.push_back(a < 0 ? -1 : h2d[a]);
      }
    }
  }
  c->res_off.push_back((int)res_ci.size());
  c->pro_off.push_back((int)pro_ci.size());
  // res_off[l] .. res_off[l+1] must be the pairs of coarse level l
  {
    std::vector<int> ro(L + 1), po(L + 1);
    for (int l = 0; l <= L; ++l) { ro[l] = c->res_off[l]; po[l] = c->pro_off[l]; }
    c->res_off = ro;
    c->pro_off = po;
  }
  // ---- IB
  const int nib = m->nib;
  c->nib = nib;
  std::vector<int> ib_cell(nib), ib_intp((size_t)nib * 8), ib_body(nib), ib_case(nib);
  std::vector<double> ib_coef((size_t)nib * 8), ib_hgdis(nib), ib_att((size_t)nib * 3);
  for (int i = 0; i < nib; ++i) {
    const int g = m->ib_cell[i];
    ib_cell[i] = h2d[g];
    ib_body[i] = m->ib_body[i];
    ib_case[i] = hblk[g] / L;
    ib_hgdis[i] = m->ib_hgdis[i];
    for (int k = 0; k < 3; ++k) ib_att[(size_t)i * 3 + k] = m->ib_attach[(size_t)i * 3 + k];
    for (int q = 0; q < 8; ++q) {
      int a = m->ib_intp[(size_t)i * 8 + q];
      double w = m->ib_coef[(size_t)i * 8 + q];
      // the zero-weight z layer of a 2D image point is skipped
      if (D == 2 && (q & 1) == 0) { a = -1; w = 0.0; }
      if (a >= 0 && m->ibidx[a] >= 0) {
        c->errmsg = "an image point interpolates from an IB cell";
        return IBG_ERR_ARG;
      }
      ib_intp[(size_t)i * 8 + q] = a < 0 ? -1 : h2d[a];
      ib_coef[(size_t)i * 8 + q] = w;
    }
  }
  // ---- patches
  const int np = m->npatch;
  c->npatch = np;
  std::vector<int> pat_intp((size_t)np * 8), pat_case(np);
  std::vector<double> pat_coef((size_t)np * 8), pat_hgdis(np), pat_nv((size_t)np * 3),
      pat_pc((size_t)np * 3), pat_area(np);
  for (int i = 0; i < np; ++i) {
    pat_case[i] = m->pat_case[i];
    pat_hgdis[i] = m->pat_hgdis[i];
    pat_area[i] = m->pat_area[i];
    for (int k = 0; k < 3; ++k) {
      pat_nv[(size_t)i * 3 + k] = m->pat_nv[(size_t)i * 3 + k];
      pat_pc[(size_t)i * 3 + k] = m->pat_pc[(size_t)i * 3 + k];
    }
    for (int q = 0; q < 8; ++q) {
      int a = m->pat_intp[(size_t)i * 8 + q];
      double w = m->pat_coef[(size_t)i * 8 + q];
      if (D == 2 && (q & 1) == 0) { a = -1; w = 0.0; }
      pat_intp[(size_t)i * 8 + q] = a < 0 ? -1 : h2d[a];
      pat_coef[(size_t)i * 8 + q] = w;
    }
  }
  c->nbody = m->nbody;
  std::vector<double> case_center((size_t)nc * 3, 0.0);
  for (int bi = m->nbody - 1; bi >= 0; --bi)
    for (int k = 0; k < 3; ++k) case_center[(size_t)m->body_case[bi] * 3 + k] = m->body_center[(size_t)bi * 3 + k];
  c->body_center.assign(m->body_center, m->body_center + (size_t)m->nbody * 3);
  c->hbody_case.assign(m->body_case, m->body_case + m->nbody);
  // ---- device allocation and upload
  ibg_status st;
#define UP(dst, vec) if ((st = dupload(c, &dst, vec))) return st
#define AL(dst, n) if ((st = dalloc(c, &dst, n))) return st
  AL(c->W64.r, N); AL(c->W64.u, N); AL(c->W64.v, N); AL(c->W64.p, N);
  if (D == 3) { AL(c->W64.w, N); } else c->W64.w = nullptr;
  if (c->mixed) {
    AL(c->W32.r, N); AL(c->W32.u, N); AL(c->W32.v, N); AL(c->W32.p, N);
    if (D == 3) { AL(c->W32.w, N); } else c->W32.w = nullptr;
  }
  AL(c->U0, (size_t)NV * Nn);
  {
    const size_t rs = c->mixed ? sizeof(float) : sizeof(double);
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(rs * NV * (size_t)Nn, 16))); c->allocs.push_back(p); c->Rv = p;
    HIPCHK(hipMemset(p, 0, std::max<size_t>(rs * NV * (size_t)Nn, 16)));
    HIPCHK(hipMalloc(&p, std::max<size_t>(rs * NV * (size_t)F, 16))); c->allocs.push_back(p); c->vs = p;
    HIPCHK(hipMemset(p, 0, std::max<size_t>(rs * NV * (size_t)F, 16)));
  }
  UP(c->flags, flags);
  UP(c->cellblk, cellblk);
  c->hcellblk = cellblk;
  UP(c->nb27, nb27);
  UP(c->cfaces, cfaces);
  UP(c->face_lohi, face_lohi);
  UP(c->face_dir, face_dir);
  UP(c->face_blk, face_blk);
  UP(c->face_st, face_st);
  UP(c->tiles, tiles);
  UP(c->tst, tst);
  UP(c->tdir, tdir);
  UP(c->slots, slots);
  UP(c->blk, hb);
  UP(c->phys, c->hphys);
  UP(c->dmg, dmg);
  UP(c->res_ci, res_ci); UP(c->res_fi, res_fi); UP(c->pro_ci, pro_ci); UP(c->pro_fi, pro_fi);
  UP(c->ib_cell, ib_cell); UP(c->ib_intp, ib_intp); UP(c->ib_body, ib_body); UP(c->ib_case, ib_case);
  UP(c->ib_coef, ib_coef); UP(c->ib_hgdis, ib_hgdis); UP(c->ib_attach, ib_att);
  UP(c->signdis, sd);
  UP(c->pat_intp, pat_intp); UP(c->pat_case, pat_case); UP(c->pat_coef, pat_coef);
  UP(c->pat_hgdis, pat_hgdis); UP(c->pat_nv, pat_nv); UP(c->pat_pc, pat_pc); UP(c->pat_area, pat_area);
  UP(c->case_center, case_center);
  AL(c->motion, (size_t)std::max(m->nbody, nc) + 1);
  std::vector<double> dt0(nc);
  for (int cs = 0; cs < nc; ++cs) dt0[cs] = ph[cs].dt_init;
  UP(c->dt, dt0);
  AL(c->dtred, (size_t)nc);
  AL(c->err, 1);
  AL(c->rpart, (size_t)std::max(c->ntiles, 1) * 10);
  AL(c->resid, (size_t)nc * 10);
  AL(c->probe, (size_t)std::max(N, F) * 5);
  HIPCHK(hipMemset(c->rpart, 0, sizeof(double) * std::max(c->ntiles, 1) * 10));
#undef UP
#undef AL
  // cell-stage update counts per step for each lmin (UPDATE cells marched)
  c->upd_per_lmin.assign(L, 0);
  {
    std::vector<int64_t> upd_lvl(L, 0);
    for (int i = 0; i < Nn; ++i)
      if (flags[i] & F_UPDATE) upd_lvl[hb[cellblk[i]].level]++;
    for (int lm = 0; lm < L; ++lm) {
      int64_t s = 0;
      for (int l = lm; l < L; ++l) s += upd_lvl[l];
      c->upd_per_lmin[lm] = s;
    }
  }
  HIPCHK(hipDeviceSynchronize());
  if ((st = reset_err(c))) return st;
  HIPCHK(hipStreamSynchronize(c->s));
  c->have_mesh = true;
  return IBG_OK;
}

ibg_status ibg_set_state(ibg_ctx *c, const double *fv, int64_t n) {
  if (!c || !fv || !c->have_mesh || n != c->N) return IBG_ERR_ARG;
  const int N = c->N;
  std::vector<double> r(N), u(N), v(N), w(N), p(N);
  for (int i = 0; i < N; ++i) {
    const double *a = fv + (size_t)c->d2h[i] * IBG_NFV;
    r[i] = a[0]; u[i] = a[1]; v[i] = a[2]; w[i] = a[3]; p[i] = a[4];
  }
  HIPCHK(hipMemcpy(c->W64.r, r.data(), N * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->W64.u, u.data(), N * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->W64.v, v.data(), N * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->W64.p, p.data(), N * 8, hipMemcpyHostToDevice));
  if (c->D == 3) HIPCHK(hipMemcpy(c->W64.w, w.data(), N * 8, hipMemcpyHostToDevice));
  if (c->mixed) {
    std::vector<float> f(N);
    auto put = [&](float *dst, const std::vector<double> &src) -> ibg_status {
      for (int i = 0; i < N; ++i) f[i] = (float)src[i];
      HIPCHK(hipMemcpy(dst, f.data(), N * 4, hipMemcpyHostToDevice));
      return IBG_OK;
    };
    ibg_status st;
    if ((st = put(c->W32.r, r))) return st;
    if ((st = put(c->W32.u, u))) return st;
    if ((st = put(c->W32.v, v))) return st;
    if ((st = put(c->W32.p, p))) return st;
    if (c->D == 3 && (st = put(c->W32.w, w))) return st;
  }
  return IBG_OK;
}

ibg_status ibg_get_state(ibg_ctx *c, double *fv, int64_t n) {
  if (!c || !fv || !c->have_mesh || n != c->N) return IBG_ERR_ARG;
  const int N = c->N;
  HIPCHK(hipStreamSynchronize(c->s));
  std::vector<double> r(N), u(N), v(N), w(N, 0.0), p(N);
  HIPCHK(hipMemcpy(r.data(), c->W64.r, N * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(u.data(), c->W64.u, N * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(v.data(), c->W64.v, N * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(p.data(), c->W64.p, N * 8, hipMemcpyDeviceToHost));
  if (c->D == 3) HIPCHK(hipMemcpy(w.data(), c->W64.w, N * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < N; ++i) {
    const int g = c->d2h[i];
    double *a = fv + (size_t)g * IBG_NFV;
    const DevPhys &P = c->hphys[c->hcellblk[i] / c->L];
    a[0] = r[i]; a[1] = u[i]; a[2] = v[i]; a[3] = w[i]; a[4] = p[i];
    // derived: Get_T_ideal_gas and Get_E (FlowVariables.H:23,33)
    a[5] = p[i] / (r[i] * P.Rg);
    a[6] = P.Cv * a[5] + 0.5 * (u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
  }
  return IBG_OK;
}

ibg_status ibg_bc_fill(ibg_ctx *c, int32_t ts) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  return run_bc(c, ts);
}

ibg_status ibg_interface_exchange(ibg_ctx *c, int32_t ts) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  return run_exchange(c, ts);
}

ibg_status ibg_ib_fill(ibg_ctx *c, const ibg_body_motion *mo) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (mo) {
    std::vector<Motion> hm(std::max(c->nbody, 1));
    for (int b = 0; b < c->nbody; ++b) {
      for (int k = 0; k < 3; ++k) {
        hm[b].tra[k] = mo[b].tra_vel[k];
        hm[b].rot[k] = mo[b].rot_vel[k];
        hm[b].center[k] = c->body_center[(size_t)b * 3 + k];
      }
    }
    HIPCHK(hipMemcpyAsync(c->motion, hm.data(), sizeof(Motion) * hm.size(), hipMemcpyHostToDevice, c->s));
    c->have_motion = 1;
  } else {
    c->have_motion = 0;
  }
  return run_ib(c);
}

ibg_status ibg_compute_dt(ibg_ctx *c, int32_t ts, double *dt_out) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  (void)ts;
  ibg_status st = run_dt(c);
  if (st) return st;
  if (dt_out) {
    HIPCHK(hipMemcpyAsync(dt_out, c->dt, sizeof(double) * c->ncases, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
  }
  return IBG_OK;
}

ibg_status ibg_old_cons(ibg_ctx *c, int32_t ts) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  // U_old is produced together with the frozen viscous residual
  return run_viscous_and_prep(c, ts);
}

ibg_status ibg_viscous_flux(ibg_ctx *c, int32_t ts) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  return run_viscous_and_prep(c, ts);
}

ibg_status ibg_stage(ibg_ctx *c, int32_t ts, int32_t subt) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (subt < 0 || subt >= c->cfg.nstage) return IBG_ERR_ARG;
  const bool last = (subt == c->cfg.nstage - 1);
  return run_stage(c, ts, subt, last ? c->probe : nullptr);
}

ibg_status ibg_time_march(ibg_ctx *c, int32_t ts0, int32_t nsteps, int32_t use_graph) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (nsteps < 0) return IBG_ERR_ARG;
  ibg_status st;
  if ((st = reset_err(c))) return st;
  c->have_motion = 0;
  const int P = c->cfg.temporal_refine ? (1 << (c->L - 1)) : 1;
  int64_t upd = 0;
  for (int k = 0; k < nsteps; ++k) upd += c->upd_per_lmin[lmin_of(c, ts0 + k)] * c->cfg.nstage;
  c->last_updates = upd;
  HIPCHK(hipEventRecord(c->ev0, c->s));
  int ts = ts0;
  const int end = ts0 + nsteps;
  if (use_graph && P > 0) {
    // run eagerly up to a period boundary, then replay one graph per period
    while (ts < end && ts % P != 0) { if ((st = run_step(c, ts))) return st; ++ts; }
    while (end - ts >= P) {
      const int with_dt = (c->cfg.dt_var_step > 0 && ts % c->cfg.dt_var_step == 0) ? 1 : 0;
      bool ok = true;
      for (int k = 1; k < P; ++k)
        if (c->cfg.dt_var_step > 0 && (ts + k) % c->cfg.dt_var_step == 0) ok = false;
      if (!ok) break;
      if (!c->gexec[with_dt]) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < P; ++k) {
          st = run_step(c, ts + k);
          if (st) { hipGraph_t tmp; (void)hipStreamEndCapture(c->s, &tmp); return st; }
        }
        HIPCHK(hipStreamEndCapture(c->s, &g));
        HIPCHK(hipGraphInstantiate(&c->gexec[with_dt], g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
      }
      HIPCHK(hipGraphLaunch(c->gexec[with_dt], c->s));
      c->last_ntiles = c->tile_end[lmin_of(c, ts + P - 1)];
      ts += P;
    }
  }
  while (ts < end) { if ((st = run_step(c, ts))) return st; ++ts; }
  HIPCHK(hipEventRecord(c->ev1, c->s));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_total_ms = ms;
  int64_t cell;
  int32_t code;
  if ((st = read_err(c, &cell, &code))) return st;
  if (code) return (ibg_status)code;
  return IBG_OK;
}

ibg_status ibg_residual(ibg_ctx *c, double *linf, double *l2) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(64), 0, c->s, c->last_ntiles, c->tiles, c->blk,
                     c->rpart, c->resid, c->ncases);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)c->ncases * 10);
  HIPCHK(hipMemcpyAsync(h.data(), c->resid, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int cs = 0; cs < c->ncases; ++cs)
    for (int k = 0; k < 5; ++k) {
      if (linf) linf[cs * 5 + k] = h[cs * 10 + k];
      if (l2) l2[cs * 5 + k] = h[cs * 10 + 5 + k];
    }
  return IBG_OK;
}

ibg_status ibg_forces(ibg_ctx *c, const ibg_body_motion *mo, double *out) {
  if (!c || !c->have_mesh || !out) return IBG_ERR_ARG;
  int hm = 0;
  if (mo) {
    // the bodies of one airfoil move together: the motion of the first body
    // of each case, about its center (Surface_Point_Velocity Body.H:709-721)
    std::vector<Motion> h(c->ncases);
    std::vector<int> first(c->ncases, -1);
    for (int b = 0; b < c->nbody; ++b) {
      const int cs = c->hbody_case[b];
      if (first[cs] < 0) first[cs] = b;
    }
    for (int cs = 0; cs < c->ncases; ++cs)
      for (int k = 0; k < 3; ++k) {
        const int b = first[cs];
        h[cs].tra[k] = b >= 0 ? mo[b].tra_vel[k] : 0.0;
        h[cs].rot[k] = b >= 0 ? mo[b].rot_vel[k] : 0.0;
        h[cs].center[k] = b >= 0 ? c->body_center[(size_t)b * 3 + k] : 0.0;
      }
    HIPCHK(hipMemcpyAsync(c->motion, h.data(), sizeof(Motion) * h.size(), hipMemcpyHostToDevice, c->s));
    hm = 1;
  }
  double *d = c->resid;   // reuse: ncases*10 >= ncases*6
  hipLaunchKernelGGL(k_forces, dim3(c->ncases), dim3(256), 0, c->s, c->npatch, c->pat_case,
                     c->pat_intp, c->pat_coef, c->pat_hgdis, c->pat_nv, c->pat_pc, c->pat_area,
                     c->case_center, c->motion, hm, c->W64, c->D, c->phys, d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, d, sizeof(double) * 6 * c->ncases, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_get_faceflux(ibg_ctx *c, double *out, int64_t nf) {
  if (!c || !c->have_mesh || !out || nf != c->F) return IBG_ERR_ARG;
  // fluxes of all faces (every level) from the current state
  const int F = c->F;
  if (F > 0) {
    if (c->D == 2) {
      if (c->mixed) {
        hipLaunchKernelGGL((k_face_probe<float, 2>), dim3(grid(F, 256)), dim3(256), 0, c->s, F, c->face_st, c->face_dir, c->face_blk, c->W32, c->blk, c->phys, c->cfg.limiter, c->probe);
      } else {
        WPtr<double> W{c->W64.r, c->W64.u, c->W64.v, c->W64.w, c->W64.p};
        hipLaunchKernelGGL((k_face_probe<double, 2>), dim3(grid(F, 256)), dim3(256), 0, c->s, F, c->face_st, c->face_dir, c->face_blk, W, c->blk, c->phys, c->cfg.limiter, c->probe);
      }
    } else {
      if (c->mixed) {
        hipLaunchKernelGGL((k_face_probe<float, 3>), dim3(grid(F, 256)), dim3(256), 0, c->s, F, c->face_st, c->face_dir, c->face_blk, c->W32, c->blk, c->phys, c->cfg.limiter, c->probe);
      } else {
        WPtr<double> W{c->W64.r, c->W64.u, c->W64.v, c->W64.w, c->W64.p};
        hipLaunchKernelGGL((k_face_probe<double, 3>), dim3(grid(F, 256)), dim3(256), 0, c->s, F, c->face_st, c->face_dir, c->face_blk, W, c->blk, c->phys, c->cfg.limiter, c->probe);
      }
    }
    HIPCHK(hipGetLastError());
  }
  std::vector<double> h((size_t)F * 5);
  HIPCHK(hipMemcpyAsync(h.data(), c->probe, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int hf = 0; hf < F; ++hf)
    for (int k = 0; k < 5; ++k) out[(size_t)hf * 5 + k] = h[(size_t)c->fh2d[hf] * 5 + k];
  return IBG_OK;
}

ibg_status ibg_get_vsflux(ibg_ctx *c, double *out, int64_t nf) {
  if (!c || !c->have_mesh || !out || nf != c->F) return IBG_ERR_ARG;
  const int F = c->F, NV = c->NV;
  HIPCHK(hipStreamSynchronize(c->s));
  std::vector<double> h((size_t)F * NV);
  if (c->mixed) {
    std::vector<float> t((size_t)F * NV);
    HIPCHK(hipMemcpy(t.data(), c->vs, t.size() * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < t.size(); ++i) h[i] = t[i];
  } else {
    HIPCHK(hipMemcpy(h.data(), c->vs, h.size() * 8, hipMemcpyDeviceToHost));
  }
  for (int hf = 0; hf < F; ++hf) {
    const int df = c->fh2d[hf];
    double *o = out + (size_t)hf * 5;
    if (c->D == 3) { for (int k = 0; k < 5; ++k) o[k] = h[(size_t)k * F + df]; }
    else { o[0] = h[df]; o[1] = h[(size_t)F + df]; o[2] = h[(size_t)2 * F + df]; o[3] = 0.0; o[4] = h[(size_t)3 * F + df]; }
  }
  return IBG_OK;
}

ibg_status ibg_get_source(ibg_ctx *c, double *out, int64_t n) {
  if (!c || !c->have_mesh || !out || n != c->N) return IBG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  std::vector<double> h((size_t)c->N * 5);
  HIPCHK(hipMemcpy(h.data(), c->probe, h.size() * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < c->N; ++i)
    for (int k = 0; k < 5; ++k) out[(size_t)c->d2h[i] * 5 + k] = (i < c->Nn) ? h[(size_t)i * 5 + k] : 0.0;
  return IBG_OK;
}

ibg_status ibg_get_dt(ibg_ctx *c, double *dt) {
  if (!c || !c->have_mesh || !dt) return IBG_ERR_ARG;
  HIPCHK(hipMemcpyAsync(dt, c->dt, sizeof(double) * c->ncases, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_set_dt(ibg_ctx *c, const double *dt) {
  if (!c || !c->have_mesh || !dt) return IBG_ERR_ARG;
  HIPCHK(hipMemcpyAsync(c->dt, dt, sizeof(double) * c->ncases, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_last_error(ibg_ctx *c, int64_t *cell, int32_t *code) {
  if (!c || !cell || !code) return IBG_ERR_ARG;
  if (!c->have_mesh) { *cell = -1; *code = 0; return IBG_OK; }
  return read_err(c, cell, code);
}

ibg_status ibg_synchronize(ibg_ctx *c) {
  if (!c) return IBG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_last_timing(ibg_ctx *c, double *total_ms, double *stage_ms, int64_t *launches) {
  if (!c) return IBG_ERR_ARG;
  if (total_ms) *total_ms = c->last_total_ms;
  if (stage_ms) *stage_ms = c->last_stage_ms;
  if (launches) *launches = c->last_stage_launches;
  return IBG_OK;
}

ibg_status ibg_cell_stage_updates(ibg_ctx *c, int64_t *count) {
  if (!c || !count) return IBG_ERR_ARG;
  *count = c->last_updates;
  return IBG_OK;
}

}  // extern "C"
