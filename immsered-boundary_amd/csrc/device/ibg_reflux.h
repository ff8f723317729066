// Optional coarse/fine flux balance (refluxing), default off.
//
// BalanceFaceFlux Fluxmodification.cpp:291-358 over the faces that
// AMR::FindModifyFace (:5-289) pairs: a coarse face between a normal coarse
// cell and a coarse ghost whose sons are normal cells, and the faces of the
// coarse cell's sons on that side.  The reference compiles the balance only
// without TEMPORAL_REFINE and leaves its two call sites commented
// (NS_Solver.H:672 on the viscous fluxes after ComptViscousStress, :686 on
// the inviscid fluxes of the last stage), so the default keeps it off.
//
// Each modified face appears in the face list of exactly one tile (its
// other cell is a ghost).  A pre-pass computes, per pair, the coarse and the
// fine fluxes with the same routines and operands as the tile kernels,
// balances them and writes the results; the tile kernels then take the
// balanced flux for those faces (tmod maps a tile face to its entry).
#pragma once

#include "ibg_kernels.h"

namespace ibg {

struct RfPair {
  int mode;       // 0: fine level refined in x, y, z (four fine faces);
                  // 1: fine level refined in x, y only, x or y face (two);
                  // 2: the same, z face (four)
  int cslot;      // tile face of the coarse face
  int cblk, fblk; // block of the coarse face's tile / of the fine faces' tiles
  int fslot[4];   // tile faces of the fine faces (-1 past the count)
};

__host__ __device__ inline int rf_nfine(int mode) { return mode == 1 ? 2 : 4; }

// F[0] the coarse face, F[1..nf] the fine faces, in place:
//   c_old = F_c (x 0.25 when the fine level is refined in z as well)
//   F_c *= 0.5;  per fine face: F_f *= 0.5;  F_c += F_f;  F_f += c_old * w
// with w = 0.5 (three-direction refinement), 0.25 (two fine faces) or 0.125
// (four fine faces of a z face), in the reference's order.
template <typename R, int NV>
__device__ __forceinline__ void balance_faces(int mode, R (&F)[5][NV]) {
#pragma clang fp contract(off)
  const int nf = rf_nfine(mode);
  const R w = mode == 0 ? R(0.5) : (mode == 1 ? R(0.25) : R(0.125));
  R cold[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    cold[v] = mode == 0 ? F[0][v] * R(0.25) : F[0][v];
    F[0][v] *= R(0.5);
  }
#pragma unroll
  for (int k = 1; k <= 4; ++k) {
    if (k > nf) break;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      F[k][v] *= R(0.5);
      F[0][v] += F[k][v];
      F[k][v] += cold[v] * w;
    }
  }
}

// Inviscid fluxes of the pairs' faces from the stage's input state, balanced
// (the last stage only).  rbal [n][5][NV].
template <typename R, int D>
__global__ void __launch_bounds__(256)
k_balance_inv(const RfPair *pairs, int n, const int4 *tst, const int8_t *tdir, Pack<R, D> Win,
              const BlkInfo *blk, const DevPhys *phys, int limiter, R *rbal,
              const unsigned long long *err) {
  constexpr int NV = D + 2;
  constexpr bool kMixed = std::is_same<R, float>::value;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || halted(err)) return;
  const RfPair p = pairs[j];
  const int nf = rf_nfine(p.mode);
  R F[5][NV];
#pragma unroll
  for (int k = 0; k <= 4; ++k) {
    if (k > nf) break;
    const int s = k == 0 ? p.cslot : p.fslot[k - 1];
    const BlkInfo &bi = blk[k == 0 ? p.cblk : p.fblk];
    const DevPhys &P = phys[bi.case_id];
    PhysR<R> ph;
    ph.gama = (R)P.gama; ph.Rg = (R)P.Rg; ph.Cv = (R)P.Cv; ph.Ma = (R)P.Ma;
    const int4 st = tst[s];
    const int dir = tdir[s];
    const Prim<R, D> q0 = pload<R, D>(Win, st.x);
    const Prim<R, D> q1 = pload<R, D>(Win, st.y);
    const Prim<R, D> q2 = pload<R, D>(Win, st.z);
    const Prim<R, D> q3 = pload<R, D>(Win, st.w);
    const R area = (R)bi.area[dir];
    if constexpr (kMixed)
      roe_flux_fast<D>(q0, q1, q2, q3, dir, area, ph, limiter, F[k]);
    else
      roe_flux<R, D>(q0, q1, q2, q3, dir, area, ph, limiter, F[k]);
  }
  balance_faces<R, NV>(p.mode, F);
#pragma unroll
  for (int k = 0; k <= 4; ++k) {
    if (k > nf) break;
#pragma unroll
    for (int v = 0; v < NV; ++v) rbal[((size_t)j * 5 + k) * NV + v] = F[k][v];
  }
}

// Viscous fluxes of the pairs' faces from the step-start state, balanced
// (once per step, before the viscous residual).  rbalv [n][5][NV].
template <typename R, int D>
__global__ void __launch_bounds__(256)
k_balance_visc(ViscArgs<R, D> a, const RfPair *pairs, int n, R *rbalv) {
  constexpr int NV = D + 2;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || halted(a.err)) return;
  const RfPair p = pairs[j];
  const int nf = rf_nfine(p.mode);
  R F[5][NV];
#pragma unroll
  for (int k = 0; k <= 4; ++k) {
    if (k > nf) break;
    const BlkInfo &bi = a.blk[k == 0 ? p.cblk : p.fblk];
    const ViscPhys<R, D> vp = visc_phys<R, D>(bi, a.phys[bi.case_id]);
    R mf;
    visc_face<R, D>(a, vp, k == 0 ? p.cslot : p.fslot[k - 1], F[k], mf);
  }
  balance_faces<R, NV>(p.mode, F);
#pragma unroll
  for (int k = 0; k <= 4; ++k) {
    if (k > nf) break;
#pragma unroll
    for (int v = 0; v < NV; ++v) rbalv[((size_t)j * 5 + k) * NV + v] = F[k][v];
  }
}

}  // namespace ibg
