// ibg_common.h — device-side types of the residual engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ibgpu.h"

namespace ibg {

struct __align__(16) F4 { float x, y, z, w; };
struct __align__(32) D4 { double x, y, z, w; };

template <typename R> struct V4;
template <> struct V4<float> { using T = F4; };
template <> struct V4<double> { using T = D4; };

// Packed per-cell record of NV = D+2 values: 2D -> one 4-vector;
// 3D -> a 4-vector (first four values) plus a scalar plane (fifth value).
// One stencil cell is one 16-byte (FP32) or two 16-byte (FP64) loads.
template <typename R, int D>
struct Pack {
  typename V4<R>::T *a;
  R *e;   // 3D only
};

enum CellFlag : uint8_t {
  F_UPDATE = 1,   // UpdateConsVars touches the cell
  F_PRIM = 2,     // ComptNewPrimVar recovers W from U
  F_FREE = 4,     // ComptNewPrimVar resets W to initvar
  F_RESID = 8,    // counted by ResinDomain
  F_FINEST = 16,  // finest level with bodies (abort code 611 vs 523)
  F_SA = 32       // signdis >= 0: ComptMutSource evaluates the cell
};

// Tile of this workgroup for the tiled stage and viscous kernels: the
// dispatch order (workgroup b takes tile b), so the whole device works on a
// narrow window of consecutive (Morton-ordered) tiles at any time and the
// neighbour tiles a stencil reaches were read moments before (in this XCD's
// L2 or in the shared last-level cache).  Giving each of the 8 XCDs one
// contiguous run of the tiles instead (workgroups are dealt to the XCDs
// round-robin) spreads the device over 8 windows and measured slower: the
// learned-closure stage 94.2 -> 90.9 us per launch on the 2.6M-cell moving
// section, the mixed stage 1529 -> 1456 us on the 61.8M-cell ensemble, and
// the plain mixed stage even.
__device__ __forceinline__ int dispatch_tile() { return (int)blockIdx.x; }

struct DevPhys {
  double gama, Ma, Re, Rg, Cv, Cp, mu, qk, SoT;
  double init[IBG_NFV];
  double inlet_T0, inlet_p0, inlet_cos, inlet_sin, outlet_p;
  double cfl, dt_ratio;
  double force_ref, p_ref_area, shear_ref_area;
  // SA closure: freestream and inlet working variable, turbulent heat
  // conduction denominator Re*Pr_tur*(gama-1)*Ma^2 (NS_Solver.cpp:540),
  // Pr^(1/3) of the wall-temperature relation (FlowConstants.cpp:44)
  double nut_inf, nut_inlet, qkt_den, pr_cuberoot;
};

// Every copy of the cell state that a once-per-step ghost fill must keep
// consistent: the FP64 master and the stencil buffers the stages rotate
// through.  Plain mode: stencil buffer 0 is the master; d1, d2 are extra
// FP64 buffers.  Mixed mode: f0 (home), f1, f2 are FP32 buffers.  The SA
// working variable has the same set of copies (nm, nd*, nf*; nm == nullptr
// when the closure is off).
template <int D>
struct StateOut {
  Pack<double, D> m, d1, d2;
  Pack<float, D> f0, f1, f2;
  double *nm, *nd1, *nd2;
  float *nf0, *nf1, *nf2;
  int nd, nf;
  const unsigned long long *err;   // error word: a fill does nothing once it is set
};

// A march stops at its first error: every kernel that writes the state,
// the time or dt returns at once when an earlier launch has set the error
// word (~0 when clear), so the state is the one of the failing launch (the
// reference aborts there, NS_Solver.cpp:888-894, NS_Solver.H:355-375).
__device__ __forceinline__ bool halted(const unsigned long long *err) {
  return err && *err != ~0ull;
}
// The same decision for a whole workgroup (one read, shared through LDS),
// for kernels whose waves cooperate through LDS: a launch that sets the
// word while some of the group's waves have read it and others not would
// otherwise leave a cooperative phase half done.
__device__ __forceinline__ bool wg_halted(const unsigned long long *err) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = halted(err) ? 1 : 0;
  __syncthreads();
  return flag != 0;
}

struct BlkInfo {
  double h[3], invh[3], area[3], vol, step_num;
  int case_id, level, finest, pad;
};

struct TileInfo {
  int cell0, ncell, face0, nface, blk, pad0, pad1, pad2;
};

struct Motion {
  double tra[3], rot[3], center[3];
};

// A domain-wall ghost (BC -81): the reference cell's inward neighbour, the
// wall normal and the wall distances of the reference cell and neighbour.
struct DmgWall {
  double nmv[3], dref, dnbr;
  int nbr, pad;
};

}  // namespace ibg
