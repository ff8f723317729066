// ibg_move.h — per-step immersed-boundary re-tagging of moving bodies on a
// fixed level set (the part of Body::MovingBody_Image, Bodymove.cpp:37-125,
// that runs between regrids), 2D sections.  Five launches per move:
//
//   k_move_front   pose of every case (given, or from the pitch/heave law at
//                  the case's time, with the wall motion of the IB fill and
//                  the wall force), new surface points and patch geometry
//                  (ComptNewLocation Bodymove.cpp:3-23, ComptPatchParams
//                  :25-35, the segment rules of the mesh builder)
//   k_move_dist    closest patch and signed distance of every cell of the
//                  marched levels: the local walk from the previous patch
//                  (RenewLevelDistance :136-155, FindClosePatchtoCell
//                  Body.H:505-542); finest-level solid and in->out cells go
//                  to lists, cells with a neighbour of another body to
//   k_move_mixed   the same with the cross-body cycle (:156-223), eight
//                  lanes per cell
//   k_move_mark    the IB set: the non-solid neighbours of solid cells
//                  (layer 1) and their non-solid neighbours (RemoveIBCells /
//                  InfectNewIBCells :345-767 keep, after every move, the set
//                  of the initial rule of Body.cpp:1016-1082); the
//                  mixed-body neighbourhood flags around changed cells
//   k_move_finish  the in->out check (abort 314); image points and weights
//                  of every IB cell (FindImageCell
//                  Body.cpp:999-1190, Mesh.cpp:1403-1442) and of the
//                  wall-force patches (Body.cpp:1506-1575); new distances in
//                  place, the stage masks of the marched cells
//                  (NS_Solver.cpp:856-983, 19-179), update count
//
// A move of the prescribed law does not depend on the flow: the time march
// runs the next step's move ahead, beside the current step's stages, on a
// second stream and with its own error word (`ahead`: the front takes the
// law at t + dt, k_move_finish only searches and checks), and the next step
// begins with k_move_join: the time advance of the step before, the move's
// error (if any) made the march's, the commit.  The results are those of
// the move run in place.
//
// The IB set is stamped with a move counter (a cell is an IB cell when its
// stamp equals the current move), so marking grows from the solid cells of
// the finest level and costs what the wall region holds, not the grid.
// The list order follows the atomic appends; every IB cell is filled
// independently of the others, so results do not depend on it.
//
// Bit-exactness.  Every expression is evaluated in the order the host mesh
// builder and the reference evaluate it, with floating-point contraction
// off, so the results equal a host evaluation bit for bit (FP64 add, mul,
// div and sqrt are correctly rounded on both sides).
#pragma once
#include "ibg_common.h"

namespace ibg {

// Rigid pose of the bodies of one case: point = pivot + R * offset with
// R = [[c, -s], [s, c]] (the rotation of the mesh builder's NACA sections).
struct Pose {
  double px, py, c, s;
};

// Prescribed pitch and heave of the bodies of one case about a pivot:
// angle of attack aoa(t) = aoa0 + amp sin(omega t) (nose up), pivot
// y(t) = y0 + heave sin(omega t).  The rotation applied to the zero-pose
// offsets is a = -aoa (the section rotates clockwise for a nose-up angle).
struct PitchHeave {
  double x0, y0, aoa0, amp, omega, heave;   // angles in radians
};

// The update count of a move is summed into kUpdSlots counters one 128-byte
// line apart (workgroup b adds into slot b % kUpdSlots; the reader sums the
// slots): thousands of workgroups adding into one address serialise on it.
constexpr int kUpdSlots = 64, kUpdSlotStride = 16;

// In-plane neighbour slots of Point_iterator_2d (i outer, j inner, k = 1)
// without the cell itself: nb27 slots 1, 4, 7, 10, 16, 19, 22, 25.
constexpr int kNb8 = 8;

// Cell sets of one move: the normal cells [0, nw) of the device order plus
// the ghosts gidx[0, ngw) (ghosts sorted by level, finest first) -- the
// levels the step marches, i.e. every level the step's kernels read.
struct MoveArgs {
  int N, Nn, npatch, ncases, nbody;
  int nw, ngw, nflag;
  const int *gidx;
  const int *key;        // [N][3]
  const int8_t *kind;    // [N] ibg_cell_kind
  const int8_t *fin;     // [N] finest level of a case with bodies
  const int16_t *cblk;   // [N] block (case, level)
  const BlkInfo *blk;
  const double *origin;  // [ncases][3]
  const double *borg;    // [nblk][3] the origin of each block's case
  const int *nb8;        // [8][N] in-plane neighbours, SoA
  const int *nb27;       // [N][27]
  const double *dist_dm; // [N] distance to the domain walls, or null
  double *signdis;
  float *sd32;
  int *cpatch;           // [N] closest patch (global index), -1: no body
  int *cp1;              // [N] new closest patch
  double *sd1;           // [N] new signed distance
  int *mk1, *mkib;       // [N] move stamps: layer 1, IB set
  int8_t *cmix;          // [N] an in-plane neighbour's closest body differs
  int *chg;              // cells whose closest body changed in this move
  int cap_chg;
  int *mixlist;          // cells left to k_move_mixed (their neighbours' bodies differ)
  int cap_mix;
  int *epoch;            // [1] move counter
  int *cnt;              // [6] list lengths: solid, (spare), in->out, IB, changed, mixed
  int cap_s, cap_io, ib_cap;
  int *slist, *iolist;
  uint8_t *flags;        // [Nn]
  // surface
  const double *pt_off;  // [npoint][2]
  const int *pat_pt0;    // [npatch] first point of the segment
  const int *pat_body;   // [npatch] global body
  const int *pat_case;   // [npatch]
  const int8_t *pat_flip;// [npatch] odd body: normal reversed
  const double *pat_dz;  // [npatch] z extent
  const int *body_p0;    // [nbody+1] patches of body b: [body_p0[b], body_p0[b+1])
  const int2 *pat_lh;    // [npatch] {body_p0[b], body_p0[b+1]} of the patch's body b
  const int4 *fbox;      // [ncases] finest-level key box {ix0, iy0, nx, ny}
  const long long *foff; // [ncases] first entry of the case's key grid
  const int *fgrid;      // cell of each in-plane key of the finest level, -1 for none
  const int *body_case;  // [nbody]
  Pose *pose;            // [ncases]
  int use_law;
  const PitchHeave *law; // [ncases]
  const double *t;       // [ncases]
  const double *dt;      // [ncases] (ahead: the law at t + dt)
  int ahead;             // the next step's move, run before the time advance
  Motion *mot, *motc;    // per body (IB fill), per case (wall force)
  double *case_center;   // [ncases][3]
  double *pat_pc, *pat_nv, *pat_area;
  int *pat_intp, *pat_cc;
  double *pat_coef, *pat_hgdis;
  // IB list
  int *ib_cell, *ib_intp, *ib_body, *ib_case;
  double *ib_coef, *ib_hgdis, *ib_attach, *ib_nv;
  unsigned long long *err;
  unsigned long long *upd;     // [1] += stage updates of this step (may be null)
  unsigned long long *lvlcnt;  // [L] per-level update counts
  int nstage;
};

constexpr double kCriticDis = 9.5;   // critic_dis, Bodymove.cpp:132

__device__ __forceinline__ int set_cell(const MoveArgs &a, int t, int n) {
  return t < n ? t : a.gidx[t - n];
}

// Append val to list (wave-aggregated: one atomic per wave); every lane of
// the wave calls it.  A full list is the capacity error 1090.
__device__ __forceinline__ void wave_append(int *cnt, bool v, int val, int *list, int cap,
                                            unsigned long long *err) {
  const unsigned long long m = __ballot(v);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader);
  if (v) {
    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
    if (pos < cap) list[pos] = val;
    else report_error(err, val, 1090);
  }
}

// Append up to K values per lane (bit k of set: val[k]) with one atomic per
// wave.  The lanes of the wave that reach the call take part; the counts and
// offsets are wave-uniform ballots.
template <int K>
__device__ __forceinline__ void wave_append_set(int *cnt, unsigned set, const int *val, int *list,
                                                int cap, unsigned long long *err) {
  unsigned long long m[K];
  int total = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    m[k] = __ballot((set >> k) & 1u);
    total += __popcll(m[k]);
  }
  if (total == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)__ballot(1)) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, total);
  base = __shfl(base, leader);
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if ((set >> k) & 1u) {
      const int pos = base + __popcll(m[k] & below);
      if (pos < cap) list[pos] = val[k];
      else report_error(err, val[k], 1090);
    }
    base += __popcll(m[k]);
  }
}

__device__ __forceinline__ void cell_center_dev(const MoveArgs &a, int i, double *x) {
#pragma clang fp contract(off)
  const BlkInfo &B = a.blk[a.cblk[i]];
  const double *o = a.origin + (size_t)B.case_id * 3;
  for (int d = 0; d < 3; ++d) x[d] = o[d] + ((double)a.key[(size_t)i * 3 + d] + 0.5) * B.h[d];
}

// PatchDistoBox Body.H:563-580
__device__ __forceinline__ void patch_dist(const MoveArgs &a, const double *x, int p, int &best,
                                           double &dist) {
#pragma clang fp contract(off)
  const double *pc = a.pat_pc + (size_t)p * 3;
  const double dx = x[0] - pc[0], dy = x[1] - pc[1], dz = x[2] - pc[2];
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  if (d < dist) { dist = d; best = p; }
}

// FindClosePatchtoCell Body.H:505-542 on the body that owns `patch`: the
// neighbours of a patch of a 2D section are the previous and the next
// segment of the same body (CollectPatchNeib Body.cpp:822-892).  The walk
// never leaves the body it starts on.
// The patch data of the first step of the walk (the start patch, its two
// neighbour segments and the start patch's normal) are loaded together with
// the body's patch range, so that a walk that stays where it starts (most
// cells, most moves) costs one round trip to the cache after its start
// patch; the comparisons are the same and in the same order as the plain
// walk, so are the results.
__device__ __forceinline__ double patch_dist_at(const double *x, double px, double py, double pz) {
#pragma clang fp contract(off)
  const double dx = x[0] - px, dy = x[1] - py, dz = x[2] - pz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// The patch data of the first step of a walk from patch p0 (its two
// neighbour segments, clamped to the patch range, and its normal) and the
// body's patch range.
struct WalkStart {
  int2 lh;
  double c0x, c0y, c0z, cmx, cmy, cmz, cqx, cqy, cqz, n0x, n0y, n0z;
};

__device__ __forceinline__ void walk_start(const MoveArgs &a, int p0, WalkStart &w) {
  const int pm = p0 > 0 ? p0 - 1 : p0, pq = p0 + 1 < a.npatch ? p0 + 1 : p0;   // in range
  w.lh = a.pat_lh[p0];
  const double *c0 = a.pat_pc + (size_t)p0 * 3, *cm = a.pat_pc + (size_t)pm * 3;
  const double *cq = a.pat_pc + (size_t)pq * 3, *n0 = a.pat_nv + (size_t)p0 * 3;
  w.c0x = c0[0]; w.c0y = c0[1]; w.c0z = c0[2];
  w.cmx = cm[0]; w.cmy = cm[1]; w.cmz = cm[2];
  w.cqx = cq[0]; w.cqy = cq[1]; w.cqz = cq[2];
  w.n0x = n0[0]; w.n0y = n0[1]; w.n0z = n0[2];
}

// The walk from `patch` given its preloaded start w.
__device__ __forceinline__ void walk_from(const MoveArgs &a, const double *x, int &patch,
                                          double &dist, double &sd, const WalkStart &w) {
#pragma clang fp contract(off)
  const int p0 = patch;
  const int lo = w.lh.x, hi = w.lh.y;
  dist = 999.0;
  {
    const double d = patch_dist_at(x, w.c0x, w.c0y, w.c0z);
    if (d < dist) { dist = d; patch = p0; }
  }
  // first step of the walk from the preloaded data
  if (p0 - 1 >= lo) {
    const double d = patch_dist_at(x, w.cmx, w.cmy, w.cmz);
    if (d < dist) { dist = d; patch = p0 - 1; }
  }
  if (p0 + 1 < hi) {
    const double d = patch_dist_at(x, w.cqx, w.cqy, w.cqz);
    if (d < dist) { dist = d; patch = p0 + 1; }
  }
  if (patch == p0) {
    sd = w.n0x * (x[0] - w.c0x) + w.n0y * (x[1] - w.c0y) + w.n0z * (x[2] - w.c0z);
    return;
  }
  // the walk moved: its remaining steps as in FindClosePatchtoCell
  int old = p0, ps0 = patch;
  for (int guard = 0; old != ps0 && guard < 1 << 20; ++guard) {
    old = patch;
    if (ps0 - 1 >= lo) patch_dist(a, x, ps0 - 1, patch, dist);
    if (ps0 + 1 < hi) patch_dist(a, x, ps0 + 1, patch, dist);
    ps0 = patch;
  }
  const double *pc = a.pat_pc + (size_t)patch * 3;
  const double *nv = a.pat_nv + (size_t)patch * 3;
  sd = nv[0] * (x[0] - pc[0]) + nv[1] * (x[1] - pc[1]) + nv[2] * (x[2] - pc[2]);
}

__device__ void walk_body(const MoveArgs &a, const double *x, int &patch, double &dist,
                          double &sd) {
  WalkStart w;
  walk_start(a, patch, w);
  walk_from(a, x, patch, dist, sd, w);
}

__device__ __forceinline__ Motion law_motion(const PitchHeave &Lw, double t, double zc) {
  const double wt = Lw.omega * t;
  const double sw = sin(wt), cw = cos(wt);
  Motion M;
  M.tra[0] = 0.0; M.tra[1] = Lw.heave * Lw.omega * cw; M.tra[2] = 0.0;
  M.rot[0] = 0.0; M.rot[1] = 0.0; M.rot[2] = -Lw.amp * Lw.omega * cw;
  M.center[0] = Lw.x0; M.center[1] = Lw.y0 + Lw.heave * sw; M.center[2] = zc;
  return M;
}

// One workgroup: move counter and list lengths, the poses (and, under the
// law, the rigid-body velocity of Body::Surface_Point_Velocity Body.H:
// 709-721 about the pivot), the reference point of each case (the pivot),
// then the patches.
// The time of case k the move takes the law at: t, or t + dt (the sum the
// time advance forms) for the next step's move run ahead of the advance.
__device__ __forceinline__ double case_time(const MoveArgs &a, int k) {
  return a.ahead ? a.t[k] + a.dt[k] : a.t[k];
}

__global__ void __launch_bounds__(256) k_move_front(MoveArgs a) {
#pragma clang fp contract(off)
  __shared__ Pose P[IBG_MAX_CASES];
  if (wg_halted(a.err)) return;
  if (threadIdx.x == 0) {
    a.epoch[0] += 1;
    for (int k = 0; k < 6; ++k) a.cnt[k] = 0;
  }
  for (int k = threadIdx.x; k < a.ncases; k += blockDim.x) {
    Pose q;
    if (a.use_law) {
      const PitchHeave Lw = a.law[k];
      const double tk = case_time(a, k);
      const double sw = sin(Lw.omega * tk);
      const double ang = -(Lw.aoa0 + Lw.amp * sw);
      q.px = Lw.x0;
      q.py = Lw.y0 + Lw.heave * sw;
      q.c = cos(ang);
      q.s = sin(ang);
      a.pose[k] = q;
      a.motc[k] = law_motion(Lw, tk, a.case_center[(size_t)k * 3 + 2]);
    } else {
      q = a.pose[k];
    }
    P[k] = q;
    a.case_center[(size_t)k * 3] = q.px;
    a.case_center[(size_t)k * 3 + 1] = q.py;
  }
  if (a.use_law)
    for (int b = threadIdx.x; b < a.nbody; b += blockDim.x) {
      const int cs = a.body_case[b];
      a.mot[b] = law_motion(a.law[cs], case_time(a, cs), a.case_center[(size_t)cs * 3 + 2]);
    }
  __syncthreads();
  for (int p = threadIdx.x; p < a.npatch; p += blockDim.x) {
    const Pose Q = P[a.pat_case[p]];
    const int i0 = a.pat_pt0[p];
    const double ox0 = a.pt_off[2 * i0], oy0 = a.pt_off[2 * i0 + 1];
    const double ox1 = a.pt_off[2 * i0 + 2], oy1 = a.pt_off[2 * i0 + 3];
    const double x0 = Q.px + (Q.c * ox0 - Q.s * oy0), y0 = Q.py + (Q.s * ox0 + Q.c * oy0);
    const double x1 = Q.px + (Q.c * ox1 - Q.s * oy1), y1 = Q.py + (Q.s * ox1 + Q.c * oy1);
    const double dx = x1 - x0, dy = y1 - y0;
    const double len = sqrt(dx * dx + dy * dy);
    double nx = -dy / len, ny = dx / len;
    if (a.pat_flip[p]) { nx = -nx; ny = -ny; }
    a.pat_pc[(size_t)p * 3] = 0.5 * (x0 + x1);
    a.pat_pc[(size_t)p * 3 + 1] = 0.5 * (y0 + y1);
    a.pat_nv[(size_t)p * 3] = nx;
    a.pat_nv[(size_t)p * 3 + 1] = ny;
    a.pat_area[p] = len * a.pat_dz[p];
  }
}

// RenewLevelDistance (Bodymove.cpp:136-249) for one cell: the walk from its
// previous patch for cells closer than critic_dis, then its entries of the
// cycle list -- the neighbours whose closest body differs, in neighbour
// order, each starting from that neighbour's patch after the neighbour's
// own walk (recomputed here: a walk keeps the body, so the bodies of the
// previous step decide which neighbours enter the list), then signdis =
// min(signdis, distance_to_dm) for walls on the domain.  On the finest
// level: solid cells to
// the solid list, in->out cells to their list, and an out->in cell that was
// not an IB cell is the reference's abort 263 (Bodymove.cpp:494-544).
// Held to 64 registers (eight waves per SIMD): the latency of the two
// round trips is hidden by the other waves, not by more registers.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_move_dist(MoveArgs a) {
#pragma clang fp contract(off)
  if (halted(a.err)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = t < a.nw + a.ngw;
  const int i = in ? set_cell(a, t, a.nw) : 0;
  bool solid = false, io = false, chg = false, defer = false;
  if (in) {
    // Two memory round trips for most cells: every per-cell operand first,
    // unconditionally (a load placed under a condition or after another's
    // use waits for its own round trip), then the block, the case origin
    // and the start of the walk, which depend only on those.
#pragma clang fp contract(off)
    const int p0 = a.cpatch[i];
    const double s0 = a.signdis[i];
    const double dm = a.dist_dm ? a.dist_dm[i] : 0.0;
    const bool mixed = a.cmix[i] != 0;
    const int8_t fn = a.fin[i], kd = a.kind[i];
    const int mk = a.mkib[i];
    const int cb = a.cblk[i];
    const int k0 = a.key[(size_t)i * 3], k1 = a.key[(size_t)i * 3 + 1], k2 = a.key[(size_t)i * 3 + 2];
    const bool fin = (fn != 0) & (kd != IBG_DMGHOST);
    const int ps = p0 >= 0 ? p0 : 0;
    WalkStart w;
    walk_start(a, ps, w);
    const int bs = a.pat_body[ps];
    const double *bh = a.blk[cb].h;
    const double *o = a.borg + (size_t)cb * 3;
    // the cell centre as cell_center_dev forms it
    double x[3];
    x[0] = o[0] + ((double)k0 + 0.5) * bh[0];
    x[1] = o[1] + ((double)k1 + 0.5) * bh[1];
    x[2] = o[2] + ((double)k2 + 0.5) * bh[2];
    const int b0 = p0 >= 0 ? bs : -1;
    int p = p0;
    double sd = s0;
    if (mixed && p0 >= 0 && s0 < kCriticDis) {
      // a neighbour's closest body differs: the cycle list's chains of
      // dependent walks go to k_move_mixed, which runs them side by side
      defer = true;
    } else {
      if (p >= 0 && s0 < kCriticDis) {
        double dist;
        walk_from(a, x, p, dist, sd, w);
        // solid loop (:224-249): signdis = min(signdis, distance_to_dm)
        if (a.dist_dm && sd < kCriticDis) sd = (dm < sd) ? dm : sd;
      }
      a.cp1[i] = p;
      a.sd1[i] = sd;
      chg = p0 >= 0 && a.pat_body[p] != b0;
      if (fin) {
        solid = sd < 0.0;
        io = s0 < 0.0 && !(sd < 0.0);
        if (!(s0 < 0.0) && sd < 0.0 && mk != a.epoch[0] - 1) report_error(a.err, i, 263);
      }
    }
  }
  wave_append(a.cnt + 0, solid, i, a.slist, a.cap_s, a.err);
  wave_append(a.cnt + 2, io, i, a.iolist, a.cap_io, a.err);
  wave_append(a.cnt + 4, chg, i, a.chg, a.cap_chg, a.err);
  wave_append(a.cnt + 5, defer, i, a.mixlist, a.cap_mix, a.err);
}

// The cells k_move_dist deferred: RenewLevelDistance with the cycle list
// (Bodymove.cpp:136-249).  Eight lanes per cell, one per in-plane
// neighbour: each lane walks its neighbour from that neighbour's patch and
// the cell from the result (the two walks of the list entry, independent of
// the list's state), then every lane applies the list in neighbour order
// to the gathered results -- the reference's sequence and change rule.  A
// fixed grid loops over the list (its length is known on the device only).
__global__ void __launch_bounds__(256) k_move_mixed(MoveArgs a) {
#pragma clang fp contract(off)
  if (halted(a.err)) return;
  const int n = min(a.cnt[5], a.cap_mix);
  const int k = threadIdx.x & 7;                  // this lane's neighbour
  const int g0 = (threadIdx.x & 63) & ~7;         // first lane of the cell's group
  const int per = (int)blockDim.x / 8;
  for (int base = blockIdx.x * per; base < n; base += gridDim.x * per) {
    const int j = base + (int)(threadIdx.x >> 3);
    const bool in = j < n;
    const int i = in ? a.mixlist[j] : 0;
    bool solid = false, io = false, chg = false;
    // the cell's own walk (every lane of the group, the same result)
    int p0 = -1, p = -1, myb0 = -1, mk = 0;
    double s0 = 0.0, sd = 0.0, dist = 0.0, dm = 0.0, x[3] = {0.0, 0.0, 0.0};
    bool fin = false, cyc = false;
    if (in) {
      p0 = a.cpatch[i];
      s0 = a.signdis[i];
      dm = a.dist_dm ? a.dist_dm[i] : 0.0;
      fin = a.fin[i] && a.kind[i] != IBG_DMGHOST;
      mk = a.mkib[i];
      p = p0;
      sd = s0;
      cell_center_dev(a, i, x);
      walk_body(a, x, p, dist, sd);   // deferred cells have p0 >= 0, s0 < critic_dis
      cyc = sd < kCriticDis;
      if (cyc) myb0 = a.pat_body[p];
    }
    // this lane's list entry
    bool ok = false;
    int nbb = -1, q1 = -1;
    double dn = 0.0, sn = 0.0;
    if (in && cyc) {
      const int an = a.nb8[(size_t)k * a.N + i];
      const int q0 = an >= 0 ? a.cpatch[an] : -1;
      if (q0 >= 0) {
        ok = true;
        nbb = a.pat_body[q0];
        if (nbb != myb0) {
          q1 = q0;
          if (a.signdis[an] < kCriticDis) {
            double xa[3], da, sa;
            cell_center_dev(a, an, xa);
            walk_body(a, xa, q1, da, sa);
          }
          walk_body(a, x, q1, dn, sn);
        }
      }
    }
    // the list in neighbour order (every lane of the wave takes part in
    // the exchanges)
    int myb = myb0;
#pragma unroll
    for (int e = 0; e < kNb8; ++e) {
      const int src = g0 + e;
      const bool ok_e = __shfl(ok ? 1 : 0, src) != 0;
      const int nbb_e = __shfl(nbb, src);
      const int q1_e = __shfl(q1, src);
      const double dn_e = __shfl(dn, src);
      const double sn_e = __shfl(sn, src);
      if (!cyc || !ok_e) continue;
      if (nbb_e == myb0) continue;      // pushed only when it differs (:172-179)
      if (nbb_e == myb) continue;       // processed only when it still differs (:193)
      bool change = false;
      if (sn_e > 0.0 && sd < 0.0) change = true;
      else if (dn_e < dist) {
        if (sd * sn_e > 0.0) change = true;
        else if (sn_e > 0.0) change = true;
      }
      if (change) { myb = nbb_e; p = q1_e; dist = dn_e; sd = sn_e; }
    }
    const bool lead = in && k == 0;
    if (lead) {
      // solid loop (:224-249): signdis = min(signdis, distance_to_dm)
      if (a.dist_dm && sd < kCriticDis) sd = (dm < sd) ? dm : sd;
      a.cp1[i] = p;
      a.sd1[i] = sd;
      chg = a.pat_body[p] != a.pat_body[p0];
      if (fin) {
        solid = sd < 0.0;
        io = s0 < 0.0 && !(sd < 0.0);
        if (!(s0 < 0.0) && sd < 0.0 && mk != a.epoch[0] - 1) report_error(a.err, i, 263);
      }
    }
    wave_append(a.cnt + 0, solid, i, a.slist, a.cap_s, a.err);
    wave_append(a.cnt + 2, io, i, a.iolist, a.cap_io, a.err);
    wave_append(a.cnt + 4, chg, i, a.chg, a.cap_chg, a.err);
  }
}

// cmix of cell x from the new closest patches of x and its neighbours
__device__ __forceinline__ void refresh_cmix(const MoveArgs &a, int x) {
  const int px = a.cp1[x];
  int8_t v = 0;
  if (px >= 0) {
    const int bx = a.pat_body[px];
    for (int k = 0; k < kNb8 && !v; ++k) {
      const int y = a.nb8[(size_t)k * a.N + x];
      if (y < 0) continue;
      const int py = a.cp1[y];
      if (py >= 0 && a.pat_body[py] != bx) v = 1;
    }
  }
  a.cmix[x] = v;
}

// The IB set of the move in one pass: eight lanes per solid cell of the
// finest level, one per in-plane neighbour x.  An eligible x (not a domain
// ghost, not solid) is a layer-1 cell -- stamped for the in->out check of
// k_move_finish -- and an IB cell, and so is every eligible in-plane
// neighbour of x: the set of the initial rule (Body.cpp:1016-1082), which
// RemoveIBCells / InfectNewIBCells (Bodymove.cpp:345-767) keep after every
// move.  Stamps by atomicMax of the move counter; the first IB stamp of a
// cell appends it to the IB list (one atomic per wave).  Threads past the
// solid lanes refresh the mixed-body flags around cells whose body changed.
__global__ void __launch_bounds__(256) k_move_mark(MoveArgs a) {
  if (halted(a.err)) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int nsl = kNb8 * a.cap_s;   // lanes of the solid list
  if (j >= nsl) {
    // the neighbourhood flags around cells whose body changed
    const int r = j - nsl;
    if (r >= min(a.cnt[4], a.cap_chg)) return;
    const int c = a.chg[r];
    refresh_cmix(a, c);
    for (int k = 0; k < kNb8; ++k) {
      const int y = a.nb8[(size_t)k * a.N + c];
      if (y >= 0) refresh_cmix(a, y);
    }
    return;
  }
  // (the lanes past the solid lanes have left: those that reach the append
  // are the low lanes of the wave)
  const int ep = a.epoch[0];
  const int js = j / kNb8, k = j % kNb8;
  int c[kNb8 + 1];
  unsigned fresh = 0;
  if (js < min(a.cnt[0], a.cap_s)) {
    const int sc = a.slist[js];
    const int x = a.nb8[(size_t)k * a.N + sc];
    if (x >= 0 && a.kind[x] != IBG_DMGHOST && !(a.sd1[x] < 0.0)) {
      bool ok[kNb8 + 1];
      c[0] = x;
      ok[0] = true;
#pragma unroll
      for (int kk = 0; kk < kNb8; ++kk) {
        const int m = a.nb8[(size_t)kk * a.N + x];
        c[kk + 1] = m;
        ok[kk + 1] = m >= 0 && a.kind[m] != IBG_DMGHOST && !(a.sd1[m] < 0.0);
      }
      atomicMax(&a.mk1[x], ep);
      // the nine stamps are independent: issued together
#pragma unroll
      for (int kk = 0; kk <= kNb8; ++kk)
        if (ok[kk] && atomicMax(&a.mkib[c[kk]], ep) < ep) {
          fresh |= 1u << kk;
          if (a.kind[c[kk]] != IBG_NORMAL) report_error(a.err, c[kk], 1082);
        }
    }
  }
  wave_append_set<kNb8 + 1>(a.cnt + 3, fresh, c, a.ib_cell, a.ib_cap, a.err);
}

// Finest-level cell of case cs with in-plane key (ix, iy), -1 for none.
__device__ __forceinline__ int fine_cell(const MoveArgs &a, int cs, int ix, int iy) {
  const int4 bx = a.fbox[cs];
  const int rx = ix - bx.x, ry = iy - bx.y;
  if (rx < 0 || ry < 0 || rx >= bx.z || ry >= bx.w) return -1;
  return a.fgrid[a.foff[cs] + (long long)ry * bx.z + rx];
}

// Cell of the finest level holding key ik, reached from `home` (key k0, k1,
// kz) along the walk of the neighbour table the reference's search takes
// (one step per axis toward the target at a time, at most 256 steps); -1
// when a key on that path is not a cell of the level.  The keys of the path
// are known in advance, so their cells are read from the key grid together
// rather than hop by hop.
__device__ int locate_dev(const MoveArgs &a, int cs, int home, int k0, int k1, int kz,
                          const int *ik) {
  if (ik[2] != kz) return -1;
  const int d0 = ik[0] - k0, d1 = ik[1] - k1;
  const int a0 = d0 < 0 ? -d0 : d0, a1 = d1 < 0 ? -d1 : d1;
  const int s0 = (d0 > 0) - (d0 < 0), s1 = (d1 > 0) - (d1 < 0);
  const int steps = a0 > a1 ? a0 : a1;
  if (steps == 0) return home;
  if (steps > 255) return -1;   // the walk gives up after 256 checks
  constexpr int kNear = 4;
  int c[kNear];
#pragma unroll
  for (int t = 1; t <= kNear; ++t)
    c[t - 1] = (t <= steps) ? fine_cell(a, cs, k0 + s0 * (t < a0 ? t : a0), k1 + s1 * (t < a1 ? t : a1))
                            : 0;
#pragma unroll
  for (int t = 1; t <= kNear; ++t)
    if (t <= steps && c[t - 1] < 0) return -1;
  if (steps <= kNear) return c[steps - 1];
  int cur = c[kNear - 1];
  for (int t = kNear + 1; t <= steps; ++t) {
    cur = fine_cell(a, cs, k0 + s0 * (t < a0 ? t : a0), k1 + s1 * (t < a1 ? t : a1));
    if (cur < 0) return -1;
  }
  return cur;
}

// Image-point search of the host mesh builder (Body.cpp:1044-1183,
// 1527-1575): march from `start` along nv in steps of one cell in the
// dominant direction until the 2x2 bracketing cells are neither solid nor
// IB cells, then the weights of Mesh::ComptHgIntpCoef (Mesh.cpp:1403-1442).
// `home` is a finest-level cell near start (case cs); on success it is the
// cell holding the image point.  Cells are found in the finest level's key
// grid: the located cell's key is the target key, its centre follows from
// it, and the bracketing cells are its in-plane neighbours by key (the z
// neighbours of a one-layer 2D level are the cell itself).
__device__ bool search_image_dev(const MoveArgs &a, int ep, int cs, const double *start,
                                 const double *nv, const double *h, const double *org, int &home,
                                 int *intp, double *coef, double &hgdis_out) {
#pragma clang fp contract(off)
  double dk[3];
  for (int d = 0; d < 3; ++d) dk[d] = nv[d] / h[d];
  double dmax = fabs(dk[0]) < fabs(dk[1]) ? fabs(dk[1]) : fabs(dk[0]);
  dmax = dmax < fabs(dk[2]) ? fabs(dk[2]) : dmax;
  const double dxyz[3] = {nv[0] / dmax, nv[1] / dmax, nv[2] / dmax};
  const double steplen = sqrt(dxyz[0] * dxyz[0] + dxyz[1] * dxyz[1] + dxyz[2] * dxyz[2]);
  const double ca[3] = {0.0, 1.0, 0.0};
  const double cb[3] = {1.0, -1.0, 1.0};
  int hk0 = a.key[(size_t)home * 3], hk1 = a.key[(size_t)home * 3 + 1];
  int hkz = a.key[(size_t)home * 3 + 2];
  double hgdis = 0.0;
  for (int it = 0; it < 64; ++it) {
    hgdis += steplen;
    double pt[3];
    for (int d = 0; d < 3; ++d) pt[d] = start[d] + nv[d] * hgdis;
    int ik[3];
    for (int d = 0; d < 3; ++d) ik[d] = (int)floor((pt[d] - org[d]) / h[d]);
    double cen[3];
    for (int d = 0; d < 3; ++d) cen[d] = org[d] + ((double)ik[d] + 0.5) * h[d];
    double dkt[3];
    for (int d = 0; d < 3; ++d) dkt[d] = (pt[d] - cen[d]) / h[d];
    dkt[2] = 0.0;   // one z layer: the image point is in it
    int hgdir[3];
    for (int d = 0; d < 3; ++d) hgdir[d] = dkt[d] > 0.0 ? 1 : 0;
    // the bracketing cells by key, read with the path (independent reads)
    int br[2][2];
    for (int pi = 0; pi < 2; ++pi)
      for (int pj = 0; pj < 2; ++pj)
        br[pi][pj] = fine_cell(a, cs, ik[0] + pi + hgdir[0] - 1, ik[1] + pj + hgdir[1] - 1);
    const int cc = locate_dev(a, cs, home, hk0, hk1, hkz, ik);
    if (cc < 0) return false;
    if (a.kind[cc] == IBG_DMGHOST) return false;
    home = cc;
    hk0 = ik[0]; hk1 = ik[1]; hkz = ik[2];
    bool ok = true;
    for (int pi = 0; pi < 2; ++pi)
      for (int pj = 0; pj < 2; ++pj)
        for (int pk = 0; pk < 2; ++pk) {
          const int nx = pi + hgdir[0], ny = pj + hgdir[1], nz = pk + hgdir[2];
          const int an = br[pi][pj];
          const int q = pi * 4 + pj * 2 + pk;
          intp[q] = an;
          coef[q] = 0.0;
          if (an < 0) { ok = false; continue; }
          if (a.sd1[an] < 0.0 || a.mkib[an] == ep) ok = false;
          coef[q] = (ca[nx] + cb[nx] * fabs(dkt[0])) * (ca[ny] + cb[ny] * fabs(dkt[1])) *
                    (ca[nz] + cb[nz] * fabs(dkt[2]));
        }
    if (ok) { hgdis_out = hgdis; return true; }
  }
  return false;
}

// Image point, weights, attach point (Body.cpp:1073) and normal of every IB
// cell, then (threads past the list capacity) the image points of the
// wall-force patches, searched from each patch centre (pat_cc holds a
// finest-level cell near the previous image point).  Failure is the
// reference's abort of FindImageCell (1183) / FindPatchHangingNode (1575).
__device__ void move_image(const MoveArgs &a, int j) {
#pragma clang fp contract(off)
  const int ep = a.epoch[0];
  int intp[8];
  double coef[8], hgdis = 0.0;
  if (j < a.ib_cap) {
    if (j >= min(a.cnt[3], a.ib_cap)) return;
    const int g = a.ib_cell[j];
    const int p = a.cp1[g];
    if (p < 0) { report_error(a.err, g, 1183); return; }
    const BlkInfo &B = a.blk[a.cblk[g]];
    const double *org = a.origin + (size_t)B.case_id * 3;
    const double nv[3] = {a.pat_nv[(size_t)p * 3], a.pat_nv[(size_t)p * 3 + 1],
                          a.pat_nv[(size_t)p * 3 + 2]};
    const double sd = a.sd1[g];
    double cen[3];
    cell_center_dev(a, g, cen);
    int home = g;
    if (!search_image_dev(a, ep, B.case_id, cen, nv, B.h, org, home, intp, coef, hgdis)) {
      report_error(a.err, g, 1183);
      for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
    }
    for (int q = 0; q < 8; ++q) {
      const bool z0 = (q & 1) == 0;   // zero-weight z layer of the 2D stencil
      a.ib_intp[(size_t)j * 8 + q] = z0 ? -1 : intp[q];
      a.ib_coef[(size_t)j * 8 + q] = z0 ? 0.0 : coef[q];
    }
    a.ib_hgdis[j] = hgdis;
    for (int d = 0; d < 3; ++d) {
      a.ib_nv[(size_t)j * 3 + d] = nv[d];
      a.ib_attach[(size_t)j * 3 + d] = cen[d] - nv[d] * sd;
    }
    a.ib_body[j] = a.pat_body[p];
    a.ib_case[j] = B.case_id;
    return;
  }
  const int p = j - a.ib_cap;
  if (p >= a.npatch) return;
  int home = a.pat_cc[p];
  const BlkInfo &B = a.blk[a.cblk[home]];
  const double *org = a.origin + (size_t)B.case_id * 3;
  const double pc[3] = {a.pat_pc[(size_t)p * 3], a.pat_pc[(size_t)p * 3 + 1],
                        a.pat_pc[(size_t)p * 3 + 2]};
  const double nv[3] = {a.pat_nv[(size_t)p * 3], a.pat_nv[(size_t)p * 3 + 1],
                        a.pat_nv[(size_t)p * 3 + 2]};
  if (!search_image_dev(a, ep, B.case_id, pc, nv, B.h, org, home, intp, coef, hgdis)) {
    report_error(a.err, a.pat_cc[p], 1575);
    for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
  } else {
    a.pat_cc[p] = home;
  }
  for (int q = 0; q < 8; ++q) {
    const bool z0 = (q & 1) == 0;
    a.pat_intp[(size_t)p * 8 + q] = z0 ? -1 : intp[q];
    a.pat_coef[(size_t)p * 8 + q] = z0 ? 0.0 : coef[q];
  }
  a.pat_hgdis[p] = hgdis;
}

// Every in->out cell must be a layer-1 cell (InfectNewIBCells
// Bodymove.cpp:575-650, abort 314: the move step is too large for the
// incremental update).  Workgroup cb of nblk, grid-stride.
__device__ void move_io_check(const MoveArgs &a, int cb, int nblk) {
  const int ep = a.epoch[0];
  for (int r = cb * blockDim.x + threadIdx.x; r < min(a.cnt[2], a.cap_io);
       r += nblk * blockDim.x) {
    const int c = a.iolist[r];
    if (a.mk1[c] != ep) report_error(a.err, c, 314);
  }
}

// New distances and patches in place (the FP32 copy for the mixed SA
// source), the stage masks of the marched normal cells as at upload; with
// `upd` set, the stage updates of this step are accumulated (one atomic
// per workgroup, grid-stride).  Workgroup cb of the nblk committing ones.
__device__ void move_commit(const MoveArgs &a, int cb, int nblk) {
  const int ep = a.epoch[0];
  const int n = a.nw + a.ngw;
  int cnt = 0;
  for (int t = cb * blockDim.x + threadIdx.x; t < n; t += nblk * blockDim.x) {
    const int i = set_cell(a, t, a.nw);
    const double sd = a.sd1[i];
    a.cpatch[i] = a.cp1[i];
    a.signdis[i] = sd;
    a.sd32[i] = (float)sd;
    if (t < a.nflag) {
      const bool ib = a.mkib[i] == ep;
      const bool solid = sd < 0.0;
      uint8_t f;
      if (!a.fin[i]) {
        f = F_UPDATE | F_RESID | ((sd > 0.0) ? F_PRIM : F_FREE);
      } else {
        f = F_FINEST;
        if (!solid && !ib) f |= F_UPDATE;
        if (sd > 0.0 && !ib) f |= F_PRIM | F_RESID;
        else if (!ib) f |= F_FREE;
      }
      if (sd >= 0.0) f |= F_SA;
      a.flags[i] = f;
      cnt += (f & F_UPDATE) ? 1 : 0;
    }
  }
  if (a.upd) {
    __shared__ int part[4];
    int v = cnt;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int s = part[0] + part[1] + part[2] + part[3];
      if (s)
        atomicAdd(a.upd + (size_t)(cb % kUpdSlots) * kUpdSlotStride,
                  (unsigned long long)s * a.nstage);
    }
  }
}

// Image search, in->out check and commit in one launch: workgroups
// [0, nimage) search (they read only the new arrays, not the committed
// ones), the others check and, with `commit`, commit.  The searches are a
// few thousand dependent chains of loads, the commit a stream over every
// marched cell: dispatched first, the searches run beside the stream
// instead of after it.  A move run ahead (commit 0) is committed by
// k_move_join.
__global__ void __launch_bounds__(256) k_move_finish(MoveArgs a, int nimage, int commit) {
  if (wg_halted(a.err)) return;
  const int b = (int)blockIdx.x, nb = (int)gridDim.x - nimage;
  if (b < nimage) {
    move_image(a, b * blockDim.x + threadIdx.x);
    return;
  }
  move_io_check(a, b - nimage, nb);
  if (commit) move_commit(a, b - nimage, nb);
}

// The start of a step whose move ran ahead (err2: that move's error word):
// the time advance of the step before, unless the march stopped in it (as
// k_motion_advance); then the move's error, if it failed, becomes the
// march's (the state stays the one after the step before, as when the
// move fails in place); else the commit.  Only workgroup 0 writes the error
// word, after reading it, so the advance sees the word the step before
// left.
__global__ void __launch_bounds__(256) k_move_join(MoveArgs a, const unsigned long long *err2,
                                                   double *t) {
  __shared__ int st;
  if (threadIdx.x == 0) st = halted(a.err) ? 1 : (halted(err2) ? 2 : 0);
  __syncthreads();
  if (st == 1) return;
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < a.ncases; k += blockDim.x) t[k] += a.dt[k];
  if (st == 2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMin(a.err, *err2);
    return;
  }
  move_commit(a, (int)blockIdx.x, (int)gridDim.x);
}

// Per-level count of the cells the stage update touches.
__global__ void __launch_bounds__(256) k_move_levelcount(MoveArgs a, int L) {
  __shared__ unsigned cnt[IBG_MAX_LEVELS];
  if (threadIdx.x < IBG_MAX_LEVELS) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.Nn && (a.flags[i] & F_UPDATE)) atomicAdd(&cnt[a.blk[a.cblk[i]].level], 1u);
  __syncthreads();
  if (threadIdx.x < L && cnt[threadIdx.x]) atomicAdd(&a.lvlcnt[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// t += dt of the step just marched.
__global__ void k_motion_advance(int ncases, double *t, const double *dt,
                                 const unsigned long long *err) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ncases && !halted(err)) t[k] += dt[k];
}

}  // namespace ibg
