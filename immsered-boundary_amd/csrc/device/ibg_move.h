// ibg_move.h — per-step immersed-boundary re-tagging of moving bodies on a
// fixed level set (the part of Body::MovingBody_Image, Bodymove.cpp:37-125,
// that runs between regrids).
//
//   k_move_patches   new surface points and patch geometry
//                    (ComptNewLocation Bodymove.cpp:3-23, ComptPatchParams
//                    :25-35, the segment rules of the mesh builder)
//   k_move_walk      closest patch by a local walk from the previous one
//                    (RenewLevelDistance :136-155, FindClosePatchtoCell
//                    Body.H:505-542)
//   k_move_cycle     the closest patch of the other bodies of the case
//                    (:156-223) and the solid flag (:224-249)
//   k_move_l1 / k_move_ib / k_move_scan / k_move_list
//                    the IB set.  RemoveIBCells / InfectNewIBCells
//                    (:345-767) keep, after every move, the set of the
//                    initial rule InfectWallBox_Init_Negative
//                    (Body.cpp:1016-1082): non-solid cells with a solid
//                    neighbour and their non-solid neighbours.  The list is
//                    rebuilt in cell order.
//   k_move_image     image points and multilinear weights of every IB cell
//                    (FindImageCell Body.cpp:999-1190, ComptHgIntpCoef
//                    Mesh.cpp:1403-1442)
//   k_move_patch_image  the same for the wall-force patches
//                    (Body.cpp:1506-1575)
//   k_move_flags     the stage masks of the new solid / IB sets
//                    (NS_Solver.cpp:856-983, 19-179)
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

struct MoveArgs {
  int N, Nn, npatch;
  const int *key;        // [N][3]
  const int8_t *kind;    // [N] ibg_cell_kind
  const int8_t *fin;     // [N] finest level of a case with bodies
  const int16_t *cblk;   // [N] block (case, level)
  const BlkInfo *blk;
  const double *origin;  // [ncases][3]
  const int *nb27;
  double *signdis;
  float *sd32;
  int *cpatch;           // [N] closest patch (global index), -1: no body
  int *cp1;              // [N] after the walk
  double *dist1, *sd1;   // [N]
  int8_t *l1, *isib;     // [N]
  uint8_t *flags;        // [Nn]
  // surface
  const double *pt_off;  // [npoint][2]
  const int *pat_pt0;    // [npatch] first point of the segment
  const int *pat_body;   // [npatch] global body
  const int *pat_case;   // [npatch]
  const int8_t *pat_flip;// [npatch] odd body: normal reversed
  const double *pat_dz;  // [npatch] z extent
  const int *body_p0;    // [nbody+1] patches of body b: [body_p0[b], body_p0[b+1])
  const Pose *pose;      // [ncases]
  double *pat_pc, *pat_nv, *pat_area;
  int *pat_intp, *pat_cc;
  double *pat_coef, *pat_hgdis;
  // IB list
  int ib_cap;
  int *ib_n;             // [1] cells in the list (clamped to ib_cap)
  int *bcount, *boff;    // [ceil(N/256)] per-block IB counts / offsets
  int *ib_cell, *ib_intp, *ib_body, *ib_case;
  double *ib_coef, *ib_hgdis, *ib_attach, *ib_nv;
  unsigned long long *err;
  unsigned long long *upd;     // [1] += stage updates of this step (may be null)
  unsigned long long *lvlcnt;  // [L] per-level update counts
  int upd_lmin, nstage;
};

constexpr double kCriticDis = 9.5;   // critic_dis, Bodymove.cpp:132

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
// segment of the same body (CollectPatchNeib Body.cpp:822-892).
__device__ void walk_body(const MoveArgs &a, const double *x, int &patch, double &dist,
                          double &sd) {
#pragma clang fp contract(off)
  const int b = a.pat_body[patch];
  const int lo = a.body_p0[b], hi = a.body_p0[b + 1];
  dist = 999.0;
  patch_dist(a, x, patch, patch, dist);
  int old = -1, ps0 = patch;
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

// New surface: points of the patch segment at the case's pose, then the
// segment centre, outward normal and area of the host mesh builder
// (ib_mesh.cpp build_segments / build_patches).
__global__ void k_move_patches(MoveArgs a) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npatch) return;
  const Pose P = a.pose[a.pat_case[p]];
  const int i0 = a.pat_pt0[p];
  const double ox0 = a.pt_off[2 * i0], oy0 = a.pt_off[2 * i0 + 1];
  const double ox1 = a.pt_off[2 * i0 + 2], oy1 = a.pt_off[2 * i0 + 3];
  const double x0 = P.px + (P.c * ox0 - P.s * oy0), y0 = P.py + (P.s * ox0 + P.c * oy0);
  const double x1 = P.px + (P.c * ox1 - P.s * oy1), y1 = P.py + (P.s * ox1 + P.c * oy1);
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

// RenewLevelDistance first loop (Bodymove.cpp:136-155): cells closer than
// critic_dis walk from their previous closest patch.
__global__ void k_move_walk(MoveArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  int p = a.cpatch[i];
  const double sd0 = a.signdis[i];
  if (p < 0 || !(sd0 < kCriticDis)) {
    a.cp1[i] = p; a.sd1[i] = sd0; a.dist1[i] = 999.0;
    return;
  }
  double x[3], dist, sd;
  cell_center_dev(a, i, x);
  walk_body(a, x, p, dist, sd);
  a.cp1[i] = p; a.sd1[i] = sd; a.dist1[i] = dist;
}

// RenewLevelDistance cycle and solid loops (Bodymove.cpp:156-249).  The
// cycle list holds, per cell, its neighbours whose closest body differs, in
// neighbour order; a cell's entries only change that cell, so one thread
// per cell processes its entries in order.  distance_to_dm is 9999 without
// wall boundaries on the domain, so signdis = min(signdis, distance_to_dm)
// leaves it unchanged.  An in->out cell of the finest level needs a solid
// neighbour (InfectNewIBCells :575-650, abort 314).
template <int D>
__global__ void k_move_cycle(MoveArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  int p = a.cp1[i];
  double sd = a.sd1[i], dist = a.dist1[i];
  const double sd_old = a.signdis[i];
  if (p >= 0 && sd < kCriticDis) {
    const int myb0 = a.pat_body[p];
    int myb = myb0;
    double x[3];
    bool have_x = false;
    for (int q = 0; q < 27; ++q) {
      if (D == 2 && q % 3 != 1) continue;
      const int an = a.nb27[(size_t)i * 27 + q];
      if (an < 0) continue;
      const int np = a.cp1[an];
      if (np < 0) continue;
      const int nbb = a.pat_body[np];
      if (nbb == myb0) continue;      // pushed only when it differs (:172-179)
      if (myb == nbb) continue;       // processed only when it still differs (:193)
      if (!have_x) { cell_center_dev(a, i, x); have_x = true; }
      int pn = np;
      double dn, sn;
      walk_body(a, x, pn, dn, sn);
      bool change = false;
      if (sn > 0.0 && sd < 0.0) change = true;
      else if (dn < dist) {
        if (sd * sn > 0.0) change = true;
        else if (sn > 0.0) change = true;
      }
      if (change) { myb = nbb; p = pn; dist = dn; sd = sn; }
    }
  }
  a.cpatch[i] = p;
  a.signdis[i] = sd;
  a.sd32[i] = (float)sd;
  // out->in cells of the finest level were IB cells (RemoveIBCells abort
  // 263, Bodymove.cpp:494-544); in->out cells are checked by k_move_l1
  const bool chk = a.fin[i] && a.kind[i] != IBG_DMGHOST;
  if (chk && !(sd_old < 0.0) && sd < 0.0 && a.isib[i] != 1) report_error(a.err, i, 263);
  a.isib[i] = (chk && sd_old < 0.0 && !(sd < 0.0)) ? 2 : 0;
}

// IB layer 1: non-solid, non-domain-ghost finest cells with a solid
// non-domain-ghost neighbour (2D: the in-plane 3x3; 3D: 27).  An in->out
// cell without one is the reference's abort 314 (the move step is too
// large for the incremental update, Bodymove.cpp:604-650).
template <int D>
__global__ void k_move_l1(MoveArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  int8_t v = 0;
  if (a.fin[i] && a.kind[i] != IBG_DMGHOST && !(a.signdis[i] < 0.0)) {
    for (int q = 0; q < 27 && !v; ++q) {
      if (D == 2 && q % 3 != 1) continue;
      const int an = a.nb27[(size_t)i * 27 + q];
      if (an >= 0 && a.signdis[an] < 0.0 && a.kind[an] != IBG_DMGHOST) v = 1;
    }
  }
  a.l1[i] = v;
  if (a.isib[i] == 2 && !v) report_error(a.err, i, 314);
}

// IB set: layer-1 cells and the non-solid, non-domain-ghost neighbours of
// layer-1 cells; per-block counts for the compaction.
template <int D>
__global__ void __launch_bounds__(256) k_move_ib(MoveArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int v = 0;
  if (i < a.N && a.fin[i] && a.kind[i] != IBG_DMGHOST && !(a.signdis[i] < 0.0)) {
    if (a.l1[i]) v = 1;
    for (int q = 0; q < 27 && !v; ++q) {
      if (D == 2 && q % 3 != 1) continue;
      const int an = a.nb27[(size_t)i * 27 + q];
      if (an >= 0 && a.l1[an]) v = 1;
    }
    if (v && a.kind[i] != IBG_NORMAL) report_error(a.err, i, 1082);
  }
  if (i < a.N) a.isib[i] = (int8_t)v;
  const int n = __syncthreads_count(v);
  if (threadIdx.x == 0) a.bcount[blockIdx.x] = n;
}

// Exclusive scan of the per-block counts (one workgroup).
__global__ void __launch_bounds__(1024) k_move_scan(MoveArgs a, int nblocks) {
  __shared__ int part[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nblocks; base += 1024) {
    const int j = base + (int)threadIdx.x;
    const int v = j < nblocks ? a.bcount[j] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (j < nblocks) a.boff[j] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (carry > a.ib_cap) report_error(a.err, a.N, 1090);
    a.ib_n[0] = carry < a.ib_cap ? carry : a.ib_cap;
  }
}

// IB list in cell order.
__global__ void __launch_bounds__(256) k_move_list(MoveArgs a) {
  __shared__ int wsum[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool v = i < a.N && a.isib[i];
  const unsigned long long m = __ballot(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  if (!v) return;
  int pos = a.boff[blockIdx.x];
  for (int k = 0; k < w; ++k) pos += wsum[k];
  pos += __popcll(m & ((1ull << lane) - 1ull));
  if (pos < a.ib_cap) a.ib_cell[pos] = i;
}

// Cell of the finest level holding key ik, reached from `home` through the
// neighbour table (one step per axis at a time); -1 when the path leaves the
// level.
template <int D>
__device__ int locate_dev(const MoveArgs &a, int home, const int *ik) {
  int cur = home;
  int k0 = a.key[(size_t)cur * 3], k1 = a.key[(size_t)cur * 3 + 1], k2 = a.key[(size_t)cur * 3 + 2];
  if (D == 2 && ik[2] != k2) return -1;
  for (int guard = 0; guard < 256; ++guard) {
    if (k0 == ik[0] && k1 == ik[1] && k2 == ik[2]) return cur;
    const int sx = (ik[0] > k0) - (ik[0] < k0), sy = (ik[1] > k1) - (ik[1] < k1);
    const int sz = D == 2 ? 0 : (ik[2] > k2) - (ik[2] < k2);
    const int nx = a.nb27[(size_t)cur * 27 + (1 + sx) * 9 + (1 + sy) * 3 + (1 + sz)];
    if (nx < 0) return -1;
    cur = nx;
    k0 += sx; k1 += sy; k2 += sz;
  }
  return -1;
}

// Image-point search of the host mesh builder (Body.cpp:1044-1183,
// 1527-1575): march from `start` along nv in steps of one cell in the
// dominant direction until the 2x2(x2) bracketing cells are neither solid
// nor IB cells, then the weights of Mesh::ComptHgIntpCoef (Mesh.cpp:
// 1403-1442).  `home` is a finest-level cell near start; on success it is
// the cell holding the image point.
template <int D>
__device__ bool search_image_dev(const MoveArgs &a, const double *start, const double *nv,
                                 const double *h, const double *org, int &home, int *intp,
                                 double *coef, double &hgdis_out) {
#pragma clang fp contract(off)
  double dk[3];
  for (int d = 0; d < 3; ++d) dk[d] = nv[d] / h[d];
  double dmax = fabs(dk[0]) < fabs(dk[1]) ? fabs(dk[1]) : fabs(dk[0]);
  dmax = dmax < fabs(dk[2]) ? fabs(dk[2]) : dmax;
  const double dxyz[3] = {nv[0] / dmax, nv[1] / dmax, nv[2] / dmax};
  const double steplen = sqrt(dxyz[0] * dxyz[0] + dxyz[1] * dxyz[1] + dxyz[2] * dxyz[2]);
  const double ca[3] = {0.0, 1.0, 0.0};
  const double cb[3] = {1.0, -1.0, 1.0};
  double hgdis = 0.0;
  for (int it = 0; it < 64; ++it) {
    hgdis += steplen;
    double pt[3];
    for (int d = 0; d < 3; ++d) pt[d] = start[d] + nv[d] * hgdis;
    int ik[3];
    for (int d = 0; d < 3; ++d) ik[d] = (int)floor((pt[d] - org[d]) / h[d]);
    const int cc = locate_dev<D>(a, home, ik);
    if (cc < 0) return false;
    if (a.kind[cc] == IBG_DMGHOST) return false;
    home = cc;
    double cen[3];
    cell_center_dev(a, cc, cen);
    double dkt[3];
    for (int d = 0; d < 3; ++d) dkt[d] = (pt[d] - cen[d]) / h[d];
    if (D == 2) dkt[2] = 0.0;
    int hgdir[3];
    for (int d = 0; d < 3; ++d) hgdir[d] = dkt[d] > 0.0 ? 1 : 0;
    bool ok = true;
    for (int pi = 0; pi < 2; ++pi)
      for (int pj = 0; pj < 2; ++pj)
        for (int pk = 0; pk < 2; ++pk) {
          const int nx = pi + hgdir[0], ny = pj + hgdir[1], nz = pk + hgdir[2];
          const int an = a.nb27[(size_t)cc * 27 + nx * 9 + ny * 3 + nz];
          const int q = pi * 4 + pj * 2 + pk;
          intp[q] = an;
          coef[q] = 0.0;
          if (an < 0) { ok = false; continue; }
          if (a.signdis[an] < 0.0 || a.isib[an]) ok = false;
          coef[q] = (ca[nx] + cb[nx] * fabs(dkt[0])) * (ca[ny] + cb[ny] * fabs(dkt[1])) *
                    (ca[nz] + cb[nz] * fabs(dkt[2]));
        }
    if (ok) { hgdis_out = hgdis; return true; }
  }
  return false;
}

// Image point, weights, attach point (Body.cpp:1073) and normal of every IB
// cell; failure is the reference's abort of FindImageCell.
template <int D>
__global__ void k_move_image(MoveArgs a) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ib_n[0]) return;
  const int g = a.ib_cell[j];
  const int p = a.cpatch[g];
  if (p < 0) { report_error(a.err, g, 1183); return; }
  const BlkInfo &B = a.blk[a.cblk[g]];
  const double *org = a.origin + (size_t)B.case_id * 3;
  const double nv[3] = {a.pat_nv[(size_t)p * 3], a.pat_nv[(size_t)p * 3 + 1],
                        a.pat_nv[(size_t)p * 3 + 2]};
  const double sd = a.signdis[g];
  double cen[3];
  cell_center_dev(a, g, cen);
  int intp[8];
  double coef[8], hgdis = 0.0;
  int home = g;
  if (!search_image_dev<D>(a, cen, nv, B.h, org, home, intp, coef, hgdis)) {
    report_error(a.err, g, 1183);
    for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
  }
  for (int q = 0; q < 8; ++q) {
    int an = intp[q];
    double w = coef[q];
    if (D == 2 && (q & 1) == 0) { an = -1; w = 0.0; }   // zero-weight z layer
    a.ib_intp[(size_t)j * 8 + q] = an;
    a.ib_coef[(size_t)j * 8 + q] = w;
  }
  a.ib_hgdis[j] = hgdis;
  for (int d = 0; d < 3; ++d) {
    a.ib_nv[(size_t)j * 3 + d] = nv[d];
    a.ib_attach[(size_t)j * 3 + d] = cen[d] - nv[d] * sd;
  }
  a.ib_body[j] = a.pat_body[p];
  a.ib_case[j] = B.case_id;
}

// Image points of the wall-force patches, searched from each patch centre;
// pat_cc holds a finest-level cell near the previous image point.
template <int D>
__global__ void k_move_patch_image(MoveArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npatch) return;
  int home = a.pat_cc[p];
  const BlkInfo &B = a.blk[a.cblk[home]];
  const double *org = a.origin + (size_t)B.case_id * 3;
  const double pc[3] = {a.pat_pc[(size_t)p * 3], a.pat_pc[(size_t)p * 3 + 1],
                        a.pat_pc[(size_t)p * 3 + 2]};
  const double nv[3] = {a.pat_nv[(size_t)p * 3], a.pat_nv[(size_t)p * 3 + 1],
                        a.pat_nv[(size_t)p * 3 + 2]};
  int intp[8];
  double coef[8], hgdis = 0.0;
  if (!search_image_dev<D>(a, pc, nv, B.h, org, home, intp, coef, hgdis)) {
    report_error(a.err, a.pat_cc[p], 1575);
    for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
  } else {
    a.pat_cc[p] = home;
  }
  for (int q = 0; q < 8; ++q) {
    int an = intp[q];
    double w = coef[q];
    if (D == 2 && (q & 1) == 0) { an = -1; w = 0.0; }
    a.pat_intp[(size_t)p * 8 + q] = an;
    a.pat_coef[(size_t)p * 8 + q] = w;
  }
  a.pat_hgdis[p] = hgdis;
}

// Stage masks of the new solid / IB sets, as at upload; with `upd` set,
// the stage updates of this step (levels >= upd_lmin) are accumulated.
__global__ void __launch_bounds__(256) k_move_flags(MoveArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cnt = 0;
  if (i < a.Nn) {
    const double sd = a.signdis[i];
    const bool ib = a.isib[i] != 0;
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
    cnt = (f & F_UPDATE) && a.blk[a.cblk[i]].level >= a.upd_lmin;
  }
  if (a.upd) {
    const int n = __syncthreads_count(cnt);
    if (threadIdx.x == 0 && n) atomicAdd(a.upd, (unsigned long long)n * a.nstage);
  }
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

// Prescribed pitch and heave of the bodies of one case about a pivot:
// angle of attack aoa(t) = aoa0 + amp sin(omega t) (nose up), pivot
// y(t) = y0 + heave sin(omega t).  The rotation applied to the zero-pose
// offsets is a = -aoa (the section rotates clockwise for a nose-up angle).
struct PitchHeave {
  double x0, y0, aoa0, amp, omega, heave;   // angles in radians
};

// Pose, reference point and wall motion of every case at its time t
// (the rigid-body velocity of Body::Surface_Point_Velocity, Body.H:709-721:
// translation plus rotation about the body centre, here the pivot).
__device__ __forceinline__ Motion law_motion(const PitchHeave &Lw, double t, double zc) {
  const double wt = Lw.omega * t;
  const double sw = sin(wt), cw = cos(wt);
  Motion M;
  M.tra[0] = 0.0; M.tra[1] = Lw.heave * Lw.omega * cw; M.tra[2] = 0.0;
  M.rot[0] = 0.0; M.rot[1] = 0.0; M.rot[2] = -Lw.amp * Lw.omega * cw;
  M.center[0] = Lw.x0; M.center[1] = Lw.y0 + Lw.heave * sw; M.center[2] = zc;
  return M;
}

// mot: per body (the IB fill); motc: per case (the wall force); case_center
// z is kept.
__global__ void k_motion_pose(int ncases, const PitchHeave *law, const double *t, Pose *pose,
                              int nbody, const int *body_case, Motion *mot, Motion *motc,
                              const double *case_center) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nbody) {
    const int cs = body_case[k];
    mot[k] = law_motion(law[cs], t[cs], case_center[(size_t)cs * 3 + 2]);
  }
  if (k < ncases) {
    const PitchHeave Lw = law[k];
    const double sw = sin(Lw.omega * t[k]);
    const double a = -(Lw.aoa0 + Lw.amp * sw);
    Pose P;
    P.px = Lw.x0;
    P.py = Lw.y0 + Lw.heave * sw;
    P.c = cos(a);
    P.s = sin(a);
    pose[k] = P;
    motc[k] = law_motion(Lw, t[k], case_center[(size_t)k * 3 + 2]);
  }
}

// Reference point of each case after the bodies moved: the pivot.  Runs
// after k_motion_pose (one launch per step, ordered on the stream).
__global__ void k_motion_center(int ncases, const Pose *pose, double *case_center) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ncases) {
    case_center[(size_t)k * 3] = pose[k].px;
    case_center[(size_t)k * 3 + 1] = pose[k].py;
  }
}

// t += dt of the step just marched.
__global__ void k_motion_advance(int ncases, double *t, const double *dt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ncases) t[k] += dt[k];
}

}  // namespace ibg
