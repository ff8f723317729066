/*
 * retag.c — CPU restatement of the per-step immersed-boundary update of
 * moving bodies between regrids (Body::MovingBody_Image, Bodymove.cpp:37-125).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  It follows the reference's own
 * incremental sequence, not the from-scratch rule of the GPU kernels:
 *   ComptNewLocation / ComptPatchParams      Bodymove.cpp:3-35
 *   RenewLevelDistance (walk, cycle, solid)  Bodymove.cpp:127-258
 *   RemoveIBCells                            Bodymove.cpp:345-568
 *   InfectNewIBCells                         Bodymove.cpp:570-767
 *   FindImageCell (+ weights)                Body.cpp:999-1190, Mesh.cpp:1403-1442
 *   FindPatchHangingNode                     Body.cpp:1506-1575
 * so that agreement with the GPU (which rebuilds the IB set with the initial
 * rule of Body.cpp:1016-1082 every step) also checks that the two rules
 * give the same set.  2D sections only (the patches of a body form one
 * polyline).  Cell lookups by key use a sorted key table of the finest
 * level, as the host mesh builder does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define CRITIC_DIS 9.5   /* Bodymove.cpp:132 */

typedef struct {
  const ibg_mesh_soa *m;
  const orc_surface *s;
  int64_t *keys;     /* sorted packed keys of one finest block */
  int32_t *kcell;
  int32_t nk;
} ctx_t;

static int64_t pack(int ix, int iy, int iz) {
  return (((int64_t)(ix + (1 << 20))) << 42) | (((int64_t)(iy + (1 << 20))) << 21) |
         (int64_t)(iz + (1 << 20));
}

static int blk_of(const ibg_mesh_soa *m, int g) {
  int lo = 0, hi = m->ncases * m->nlevels;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (m->blk_start[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

static void center(const ctx_t *C, int g, double *x) {
  const ibg_mesh_soa *m = C->m;
  const int b = blk_of(m, g), cs = b / m->nlevels;
  const double *o = C->s->case_origin + (size_t)cs * 3;
  for (int d = 0; d < 3; ++d)
    x[d] = o[d] + ((double)m->key[(size_t)g * 3 + d] + 0.5) * m->blk_h[(size_t)b * 3 + d];
}

/* PatchDistoBox Body.H:563-580 */
static void patch_dis(const double *pat_pc, const double *x, int p, int *best, double *dist) {
  const double *pc = pat_pc + (size_t)p * 3;
  double dx = x[0] - pc[0], dy = x[1] - pc[1], dz = x[2] - pc[2];
  double d = sqrt(dx * dx + dy * dy + dz * dz);
  if (d < *dist) { *dist = d; *best = p; }
}

/* FindClosePatchtoCell Body.H:505-542; patch neighbours of a section:
   previous and next segment of the same body */
static void find_close(const orc_surface *s, const double *pat_pc, const double *pat_nv,
                       const double *x, int *patch, double *dist, double *sd) {
  const int b = s->pat_body[*patch];
  const int lo = s->body_p0[b], hi = s->body_p0[b + 1];
  *dist = 999.0;
  patch_dis(pat_pc, x, *patch, patch, dist);
  int oldpatch = -1, ps0 = *patch;
  while (oldpatch != ps0) {
    oldpatch = *patch;
    int nbr[2], nn = 0;
    if (ps0 - 1 >= lo) nbr[nn++] = ps0 - 1;
    if (ps0 + 1 < hi) nbr[nn++] = ps0 + 1;
    for (int k = 0; k < nn; ++k) patch_dis(pat_pc, x, nbr[k], patch, dist);
    ps0 = *patch;
  }
  const double *pc = pat_pc + (size_t)*patch * 3, *nv = pat_nv + (size_t)*patch * 3;
  *sd = nv[0] * (x[0] - pc[0]) + nv[1] * (x[1] - pc[1]) + nv[2] * (x[2] - pc[2]);
}

static int cmp_key(const void *a, const void *b) {
  const int64_t x = ((const int64_t *)a)[0], y = ((const int64_t *)b)[0];
  return (x > y) - (x < y);
}

static int find_cell(const ctx_t *C, int ix, int iy, int iz) {
  const int64_t k = pack(ix, iy, iz);
  int lo = 0, hi = C->nk - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (C->keys[mid] == k) return C->kcell[mid];
    if (C->keys[mid] < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static void build_finder(ctx_t *C, int b) {
  const ibg_mesh_soa *m = C->m;
  const int s = m->blk_start[b], e = m->blk_start[b + 1];
  C->nk = e - s;
  int64_t *pairs = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(C->nk > 0 ? C->nk : 1));
  for (int g = s; g < e; ++g) {
    pairs[2 * (g - s)] = pack(m->key[(size_t)g * 3], m->key[(size_t)g * 3 + 1], m->key[(size_t)g * 3 + 2]);
    pairs[2 * (g - s) + 1] = g;
  }
  qsort(pairs, (size_t)C->nk, sizeof(int64_t) * 2, cmp_key);
  free(C->keys); free(C->kcell);
  C->keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)(C->nk > 0 ? C->nk : 1));
  C->kcell = (int32_t *)malloc(sizeof(int32_t) * (size_t)(C->nk > 0 ? C->nk : 1));
  for (int i = 0; i < C->nk; ++i) { C->keys[i] = pairs[2 * i]; C->kcell[i] = (int32_t)pairs[2 * i + 1]; }
  free(pairs);
}

/* image-point march of Body.cpp:1044-1183 / 1527-1575 with the weights of
   Mesh::ComptHgIntpCoef Mesh.cpp:1403-1442 */
static int search(const ctx_t *C, const double *h, const double *org, const double *start,
                  const double *nv, const double *sdis, const int8_t *ib, int32_t *intp,
                  double *coef, double *hgdis_out, int *close) {
  const ibg_mesh_soa *m = C->m;
  double dk[3];
  for (int d = 0; d < 3; ++d) dk[d] = nv[d] / h[d];
  double dmax = fabs(dk[0]) < fabs(dk[1]) ? fabs(dk[1]) : fabs(dk[0]);
  dmax = dmax < fabs(dk[2]) ? fabs(dk[2]) : dmax;
  double dxyz[3] = {nv[0] / dmax, nv[1] / dmax, nv[2] / dmax};
  double steplen = sqrt(dxyz[0] * dxyz[0] + dxyz[1] * dxyz[1] + dxyz[2] * dxyz[2]);
  const double ca[3] = {0.0, 1.0, 0.0}, cb[3] = {1.0, -1.0, 1.0};
  double hgdis = 0.0;
  for (int it = 0; it < 64; ++it) {
    hgdis += steplen;
    double pt[3];
    int ik[3];
    for (int d = 0; d < 3; ++d) pt[d] = start[d] + nv[d] * hgdis;
    for (int d = 0; d < 3; ++d) ik[d] = (int)floor((pt[d] - org[d]) / h[d]);
    int cc = find_cell(C, ik[0], ik[1], ik[2]);
    if (cc < 0 || m->kind[cc] == IBG_DMGHOST) return 0;
    double cen[3], dkt[3];
    center(C, cc, cen);
    for (int d = 0; d < 3; ++d) dkt[d] = (pt[d] - cen[d]) / h[d];
    if (m->dim == 2) dkt[2] = 0.0;
    int hgdir[3];
    for (int d = 0; d < 3; ++d) hgdir[d] = dkt[d] > 0.0 ? 1 : 0;
    int ok = 1;
    for (int pi = 0; pi < 2; ++pi)
      for (int pj = 0; pj < 2; ++pj)
        for (int pk = 0; pk < 2; ++pk) {
          int nx = pi + hgdir[0], ny = pj + hgdir[1], nz = pk + hgdir[2];
          int a = m->nb[(size_t)cc * 27 + nx * 9 + ny * 3 + nz];
          int q = pi * 4 + pj * 2 + pk;
          intp[q] = a;
          coef[q] = 0.0;
          if (a < 0) { ok = 0; continue; }
          if (sdis[a] < 0.0 || ib[a]) ok = 0;
          coef[q] = (ca[nx] + cb[nx] * fabs(dkt[0])) * (ca[ny] + cb[ny] * fabs(dkt[1])) *
                    (ca[nz] + cb[nz] * fabs(dkt[2]));
        }
    if (ok) { *hgdis_out = hgdis; *close = cc; return 1; }
  }
  return 0;
}

static void set_err(orc_retag_out *o, int32_t code, int64_t cell) {
  if (o->err_code == 0) { o->err_code = code; o->err_cell = cell; }
}

/* in-plane 3x3 (2D, Point_iterator_2d) or 27 (3D) neighbour slots */
static int nslots(int dim, int *q) {
  int n = 0;
  for (int i = 0; i < 27; ++i)
    if (dim == 3 || i % 3 == 1) q[n++] = i;
  return n;
}

int32_t orc_move_bodies(const ibg_mesh_soa *m, const orc_surface *s, const double *pose,
                        orc_retag_out *o) {
  const int N = m->ncells, L = m->nlevels, np = s->npatch;
  ctx_t C = {m, s, NULL, NULL, 0};
  int Q[27];
  const int nq = nslots(m->dim, Q);
  o->err_code = 0; o->err_cell = -1; o->nib = 0;
  /* ComptNewLocation / ComptPatchParams: points at the pose, then the
     patch centre, normal and area of each segment */
  for (int p = 0; p < np; ++p) {
    const double *P = pose + (size_t)s->pat_case[p] * 4;
    const int i0 = s->pat_pt0[p];
    const double ox0 = s->pt_off[2 * i0], oy0 = s->pt_off[2 * i0 + 1];
    const double ox1 = s->pt_off[2 * i0 + 2], oy1 = s->pt_off[2 * i0 + 3];
    const double x0 = P[0] + (P[2] * ox0 - P[3] * oy0), y0 = P[1] + (P[3] * ox0 + P[2] * oy0);
    const double x1 = P[0] + (P[2] * ox1 - P[3] * oy1), y1 = P[1] + (P[3] * ox1 + P[2] * oy1);
    const double dx = x1 - x0, dy = y1 - y0;
    const double len = sqrt(dx * dx + dy * dy);
    double nx = -dy / len, ny = dx / len;
    if (s->pat_flip[p]) { nx = -nx; ny = -ny; }
    o->pat_pc[(size_t)p * 3] = 0.5 * (x0 + x1);
    o->pat_pc[(size_t)p * 3 + 1] = 0.5 * (y0 + y1);
    o->pat_nv[(size_t)p * 3] = nx;
    o->pat_nv[(size_t)p * 3 + 1] = ny;
    o->pat_area[p] = len * s->pat_dz[p];
  }
  /* RenewLevelDistance, per level: walk (:136-155) */
  int8_t *solid_old = (int8_t *)malloc((size_t)N);
  double *dist = (double *)malloc(sizeof(double) * (size_t)N);
  for (int g = 0; g < N; ++g) {
    solid_old[g] = o->signdis[g] < 0.0;
    dist[g] = 999.0;
  }
  for (int g = 0; g < N; ++g) {
    if (o->cell_patch[g] < 0 || !(o->signdis[g] < CRITIC_DIS)) continue;
    double x[3];
    center(&C, g, x);
    int p = o->cell_patch[g];
    find_close(s, o->pat_pc, o->pat_nv, x, &p, &dist[g], &o->signdis[g]);
    o->cell_patch[g] = p;
  }
  /* cycle list (:156-188), then processed in order (:189-223); the list is
     collected before any entry is processed */
  {
    int64_t cap = 1024, n = 0;
    int32_t *cb = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)cap);
    for (int g = 0; g < N; ++g) {
      if (o->cell_patch[g] < 0 || !(o->signdis[g] < CRITIC_DIS)) continue;
      const int myb = s->pat_body[o->cell_patch[g]];
      for (int k = 0; k < nq; ++k) {
        const int an = m->nb[(size_t)g * 27 + Q[k]];
        if (an < 0 || o->cell_patch[an] < 0) continue;
        const int nb = s->pat_body[o->cell_patch[an]];
        if (nb == myb) continue;
        if (n == cap) { cap *= 2; cb = (int32_t *)realloc(cb, sizeof(int32_t) * 3 * (size_t)cap); }
        cb[3 * n] = g; cb[3 * n + 1] = nb; cb[3 * n + 2] = o->cell_patch[an];
        ++n;
      }
    }
    for (int64_t e = 0; e < n; ++e) {
      const int g = cb[3 * e];
      if (s->pat_body[o->cell_patch[g]] == cb[3 * e + 1]) continue;
      double x[3], dn, sn;
      center(&C, g, x);
      int pn = cb[3 * e + 2];
      find_close(s, o->pat_pc, o->pat_nv, x, &pn, &dn, &sn);
      const double sd = o->signdis[g];
      int change = 0;
      if (sn > 0.0 && sd < 0.0) change = 1;
      else if (dn < dist[g]) {
        if (sd * sn > 0.0) change = 1;
        else if (sn > 0.0) change = 1;
      }
      if (change) { o->cell_patch[g] = pn; dist[g] = dn; o->signdis[g] = sn; }
    }
    free(cb);
  }
  /* solid loop (:224-249): signdis = min(signdis, distance_to_dm) */
  for (int g = 0; g < N; ++g) {
    if (!m->dist_dm || o->cell_patch[g] < 0 || !(o->signdis[g] < CRITIC_DIS)) continue;
    const double dm = m->dist_dm[g];
    if (dm < o->signdis[g]) o->signdis[g] = dm;
  }
  int8_t *solid = (int8_t *)malloc((size_t)N);
  for (int g = 0; g < N; ++g) solid[g] = o->signdis[g] < 0.0;
  /* RemoveIBCells / InfectNewIBCells on the finest level of each case */
  for (int cs = 0; cs < m->ncases; ++cs) {
    const int b = cs * L + L - 1;
    const int bs = m->blk_start[b], be = m->blk_start[b + 1];
    int8_t *ib = o->ib;
    int8_t *mark = (int8_t *)calloc((size_t)N, 1);
    /* RemoveIBCells: IB cells near in->out cells that lost every solid cell
       of their two-ring (:349-488) */
    for (int g = bs; g < be; ++g) {
      if (m->kind[g] == IBG_DMGHOST) continue;
      if (!(solid_old[g] && !solid[g])) continue;           /* in -> out */
      for (int k0 = 0; k0 < nq; ++k0) {
        const int a0 = m->nb[(size_t)g * 27 + Q[k0]];
        if (a0 < 0 || solid[a0] || m->kind[a0] == IBG_DMGHOST) continue;
        for (int k1 = 0; k1 < nq; ++k1) {
          const int a1 = m->nb[(size_t)a0 * 27 + Q[k1]];
          if (a1 < 0 || !ib[a1] || m->kind[a1] == IBG_DMGHOST) continue;
          int keep = 0;
          for (int k2 = 0; k2 < nq && !keep; ++k2) {
            const int a2 = m->nb[(size_t)a1 * 27 + Q[k2]];
            if (a2 < 0 || m->kind[a2] == IBG_DMGHOST) continue;
            if (solid[a2]) { keep = 1; break; }
            for (int k3 = 0; k3 < nq; ++k3) {
              const int a3 = m->nb[(size_t)a2 * 27 + Q[k3]];
              if (a3 >= 0 && m->kind[a3] != IBG_DMGHOST && solid[a3]) { keep = 1; break; }
            }
          }
          if (!keep) mark[a1] = 1;
        }
      }
    }
    for (int g = bs; g < be; ++g) if (mark[g]) ib[g] = 0;
    /* out -> in cells were IB cells (abort 263, :494-544) and leave the set */
    for (int g = bs; g < be; ++g) {
      if (m->kind[g] == IBG_DMGHOST) continue;
      if (!(!solid_old[g] && solid[g])) continue;
      if (!ib[g]) set_err(o, 263, g);
      ib[g] = 0;
    }
    /* InfectNewIBCells: in -> out cells next to a solid cell (:575-650) */
    for (int g = bs; g < be; ++g) {
      if (m->kind[g] == IBG_DMGHOST) continue;
      if (!(solid_old[g] && !solid[g])) continue;
      if (ib[g]) set_err(o, 289, g);
      int inf = 0;
      for (int k = 0; k < nq && !inf; ++k) {
        const int a = m->nb[(size_t)g * 27 + Q[k]];
        if (a >= 0 && m->kind[a] != IBG_DMGHOST && o->signdis[a] < 0.0) inf = 1;
      }
      if (inf) ib[g] = 1; else set_err(o, 314, g);
    }
    /* the non-solid neighbours of the non-solid neighbours of out -> in
       cells (:651-698) */
    for (int g = bs; g < be; ++g) {
      if (m->kind[g] == IBG_DMGHOST) continue;
      if (!(!solid_old[g] && solid[g])) continue;
      for (int k1 = 0; k1 < nq; ++k1) {
        const int a1 = m->nb[(size_t)g * 27 + Q[k1]];
        if (a1 < 0 || m->kind[a1] == IBG_DMGHOST || !(o->signdis[a1] >= 0.0)) continue;
        for (int k2 = 0; k2 < nq; ++k2) {
          const int a2 = m->nb[(size_t)a1 * 27 + Q[k2]];
          if (a2 >= 0 && m->kind[a2] != IBG_DMGHOST && o->signdis[a2] >= 0.0) ib[a2] = 1;
        }
      }
    }
    free(mark);
  }
  /* IB list in cell order; FindImageCell for every IB cell */
  for (int g = 0; g < N; ++g) {
    if (!o->ib[g]) continue;
    if (m->kind[g] != IBG_NORMAL) set_err(o, 1082, g);
    if (o->nib < o->ib_cap) o->ib_cell[o->nib] = g;
    o->nib++;
  }
  if (o->nib > o->ib_cap) { set_err(o, 1090, -1); o->nib = o->ib_cap; }
  int fb = -1;
  for (int j = 0; j < o->nib; ++j) {
    const int g = o->ib_cell[j];
    const int b = blk_of(m, g), cs = b / L;
    if (b != fb) { build_finder(&C, b); fb = b; }
    const int p = o->cell_patch[g];
    const double *nv = o->pat_nv + (size_t)p * 3;
    const double sd = o->signdis[g];
    double cen[3], hgdis = 0.0;
    center(&C, g, cen);
    int32_t intp[8];
    double coef[8];
    int cc = -1;
    if (!search(&C, m->blk_h + (size_t)b * 3, s->case_origin + (size_t)cs * 3, cen, nv,
                o->signdis, o->ib, intp, coef, &hgdis, &cc)) {
      set_err(o, 1183, g);
      for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
    }
    for (int q = 0; q < 8; ++q) {
      o->ib_intp[(size_t)j * 8 + q] = intp[q];
      o->ib_coef[(size_t)j * 8 + q] = coef[q];
    }
    o->ib_hgdis[j] = hgdis;
    for (int d = 0; d < 3; ++d) {
      o->ib_nv[(size_t)j * 3 + d] = nv[d];
      o->ib_attach[(size_t)j * 3 + d] = cen[d] - nv[d] * sd;
    }
    o->ib_body[j] = s->pat_body[p];
  }
  /* FindPatchHangingNode for the wall-force patches */
  for (int p = 0; p < np; ++p) {
    const int cs = s->pat_case[p], b = cs * L + L - 1;
    if (b != fb) { build_finder(&C, b); fb = b; }
    int32_t intp[8];
    double coef[8], hgdis = 0.0;
    int cc = -1;
    if (!search(&C, m->blk_h + (size_t)b * 3, s->case_origin + (size_t)cs * 3,
                o->pat_pc + (size_t)p * 3, o->pat_nv + (size_t)p * 3, o->signdis, o->ib, intp,
                coef, &hgdis, &cc)) {
      set_err(o, 1575, p);
      for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
    }
    for (int q = 0; q < 8; ++q) {
      o->pat_intp[(size_t)p * 8 + q] = intp[q];
      o->pat_coef[(size_t)p * 8 + q] = coef[q];
    }
    o->pat_hgdis[p] = hgdis;
    o->pat_close[p] = cc;
  }
  free(C.keys); free(C.kcell);
  free(solid_old); free(solid); free(dist);
  return o->err_code;
}
