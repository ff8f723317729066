// ib_mesh.cpp — host AMR + immersed-boundary mesh builder.
//
// Produces the flattened level set of ibgpu.h (ibg_mesh_soa) for one or more
// airfoil cases.  The rules are those of the reference's initial refinement
// and IB setup; each step cites the reference code whose behaviour it
// restates.  The reference builds the same sets incrementally across MPI
// ranks (AMR.cpp:6-209); here every case is built at once on the host, and
// the raw index order is this builder's own (canonical keys are what match).
#include "ib_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <climits>
#include <memory>
#include <parallel/algorithm>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

std::string g_last_error;

enum { NORMAL = IBG_NORMAL, DMGHOST = IBG_DMGHOST, RES1 = IBG_RES1,
       RES2 = IBG_RES2, PRO = IBG_PRO };

// An allocator that leaves new elements uninitialised: for the per-cell
// arrays that the flattening loop writes in full (in parallel, so the pages
// are first touched by the threads that fill them).
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U> struct rebind { using other = NoInitAlloc<U>; };
  NoInitAlloc() = default;
  template <class U> NoInitAlloc(const NoInitAlloc<U> &) {}
  template <class U> void construct(U *p) noexcept { ::new ((void *)p) U; }
  template <class U, class... A> void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
template <class T> using fvec = std::vector<T, NoInitAlloc<T>>;

inline int64_t pack_key(int ix, int iy, int iz) {
  // 21 bits per coordinate, offset so that ghosts (>= -4) stay positive
  return ((int64_t)(ix + 8) << 42) | ((int64_t)(iy + 8) << 21) | (int64_t)iz;
}

// Index of the cells of one level by integer key: a dense table over the
// bounding box of the keys (-1 where no cell is).  A refined level covers a
// band around the bodies, so its box holds a small multiple of its cells;
// a lookup is one load where a hash map costs a probe chain.
struct KeyGrid {
  int x0 = 0, y0 = 0, z0 = 0, wx = 0, wy = 0, wz = 0;
  std::vector<int32_t> idx;
  template <class KeyOf>
  void build(int n, KeyOf key) {
    idx.clear();
    wx = wy = wz = 0;
    if (n <= 0) return;
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (int i = 0; i < n; ++i) {
      int k[3];
      key(i, k);
      for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], k[d]); hi[d] = std::max(hi[d], k[d]); }
    }
    x0 = lo[0]; y0 = lo[1]; z0 = lo[2];
    wx = hi[0] - lo[0] + 1; wy = hi[1] - lo[1] + 1; wz = hi[2] - lo[2] + 1;
    idx.assign((size_t)wx * wy * wz, -1);
    for (int i = 0; i < n; ++i) {
      int k[3];
      key(i, k);
      idx[((size_t)(k[2] - z0) * wy + (k[1] - y0)) * wx + (k[0] - x0)] = i;
    }
  }
  int find(int ix, int iy, int iz) const {
    const unsigned x = (unsigned)(ix - x0), y = (unsigned)(iy - y0), z = (unsigned)(iz - z0);
    if (x >= (unsigned)wx || y >= (unsigned)wy || z >= (unsigned)wz) return -1;
    return idx[((size_t)z * wy + y) * wx + x];
  }
};

inline uint64_t spread2(uint32_t v) {
  uint64_t x = v & 0xffffffffu;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
inline uint64_t morton(int dim, int ix, int iy, int iz) {
  uint32_t x = (uint32_t)(ix + 8), y = (uint32_t)(iy + 8), z = (uint32_t)iz;
  if (dim == 2) return spread2(x) | (spread2(y) << 1);
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

struct LCell {
  int ix, iy, iz;
  int8_t kind;
};

struct Level {
  std::vector<LCell> cells;
  KeyGrid map;
  int nzl = 1;          // z cells on this level (periodic)
  double h[3];
  bool twod = true;     // refined from the coarser level in x,y only
  void rebuild_map() {
    map.build((int)cells.size(), [&](int i, int *k) {
      k[0] = cells[i].ix; k[1] = cells[i].iy; k[2] = cells[i].iz;
    });
  }
  int find(int ix, int iy, int iz) const {
    int zz = ((iz % nzl) + nzl) % nzl;   // z is periodic (Userfunction_cy_naca15.cpp:176)
    return map.find(ix, iy, zz);
  }
};

struct Patch {
  double pc[3], nv[3], area;
  int body;
};

// CtoFPair with level-local indices, resolved to global after sorting
struct LPair {
  int ci;            // level-local coarse cell (level l)
  int fi[8];         // level-local sons (level l+1), -1 where absent
  int layer;         // 0: f_res_ghost[0], 1: f_res_ghost[1], 2: f_pro_ghost
};

struct CaseMesh {
  ibm_case cs;
  std::vector<Level> lv;
  std::vector<std::vector<LPair>> pairs;   // per coarse level
  std::vector<Patch> patches;
  std::vector<std::vector<double>> segxy;  // per body: points
  std::vector<int> seg_body, seg_first;    // segment -> body, point index
  std::vector<double> seg_cx, seg_cy, seg_nx, seg_ny, seg_len;
  // runs of kSegRun consecutive segment centers with their bounding boxes
  std::vector<double> run_x0, run_x1, run_y0, run_y1;
};

constexpr int kSegRun = 8;

void build_segment_runs(CaseMesh &m) {
  const int ns = (int)m.seg_cx.size();
  m.run_x0.clear(); m.run_x1.clear(); m.run_y0.clear(); m.run_y1.clear();
  for (int s0 = 0; s0 < ns; s0 += kSegRun) {
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (int s = s0; s < std::min(ns, s0 + kSegRun); ++s) {
      x0 = std::min(x0, m.seg_cx[s]); x1 = std::max(x1, m.seg_cx[s]);
      y0 = std::min(y0, m.seg_cy[s]); y1 = std::max(y1, m.seg_cy[s]);
    }
    m.run_x0.push_back(x0); m.run_x1.push_back(x1);
    m.run_y0.push_back(y0); m.run_y1.push_back(y1);
  }
}

void build_segments(CaseMesh &m) {
  const ibm_case &c = m.cs;
  int off = 0;
  for (int b = 0; b < c.nbody; ++b) {
    int n = c.body_npts[b];
    for (int i = 0; i + 1 < n; ++i) {
      double x0 = c.body_xy[2 * (off + i)], y0 = c.body_xy[2 * (off + i) + 1];
      double x1 = c.body_xy[2 * (off + i + 1)], y1 = c.body_xy[2 * (off + i + 1) + 1];
      double dx = x1 - x0, dy = y1 - y0;
      double len = std::sqrt(dx * dx + dy * dy);
      if (len <= 0) continue;
      // Outward normal: the upper surface (even body index) is given LE->TE
      // with the fluid on its left; the lower surface (odd) on its right.
      double nx = -dy / len, ny = dx / len;
      if (b % 2 == 1) { nx = -nx; ny = -ny; }
      m.seg_cx.push_back(0.5 * (x0 + x1));
      m.seg_cy.push_back(0.5 * (y0 + y1));
      m.seg_nx.push_back(nx);
      m.seg_ny.push_back(ny);
      m.seg_len.push_back(len);
      m.seg_body.push_back(b);
      m.seg_first.push_back(off + i);
    }
    off += n;
  }
}

// Nearest segment center to (x,y): the smallest dx^2 + dy^2, the lowest
// index among equal ones (what a scan with a strict < in index order keeps).
// The runs' boxes bound the distance of their centers from below (the same
// rounded differences, monotone), so a run whose bound exceeds the best
// distance so far cannot hold the answer and is skipped; the run with the
// smallest bound is scanned first.
int nearest_segment(const CaseMesh &m, double x, double y) {
  const int ns = (int)m.seg_cx.size();
  const int nr = (int)m.run_x0.size();
  if (ns == 0) return -1;
  const double *cx = m.seg_cx.data(), *cy = m.seg_cy.data();
  int best = -1;
  double bd = 1e300;
  auto scan = [&](int r) {
    const int s1 = std::min(ns, (r + 1) * kSegRun);
    for (int s = r * kSegRun; s < s1; ++s) {
      const double dx = x - cx[s], dy = y - cy[s];
      const double d = dx * dx + dy * dy;
      if (d < bd || (d == bd && s < best)) { bd = d; best = s; }
    }
  };
  double lbs[4096];
  double *lb = nr <= 4096 ? lbs : new double[nr];
  int r0 = 0;
  for (int r = 0; r < nr; ++r) {
    const double ex = std::max(0.0, std::max(m.run_x0[r] - x, x - m.run_x1[r]));
    const double ey = std::max(0.0, std::max(m.run_y0[r] - y, y - m.run_y1[r]));
    lb[r] = ex * ex + ey * ey;
    if (lb[r] < lb[r0]) r0 = r;
  }
  scan(r0);
  for (int r = 0; r < nr; ++r)
    if (r != r0 && lb[r] <= bd) scan(r);
  if (lb != lbs) delete[] lb;
  return best;
}

void cell_center(const CaseMesh &m, const Level &L, const LCell &c, double *xyz) {
  // ComptCellParams MeshOp.H:4-9: boxcenter = (i+0.5)*dh
  xyz[0] = m.cs.origin[0] + ((double)c.ix + 0.5) * L.h[0];
  xyz[1] = m.cs.origin[1] + ((double)c.iy + 0.5) * L.h[1];
  xyz[2] = m.cs.origin[2] + ((double)c.iz + 0.5) * L.h[2];
}

// Closest patch and signed distance: BoxtoWall of the closest patch center,
// signdis = nv . (boxcenter - pc) (Body.cpp:791-821).
void closest_patch(const CaseMesh &m, const double *xyz, int *patch, double *signdis) {
  int s = nearest_segment(m, xyz[0], xyz[1]);
  if (s < 0) { *patch = -1; *signdis = 9999.0; return; }
  int sp = std::max(m.cs.span_patches, 1);
  double dz = m.cs.span / sp;
  int zl = (int)std::floor((xyz[2] - m.cs.origin[2]) / dz);
  zl = std::min(std::max(zl, 0), sp - 1);
  int p = s * sp + zl;
  const Patch &P = m.patches[p];
  *patch = p;
  *signdis = P.nv[0] * (xyz[0] - P.pc[0]) + P.nv[1] * (xyz[1] - P.pc[1]) +
             P.nv[2] * (xyz[2] - P.pc[2]);
}

void build_patches(CaseMesh &m) {
  int sp = std::max(m.cs.span_patches, 1);
  double dz = m.cs.span / sp;
  for (size_t s = 0; s < m.seg_cx.size(); ++s) {
    for (int k = 0; k < sp; ++k) {
      Patch P;
      P.pc[0] = m.seg_cx[s]; P.pc[1] = m.seg_cy[s];
      P.pc[2] = m.cs.origin[2] + (k + 0.5) * dz;
      P.nv[0] = m.seg_nx[s]; P.nv[1] = m.seg_ny[s]; P.nv[2] = 0.0;
      P.area = m.seg_len[s] * dz;
      P.body = m.seg_body[s];
      m.patches.push_back(P);
    }
  }
}

// ---------------------------------------------------------------- levels
void build_levels(CaseMesh &m) {
  const ibm_case &c = m.cs;
  const int L = c.nlevels;
  const bool d3 = (c.dim == 3);
  m.lv.resize(L);
  m.pairs.resize(L);
  for (int l = 0; l < L; ++l) {
    Level &lv = m.lv[l];
    lv.twod = (l == 0) ? false : (!d3 || l >= c.twod_from);
    double rz = 1.0;
    if (l > 0) { rz = m.lv[l - 1].h[2] / (lv.twod ? 1.0 : 2.0); }
    // dh[i] (mpiparams.cpp:134-160): x,y halve per level, z by power ratio
    lv.h[0] = c.dh0[0] / std::pow(2.0, l);
    lv.h[1] = c.dh0[1] / std::pow(2.0, l);
    lv.h[2] = (l == 0) ? c.dh0[2] : rz;
    lv.nzl = (l == 0) ? c.nz : m.lv[l - 1].nzl * (lv.twod ? 1 : 2);
  }
  // Level 0: background grid plus two layers (ighost, Constants.H:7) of
  // domain ghosts around x and y; z is periodic (no z ghosts).
  Level &l0 = m.lv[0];
  for (int k = 0; k < c.nz; ++k)
    for (int j = -2; j < c.ny + 2; ++j)
      for (int i = -2; i < c.nx + 2; ++i) {
        bool in = (i >= 0 && i < c.nx && j >= 0 && j < c.ny);
        l0.cells.push_back({i, j, k, (int8_t)(in ? NORMAL : DMGHOST)});
      }
  l0.rebuild_map();

  for (int l = 0; l + 1 < L; ++l) {
    Level &lc = m.lv[l];
    Level &lf = m.lv[l + 1];
    const int n = (int)lc.cells.size();
    std::vector<double> sd(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double xyz[3]; int p;
      cell_center(m, lc, lc.cells[i], xyz);
      closest_patch(m, xyz, &p, &sd[i]);
    }
    // tag = 0: band rule of TagtheBox (Userfunction_cy_naca15.cpp:589-600)
    std::vector<int8_t> tag(n, -1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      if (lc.cells[i].kind != NORMAL) continue;
      if (sd[i] > c.band_lo[l] && sd[i] < c.band_hi[l]) tag[i] = 0;
    }
    // Proper nesting: a tagged cell keeps its tag only when every cell within
    // two cells of it (x,y) is a normal cell of this level, so that the new
    // prolongation ring and its gradient stencil lie on normal cells.
    std::vector<int8_t> keep(n, 0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      if (tag[i] != 0) continue;
      const LCell &cc = lc.cells[i];
      bool ok = true;
      for (int dj = -2; dj <= 2 && ok; ++dj)
        for (int di = -2; di <= 2 && ok; ++di) {
          int a = lc.find(cc.ix + di, cc.iy + dj, cc.iz);
          if (a < 0 || lc.cells[a].kind != NORMAL) ok = false;
        }
      keep[i] = ok;
    }
    for (int i = 0; i < n; ++i) if (tag[i] == 0 && !keep[i]) tag[i] = -1;
    // TagRefineGhost AMRGhost_refine.cpp:30-98: tag 1 for tagged cells with
    // an untagged 26-neighbour (which become prolongation cells, tag -2),
    // tag 2 for tagged cells with an untagged cell two cells away on an axis.
    // Each cell's outcome depends on the tags before the pass only (the
    // pass's own writes never turn a cell into or out of the "untagged"
    // class), so each cell gathers from its neighbours.
    const int zr = d3 ? 1 : 0;
    std::vector<int8_t> t2(tag);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const LCell &cc = lc.cells[i];
      const bool tagged = tag[i] == 0;
      bool hit = false, ring = false;
      for (int dk = -zr; dk <= zr; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
          for (int di = -1; di <= 1; ++di) {
            int a = lc.find(cc.ix + di, cc.iy + dj, cc.iz + dk);
            if (a < 0) continue;
            if (tagged && tag[a] == -1 && lc.cells[a].kind != DMGHOST) hit = true;
            // the neighbour relation is symmetric (offsets and periodic z)
            if (!tagged && tag[a] == 0) ring = true;
          }
      if (tagged && hit) t2[i] = 1;
      if (!tagged && ring) t2[i] = -2;
    }
    tag.swap(t2);
    t2.assign(tag.begin(), tag.end());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      if (tag[i] != 0) continue;
      const LCell &cc = lc.cells[i];
      const int dd[6][3] = {{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},{0,0,-1},{0,0,1}};
      for (int q = 0; q < (d3 ? 6 : 4); ++q) {
        int a = lc.find(cc.ix + dd[q][0], cc.iy + dd[q][1], cc.iz + dd[q][2]);
        if (a < 0) continue;
        int a2 = lc.find(cc.ix + 2 * dd[q][0], cc.iy + 2 * dd[q][1], cc.iz + 2 * dd[q][2]);
        if (a2 < 0) continue;
        if ((tag[a2] == -1 || tag[a2] == -2) && lc.cells[a2].kind != DMGHOST) t2[i] = 2;
      }
    }
    tag.swap(t2);
    // the prolongation ring must be made of normal cells
    for (int i = 0; i < n; ++i)
      if (tag[i] == -2 && lc.cells[i].kind != NORMAL)
        throw std::runtime_error("refinement ring reaches a ghost cell on level " + std::to_string(l));
    // MakeSmallBox AMRLevel.cpp:413: sons of tagged and ring cells
    const int kz = lf.twod ? 1 : 2;
    const int zf = lf.twod ? 1 : 2;
    std::vector<LPair> &pl = m.pairs[l];
    for (int pass = 0; pass < 2; ++pass) {
      for (int i = 0; i < n; ++i) {
        bool tagged = (tag[i] >= 0);
        bool ring = (tag[i] == -2);
        if ((pass == 0 && !tagged) || (pass == 1 && !ring)) continue;
        const LCell &cc = lc.cells[i];
        LPair pr;
        pr.ci = i;
        for (int q = 0; q < 8; ++q) pr.fi[q] = -1;
        for (int si = 0; si < 2; ++si)
          for (int sj = 0; sj < 2; ++sj)
            for (int sk = 0; sk < kz; ++sk) {
              int fx = 2 * cc.ix + si, fy = 2 * cc.iy + sj, fz = cc.iz * zf + sk;
              pr.fi[si * 4 + sj * 2 + sk] = (int)lf.cells.size();
              lf.cells.push_back({fx, fy, fz, (int8_t)(pass == 0 ? NORMAL : PRO)});
            }
        if (pass == 1) { pr.layer = 2; pl.push_back(pr); }
        else if (tag[i] == 1) { pr.layer = 0; pl.push_back(pr); }
        else if (tag[i] == 2) { pr.layer = 1; pl.push_back(pr); }
      }
    }
    lf.rebuild_map();
    // coarse cells: tag 1 / 2 become restriction ghosts, tag 0 are removed
    // (RemoveGhostandNormal_Refine AMRGhostcells.cpp:265-330)
    std::vector<int> newidx(n, -1);
    std::vector<LCell> kept;
    kept.reserve(n);
    for (int i = 0; i < n; ++i) {
      if (tag[i] == 0) continue;
      LCell cc = lc.cells[i];
      if (tag[i] == 1) cc.kind = RES1;
      if (tag[i] == 2) cc.kind = RES2;
      newidx[i] = (int)kept.size();
      kept.push_back(cc);
    }
    lc.cells.swap(kept);
    lc.rebuild_map();
    for (auto &pr : pl) pr.ci = newidx[pr.ci];
    // pairs of the previous coarser level reference this level's sons; a son
    // removed here would break a restriction pair
    if (l > 0) {
      for (auto &pr : m.pairs[l - 1])
        for (int q = 0; q < 8; ++q)
          if (pr.fi[q] >= 0) {
            int nidx = newidx[pr.fi[q]];
            if (nidx < 0) throw std::runtime_error("a restriction son was removed by refinement");
            pr.fi[q] = nidx;
          }
    }
  }
}

}  // namespace

// ----------------------------------------------------------------- handle
struct ibm_handle {
  ibg_mesh_soa soa;
  int ncases = 0, L = 0, dim = 2;
  std::vector<int32_t> blk_start, blk_nnormal, blk_twod;
  std::vector<double> blk_h;
  fvec<int32_t> key, nb, cell_faces;
  std::vector<int32_t> ibidx;
  fvec<int8_t> kind;
  fvec<double> signdis, center;
  std::vector<int32_t> blk_face_start;
  fvec<int32_t> face_cells;
  fvec<int8_t> face_dir;
  std::vector<int32_t> dmg_cell, dmg_ref, dmg_type, dmg_nbr;
  std::vector<double> dmg_nmv, dist_dm;
  bool walls = false;
  std::vector<int32_t> res_ci, res_fi, res_blk, pro_ci, pro_fi, pro_blk;
  std::vector<int32_t> mdy;   // [nmdy][6] flux-balance faces
  std::vector<int32_t> ib_cell, ib_intp, ib_body;
  std::vector<double> ib_coef, ib_hgdis, ib_nv, ib_attach;
  std::vector<int32_t> pat_intp, pat_case, pat_close;
  std::vector<double> pat_coef, pat_hgdis, pat_nv, pat_pc, pat_area;
  std::vector<int32_t> cell_patch;   // BoxtoWall::patch, global patch index
  // surface of the moving-body update (ibm_surface)
  std::vector<int32_t> sf_pt0, sf_body, sf_body_p0;
  std::vector<int8_t> sf_flip;
  std::vector<double> sf_dz, sf_origin;
  int sf_npoint = 0;
  bool sf_ok = true;
  std::vector<int32_t> body_case;
  std::vector<double> body_center;
  std::vector<ibm_case> cases;
};

namespace {

// Image-point search (Body.cpp:1044-1183 for IB cells, Body.cpp:1527-1575 for
// patches): march from `start` along nv in steps of one cell in the dominant
// direction until the 2x2(x2) bracketing cells are neither solid nor IB cells
// (WALLINTPFLAG_NOTINFECTED, Body.H:17), then the multilinear weights of
// Mesh::ComptHgIntpCoef (Mesh.cpp:1403-1442).

struct Finder {
  const Level *lv;
  int gbase;                       // global index of level-local cell 0
  const std::vector<int> *perm;    // level-local -> position in sorted block
  int find(int ix, int iy, int iz) const {
    int a = lv->find(ix, iy, iz);
    return a < 0 ? -1 : gbase + (*perm)[a];
  }
};

bool search_image(const Finder &F, const ibm_handle &H, const double *h,
                  const double *origin, const double *start, const double *nv,
                  const fvec<int8_t> &solid_g, const std::vector<int32_t> &ibidx_g,
                  int32_t *intp, double *coef, double *hgdis_out, int *closecell) {
  double dk[3];
  for (int d = 0; d < 3; ++d) dk[d] = nv[d] / h[d];
  double dmax = std::max(std::fabs(dk[0]), std::fabs(dk[1]));
  dmax = std::max(dmax, std::fabs(dk[2]));
  double dxyz[3] = {nv[0] / dmax, nv[1] / dmax, nv[2] / dmax};
  double steplen = std::sqrt(dxyz[0] * dxyz[0] + dxyz[1] * dxyz[1] + dxyz[2] * dxyz[2]);
  const double ca[3] = {0.0, 1.0, 0.0};
  const double cb[3] = {1.0, -1.0, 1.0};
  double hgdis = 0.0;
  for (int it = 0; it < 64; ++it) {
    hgdis += steplen;
    double pt[3];
    for (int d = 0; d < 3; ++d) pt[d] = start[d] + nv[d] * hgdis;
    int ik[3];
    for (int d = 0; d < 3; ++d) ik[d] = (int)std::floor((pt[d] - origin[d]) / h[d]);
    int cc = F.find(ik[0], ik[1], ik[2]);
    if (cc < 0) return false;
    if (H.kind[cc] == IBG_DMGHOST) return false;
    const double *cen = &H.center[3 * (size_t)cc];
    double dkt[3];
    for (int d = 0; d < 3; ++d) dkt[d] = (pt[d] - cen[d]) / h[d];
    if (H.dim == 2) dkt[2] = 0.0;   // one z layer: the image point is in it
    int hgdir[3];
    for (int d = 0; d < 3; ++d) hgdir[d] = dkt[d] > 0.0 ? 1 : 0;
    bool ok = true;
    for (int pi = 0; pi < 2; ++pi)
      for (int pj = 0; pj < 2; ++pj)
        for (int pk = 0; pk < 2; ++pk) {
          int nx = pi + hgdir[0], ny = pj + hgdir[1], nz = pk + hgdir[2];
          int a = H.nb[(size_t)cc * 27 + nx * 9 + ny * 3 + nz];
          int q = pi * 4 + pj * 2 + pk;
          intp[q] = a;
          if (a < 0) { ok = false; continue; }
          if (solid_g[a] || ibidx_g[a] >= 0) ok = false;
          coef[q] = (ca[nx] + cb[nx] * std::fabs(dkt[0])) *
                    (ca[ny] + cb[ny] * std::fabs(dkt[1])) *
                    (ca[nz] + cb[nz] * std::fabs(dkt[2]));
        }
    if (ok) { *hgdis_out = hgdis; *closecell = cc; return true; }
  }
  return false;
}

void build_all(ibm_handle &H, std::vector<CaseMesh> &cms) {
  const int nc = (int)cms.size();
  const int L = cms[0].cs.nlevels;
  const int dim = cms[0].cs.dim;
  H.ncases = nc; H.L = L; H.dim = dim;
  const int nblk = nc * L;
  H.blk_start.assign(nblk + 1, 0);
  H.blk_nnormal.assign(nblk, 0);
  H.blk_twod.assign(nblk, 0);
  H.blk_h.assign((size_t)nblk * 3, 0.0);
  // sorted order per block: normal cells in Morton order, then ghosts by
  // (kind, Morton)
  std::vector<std::vector<std::vector<int>>> order(nc), perm(nc);
  int64_t total = 0;
  for (int c = 0; c < nc; ++c) {
    order[c].resize(L); perm[c].resize(L);
    for (int l = 0; l < L; ++l) {
      const Level &lv = cms[c].lv[l];
      int n = (int)lv.cells.size();
      std::vector<int> &o = order[c][l];
      o.resize(n);
      for (int i = 0; i < n; ++i) o[i] = i;
      std::vector<uint64_t> mk(n);
      for (int i = 0; i < n; ++i) mk[i] = morton(dim, lv.cells[i].ix, lv.cells[i].iy, lv.cells[i].iz);
      // keys are unique, so the order is too, whatever the sort
      __gnu_parallel::sort(o.begin(), o.end(), [&](int a, int b) {
        int ka = lv.cells[a].kind == NORMAL ? 0 : lv.cells[a].kind;
        int kb = lv.cells[b].kind == NORMAL ? 0 : lv.cells[b].kind;
        if (ka != kb) return ka < kb;
        return mk[a] < mk[b];
      });
      perm[c][l].resize(n);
      for (int i = 0; i < n; ++i) perm[c][l][o[i]] = i;
      int b = c * L + l;
      int nn = 0;
      for (int i = 0; i < n; ++i) nn += (lv.cells[i].kind == NORMAL);
      H.blk_nnormal[b] = nn;
      H.blk_twod[b] = lv.twod ? 1 : 0;
      for (int d = 0; d < 3; ++d) H.blk_h[(size_t)b * 3 + d] = lv.h[d];
      H.blk_start[b] = (int32_t)total;
      total += n;
    }
  }
  if (total > 2000000000LL) throw std::runtime_error("mesh exceeds int32 cell indices");
  H.blk_start[nblk] = (int32_t)total;
  const size_t N = (size_t)total;
  // written in full by the loop below (every cell of every block)
  H.key.resize(N * 3);
  H.kind.resize(N);
  H.nb.resize(N * 27);
  H.signdis.resize(N);
  H.center.resize(N * 3);
  fvec<int32_t> cpatch(N);
  fvec<int8_t> solid(N);
  H.cell_faces.resize(N * 6);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)N * 6; ++i) H.cell_faces[i] = -1;
  H.ibidx.assign(N, -1);
  const int dmax = (dim == 3) ? 3 : 2;

  for (int c = 0; c < nc; ++c) {
    CaseMesh &cm = cms[c];
    for (int l = 0; l < L; ++l) {
      const Level &lv = cm.lv[l];
      const int b = c * L + l;
      const int base = H.blk_start[b];
      const std::vector<int> &o = order[c][l];
      const std::vector<int> &pm = perm[c][l];
#pragma omp parallel for schedule(static)
      for (long p = 0; p < (long)o.size(); ++p) {
        const LCell &cc = lv.cells[o[p]];
        size_t g = base + p;
        H.key[g * 3] = cc.ix; H.key[g * 3 + 1] = cc.iy; H.key[g * 3 + 2] = cc.iz;
        H.kind[g] = cc.kind;
        cell_center(cm, lv, cc, &H.center[g * 3]);
        int pt; double sd;
        closest_patch(cm, &H.center[g * 3], &pt, &sd);
        H.signdis[g] = sd;
        cpatch[g] = pt;
        solid[g] = (sd < 0.0) ? 1 : 0;   // DistanceSign_boxtowall Body.cpp:810-817
        // Box::neib[3][3][3]: same-level neighbours, z periodic
        for (int di = 0; di < 3; ++di)
          for (int dj = 0; dj < 3; ++dj)
            for (int dk = 0; dk < 3; ++dk) {
              int a = lv.find(cc.ix + di - 1, cc.iy + dj - 1, cc.iz + dk - 1);
              H.nb[g * 27 + di * 9 + dj * 3 + dk] = (a < 0) ? -1 : base + pm[a];
            }
      }
    }
  }
  // faces (Mesh::ConstructBoxFace Mesh.cpp:18-83): every normal cell gets its
  // 2*dim faces; a face is shared with the neighbour on that side, bi[0] is
  // the low side.  2D builds no z faces (ComputeFaceFlux skips fnv==2,
  // NS_Solver.H:728-730).
  // Face numbering: per block, per direction, cells in index order, low side
  // then high side; a face between two normal cells of the block is numbered
  // at the lower-indexed of them.  Counted, scanned, then filled in parallel.
  H.blk_face_start.assign(nblk + 1, 0);
  {
    size_t nf = 0;
    std::vector<int32_t> cnt;
    for (int c = 0; c < nc; ++c)
      for (int l = 0; l < L; ++l) {
        const int b = c * L + l;
        H.blk_face_start[b] = (int32_t)nf;
        const int s = H.blk_start[b], e = s + H.blk_nnormal[b], n = e - s;
        for (int d = 0; d < dmax; ++d) {
          int o0[3] = {1, 1, 1}, o1[3] = {1, 1, 1};
          o0[d] = 0; o1[d] = 2;
          const int q0 = o0[0] * 9 + o0[1] * 3 + o0[2], q1 = o1[0] * 9 + o1[1] * 3 + o1[2];
          auto nbr = [&](int g, int sd) { return H.nb[(size_t)g * 27 + (sd ? q1 : q0)]; };
          auto owns = [&](int g, int a) { return !(a >= s && a < e && a < g); };
          cnt.assign((size_t)n + 1, 0);
          int missing = 0;
#pragma omp parallel for schedule(static) reduction(+ : missing)
          for (int i = 0; i < n; ++i) {
            const int g = s + i;
            int k = 0;
            for (int sd = 0; sd < 2; ++sd) {
              const int a = nbr(g, sd);
              if (a < 0) { ++missing; continue; }
              k += owns(g, a) ? 1 : 0;
            }
            cnt[i] = k;
          }
          if (missing) throw std::runtime_error("normal cell without a face neighbour");
          int32_t run = 0;
          for (int i = 0; i < n; ++i) { const int32_t k = cnt[i]; cnt[i] = run; run += k; }
          if (nf + (size_t)run > (size_t)INT32_MAX) throw std::runtime_error("mesh exceeds int32 face indices");
          H.face_cells.resize((nf + run) * 2);
          H.face_dir.resize(nf + run);
#pragma omp parallel for schedule(static)
          for (int i = 0; i < n; ++i) {
            const int g = s + i;
            int f = (int)nf + cnt[i];
            for (int sd = 0; sd < 2; ++sd) {
              const int a = nbr(g, sd);
              if (!owns(g, a)) continue;
              H.face_cells[(size_t)f * 2] = sd == 0 ? a : g;
              H.face_cells[(size_t)f * 2 + 1] = sd == 0 ? g : a;
              H.face_dir[f] = (int8_t)d;
              H.cell_faces[(size_t)g * 6 + d * 2 + sd] = f;
              H.cell_faces[(size_t)a * 6 + d * 2 + (1 - sd)] = f;
              ++f;
            }
          }
          nf += (size_t)run;
        }
      }
    H.blk_face_start[nblk] = (int32_t)nf;
  }

  // domain ghosts (Domain::FaceProperty Domain.H:372-446): the first matching
  // domain face in the order x-, x+, y-, y+ (Userfunction_cy_naca15.cpp:
  // 283-346), refcell one step inward (two for the outer layer).  Corner
  // ghosts come after edge ghosts so that their reference cell is filled
  // first.
  for (int c = 0; c < nc; ++c) {
    const ibm_case &cs = cms[c].cs;
    int b = c * L + 0;
    int s = H.blk_start[b], e = H.blk_start[b + 1];
    const int fdefault[4] = {-13, -12, -50, -50};
    int ftype[4];
    for (int j = 0; j < 4; ++j) ftype[j] = cs.face_bc[j] ? cs.face_bc[j] : fdefault[j];
    const int fdir[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    for (int pass = 0; pass < 2; ++pass)
      for (int g = s; g < e; ++g) {
        if (H.kind[g] != IBG_DMGHOST) continue;
        int ix = H.key[(size_t)g * 3], iy = H.key[(size_t)g * 3 + 1];
        bool xo = (ix < 0 || ix >= cs.nx), yo = (iy < 0 || iy >= cs.ny);
        bool corner = xo && yo;
        if ((pass == 0) == corner) continue;
        int j;
        if (ix < 0) j = 0; else if (ix >= cs.nx) j = 1; else if (iy < 0) j = 2; else j = 3;
        auto inface = [&](int x, int y) {
          if (j == 0) return x < 0;
          if (j == 1) return x >= cs.nx;
          if (j == 2) return y < 0;
          return y >= cs.ny;
        };
        int ref = H.nb[(size_t)g * 27 + (1 + fdir[j][0]) * 9 + (1 + fdir[j][1]) * 3 + 1];
        if (ref >= 0 && inface(H.key[(size_t)ref * 3], H.key[(size_t)ref * 3 + 1]))
          ref = H.nb[(size_t)ref * 27 + (1 + fdir[j][0]) * 9 + (1 + fdir[j][1]) * 3 + 1];
        if (ref < 0) throw std::runtime_error("domain ghost without a reference cell");
        H.dmg_cell.push_back(g);
        H.dmg_ref.push_back(ref);
        H.dmg_type.push_back(ftype[j]);
        // a wall ghost's reference neighbour: one step further inside from
        // the reference cell (BCValues.H:98-104), and the inward wall normal
        int nbr = -1;
        if (ftype[j] == -81) {
          nbr = H.nb[(size_t)ref * 27 + (1 + fdir[j][0]) * 9 + (1 + fdir[j][1]) * 3 + 1];
          if (nbr < 0) throw std::runtime_error("wall ghost without a reference neighbour");
          H.walls = true;
        }
        H.dmg_nbr.push_back(nbr);
        H.dmg_nmv.push_back((double)fdir[j][0]);
        H.dmg_nmv.push_back((double)fdir[j][1]);
        H.dmg_nmv.push_back(0.0);
      }
  }
  // coarse/fine pairs, global indices: per coarse level f_res_ghost[0],
  // f_res_ghost[1], then f_pro_ghost
  for (int c = 0; c < nc; ++c)
    for (int l = 0; l + 1 < L; ++l) {
      const int bc = c * L + l, bf = c * L + l + 1;
      for (int layer = 0; layer < 3; ++layer)
        for (const LPair &pr : cms[c].pairs[l]) {
          if (pr.layer != layer) continue;
          int gci = H.blk_start[bc] + perm[c][l][pr.ci];
          int fi[8];
          for (int q = 0; q < 8; ++q)
            fi[q] = pr.fi[q] < 0 ? -1 : H.blk_start[bf] + perm[c][l + 1][pr.fi[q]];
          if (layer < 2) {
            H.res_ci.push_back(gci);
            for (int q = 0; q < 8; ++q) H.res_fi.push_back(fi[q]);
            H.res_blk.push_back(bc);
          } else {
            H.pro_ci.push_back(gci);
            for (int q = 0; q < 8; ++q) H.pro_fi.push_back(fi[q]);
            H.pro_blk.push_back(bc);
          }
        }
    }
  // coarse/fine flux-balance faces (AMR::FindModifyFace Fluxmodification.cpp:
  // 5-289): for every prolongation pair -- a normal coarse cell whose sons
  // are prolongation ghosts -- each face towards a first-layer restriction
  // ghost (a coarse ghost whose sons are normal cells), with the faces of the
  // coarse cell's sons on that side.  The sons follow the reference's
  // fface[nx][ny]: first transverse axis outer, second inner.
  for (size_t i = 0; i < H.pro_ci.size(); ++i) {
    const int C = H.pro_ci[i];
    const int bf = H.pro_blk[i] + 1;   // the fine level of the same case
    const bool twod = H.blk_twod[bf] != 0;
    const int32_t *fi = &H.pro_fi[i * 8];
    for (int d = 0; d < dim; ++d)
      for (int s2 = 0; s2 < 2; ++s2) {
        int o[3] = {0, 0, 0};
        o[d] = s2 ? 1 : -1;
        const int P = H.nb[(size_t)C * 27 + (o[0] + 1) * 9 + (o[1] + 1) * 3 + (o[2] + 1)];
        if (P < 0 || H.kind[P] != IBG_RES1) continue;
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        const int n1 = 2, n2 = (twod && d != 2) ? 1 : (dim == 3 ? 2 : 1);
        const int mode = !twod ? 0 : (d == 2 ? 2 : 1);
        H.mdy.push_back(mode);
        H.mdy.push_back(H.cell_faces[(size_t)C * 6 + d * 2 + s2]);
        int nf = 0;
        for (int u = 0; u < n1; ++u)
          for (int v = 0; v < n2; ++v) {
            int idx[3];
            idx[d] = (twod && d == 2) ? 0 : s2;   // a twod son spans the coarse z extent
            idx[t1] = u;
            idx[t2] = v;
            const int son = fi[idx[0] * 4 + idx[1] * 2 + idx[2]];
            if (son < 0) throw std::runtime_error("a prolongation pair without the son of a balance face");
            const int ff = H.cell_faces[(size_t)son * 6 + d * 2 + s2];
            if (ff < 0) throw std::runtime_error("a balance face's son without that face");
            H.mdy.push_back(ff);
            ++nf;
          }
        for (; nf < 4; ++nf) H.mdy.push_back(-1);
      }
  }
  // immersed boundary on the finest level of each case
  int patch_base = 0;
  for (int c = 0; c < nc; ++c) {
    CaseMesh &cm = cms[c];
    const ibm_case &cs = cm.cs;
    const int lf = L - 1;
    const int b = c * L + lf;
    const int s = H.blk_start[b], e = H.blk_start[b + 1];
    const Level &lv = cm.lv[lf];
    Finder F{&lv, s, &perm[c][lf]};
    for (int bi = 0; bi < cs.nbody; ++bi) {
      H.body_case.push_back(c);
      for (int d = 0; d < 3; ++d) H.body_center.push_back(cs.body_center[d]);
    }
    if (cs.nbody > 0) {
      // InfectWallBox_Init_Negative Body.cpp:1016-1082: non-solid cells with a
      // solid neighbour (2D: in-plane 3x3; 3D: 27) and their non-solid
      // neighbours are IB cells.
      std::vector<char> inf(e - s, 0);
      auto nbrs = [&](int g, int q) -> int {
        return H.nb[(size_t)g * 27 + q];
      };
      std::vector<int> ql;
      for (int di = 0; di < 3; ++di)
        for (int dj = 0; dj < 3; ++dj)
          for (int dk = 0; dk < 3; ++dk)
            if (dim == 3 || dk == 1) ql.push_back(di * 9 + dj * 3 + dk);
      for (int g = s; g < e; ++g) {
        if (H.kind[g] == IBG_DMGHOST || solid[g]) continue;
        bool hit = false;
        for (int q : ql) {
          int a = nbrs(g, q);
          if (a >= 0 && solid[a] && H.kind[a] != IBG_DMGHOST) { hit = true; break; }
        }
        if (!hit) continue;
        inf[g - s] = 1;
        for (int q : ql) {
          int a = nbrs(g, q);
          if (a >= 0 && !solid[a] && H.kind[a] != IBG_DMGHOST) inf[a - s] = 1;
        }
      }
      for (int g = s; g < e; ++g) {
        if (!inf[g - s]) continue;
        if (H.kind[g] != IBG_NORMAL)
          throw std::runtime_error("an immersed-boundary cell lies on a ghost of the finest level");
        H.ibidx[g] = (int32_t)H.ib_cell.size();
        H.ib_cell.push_back(g);
      }
      const double *h = &H.blk_h[(size_t)b * 3];
      for (size_t k = 0; k < H.ib_cell.size(); ++k) {
        int g = H.ib_cell[k];
        if (g < s || g >= e) continue;
        const Patch &P = cm.patches[cpatch[g]];
        double sd = H.signdis[g];
        const double *cen = &H.center[(size_t)g * 3];
        double att[3];
        for (int d = 0; d < 3; ++d) att[d] = cen[d] - P.nv[d] * sd;   // Body.cpp:1073
        int32_t intp[8]; double coef[8], hgdis; int cc;
        for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
        if (!search_image(F, H, h, cs.origin, cen, P.nv, solid, H.ibidx, intp, coef, &hgdis, &cc))
          throw std::runtime_error("no valid image point for an IB cell");
        for (int q = 0; q < 8; ++q) { H.ib_intp.push_back(intp[q]); H.ib_coef.push_back(coef[q]); }
        H.ib_hgdis.push_back(hgdis);
        for (int d = 0; d < 3; ++d) { H.ib_nv.push_back(P.nv[d]); H.ib_attach.push_back(att[d]); }
        H.ib_body.push_back(P.body + (int)(H.body_case.size() - cs.nbody));
      }
      // wall-force patches, FindPatchHangingNode Body.cpp:1506-1575
      for (size_t p = 0; p < cm.patches.size(); ++p) {
        const Patch &P = cm.patches[p];
        int32_t intp[8]; double coef[8], hgdis; int cc;
        for (int q = 0; q < 8; ++q) { intp[q] = -1; coef[q] = 0.0; }
        if (!search_image(F, H, h, cs.origin, P.pc, P.nv, solid, H.ibidx, intp, coef, &hgdis, &cc))
          throw std::runtime_error("no valid image point for a surface patch");
        H.pat_close.push_back(cc);
        for (int q = 0; q < 8; ++q) { H.pat_intp.push_back(intp[q]); H.pat_coef.push_back(coef[q]); }
        H.pat_hgdis.push_back(hgdis);
        for (int d = 0; d < 3; ++d) { H.pat_nv.push_back(P.nv[d]); H.pat_pc.push_back(P.pc[d]); }
        H.pat_area.push_back(P.area);
        H.pat_case.push_back(c);
      }
      patch_base += (int)cm.patches.size();
    }
  }
  // Mesh::ComptDistancetoDomain (Mesh.cpp:1638-1686, 1688-1770): the distance
  // of normal cells to the nearest wall face (|(c - fc) . n| on the plane
  // faces of the box), except IB cells of the finest level; signdis becomes
  // min(signdis, distance_to_dm)
  H.dist_dm.assign(N, 9999.0);
  for (int c = 0; c < nc && H.walls; ++c) {
    const ibm_case &cs = cms[c].cs;
    bool wall[4];
    bool any = false;
    for (int j = 0; j < 4; ++j) { wall[j] = cs.face_bc[j] == -81; any = any || wall[j]; }
    if (!any) continue;
    const double xlo = cs.origin[0], xhi = cs.origin[0] + cs.nx * cs.dh0[0];
    const double ylo = cs.origin[1], yhi = cs.origin[1] + cs.ny * cs.dh0[1];
    for (int l = 0; l < L; ++l) {
      const int b = c * L + l;
      for (int g = H.blk_start[b]; g < H.blk_start[b] + H.blk_nnormal[b]; ++g) {
        if (l == L - 1 && cs.nbody > 0 && H.ibidx[g] >= 0) continue;
        const double *x = &H.center[(size_t)g * 3];
        double d = 9999.0;
        if (wall[0]) d = std::min(d, std::fabs(x[0] - xlo));
        if (wall[1]) d = std::min(d, std::fabs(x[0] - xhi));
        if (wall[2]) d = std::min(d, std::fabs(x[1] - ylo));
        if (wall[3]) d = std::min(d, std::fabs(x[1] - yhi));
        H.dist_dm[g] = d;
        H.signdis[g] = std::min(H.signdis[g], d);
      }
    }
  }
  // the segment, body and flip of every patch, for the moving-body update
  // (one patch per segment of a 2D section)
  {
    int pbase = 0, gbody = 0;
    H.sf_body_p0.assign(1, 0);
    for (int c = 0; c < nc; ++c) {
      const CaseMesh &cm = cms[c];
      const ibm_case &cs = cm.cs;
      if (cs.nbody == 0) continue;
      if (std::max(cs.span_patches, 1) != 1 || dim != 2) H.sf_ok = false;
      for (int b = 0; b < cs.nbody; ++b) {
        int n = 0;
        for (size_t sgi = 0; sgi < cm.seg_body.size(); ++sgi)
          if (cm.seg_body[sgi] == b) {
            H.sf_pt0.push_back(H.sf_npoint + cm.seg_first[sgi]);
            H.sf_body.push_back(gbody + b);
            H.sf_flip.push_back((int8_t)(b % 2 == 1));
            H.sf_dz.push_back(cs.span / std::max(cs.span_patches, 1));
            ++n;
          }
        H.sf_body_p0.push_back(H.sf_body_p0.back() + n);
      }
      int npts = 0;
      for (int b = 0; b < cs.nbody; ++b) npts += cs.body_npts[b];
      H.sf_npoint += npts;
      gbody += cs.nbody;
      pbase += (int)cm.patches.size();
    }
    (void)pbase;
    for (int c = 0; c < nc; ++c)
      for (int d = 0; d < 3; ++d) H.sf_origin.push_back(cms[c].cs.origin[d]);
  }
  // closest patch of every cell as a global patch index
  H.cell_patch.assign(N, -1);
  {
    std::vector<int> base(nc + 1, 0);
    for (int c = 0; c < nc; ++c) base[c + 1] = base[c] + (cms[c].cs.nbody > 0 ? (int)cms[c].patches.size() : 0);
    for (int c = 0; c < nc; ++c) {
      if (cms[c].cs.nbody == 0) continue;
      for (int g = H.blk_start[c * L]; g < H.blk_start[(c + 1) * L]; ++g)
        if (cpatch[g] >= 0) H.cell_patch[g] = base[c] + cpatch[g];
    }
  }
}

void fill_soa(ibm_handle &H) {
  ibg_mesh_soa &m = H.soa;
  std::memset(&m, 0, sizeof(m));
  m.dim = H.dim; m.ncases = H.ncases; m.nlevels = H.L;
  m.ncells = H.blk_start.back();
  m.blk_start = H.blk_start.data(); m.blk_nnormal = H.blk_nnormal.data();
  m.blk_h = H.blk_h.data(); m.blk_twod = H.blk_twod.data();
  m.key = H.key.data(); m.kind = H.kind.data(); m.nb = H.nb.data();
  m.cell_faces = H.cell_faces.data(); m.signdis = H.signdis.data(); m.ibidx = H.ibidx.data();
  m.nfaces = (int32_t)H.face_dir.size();
  m.blk_face_start = H.blk_face_start.data(); m.face_cells = H.face_cells.data();
  m.face_dir = H.face_dir.data();
  m.ndmg = (int32_t)H.dmg_cell.size();
  m.dmg_cell = H.dmg_cell.data(); m.dmg_ref = H.dmg_ref.data(); m.dmg_type = H.dmg_type.data();
  m.nres = (int32_t)H.res_ci.size();
  m.res_ci = H.res_ci.data(); m.res_fi = H.res_fi.data(); m.res_blk = H.res_blk.data();
  m.npro = (int32_t)H.pro_ci.size();
  m.pro_ci = H.pro_ci.data(); m.pro_fi = H.pro_fi.data(); m.pro_blk = H.pro_blk.data();
  m.nib = (int32_t)H.ib_cell.size();
  m.ib_cell = H.ib_cell.data(); m.ib_intp = H.ib_intp.data(); m.ib_coef = H.ib_coef.data();
  m.ib_hgdis = H.ib_hgdis.data(); m.ib_nv = H.ib_nv.data(); m.ib_attach = H.ib_attach.data();
  m.ib_body = H.ib_body.data();
  m.npatch = (int32_t)H.pat_area.size();
  m.pat_intp = H.pat_intp.data(); m.pat_coef = H.pat_coef.data(); m.pat_hgdis = H.pat_hgdis.data();
  m.pat_nv = H.pat_nv.data(); m.pat_pc = H.pat_pc.data(); m.pat_area = H.pat_area.data();
  m.pat_case = H.pat_case.data();
  m.nbody = (int32_t)H.body_case.size();
  m.body_case = H.body_case.data(); m.body_center = H.body_center.data();
  if (H.walls) {
    m.dmg_nbr = H.dmg_nbr.data(); m.dmg_nmv = H.dmg_nmv.data(); m.dist_dm = H.dist_dm.data();
  }
  m.nmdy = (int32_t)(H.mdy.size() / 6);
  m.mdy = m.nmdy ? H.mdy.data() : nullptr;
}

}  // namespace

extern "C" {

ibm_handle *ibm_build(const ibm_case *cases, int32_t ncases) {
  try {
    if (!cases || ncases < 1 || ncases > IBG_MAX_CASES) throw std::runtime_error("bad case count");
    std::vector<CaseMesh> cms(ncases);
    for (int c = 0; c < ncases; ++c) {
      const ibm_case &cs = cases[c];
      if (cs.nlevels < 1 || cs.nlevels > IBG_MAX_LEVELS) throw std::runtime_error("bad level count");
      if (cs.nlevels != cases[0].nlevels || cs.dim != cases[0].dim)
        throw std::runtime_error("cases of one batch must share dim and level count");
      if (cs.dim == 2 && cs.nz != 1) throw std::runtime_error("2D cases have nz == 1");
      if (cs.nbody > IBM_MAX_BODIES) throw std::runtime_error("too many bodies");
      cms[c].cs = cs;
      build_segments(cms[c]);
      build_segment_runs(cms[c]);
      build_patches(cms[c]);
      build_levels(cms[c]);
      if (cs.refine_only) {
        // the bodies have placed the refinement bands; the flow sees none of
        // them (no solid or immersed-boundary cells, no wall patches)
        CaseMesh &cm = cms[c];
        cm.cs.nbody = 0;
        cm.patches.clear(); cm.segxy.clear(); cm.seg_body.clear(); cm.seg_first.clear();
        cm.seg_cx.clear(); cm.seg_cy.clear(); cm.seg_nx.clear(); cm.seg_ny.clear();
        cm.seg_len.clear();
        cm.run_x0.clear(); cm.run_x1.clear(); cm.run_y0.clear(); cm.run_y1.clear();
      }
    }
    ibm_handle *H = new ibm_handle();
    H->cases.assign(cases, cases + ncases);
    build_all(*H, cms);
    fill_soa(*H);
    return H;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return nullptr;
  }
}

const ibg_mesh_soa *ibm_mesh(const ibm_handle *h) { return h ? &h->soa : nullptr; }

const char *ibm_last_error(void) { return g_last_error.c_str(); }

void ibm_free(ibm_handle *h) { delete h; }

// AMR::ComptDatainNewBox (AMR.H:148-215) after a regrid, on canonical keys:
// a cell whose key was a normal cell of the same level keeps its data, and
// so does a ghost whose key existed; a
// cell whose old region was refined takes the restriction of the old leaves
// below it (AverageCell TwoLevelOp.H:45-87, level by level, the SA variable
// averaged); a cell whose old parent
// was a normal cell takes its prolongation (Prolongation TwoLevelOp.H:
// 183-215: conserved variables U_p +- 1/4 of their central differences over
// the parent's face neighbours, ComptFineCellData :12-44), or a copy of a
// ghost parent (AMR.H:208-213); otherwise the nearest older cell at the same
// key or above is copied.
int32_t ibm_transfer(const ibm_handle *oh, const ibg_physics *ph, const double *fv_old,
                     const double *nut_old, const ibm_handle *nh, double *fv_new,
                     double *nut_new) {
  try {
    if (!oh || !nh || !ph || !fv_old || !fv_new) throw std::runtime_error("null argument");
    if (oh->ncases != nh->ncases || oh->L != nh->L || oh->dim != nh->dim)
      throw std::runtime_error("level sets of different cases");
    if ((nut_old == nullptr) != (nut_new == nullptr)) throw std::runtime_error("nut in or out missing");
    const int L = oh->L, nc = oh->ncases, dim = oh->dim;
    const int NF = IBG_NFV;
    if (oh->blk_twod != nh->blk_twod) throw std::runtime_error("level sets refine differently");
    std::vector<KeyGrid> omap((size_t)nc * L);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nc * L; ++b) {
      const int s0 = oh->blk_start[b];
      omap[b].build(oh->blk_start[b + 1] - s0, [&](int i, int *k) {
        for (int d = 0; d < 3; ++d) k[d] = oh->key[(size_t)(s0 + i) * 3 + d];
      });
    }
    auto old_at = [&](int c, int l, int ix, int iy, int iz) -> int {
      if (l < 0 || l >= L) return -1;
      const int b = c * L + l;
      const int i = omap[(size_t)b].find(ix, iy, iz);
      return i < 0 ? -1 : oh->blk_start[b] + i;
    };
    auto fdiv2 = [](int k) { return k >= 0 ? k / 2 : -((-k + 1) / 2); };
    auto put = [&](double *f, const ibg_physics &P, double r, double u, double v, double w, double pr) {
      // FlowVariables (roe,u,v,w,p,T,e), T and e by Get_T / Get_E (FlowVariables.H:23,33)
      f[0] = r; f[1] = u; f[2] = v; f[3] = w; f[4] = pr;
      f[5] = pr / (r * P.Rg);
      f[6] = P.Cv * f[5] + 0.5 * (u * u + v * v + w * w);
    };
    auto cons = [&](int a, double *U) {
      const double *f = fv_old + (size_t)a * NF;
      U[0] = f[0]; U[1] = f[0] * f[1]; U[2] = f[0] * f[2]; U[3] = f[0] * f[3]; U[4] = f[0] * f[6];
    };
    // value of an old region: a normal cell, or the restriction (AverageCell)
    // of the values of its sons when every son resolves
    std::function<bool(int, int, int, int, int, double &, double &, double &, double &, double &,
                       double &)>
        leafval = [&](int c, int l, int kx, int ky, int kz, double &r, double &u, double &v,
                      double &w, double &pr, double &nv) -> bool {
      const int a = old_at(c, l, kx, ky, kz);
      if (a >= 0 && oh->kind[a] == IBG_NORMAL) {
        const double *o = fv_old + (size_t)a * NF;
        r = o[0]; u = o[1]; v = o[2]; w = o[3]; pr = o[4];
        nv = nut_old ? nut_old[a] : 0.0;
        return true;
      }
      if (l + 1 >= L) return false;
      const ibg_physics &P = ph[c];
      const bool tw = oh->blk_twod[(size_t)c * L + l + 1] != 0;
      double rs = 0, rus = 0, rvs = 0, rws = 0, res = 0, ns = 0;
      int cnt = 0;
      for (int si = 0; si < 2; ++si)
        for (int sj = 0; sj < 2; ++sj)
          for (int sk = 0; sk < (tw ? 1 : 2); ++sk) {
            double r2, u2, v2, w2, p2, n2;
            if (!leafval(c, l + 1, 2 * kx + si, 2 * ky + sj, tw ? kz : 2 * kz + sk, r2, u2, v2, w2,
                         p2, n2))
              return false;
            const double T = p2 / (r2 * P.Rg);
            const double e = P.Cv * T + 0.5 * (u2 * u2 + v2 * v2 + w2 * w2);
            rs += r2; rus += u2 * r2; rvs += v2 * r2; rws += w2 * r2; res += e * r2; ns += n2;
            ++cnt;
          }
      u = rus / rs; v = rvs / rs; w = rws / rs;
      const double e = res / rs;
      r = rs / cnt;
      pr = (e - 0.5 * (u * u + v * v + w * w)) * (P.gama - 1.0) * r;
      nv = ns / cnt;
      return true;
    };
    const int N = nh->blk_start[(size_t)nc * L];
    for (int b = 0; b < nc * L; ++b) {
      const int c = b / L, l = b % L;
      const ibg_physics &P = ph[c];
      const bool twod_l = nh->blk_twod[b] != 0;                  // level l over l-1
      for (int g = nh->blk_start[b]; g < nh->blk_start[b + 1]; ++g) {
        const int kx = nh->key[(size_t)g * 3], ky = nh->key[(size_t)g * 3 + 1], kz = nh->key[(size_t)g * 3 + 2];
        double *f = fv_new + (size_t)g * NF;
        double *nu = nut_new ? nut_new + g : nullptr;
        int a = old_at(c, l, kx, ky, kz);
        // a kept cell, or a ghost whose key existed (ghosts are refilled by
        // the ghost fills of the next step)
        if (a >= 0 && (oh->kind[a] == IBG_NORMAL || nh->kind[g] != IBG_NORMAL)) {
          std::memcpy(f, fv_old + (size_t)a * NF, sizeof(double) * NF);
          if (nu) *nu = nut_old[a];
          continue;
        }
        // restriction: the average of the conserved variables over the old
        // leaves below (sons, or theirs where a son was itself refined)
        {
          double r, u, v, w, pr, nv = 0.0;
          if (leafval(c, l, kx, ky, kz, r, u, v, w, pr, nv)) {
            put(f, P, r, u, v, w, pr);
            if (nu) *nu = nv;
            continue;
          }
        }
        // prolongation of the old parent
        if (l > 0) {
          const int px = fdiv2(kx), py = fdiv2(ky), pz = twod_l ? kz : fdiv2(kz);
          const int pa = old_at(c, l - 1, px, py, pz);
          if (pa >= 0) {
            const int *nb = oh->nb.data() + (size_t)pa * 27;
            const int ax0 = nb[0 * 9 + 1 * 3 + 1], ax1 = nb[2 * 9 + 1 * 3 + 1];
            const int ay0 = nb[1 * 9 + 0 * 3 + 1], ay1 = nb[1 * 9 + 2 * 3 + 1];
            const int az0 = nb[1 * 9 + 1 * 3 + 0], az1 = nb[1 * 9 + 1 * 3 + 2];
            const bool full = ax0 >= 0 && ax1 >= 0 && ay0 >= 0 && ay1 >= 0 && az0 >= 0 && az1 >= 0;
            if (oh->kind[pa] == IBG_NORMAL && full) {
              double Uc[5], X0[5], X1[5], Y0[5], Y1[5], Z0[5], Z1[5];
              cons(pa, Uc); cons(ax0, X0); cons(ax1, X1); cons(ay0, Y0); cons(ay1, Y1);
              cons(az0, Z0); cons(az1, Z1);
              const double finc[2] = {-0.25, 0.25};
              const int si = kx - 2 * px, sj = ky - 2 * py, sk = twod_l ? 0 : kz - 2 * pz;
              double U[5];
              for (int t = 0; t < 5; ++t) {
                const double gx = (X1[t] - X0[t]) * 0.5, gy = (Y1[t] - Y0[t]) * 0.5,
                             gz = (Z1[t] - Z0[t]) * 0.5;
                U[t] = twod_l ? Uc[t] + finc[si] * gx + finc[sj] * gy
                              : Uc[t] + finc[si] * gx + finc[sj] * gy + finc[sk] * gz;
              }
              const double r = U[0], u = U[1] / U[0], v = U[2] / U[0], w = U[3] / U[0], e = U[4] / U[0];
              const double ke = (dim == 3) ? 0.5 * (u * u + v * v + w * w) : 0.5 * (u * u + v * v);
              put(f, P, r, u, v, w, (e - ke) * (P.gama - 1.0) * r);
              if (nu) {
                const double gx = (nut_old[ax1] - nut_old[ax0]) * 0.5, gy = (nut_old[ay1] - nut_old[ay0]) * 0.5,
                             gz = (nut_old[az1] - nut_old[az0]) * 0.5;
                *nu = twod_l ? nut_old[pa] + finc[si] * gx + finc[sj] * gy
                             : nut_old[pa] + finc[si] * gx + finc[sj] * gy + finc[sk] * gz;
              }
              continue;
            }
            if (a < 0) a = pa;   // a ghost parent is copied
          }
        }
        // the same key as a ghost, or the nearest older ancestor
        for (int up = 1; a < 0 && l - up >= 0; ++up) {
          int ax = kx, ay = ky, az = kz;
          for (int k = 0; k < up; ++k) {
            const bool tw = nh->blk_twod[(size_t)c * L + l - k] != 0;
            ax = fdiv2(ax); ay = fdiv2(ay); if (!tw) az = fdiv2(az);
          }
          a = old_at(c, l - up, ax, ay, az);
        }
        if (a < 0)
          throw std::runtime_error("cell (case " + std::to_string(c) + ", level " + std::to_string(l) +
                                   ", key " + std::to_string(kx) + " " + std::to_string(ky) + " " +
                                   std::to_string(kz) + ", kind " + std::to_string((int)nh->kind[g]) +
                                   ") of the new level set has no old cell above it");
        std::memcpy(f, fv_old + (size_t)a * NF, sizeof(double) * NF);
        if (nu) *nu = nut_old[a];
      }
    }
    (void)N;
    return 0;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return 1;
  }
}

// ------------------------------------------------------- Tecplot output
// Pltoutput::Output_Grid / Output_Soln (Pltoutput.cpp:3-100, 417-450,
// Pltoutput.H:354-380): one FEBRICK zone "N-<case> L-<level>" per level,
// the 8 corners of every cell as nodes (shared corners merged), the
// solution cell-centred in BLOCK packing with the reference's variables.
// Every cell of the level is an element, ghosts included (the reference's
// PRINTGHOST / INPUT_GHOST form), so a solution file is also an exact
// restart (values printed with 17 significant digits).
namespace {
struct ZoneNodes {
  std::vector<int32_t> conn;    // [ncell][8] 1-based node numbers
  std::vector<double> xyz;      // [nnode][3]
};

ZoneNodes zone_nodes(const ibm_handle &H, int b) {
  const int L = H.L, c = b / L;
  const ibm_case &cs = H.cases[c];
  const double *h = &H.blk_h[(size_t)b * 3];
  ZoneNodes z;
  std::unordered_map<int64_t, int32_t> id;
  const int s = H.blk_start[b], e = H.blk_start[b + 1];
  z.conn.reserve((size_t)(e - s) * 8);
  // FEBRICK corner order: bottom face counter-clockwise, then the top face
  const int ci[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                        {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int g = s; g < e; ++g) {
    for (int k = 0; k < 8; ++k) {
      const int x = H.key[(size_t)g * 3] + ci[k][0], y = H.key[(size_t)g * 3 + 1] + ci[k][1],
                zz = H.key[(size_t)g * 3 + 2] + ci[k][2];
      const int64_t key = pack_key(x, y, zz);
      auto it = id.find(key);
      int32_t n;
      if (it == id.end()) {
        n = (int32_t)(z.xyz.size() / 3) + 1;
        id.emplace(key, n);
        z.xyz.push_back(cs.origin[0] + x * h[0]);
        z.xyz.push_back(cs.origin[1] + y * h[1]);
        z.xyz.push_back(cs.origin[2] + zz * h[2]);
      } else {
        n = it->second;
      }
      z.conn.push_back(n);
    }
  }
  return z;
}

void print_block(FILE *f, const std::vector<double> &v) {
  for (size_t i = 0; i < v.size(); ++i)
    std::fprintf(f, (i % 5 == 4 || i + 1 == v.size()) ? "%.17g\n" : "%.17g ", v[i]);
}
}  // namespace

int32_t ibm_write_grid(const ibm_handle *h, const char *path) {
  if (!h || !path) return 1;
  FILE *f = std::fopen(path, "w");
  if (!f) { g_last_error = std::string("cannot open ") + path; return 1; }
  std::fprintf(f, "TITLE = \"Boxgrid\"\nFILETYPE = GRID\nVARIABLES = \"x\",\"y\",\"z\"\n");
  for (int b = 0; b < h->ncases * h->L; ++b) {
    const ZoneNodes z = zone_nodes(*h, b);
    const int ne = h->blk_start[b + 1] - h->blk_start[b];
    std::fprintf(f, "zone T = \"N-%d L-%d\", NODES=%d, ELEMENTS=%d, DATAPACKING=BLOCK, ZONETYPE=FEBRICK\n",
                 b / h->L, b % h->L, (int)(z.xyz.size() / 3), ne);
    for (int d = 0; d < 3; ++d) {
      std::vector<double> col(z.xyz.size() / 3);
      for (size_t n = 0; n < col.size(); ++n) col[n] = z.xyz[n * 3 + d];
      print_block(f, col);
    }
    for (int g = 0; g < ne; ++g) {
      for (int k = 0; k < 8; ++k) std::fprintf(f, k < 7 ? "%d " : "%d\n", z.conn[(size_t)g * 8 + k]);
    }
  }
  std::fclose(f);
  return 0;
}

int32_t ibm_write_solution(const ibm_handle *h, const ibg_physics *ph, const double *fv,
                           const double *nut, int32_t ts, const double *dt, const double *t,
                           const char *path) {
  if (!h || !ph || !fv || !path) return 1;
  FILE *f = std::fopen(path, "w");
  if (!f) { g_last_error = std::string("cannot open ") + path; return 1; }
  const int NF = IBG_NFV;
  std::fprintf(f, "TITLE = \"Flow Solution\"\nFILETYPE = SOLUTION\n");
  if (nut)
    std::fprintf(f, "VARIABLES = \"roe\",\"u\",\"v\",\"w\",\"p\",\"T\",\"viseddy\",\"Ma\",\"dis\",\"dmdis\",\"i_index\",\"j_index\",\"k_index\",\"Omegaz\"\n");
  else
    std::fprintf(f, "VARIABLES = \"roe\",\"u\",\"v\",\"w\",\"p\",\"T\",\"Ma\",\"dis\",\"dmdis\",\"i_index\",\"j_index\",\"k_index\",\"Omegaz\"\n");
  std::fprintf(f, "DATASETAUXDATA ts=\"%d\"\n", ts);
  for (int c = 0; c < h->ncases; ++c) {
    if (dt) std::fprintf(f, "DATASETAUXDATA dt%d=\"%.17g\"\n", c, dt[c]);
    if (t) std::fprintf(f, "DATASETAUXDATA t%d=\"%.17g\"\n", c, t[c]);
  }
  const int nvar = nut ? 14 : 13;
  for (int b = 0; b < h->ncases * h->L; ++b) {
    const int c = b / h->L;
    const ibg_physics &P = ph[c];
    const ZoneNodes z = zone_nodes(*h, b);
    const int s = h->blk_start[b], e = h->blk_start[b + 1], ne = e - s;
    std::fprintf(f, "zone T = \"N-%d L-%d\", NODES=%d, ELEMENTS=%d, DATAPACKING=BLOCK, ZONETYPE=FEBRICK\n",
                 c, b % h->L, (int)(z.xyz.size() / 3), ne);
    std::fprintf(f, "VARLOCATION = ([1-%d]=CELLCENTERED)\n", nvar);
    std::vector<double> col(ne);
    for (int v = 0; v < 6; ++v) {
      for (int i = 0; i < ne; ++i) col[i] = fv[(size_t)(s + i) * NF + v];
      print_block(f, col);
    }
    if (nut) {
      for (int i = 0; i < ne; ++i) col[i] = nut[s + i];
      print_block(f, col);
    }
    for (int i = 0; i < ne; ++i) {   // Ma = |u| / sqrt(gama Rg T)
      const double *q = fv + (size_t)(s + i) * NF;
      col[i] = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / std::sqrt(P.gama * P.Rg * q[5]);
    }
    print_block(f, col);
    for (int i = 0; i < ne; ++i) col[i] = h->signdis[s + i];
    print_block(f, col);
    for (int i = 0; i < ne; ++i) col[i] = h->dist_dm.empty() ? 9999.0 : h->dist_dm[s + i];
    print_block(f, col);
    for (int d = 0; d < 3; ++d) {
      for (int i = 0; i < ne; ++i) col[i] = h->key[(size_t)(s + i) * 3 + d];
      print_block(f, col);
    }
    // Omegaz = dv/dx - du/dy by central differences over the face
    // neighbours (0 where one is missing)
    const double *hh = &h->blk_h[(size_t)b * 3];
    for (int i = 0; i < ne; ++i) {
      const int g = s + i;
      const int *nb = &h->nb[(size_t)g * 27];
      const int xm = nb[0 * 9 + 1 * 3 + 1], xp = nb[2 * 9 + 1 * 3 + 1];
      const int ym = nb[1 * 9 + 0 * 3 + 1], yp = nb[1 * 9 + 2 * 3 + 1];
      col[i] = (xm < 0 || xp < 0 || ym < 0 || yp < 0)
                   ? 0.0
                   : (fv[(size_t)xp * NF + 2] - fv[(size_t)xm * NF + 2]) / (2.0 * hh[0]) -
                         (fv[(size_t)yp * NF + 1] - fv[(size_t)ym * NF + 1]) / (2.0 * hh[1]);
    }
    print_block(f, col);
  }
  std::fclose(f);
  return 0;
}

// InputSoln / ReadLevelSolution (input_soln.cpp:12-180): the zones of a
// solution file back into the state of the same level set (element count =
// cells of the level, abort 65 otherwise); e = Cv T + |u|^2/2.
int32_t ibm_read_solution(const ibm_handle *h, const ibg_physics *ph, const char *path, double *fv,
                          double *nut, int32_t *ts, double *dt, double *t) {
  if (!h || !ph || !path || !fv) return 1;
  FILE *f = std::fopen(path, "r");
  if (!f) { g_last_error = std::string("cannot open ") + path; return 1; }
  const int NF = IBG_NFV;
  char line[4096];
  bool has_nut = false;
  int zones = 0;
  int32_t rc = 0;
  std::vector<char> seen((size_t)h->ncases * h->L, 0);   // one zone per (case, level)
  auto fail = [&](const std::string &msg) { g_last_error = msg; rc = 1; };
  while (rc == 0 && std::fgets(line, sizeof(line), f)) {
    std::string sl(line);
    if (sl.rfind("VARIABLES", 0) == 0) { has_nut = sl.find("viseddy") != std::string::npos; continue; }
    if (sl.rfind("DATASETAUXDATA", 0) == 0) {
      int k; double v; int iv;
      if (ts && std::sscanf(line, "DATASETAUXDATA ts=\"%d\"", &iv) == 1) *ts = iv;
      else if (dt && std::sscanf(line, "DATASETAUXDATA dt%d=\"%lf\"", &k, &v) == 2 && k >= 0 && k < h->ncases) dt[k] = v;
      else if (t && std::sscanf(line, "DATASETAUXDATA t%d=\"%lf\"", &k, &v) == 2 && k >= 0 && k < h->ncases) t[k] = v;
      continue;
    }
    if (sl.rfind("zone", 0) != 0) continue;
    int c = -1, l = -1, nn = 0, ne = 0;
    if (std::sscanf(line, "zone T = \"N-%d L-%d\", NODES=%d, ELEMENTS=%d", &c, &l, &nn, &ne) != 4 ||
        c < 0 || c >= h->ncases || l < 0 || l >= h->L) {
      fail("bad zone header");
      break;
    }
    const int b = c * h->L + l, s = h->blk_start[b];
    if (seen[b]) { fail("zone N-" + std::to_string(c) + " L-" + std::to_string(l) + " repeated"); break; }
    seen[b] = 1;
    if (ne != h->blk_start[b + 1] - s) { fail("zone element count differs from the level (abort 65)"); break; }
    if (!std::fgets(line, sizeof(line), f)) { fail("truncated file"); break; }   // VARLOCATION
    const int nv = has_nut ? 7 : 6;
    std::vector<double> col(ne);
    for (int v = 0; v < nv && rc == 0; ++v) {
      for (int i = 0; i < ne; ++i)
        if (std::fscanf(f, "%lf", &col[i]) != 1) { fail("truncated zone data"); break; }
      if (rc) break;
      for (int i = 0; i < ne; ++i) {
        if (v < 6) fv[(size_t)(s + i) * NF + v] = col[i];
        else if (nut) nut[s + i] = col[i];
      }
    }
    // skip the derived variables (Ma, dis, dmdis, indices, Omegaz)
    for (int v = 0; v < 7 && rc == 0; ++v)
      for (int i = 0; i < ne; ++i)
        if (std::fscanf(f, "%lf", &col[0]) != 1) { fail("truncated zone data"); break; }
    if (rc) break;
    const ibg_physics &P = ph[c];
    for (int i = 0; i < ne; ++i) {
      double *q = fv + (size_t)(s + i) * NF;
      q[6] = P.Cv * q[5] + 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    }
    ++zones;
  }
  std::fclose(f);
  for (size_t b = 0; rc == 0 && b < seen.size(); ++b)
    if (!seen[b]) {
      g_last_error = "missing zone N-" + std::to_string(b / h->L) + " L-" + std::to_string(b % h->L);
      rc = 1;
    }
  if (rc == 0 && zones != h->ncases * h->L) { g_last_error = "missing zones"; rc = 1; }
  if (rc == 0 && nut && !has_nut) { g_last_error = "no viseddy in the file"; rc = 1; }
  return rc;
}

int32_t ibm_surface(const ibm_handle *h, const double *pt_off, int32_t npoint, ibg_surface *out) {
  if (!h || !pt_off || !out) return 1;
  if (!h->sf_ok) { g_last_error = "moving bodies: 2D sections with one patch per segment"; return 1; }
  if (npoint != h->sf_npoint) { g_last_error = "surface point count differs from the bodies"; return 1; }
  std::memset(out, 0, sizeof(*out));
  out->npoint = npoint;
  out->pt_off = pt_off;
  out->pat_pt0 = h->sf_pt0.data();
  out->pat_body = h->sf_body.data();
  out->pat_flip = h->sf_flip.data();
  out->pat_dz = h->sf_dz.data();
  out->body_p0 = h->sf_body_p0.data();
  out->case_origin = h->sf_origin.data();
  out->cell_patch = h->cell_patch.data();
  out->pat_close = h->pat_close.data();
  return 0;
}

int32_t ibm_cell_patch(const ibm_handle *h, int32_t *out) {
  if (!h || !out) return 1;
  std::memcpy(out, h->cell_patch.data(), h->cell_patch.size() * sizeof(int32_t));
  return 0;
}

int32_t ibm_patch_close(const ibm_handle *h, int32_t *out) {
  if (!h || !out) return 1;
  std::memcpy(out, h->pat_close.data(), h->pat_close.size() * sizeof(int32_t));
  return 0;
}

int32_t ibm_cell_centers(const ibm_handle *h, double *xyz) {
  if (!h || !xyz) return 1;
  std::memcpy(xyz, h->center.data(), h->center.size() * sizeof(double));
  return 0;
}

// FlowConstants.cpp:5-44 with the case globals of Userfunction_cy_naca15.cpp:
// roe_ref = u_ref = L_ref = 1, mu_ref = 1/Re, T_ref = u_ref^2/(Ma^2 Rg_ref gama);
// freestream (SetInitValues :356-371) u = 1, T = (1/Ma)^2/(gama Rg), roe = 1;
// inlet total state (SetBDFaceVars :396-420), outlet facevar[1].p + backpressure.
int32_t ibm_physics(const ibm_case *c, ibg_physics *o) {
  if (!c || !o) return 1;
  std::memset(o, 0, sizeof(*o));
  const double Rg_ref = 287.4, u_ref = 1.0, roe_ref = 1.0, L_ref = 1.0;
  o->gama = c->gama;
  o->Ma = c->Ma;
  o->Re = c->Re;
  o->Pr = c->Pr;
  const double T_ref = u_ref * u_ref / (c->Ma * c->Ma * Rg_ref * c->gama);
  o->Rg = 1.0 / (c->gama * std::pow(c->Ma, 2));
  o->Cv = o->Rg / (c->gama - 1.0);
  o->Cp = o->Cv + o->Rg;
  o->mu = 1.0;
  o->qk = o->mu / (o->Re * o->Pr * (c->gama - 1) * std::pow(c->Ma, 2));
  o->S_over_T_ref = 110.33333333 / T_ref;
  double *iv = o->initvar;   // roe u v w p T e
  iv[1] = 1.0; iv[2] = 0.0; iv[3] = 0.0;
  double speed = std::pow(std::pow(iv[1], 2) + std::pow(iv[2], 2) + std::pow(iv[3], 2), 0.5);
  iv[5] = std::pow(speed / c->Ma, 2) / (c->gama * o->Rg);
  iv[0] = 1.0;
  iv[4] = iv[0] * o->Rg * iv[5];
  iv[6] = o->Cv * iv[5] + 0.5 * (iv[1] * iv[1] + iv[2] * iv[2] + iv[3] * iv[3]);
  double ke = 0.5 * (iv[1] * iv[1] + iv[2] * iv[2] + iv[3] * iv[3]);
  o->inlet_T0 = ke * iv[0] / o->Cp + iv[5];
  o->inlet_p0 = iv[4] * std::pow(1.0 + 0.2 * std::pow(c->Ma, 2), 1.4 / 0.4);
  o->inlet_angle = 0.0;
  o->outlet_p = iv[4] + c->backpressure;
  o->cfl = c->cfl;
  o->dt_ratio = 1.0;
  o->dt_init = c->dt_init;
  const double mu_ref = 1.0 / c->Re;
  const double p_ref = roe_ref * u_ref * u_ref, area_ref = L_ref * L_ref;
  const double shear_ref = mu_ref * u_ref / L_ref;
  o->p_ref_area = p_ref * area_ref;
  o->shear_ref_area = shear_ref * area_ref;
  // F_ref = 0.5 u^2 roe (L^2 * span) (Userfunction_cy_naca15.cpp:439)
  o->force_ref = 0.5 * u_ref * u_ref * roe_ref * (L_ref * L_ref * c->span);
  o->moment_ref = o->force_ref * L_ref;
  // SA closure: initvar.viseddy = far_vis_ratio * viseddy with viseddy = 1
  // (Turbulence.cpp:14, FlowConstants.cpp:19), Pr_tur (FlowConstants.cpp:11)
  o->nut_inf = 5.0 * 1.0;
  o->nut_inlet = 0.0;
  o->Pr_tur = 0.9;
  return 0;
}

int32_t ibm_init_state(const ibm_handle *h, const ibg_physics *phys, double *fv) {
  if (!h || !phys || !fv) return 1;
  const int L = h->L;
  for (int c = 0; c < h->ncases; ++c) {
    int s = h->blk_start[c * L], e = h->blk_start[(c + 1) * L];
    for (int g = s; g < e; ++g)
      for (int v = 0; v < IBG_NFV; ++v) fv[(size_t)g * IBG_NFV + v] = phys[c].initvar[v];
  }
  return 0;
}

void ibm_naca4_offsets(double t, int32_t npts, double *off) {
  const double pi = 3.14159265358979323846;
  for (int s = 0; s < 2; ++s) {
    for (int i = 0; i < npts; ++i) {
      double x = 0.5 * (1.0 - std::cos(pi * i / (npts - 1)));
      double yt = 5.0 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x +
                             0.2843 * x * x * x - 0.1036 * x * x * x * x);
      off[2 * (s * npts + i)] = x - 0.25;    // about the quarter chord
      off[2 * (s * npts + i) + 1] = (s == 0) ? yt : -yt;
    }
  }
}

void ibm_pose(double xc, double yc, double aoa_deg, double *pose) {
  const double pi = 3.14159265358979323846;
  const double a = -aoa_deg * pi / 180.0;   // nose up: rotate clockwise
  pose[0] = xc; pose[1] = yc; pose[2] = std::cos(a); pose[3] = std::sin(a);
}

void ibm_naca4(double t, int32_t npts, double xc, double yc, double aoa_deg, double *xy) {
  std::vector<double> off((size_t)4 * npts);
  double P[4];
  ibm_naca4_offsets(t, npts, off.data());
  ibm_pose(xc, yc, aoa_deg, P);
  for (int k = 0; k < 2 * npts; ++k) {
    const double px = off[2 * k], py = off[2 * k + 1];
    const double rx = P[2] * px - P[3] * py, ry = P[3] * px + P[2] * py;
    xy[2 * k] = P[0] + rx;
    xy[2 * k + 1] = P[1] + ry;
  }
}

}  // extern "C"
