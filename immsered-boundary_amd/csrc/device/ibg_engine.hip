// ibg_engine.hip — MI355X (gfx950) residual engine behind include/ibgpu.h.
//
// Device layout.  All cells of all cases live in one index space reordered
// for the hot path: normal cells first, sorted by level from finest to
// coarsest (then case, then Morton order), ghost cells after.  Under the
// reference's temporal sub-cycling (mpiparams.cpp:186-212) the levels
// marched in a step are always the finest ones, so the marched cells, faces
// and tiles of every step are a prefix of their arrays and every kernel is
// one launch over a prefix.
//
// State is a packed per-cell record: W64 = (roe,u,v,p) [2D] or
// (roe,u,v,w)+p [3D] in FP64 is the master copy.  The flux stencils read
// one of three stencil buffers S0, S1, S2 (FP64 in the plain mode, where S0
// is the master itself; FP32 in the mixed mode).  Stage k reads one buffer
// and writes another (S0 -> S1 -> S2 -> ... -> S0, the last stage always
// back into S0), so no workgroup of a launch sees a cell another workgroup
// has already advanced.  Ghost fills write the master and every buffer.  T
// and e are derived (T = p/(roe Rg), e = Cv T + |u|^2/2, FlowVariables.H:
// 23,33).  The SA working variable, when the closure is on, has the same
// master + stencil buffer structure.
//
// Hot kernel: k_stage / k_stage_ml (ibg_kernels.h), one workgroup per tile
// of TILE Morton-consecutive normal cells (tiles given to the XCDs in
// contiguous runs); face fluxes live only in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "ibgpu.h"
#include "ibg_common.h"
#include "ibg_kernels.h"
#include "ibg_closure.h"
#include "ibg_move.h"
#include "ibg_reflux.h"

using namespace ibg;

struct ibg_ctx {
  ibg_config cfg{};
  int dev = 0;                 // HIP device current at ibg_create
  int ncu = 0;                 // its compute units
  std::map<std::pair<const void *, size_t>, int> occ;   // resident workgroups per CU
  hipStream_t s = nullptr;
  bool mixed = false;
  int D = 2, NV = 4, ncases = 0, L = 0, nblk = 0;
  int N = 0, Nn = 0, F = 0, ntiles = 0, tile = kTile, max_tile_faces = 1;
  int nbody = 0, npatch = 0, nib = 0;
  bool have_mesh = false;
  std::vector<int> h2d, d2h, fh2d;
  std::vector<int> cell_end, face_end, tile_end;
  std::vector<int> res_off, pro_off, blk_twod;
  std::vector<int16_t> hcellblk;
  std::vector<DevPhys> hphys;
  std::vector<double> body_center;
  std::vector<int> hbody_case;
  std::vector<int64_t> upd_per_lmin;
  int ndmg_edge = 0, ndmg_corner = 0;
  // device buffers (packed records)
  void *W64a = nullptr, *Rva = nullptr;
  double *W64e = nullptr;
  void *Rve = nullptr;
  int nbuf = 3;                                  // stencil buffers in use
  void *Sa[3] = {nullptr, nullptr, nullptr};     // stencil buffers (type R)
  void *Se[3] = {nullptr, nullptr, nullptr};
  bool turb = false;
  double *nu64 = nullptr;                        // SA working variable, master
  void *NUs[3] = {nullptr, nullptr, nullptr};    // its stencil buffers (type R)
  void *mutf = nullptr;                          // [Nn] (type R)
  double *nuprobe = nullptr, *mprobe = nullptr;
  // learned closure (separate allocations: they survive a mesh upload)
  struct Closure {
    bool ready = false, stats = false, on = false;
    long long cap = 0;
    float *X = nullptr, *Z = nullptr;
    int *cell = nullptr;
    unsigned long long *count = nullptr, *mlcount = nullptr;   // count[0]: rows stored
    int *tcount = nullptr;      // sampler: kept cells per tile
    long long *toff = nullptr;  // sampler: first row of each tile
    int ntc = 0;                // tiles the two arrays hold
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *lossp = nullptr, *gsum = nullptr;
    __bf16 *W1t = nullptr, *W2t = nullptr, *W2 = nullptr, *W2p = nullptr;
    float *w3 = nullptr, *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
    float *stat = nullptr;    // fmean[8] fstd[8] flo[8] fhi[8] zst[2], then 1/fstd[8]
    double *mom = nullptr;
    int step = 0, max_wg = 0;
    std::vector<void *> allocs;
  } cl;
  uint8_t *flags = nullptr;
  int *cellcase = nullptr, *nb27 = nullptr;
  TileInfo *tiles = nullptr;
  int *case_tiles = nullptr;   // [case][level] = {first tile, end tile}
  int4 *tst = nullptr, *tvs = nullptr;
  int8_t *tdir = nullptr;
  uint16_t *slots = nullptr;
  int4 *face_st = nullptr, *face_vst = nullptr;
  int8_t *face_dir = nullptr;
  int16_t *face_blk = nullptr;
  BlkInfo *blk = nullptr;
  DevPhys *phys = nullptr;
  int4 *dmg = nullptr;
  DmgWall *dmgw = nullptr;      // wall ghosts (BC -81), parallel to dmg; null without walls
  double *dist_dm = nullptr;    // [N] distance to the domain walls; null without walls
  int *res_ci = nullptr, *res_fi = nullptr, *pro_ci = nullptr, *pro_fi = nullptr;
  // optional coarse/fine flux balance: pairs, tile face -> balanced-flux
  // entry, the balanced inviscid / viscous fluxes [nrf][5][NV] (R-typed)
  int nrf = 0;
  RfPair *rfp = nullptr;
  int *tmod = nullptr;
  void *rbal = nullptr, *rbalv = nullptr;
  int *ib_cell = nullptr, *ib_intp = nullptr, *ib_body = nullptr, *ib_case = nullptr;
  double *ib_coef = nullptr, *ib_hgdis = nullptr, *ib_attach = nullptr, *ib_nv = nullptr,
         *signdis = nullptr;
  float *sd32 = nullptr;   // signdis in FP32 (the mixed SA source)
  int *pat_intp = nullptr, *pat_case = nullptr;
  double *pat_coef = nullptr, *pat_hgdis = nullptr, *pat_nv = nullptr, *pat_pc = nullptr,
         *pat_area = nullptr, *case_center = nullptr;
  Motion *motion = nullptr;
  int have_motion = 0;
  // moving bodies (2D sections on a fixed level set); device arrays live in
  // allocs with the mesh
  struct MoveState {
    bool on = false, law = false;
    int npoint = 0, cap = 0, cap_s = 0, cap_io = 0;
    int *key = nullptr, *cpatch = nullptr, *cp1 = nullptr, *nb8 = nullptr;
    int8_t *kind = nullptr, *fin = nullptr;
    int16_t *cblk = nullptr;
    double *origin = nullptr, *borg = nullptr, *sd1 = nullptr;
    double *pt_off = nullptr, *pat_dz = nullptr;
    int *pat_pt0 = nullptr, *pat_body = nullptr, *body_p0 = nullptr, *pat_cc = nullptr;
    int2 *pat_lh = nullptr;    // [npatch] patch range of the patch's body
    int4 *fbox = nullptr;      // [ncases] finest-level key box: ix0, iy0, nx, ny
    long long *foff = nullptr; // [ncases] first entry of the case's key grid
    int *fgrid = nullptr;      // finest-level cell of each key, -1 for none
    int8_t *pat_flip = nullptr;
    Pose *pose = nullptr;
    int *mk1 = nullptr, *mkib = nullptr;  // move stamps: layer 1, IB set
    int8_t *cmix = nullptr;               // an in-plane neighbour's body differs
    int *chg = nullptr;                   // cells whose body changed
    int cap_chg = 0;
    int *epoch = nullptr;                 // [1] move counter
    int *cnt = nullptr;                   // [6] solid, (spare), in->out, IB, changed, mixed
    int *mixlist = nullptr;               // [N] cells deferred to k_move_mixed
    int *slist = nullptr, *iolist = nullptr;
    int *gidx = nullptr;          // ghosts sorted by level, finest first
    std::vector<int> gend;        // [l]: ghosts of levels >= l
    unsigned long long *upd = nullptr, *lvlcnt = nullptr;
    PitchHeave *lawd = nullptr;
    Motion *motc = nullptr;       // per case, for the wall force
    int *bcase_d = nullptr;       // case of each body
    double *t = nullptr;
  } mv;
  double *dt = nullptr, *cflpart = nullptr, *rpart = nullptr, *resid = nullptr, *probe = nullptr;
  unsigned long long *err = nullptr;
  unsigned long long *err2 = nullptr;   // error word of the next step's move, run ahead
  int last_ntiles = 0;
  // captured step ranges of the time march, keyed by graph_key(): the
  // launch sequence of a range depends only on its first step's phase in
  // the level schedule, its length and whether it starts with a dt update
  std::map<long long, hipGraphExec_t> gexec;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // the next step's move runs ahead on s2, forked after the IB fill and dt
  // (evf) and joined at the next step's start (evj)
  hipStream_t s2 = nullptr;
  hipEvent_t evf = nullptr, evj = nullptr;
  double last_total_ms = 0.0, last_stage_ms = 0.0;
  int64_t last_stage_launches = 0, last_updates = 0;
  std::vector<void *> allocs;
  std::string errmsg;
  void *hstage = nullptr;      // pinned host staging of the state boundary
  size_t hstage_bytes = 0;

  template <int D> Pack<double, D> w64() const { return Pack<double, D>{(D4 *)W64a, W64e}; }
  MlpDev mlpdev() const {
    MlpDev M;
    M.W1t = cl.W1t; M.W2t = cl.W2t; M.W2 = cl.W2; M.W2p = cl.W2p; M.w3 = cl.w3;
    M.b1 = cl.b1; M.b2 = cl.b2; M.b3 = cl.b3;
    M.fmean = cl.stat; M.fstd = cl.stat + 8; M.flo = cl.stat + 16; M.fhi = cl.stat + 24;
    M.zst = cl.stat + 32;
    M.finv = cl.stat + 34;
    return M;
  }
  template <typename R, int D> Pack<R, D> rv() const {
    return Pack<R, D>{(typename V4<R>::T *)Rva, (R *)Rve};
  }
  template <typename R, int D> Pack<R, D> sbuf(int k) const {
    return Pack<R, D>{(typename V4<R>::T *)Sa[k], (R *)Se[k]};
  }
  template <typename R, int D> Pack<R, D> wstencil() const { return sbuf<R, D>(0); }
  template <typename R> R *nubuf(int k) const { return (R *)NUs[k]; }
  template <int D> StateOut<D> sout() const {
    StateOut<D> o{};
    o.err = err;
    o.m = w64<D>();
    o.nm = turb ? nu64 : nullptr;
    if (!mixed) {
      o.nd = nbuf - 1;
      if (nbuf > 1) { o.d1 = sbuf<double, D>(1); o.nd1 = (double *)NUs[1]; }
      if (nbuf > 2) { o.d2 = sbuf<double, D>(2); o.nd2 = (double *)NUs[2]; }
    } else {
      o.nf = nbuf;
      o.f0 = sbuf<float, D>(0); o.nf0 = (float *)NUs[0];
      if (nbuf > 1) { o.f1 = sbuf<float, D>(1); o.nf1 = (float *)NUs[1]; }
      if (nbuf > 2) { o.f2 = sbuf<float, D>(2); o.nf2 = (float *)NUs[2]; }
    }
    return o;
  }
};

// Staging memory of ibg_closure_infer_noalloc: one file-scope device array
// per GPU, held until the process ends, so that entry point calls no
// allocator.  Host code never passes the array's name to a kernel or copy:
// it uses the device address that ibg_create looks up once per device with
// hipGetSymbolAddress (g_scratch_dev).  Every context on a device shares
// that device's array; calls that overlap on one device race.
constexpr long long kScratchRows = 1 << 16;                  // feature rows per call
constexpr size_t kScratchFloats = (size_t)kScratchRows * (kMlpIn + 1);   // x[rows][8], then z[rows]
static __device__ __align__(256) float scratch[kScratchFloats];
constexpr int kMaxDevices = 64;
static float *g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

namespace {

// byte offset of the learned closure's weights under the persistent
// schedule: after the stage's face arrays of the largest tile
size_t ml_resident_offset(const ibg_ctx *c) {
  size_t per = 4 * sizeof(float) + (c->D == 3 ? sizeof(float) : 0);
  if (c->turb) per += 2 * sizeof(int);
  return ((size_t)c->max_tile_faces * per + 15) & ~(size_t)15;
}

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "ibgpu: %s failed: %s (%s:%d)\n", #x,              \
              hipGetErrorString(e_), __FILE__, __LINE__);                \
      return IBG_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

// Host loops over cells or faces (mesh upload, the state boundary): the
// range [0, n) split into contiguous chunks over up to 16 threads; f(lo, hi)
// must touch only its own entries.  Ranges shorter than serial_below run on
// the calling thread (per-cell loops: threads cost more than they save below
// ~64 K items; per-tile loops, 256 cells an item, pass a smaller bound).
template <typename F>
void host_parallel(int n, F f, int serial_below = 1 << 16) {
  const int nt = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
  if (n < serial_below || nt == 1) { f(0, n); return; }
  std::vector<std::thread> th;
  const int chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int a = t * chunk, b = std::min(n, a + chunk);
    if (a < b) th.emplace_back([=] { f(a, b); });
  }
  for (auto &x : th) x.join();
}

// The first index in [0, n) at which check(i) reports a failure (a non-null
// message), searched in parallel chunks: the same index and message as a
// serial scan.  Returns -1 when every index passes.
template <typename Check>
int first_failure(int n, Check check, const char **msg) {
  std::mutex mu;
  int best = -1;
  const char *bmsg = nullptr;
  host_parallel(n, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const char *e = check(i);
      if (e) {
        std::lock_guard<std::mutex> lk(mu);
        if (best < 0 || i < best) { best = i; bmsg = e; }
        return;
      }
    }
  });
  if (msg) *msg = bmsg;
  return best;
}

ibg_status dalloc_bytes(ibg_ctx *c, void **p, size_t bytes) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(bytes, 64)));
  HIPCHK(hipMemset(q, 0, std::max<size_t>(bytes, 64)));
  c->allocs.push_back(q);
  *p = q;
  return IBG_OK;
}

template <typename T>
ibg_status dalloc(ibg_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  ibg_status st = dalloc_bytes(c, &q, n * sizeof(T));
  *p = reinterpret_cast<T *>(q);
  return st;
}

// An allocator that leaves new elements uninitialised, for the large host
// staging arrays of an upload that a parallel loop fills in full.
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
template <class T> using hvec = std::vector<T, NoInitAlloc<T>>;

template <typename T, typename A>
ibg_status dupload(ibg_ctx *c, T **p, const std::vector<T, A> &v) {
  ibg_status st = dalloc(c, p, v.size());
  if (st) return st;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return IBG_OK;
}

void drop_graphs(ibg_ctx *c) {
  for (auto &g : c->gexec) (void)hipGraphExecDestroy(g.second);
  c->gexec.clear();
}

void free_all(ibg_ctx *c) {
  drop_graphs(c);
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  c->have_mesh = false;
  c->nrf = 0;
  c->rfp = nullptr; c->tmod = nullptr; c->rbal = nullptr; c->rbalv = nullptr;
  c->mv = ibg_ctx::MoveState{};
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

// LDS of the tiled kernels: face fluxes (4 values + the 3D fifth), then for
// the SA closure either the two face-cell indices (stage) or the face flux
// of the working variable (viscous pass; a fused first stage holds either).
size_t tile_smem(const ibg_ctx *c, size_t rsize, bool stage) {
  size_t per = 4 * rsize + (c->D == 3 ? rsize : 0);
  if (c->turb) per += stage ? std::max<size_t>(2 * sizeof(int), rsize) : rsize;
  const size_t b = (size_t)c->max_tile_faces * per;
  if (stage && c->cl.on) {
    // the network's weights: after every face array, resident across the
    // tiles (persistent schedule), or after the face fluxes over the
    // face-cell indices (one workgroup per tile)
    if (c->cfg.schedule == 1) return ml_resident_offset(c) + kStageMlSmem;
    const size_t m = (c->D == 2 ? ml_smem_offset<float, 2>(c->max_tile_faces)
                                : ml_smem_offset<float, 3>(c->max_tile_faces)) + kStageMlSmem;
    return std::max(b, m);
  }
  return b;
}

// Stencil buffers of stage k: S0 -> S1 -> S2 -> S1 ... with the last stage
// writing S0; a single stage writes S1 and is copied back.
void stage_bufs(int nstage, int subt, int *in, int *out) {
  int cur = 0;
  for (int k = 0;; ++k) {
    const int o = (k == nstage - 1 && nstage > 1) ? 0 : (cur == 1 ? 2 : 1);
    if (k == subt) { *in = cur; *out = o; return; }
    cur = o;
  }
}

template <typename R, int D>
ViscArgs<R, D> visc_args(ibg_ctx *c) {
  ViscArgs<R, D> a;
  a.tiles = c->tiles; a.tst = c->tst; a.tvs = c->tvs; a.tdir = c->tdir; a.slots = c->slots;
  a.blk = c->blk; a.phys = c->phys;
  a.W = c->sbuf<R, D>(0);
  a.Rv = c->rv<R, D>();
  a.nu = c->nubuf<R>(0);
  a.mutf = (R *)c->mutf;
  a.viscous = c->cfg.viscous;
  a.turb = c->turb ? 1 : 0;
  a.err = c->err;
  a.tmod = c->tmod;
  a.rbalv = (c->cfg.reflux && c->nrf > 0) ? (const R *)c->rbalv : nullptr;
  return a;
}

template <typename R, int D, bool ML>
ibg_status launch_stage(ibg_ctx *c, int ntile, int ncell, int subt, bool last, double *src,
                        double *nusrc, bool fuse) {
  if (ntile <= 0) return IBG_OK;
  int bin = 0, bout = 0;
  stage_bufs(c->cfg.nstage, subt, &bin, &bout);
  StageArgs<R, D> a;
  a.tiles = c->tiles; a.tst = c->tst; a.tdir = c->tdir; a.slots = c->slots;
  a.flags = c->flags; a.blk = c->blk; a.phys = c->phys; a.dt = c->dt;
  a.Win = c->sbuf<R, D>(bin);
  a.Wout = c->sbuf<R, D>(bout);
  a.W64 = c->w64<D>();
  a.Rv = c->rv<R, D>();
  a.turb = c->turb ? 1 : 0;
  a.nuin = c->nubuf<R>(bin);
  a.nuout = c->nubuf<R>(bout);
  a.nu64 = c->nu64;
  a.mutf = (const R *)c->mutf;
  a.signdis = c->signdis;
  a.sd32 = c->sd32;
  a.mlp = c->mlpdev();
  a.mlcount = c->cl.mlcount;
  a.alpha = c->cfg.alpha[subt];
  a.limiter = c->cfg.limiter;
  a.last = last ? 1 : 0;
  a.rpart = c->rpart;
  a.src = src;
  a.nusrc = nusrc;
  a.err = c->err;
  // the step's first stage reads the home buffer S0, the state k_visc_prep reads
  a.fuse_visc = (fuse && subt == 0 && kFuseVisc<R, D>) ? 1 : 0;
  a.va = visc_args<R, D>(c);
  a.tmod = c->tmod;
  a.rbal = nullptr;
  if (c->cfg.reflux && last && c->nrf > 0) {
    // BalanceFaceFlux on the last stage's inviscid fluxes (NS_Solver.H:686)
    hipLaunchKernelGGL((k_balance_inv<R, D>), dim3(grid(c->nrf, 256)), dim3(256), 0, c->s, c->rfp,
                       c->nrf, (const int4 *)c->tst, (const int8_t *)c->tdir, a.Win,
                       (const BlkInfo *)c->blk, (const DevPhys *)c->phys, c->cfg.limiter,
                       (R *)c->rbal, (const unsigned long long *)c->err);
    HIPCHK(hipGetLastError());
    a.rbal = (const R *)c->rbal;
  }
  const size_t smem = tile_smem(c, sizeof(R), true);
  a.ntile = ntile;
  a.mlw_off = (unsigned)ml_resident_offset(c);
  if (c->cfg.schedule == 1 && ntile >= 8) {
    // persistent schedule: as many workgroups as the device holds at once
    // (at least one per XCD, so each XCD's run of tiles has a worker)
    const void *kfn = ML ? (const void *)k_stage_ml<D, true> : (const void *)k_stage<R, D, false, true>;
    auto key = std::make_pair(kfn, smem);
    auto it = c->occ.find(key);
    if (it == c->occ.end()) {
      int nb = 0;
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, c->tile, smem));
      it = c->occ.emplace(key, std::max(nb, 1)).first;
    }
    const int grid_n = std::max(8, std::min(ntile, it->second * c->ncu));
    if constexpr (ML)
      hipLaunchKernelGGL((k_stage_ml<D, true>), dim3(grid_n), dim3(c->tile), smem, c->s, a);
    else
      hipLaunchKernelGGL((k_stage<R, D, false, true>), dim3(grid_n), dim3(c->tile), smem, c->s, a);
  } else {
    if constexpr (ML)
      hipLaunchKernelGGL((k_stage_ml<D, false>), dim3(ntile), dim3(c->tile), smem, c->s, a);
    else
      hipLaunchKernelGGL((k_stage<R, D, false, false>), dim3(ntile), dim3(c->tile), smem, c->s, a);
  }
  HIPCHK(hipGetLastError());
  if (c->cfg.nstage == 1) {
    // single-stage scheme: the marched cells go back into the home buffer
    hipLaunchKernelGGL((k_copy_back<R, D>), dim3(grid(ncell, 256)), dim3(256), 0, c->s, ncell,
                       c->sbuf<R, D>(bout), c->sbuf<R, D>(0), c->nubuf<R>(bout),
                       c->turb ? c->nubuf<R>(0) : (R *)nullptr, (const unsigned long long *)c->err);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

// fuse: the stage-0 launch also computes the frozen viscous residual (the
// time march does this; ibg_stage keeps the separate ibg_viscous_flux call)
ibg_status run_stage(ibg_ctx *c, int ts, int subt, double *src, double *nusrc, bool fuse = false) {
  const int lm = lmin_of(c, ts);
  const int nt = c->tile_end[lm], ncell = c->cell_end[lm];
  const bool last = (subt == c->cfg.nstage - 1);
  if (last) c->last_ntiles = nt;
  if (c->cl.on)   // learned closure: mixed mode with the SA closure only
    return c->D == 2 ? launch_stage<float, 2, true>(c, nt, ncell, subt, last, src, nusrc, fuse)
                     : launch_stage<float, 3, true>(c, nt, ncell, subt, last, src, nusrc, fuse);
  if (c->D == 2)
    return c->mixed ? launch_stage<float, 2, false>(c, nt, ncell, subt, last, src, nusrc, fuse)
                    : launch_stage<double, 2, false>(c, nt, ncell, subt, last, src, nusrc, fuse);
  return c->mixed ? launch_stage<float, 3, false>(c, nt, ncell, subt, last, src, nusrc, fuse)
                  : launch_stage<double, 3, false>(c, nt, ncell, subt, last, src, nusrc, fuse);
}

// whether the time march folds k_visc_prep into the first stage
inline bool fused_visc(const ibg_ctx *c) { return c->D == 2 && !c->mixed; }

// BalanceFaceFlux on the step's viscous fluxes (NS_Solver.H:672), before
// the viscous residual that reads them
template <typename R, int D>
ibg_status launch_visc_balance(ibg_ctx *c) {
  if (!c->cfg.reflux || !c->cfg.viscous || c->nrf <= 0) return IBG_OK;
  hipLaunchKernelGGL((k_balance_visc<R, D>), dim3(grid(c->nrf, 256)), dim3(256), 0, c->s,
                     visc_args<R, D>(c), (const RfPair *)c->rfp, c->nrf, (R *)c->rbalv);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

ibg_status run_visc_balance(ibg_ctx *c) {
  if (c->D == 2)
    return c->mixed ? launch_visc_balance<float, 2>(c) : launch_visc_balance<double, 2>(c);
  return c->mixed ? launch_visc_balance<float, 3>(c) : launch_visc_balance<double, 3>(c);
}

template <typename R, int D>
ibg_status launch_visc_prep(ibg_ctx *c, int ntile) {
  if (ntile <= 0) return IBG_OK;
  ibg_status st;
  if ((st = launch_visc_balance<R, D>(c))) return st;
  const ViscArgs<R, D> a = visc_args<R, D>(c);
  hipLaunchKernelGGL((k_visc_prep<R, D>), dim3(ntile), dim3(c->tile), tile_smem(c, sizeof(R), false), c->s, a);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

ibg_status run_visc_prep(ibg_ctx *c, int ts) {
  const int nt = c->tile_end[lmin_of(c, ts)];
  if (c->D == 2)
    return c->mixed ? launch_visc_prep<float, 2>(c, nt) : launch_visc_prep<double, 2>(c, nt);
  return c->mixed ? launch_visc_prep<float, 3>(c, nt) : launch_visc_prep<double, 3>(c, nt);
}

template <int D>
ibg_status launch_bc(ibg_ctx *c) {
  if (c->ndmg_edge > 0) {
    hipLaunchKernelGGL((k_bc<D>), dim3(grid(c->ndmg_edge, 256)), dim3(256), 0, c->s, c->ndmg_edge,
                       c->dmg, (const DmgWall *)c->dmgw, c->sout<D>(), c->phys);
    HIPCHK(hipGetLastError());
  }
  if (c->ndmg_corner > 0) {
    hipLaunchKernelGGL((k_bc<D>), dim3(grid(c->ndmg_corner, 256)), dim3(256), 0, c->s,
                       c->ndmg_corner, c->dmg + c->ndmg_edge,
                       c->dmgw ? (const DmgWall *)c->dmgw + c->ndmg_edge : (const DmgWall *)nullptr,
                       c->sout<D>(), c->phys);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

ibg_status run_bc(ibg_ctx *c, int ts) {
  // domain ghosts live on level 0; BCBoxTreat runs on marched levels only
  if (!marched(c, 0, ts)) return IBG_OK;
  return c->D == 2 ? launch_bc<2>(c) : launch_bc<3>(c);
}

template <int D>
ibg_status launch_exchange(ibg_ctx *c, int ts) {
  for (int l = 0; l + 1 < c->L; ++l) {
    if (!transfer_due(c, l, ts)) continue;
    const int n = c->res_off[l + 1] - c->res_off[l];
    if (n > 0) {
      hipLaunchKernelGGL((k_restrict<D>), dim3(grid(n, 256)), dim3(256), 0, c->s, n,
                         c->res_ci + c->res_off[l], c->res_fi + (size_t)8 * c->res_off[l],
                         c->sout<D>(), c->blk_twod[l + 1], c->phys, c->cellcase);
      HIPCHK(hipGetLastError());
    }
  }
  for (int l = 0; l + 1 < c->L; ++l) {
    if (!transfer_due(c, l, ts)) continue;
    const int n = c->pro_off[l + 1] - c->pro_off[l];
    if (n > 0) {
      hipLaunchKernelGGL((k_prolong<D>), dim3(grid(n, 256)), dim3(256), 0, c->s, n,
                         c->pro_ci + c->pro_off[l], c->pro_fi + (size_t)8 * c->pro_off[l],
                         c->nb27, c->sout<D>(), c->blk_twod[l + 1], c->phys, c->cellcase);
      HIPCHK(hipGetLastError());
    }
  }
  return IBG_OK;
}

ibg_status run_exchange(ibg_ctx *c, int ts) {
  return c->D == 2 ? launch_exchange<2>(c, ts) : launch_exchange<3>(c, ts);
}

template <int D>
ibg_status launch_ib(ibg_ctx *c) {
  const int n = c->mv.on ? c->mv.cap : c->nib;
  if (n > 0) {
    hipLaunchKernelGGL((k_ib<D>), dim3(grid(n, 256)), dim3(256), 0, c->s, n,
                       c->mv.on ? c->mv.cnt + 3 : (const int *)nullptr, c->ib_cell,
                       c->ib_intp, c->ib_coef, c->ib_hgdis, c->ib_attach, c->ib_nv, c->ib_body,
                       c->ib_case, c->signdis, c->motion, c->have_motion, c->sout<D>(), c->phys,
                       c->err);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

ibg_status run_ib(ibg_ctx *c) { return c->D == 2 ? launch_ib<2>(c) : launch_ib<3>(c); }

MoveArgs move_args(ibg_ctx *c, unsigned long long *upd, int lmin, bool ahead = false) {
  MoveArgs a{};
  ibg_ctx::MoveState &v = c->mv;
  a.N = c->N; a.Nn = c->Nn; a.npatch = c->npatch; a.ncases = c->ncases; a.nbody = c->nbody;
  a.nw = c->cell_end[lmin]; a.ngw = v.gend[lmin]; a.nflag = c->cell_end[lmin];
  a.gidx = v.gidx;
  a.key = v.key; a.kind = v.kind; a.fin = v.fin; a.cblk = v.cblk; a.blk = c->blk;
  a.origin = v.origin; a.borg = v.borg; a.nb8 = v.nb8; a.nb27 = c->nb27; a.dist_dm = c->dist_dm;
  a.signdis = c->signdis; a.sd32 = c->sd32; a.cpatch = v.cpatch; a.cp1 = v.cp1; a.sd1 = v.sd1;
  a.mk1 = v.mk1; a.mkib = v.mkib; a.epoch = v.epoch; a.cnt = v.cnt;
  a.cmix = v.cmix; a.chg = v.chg; a.cap_chg = v.cap_chg;
  a.mixlist = v.mixlist; a.cap_mix = c->N;
  a.cap_s = v.cap_s; a.cap_io = v.cap_io; a.ib_cap = v.cap;
  a.slist = v.slist; a.iolist = v.iolist;
  a.flags = c->flags;
  a.pt_off = v.pt_off; a.pat_pt0 = v.pat_pt0; a.pat_body = v.pat_body; a.pat_case = c->pat_case;
  a.pat_flip = v.pat_flip; a.pat_dz = v.pat_dz; a.body_p0 = v.body_p0; a.body_case = v.bcase_d;
  a.pat_lh = v.pat_lh;
  a.fbox = v.fbox; a.foff = v.foff; a.fgrid = v.fgrid;
  a.pose = v.pose; a.use_law = v.law ? 1 : 0; a.law = v.lawd; a.t = v.t; a.dt = c->dt;
  a.ahead = ahead ? 1 : 0;
  a.mot = c->motion; a.motc = v.motc; a.case_center = c->case_center;
  a.pat_pc = c->pat_pc; a.pat_nv = c->pat_nv; a.pat_area = c->pat_area;
  a.pat_intp = c->pat_intp; a.pat_cc = v.pat_cc; a.pat_coef = c->pat_coef; a.pat_hgdis = c->pat_hgdis;
  a.ib_cell = c->ib_cell; a.ib_intp = c->ib_intp; a.ib_body = c->ib_body; a.ib_case = c->ib_case;
  a.ib_coef = c->ib_coef; a.ib_hgdis = c->ib_hgdis; a.ib_attach = c->ib_attach; a.ib_nv = c->ib_nv;
  a.err = ahead ? c->err2 : c->err; a.upd = upd; a.lvlcnt = v.lvlcnt; a.nstage = c->cfg.nstage;
  return a;
}

// The launches of one move up to the commit, on stream st.
void launch_move(ibg_ctx *c, const MoveArgs &a, hipStream_t st, bool commit) {
  const int nw = a.nw + a.ngw;
  hipLaunchKernelGGL(k_move_front, dim3(1), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_move_dist, dim3(grid(nw, 256)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_move_mixed, dim3(256), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_move_mark, dim3(grid(kNb8 * a.cap_s + a.cap_chg, 256)), dim3(256), 0, st, a);
  const int nimage = (int)grid(a.ib_cap + c->npatch, 256);
  // with the commit: its workgroups also check; without: a few check
  const int nrest = commit ? (int)std::min(grid(nw, 256), 2048u)
                           : (int)std::min(grid(a.cap_io, 256), 64u);
  hipLaunchKernelGGL(k_move_finish, dim3(nimage + std::max(nrest, 1)), dim3(256), 0, st, a, nimage,
                     commit ? 1 : 0);
}

// One move of the bodies (to the poses in mv.pose, or by the law at each
// case's time): surface, distances, solid and IB sets, image points, stage
// masks of the levels marched from lmin.  Stream-ordered, capturable.
ibg_status run_move(ibg_ctx *c, unsigned long long *upd, int lmin) {
  if (c->D != 2) return IBG_ERR_STATE;
  launch_move(c, move_args(c, upd, lmin), c->s, true);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

// The move of step ts (by the law, at t + dt) ahead of the time advance of
// the step before, on s2 beside that step's remaining launches: it waits
// for the fork point c->evf, which the step records after its IB fill and
// dt, and run_move_join joins it at step ts.  Nothing the step launches
// after the fork reads what the move writes before its commit, or writes
// what the move reads.  The step issues these launches after its viscous
// pass: in a captured graph the executor then keeps the step's own chain
// on one hardware queue (a chain that changes queue waits ~10 us per
// change; 0.506 -> 0.494 ms per step of the 2.6M-cell moving section).
ibg_status run_move_ahead(ibg_ctx *c, int ts) {
  if (c->D != 2 || !c->mv.law) return IBG_ERR_STATE;
  HIPCHK(hipStreamWaitEvent(c->s2, c->evf, 0));
  launch_move(c, move_args(c, c->mv.upd, lmin_of(c, ts), true), c->s2, false);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->evj, c->s2));
  return IBG_OK;
}

// The start of step ts after run_move_ahead(ts): the time advance of the
// step before, then the commit of the move (k_move_join).
ibg_status run_move_join(ibg_ctx *c, int ts) {
  HIPCHK(hipStreamWaitEvent(c->s, c->evj, 0));
  const MoveArgs a = move_args(c, c->mv.upd, lmin_of(c, ts));
  const int ncommit = (int)std::min(grid(a.nw + a.ngw, 256), 2048u);
  hipLaunchKernelGGL(k_move_join, dim3(std::max(ncommit, 1)), dim3(256), 0, c->s, a,
                     (const unsigned long long *)c->err2, c->mv.t);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

ibg_status run_dt(ibg_ctx *c) {
  if (c->ntiles > 0) {
    if (c->D == 2)
      hipLaunchKernelGGL((k_cfl<2>), dim3(c->ntiles), dim3(c->tile), 0, c->s, c->tiles, c->w64<2>(),
                         c->blk, c->phys, c->cflpart, c->err);
    else
      hipLaunchKernelGGL((k_cfl<3>), dim3(c->ntiles), dim3(c->tile), 0, c->s, c->tiles, c->w64<3>(),
                         c->blk, c->phys, c->cflpart, c->err);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_dt_final, dim3(c->ncases), dim3(256), 0, c->s, c->L, c->case_tiles,
                     c->cflpart, c->dt, c->phys, c->ncases, (const unsigned long long *)c->err);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

// one full step: main.cpp:87-97.  With the prescribed law, `joined`: the
// step's move ran ahead during the step before (run_move_ahead), `ahead`:
// the next step's move runs ahead during this one.
ibg_status run_step(ibg_ctx *c, int ts, bool joined, bool ahead) {
  ibg_status st;
  if ((st = run_exchange(c, ts))) return st;
  if (c->mv.law) {
    // MovingBody_Image at the start of IBConditions_Image (Body.cpp:1353-1357)
    if (joined) st = run_move_join(c, ts);
    else st = run_move(c, c->mv.upd, lmin_of(c, ts));
    if (st) return st;
  }
  if ((st = run_ib(c))) return st;
  if (c->cfg.dt_var_step > 0 && ts % c->cfg.dt_var_step == 0)
    if ((st = run_dt(c))) return st;
  if (c->mv.law && ahead) HIPCHK(hipEventRecord(c->evf, c->s));   // the fork point
  if ((st = run_bc(c, ts))) return st;
  const bool fuse = fused_visc(c);
  if (!fuse) {
    if ((st = run_visc_prep(c, ts))) return st;
  } else {
    // the stage-0 launch computes the viscous residual: balance first
    if ((st = run_visc_balance(c))) return st;
  }
  if (c->mv.law && ahead)
    if ((st = run_move_ahead(c, ts + 1))) return st;
  for (int k = 0; k < c->cfg.nstage; ++k)
    if ((st = run_stage(c, ts, k, nullptr, nullptr, fuse))) return st;
  if (c->mv.law && !ahead) {
    // (with the next move ahead, its join advances the time)
    hipLaunchKernelGGL(k_motion_advance, dim3(grid(c->ncases, 64)), dim3(64), 0, c->s, c->ncases,
                       c->mv.t, (const double *)c->dt, (const unsigned long long *)c->err);
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}

ibg_status reset_err(ibg_ctx *c) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->s, c->err, 1, ~0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->s, c->err2, 1, ~0ull);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

ibg_status read_err(ibg_ctx *c, int64_t *cell, int32_t *code) {
  unsigned long long v = 0;
  HIPCHK(hipMemcpyAsync(&v, c->err, sizeof(v), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (v == ~0ull) { *cell = -1; *code = 0; return IBG_OK; }
  const int dcell = (int)(v >> 16);
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
  // the cells of the normal ranges are the first-ranked of their blocks
  // (checked above), so cell g is normal exactly when it lies below its
  // block's normal count
  std::vector<int> gblk(N);
  for (int b = 0; b < nblk; ++b)
    for (int g = m->blk_start[b]; g < m->blk_start[b + 1]; ++g) gblk[g] = b;
  const char *msg = nullptr;
  if (first_failure(N, [&](int g) -> const char * {
        const int b = gblk[g];
        if (g >= m->blk_start[b] + m->blk_nnormal[b]) return nullptr;
        for (int d = 0; d < D; ++d) {
          for (int s2 = 0; s2 < 2; ++s2) {
            const int f = m->cell_faces[(size_t)g * 6 + d * 2 + s2];
            if (f < 0 || f >= m->nfaces) return "normal cell without face";
            if (m->face_cells[2 * f + (1 - s2)] != g) return "cell_faces and face_cells disagree";
          }
          // the other cells of the two faces are the Box::neib neighbours
          // (the SA source gathers them from the face list, NS_Solver.cpp:
          // 662-666)
          int e[3] = {1, 1, 1};
          e[d] = 0;
          const int lo = nb(g, e[0], e[1], e[2]);
          e[d] = 2;
          const int hi = nb(g, e[0], e[1], e[2]);
          if (m->face_cells[2 * m->cell_faces[(size_t)g * 6 + d * 2]] != lo ||
              m->face_cells[2 * m->cell_faces[(size_t)g * 6 + d * 2 + 1] + 1] != hi)
            return "face list and neighbour table disagree";
        }
        return nullptr;
      }, &msg) >= 0)
    return fail(msg);
  if (first_failure(m->nfaces, [&](int f) -> const char * {
        const int lo = m->face_cells[2 * f], hi = m->face_cells[2 * f + 1], d = m->face_dir[f];
        if (!in(lo) || !in(hi) || d < 0 || d >= D) return "face out of range";
        int o[3] = {1, 1, 1};
        o[d] = 0;
        int LL = nb(lo, o[0], o[1], o[2]);
        o[d] = 2;
        int RR = nb(hi, o[0], o[1], o[2]);
        int LLf = LL, RRf = RR;
        if (LL < 0) LLf = RR;
        if (RR < 0) RRf = LL;
        if (!in(LLf) || !in(RRf)) return "face stencil without LL/RR";
        // viscous stencil of the low cell (NS_Solver.cpp:413-425)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
              int idx[3] = {i, j, k};
              if (idx[d] == 0) continue;
              if (D == 2 && k != 1) continue;
              if (!in(nb(lo, i, j, k))) return "viscous stencil neighbour missing";
            }
        return nullptr;
      }, &msg) >= 0)
    return fail(msg);
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
  if ((m->dmg_nbr == nullptr) != (m->dist_dm == nullptr) || (m->dmg_nbr == nullptr) != (m->dmg_nmv == nullptr))
    return fail("domain wall arrays incomplete");
  for (int i = 0; i < m->ndmg; ++i) {
    if (m->dmg_type[i] != -81) continue;
    if (!m->dmg_nbr) return fail("wall ghost without its reference neighbour");
    if (!in(m->dmg_nbr[i])) return fail("wall ghost reference neighbour out of range");
  }
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
  if (m->nmdy < 0 || (m->nmdy > 0 && !m->mdy)) return fail("balance face list");
  for (int i = 0; i < m->nmdy; ++i) {
    const int32_t *r = m->mdy + (size_t)i * 6;
    if (r[0] < 0 || r[0] > 2) return fail("balance face form");
    const int nf = rf_nfine(r[0]);
    if (r[1] < 0 || r[1] >= m->nfaces) return fail("balance coarse face out of range");
    for (int k = 0; k < 4; ++k) {
      const int f = r[2 + k];
      if (k < nf ? (f < 0 || f >= m->nfaces) : f != -1) return fail("balance fine face out of range");
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

const char *ibg_version(void) { return "ibgpu 0.2 gfx950"; }

ibg_status ibg_check_mesh(const ibg_mesh_soa *m, const char **msg) {
  ibg_status st = validate_mesh(m);
  if (msg) *msg = st ? g_check_msg.c_str() : "";
  return st;
}

ibg_status ibg_create(const ibg_config *cfg, ibg_ctx **out) {
  if (!cfg || !out) return IBG_ERR_ARG;
  if (!(cfg->nstage == 1 || cfg->nstage == 3 || cfg->nstage == 4)) return IBG_ERR_ARG;
  if (cfg->mode != IBG_MODE_FP64 && cfg->mode != IBG_MODE_MIXED) return IBG_ERR_ARG;
  if (cfg->turbulence != 0 && cfg->turbulence != 1) return IBG_ERR_ARG;
  if (cfg->turbulence && !cfg->viscous) return IBG_ERR_ARG;
  if (cfg->schedule != 0 && cfg->schedule != 1) return IBG_ERR_ARG;
  // BalanceFaceFlux exists only without TEMPORAL_REFINE (Fluxmodification.cpp:293)
  if (cfg->reflux != 0 && cfg->reflux != 1) return IBG_ERR_ARG;
  if (cfg->reflux && cfg->temporal_refine) return IBG_ERR_ARG;
  const int tile = cfg->tile > 0 ? cfg->tile : kTile;
  if (tile % 64 != 0 || tile > kTile) return IBG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    fprintf(stderr, "ibgpu: no HIP device\n");
    return IBG_ERR_HIP;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return IBG_ERR_HIP;
  {
    // the staging array's address on this device (once per device)
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        fprintf(stderr, "ibgpu: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
        return IBG_ERR_HIP;
      }
      g_scratch_dev[dev] = (float *)p;
    }
  }
  ibg_ctx *c = new ibg_ctx();
  c->cfg = *cfg;
  c->dev = dev;
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      c->ncu < 1)
    c->ncu = 1;
  c->mixed = (cfg->mode == IBG_MODE_MIXED);
  c->turb = (cfg->turbulence != 0);
  c->nbuf = (cfg->nstage == 1) ? 2 : 3;
  c->tile = tile;
  // default stream priority (no priority streams)
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) { delete c; return IBG_ERR_HIP; }
  if (hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->evf, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->evj, hipEventDisableTiming) != hipSuccess) {
    if (c->evf) (void)hipEventDestroy(c->evf);
    if (c->evj) (void)hipEventDestroy(c->evj);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    (void)hipStreamDestroy(c->s);
    delete c;
    return IBG_ERR_HIP;
  }
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  *out = c;
  return IBG_OK;
}

ibg_status ibg_destroy(ibg_ctx *c) {
  if (!c) return IBG_ERR_ARG;
  (void)hipStreamSynchronize(c->s);
  free_all(c);
  for (void *p : c->cl.allocs) (void)hipFree(p);
  c->cl.allocs.clear();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evf) (void)hipEventDestroy(c->evf);
  if (c->evj) (void)hipEventDestroy(c->evj);
  if (c->s2) (void)hipStreamDestroy(c->s2);
  if (c->s) (void)hipStreamDestroy(c->s);
  if (c->hstage) (void)hipHostFree(c->hstage);
  delete c;
  return IBG_OK;
}

ibg_status ibg_upload_mesh(ibg_ctx *c, const ibg_mesh_soa *m, const ibg_physics *ph) {
  if (!c || !m || !ph) return IBG_ERR_ARG;
  {
    ibg_status vs = validate_mesh(m);
    if (vs) {
      c->errmsg = g_check_msg;
      fprintf(stderr, "ibgpu: mesh rejected: %s\n", g_check_msg.c_str());
      return vs;
    }
  }
  if (m->ncases > IBG_MAX_CASES || m->nlevels > IBG_MAX_LEVELS) return IBG_ERR_ARG;
  if (!c->turb)
    for (int i = 0; i < m->ndmg; ++i)
      if (m->dmg_type[i] == -81) {
        // the laminar build has no wall boundary: BCBoxTreat stops with its
        // unidentified-type abort (BCValues.H:170-176)
        c->errmsg = "domain wall (-81) needs the SA closure";
        fprintf(stderr, "ibgpu: mesh rejected: %s\n", c->errmsg.c_str());
        return IBG_ERR_ARG;
      }
  if (c->turb && m->nib > 0 && !m->ib_nv) { c->errmsg = "SA closure needs ib_nv"; return IBG_ERR_ARG; }
  (void)hipStreamSynchronize(c->s);
  free_all(c);
  const int D = m->dim, NV = D + 2, L = m->nlevels, nc = m->ncases, nblk = nc * L;
  c->D = D; c->NV = NV; c->L = L; c->ncases = nc; c->nblk = nblk;
  const int N = m->ncells;
  c->N = N;
  std::vector<int> hblk(N);
  for (int b = 0; b < nblk; ++b)
    std::fill(hblk.begin() + m->blk_start[b], hblk.begin() + m->blk_start[b + 1], b);
  // ---- device cell order: normal cells by (level desc, case, Morton), then ghosts
  std::vector<int> order;
  order.reserve(N);
  c->cell_end.assign(L, 0);
  for (int l = L - 1; l >= 0; --l) {
    for (int cs = 0; cs < nc; ++cs) {
      const int b = cs * L + l;
      for (int g = m->blk_start[b]; g < m->blk_start[b] + m->blk_nnormal[b]; ++g) order.push_back(g);
    }
    c->cell_end[l] = (int)order.size();
  }
  const int Nn = (int)order.size();
  c->Nn = Nn;
  for (int b = 0; b < nblk; ++b)
    for (int g = m->blk_start[b] + m->blk_nnormal[b]; g < m->blk_start[b + 1]; ++g) order.push_back(g);
  c->d2h = order;
  c->h2d.assign(N, -1);
  host_parallel(N, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) c->h2d[order[i]] = i;
  });
  const std::vector<int> &h2d = c->h2d;
  // ---- faces: device order by (level desc, case)
  const int F = m->nfaces;
  c->F = F;
  std::vector<int> forder;
  forder.reserve(F);
  c->face_end.assign(L, 0);
  for (int l = L - 1; l >= 0; --l) {
    for (int cs = 0; cs < nc; ++cs) {
      const int b = cs * L + l;
      for (int f = m->blk_face_start[b]; f < m->blk_face_start[b + 1]; ++f) forder.push_back(f);
    }
    c->face_end[l] = (int)forder.size();
  }
  c->fh2d.assign(F, -1);
  for (int i = 0; i < F; ++i) c->fh2d[forder[i]] = i;
  // ---- blocks and physics
  std::vector<BlkInfo> hb(nblk);
  c->blk_twod.assign(L, 0);
  for (int b = 0; b < nblk; ++b) {
    BlkInfo &B = hb[b];
    const double *h = &m->blk_h[3 * b];
    for (int d = 0; d < 3; ++d) { B.h[d] = h[d]; B.invh[d] = 1.0 / h[d]; }
    // ComptFaceArea MeshOp.H:17-36, ComptCellParams MeshOp.H:4-12
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
    P.nut_inf = p.nut_inf; P.nut_inlet = p.nut_inlet;
    P.qkt_den = p.Re * p.Pr_tur * (p.gama - 1) * pow(p.Ma, 2);
    P.pr_cuberoot = pow(p.Pr, 1.0 / 3.0);
  }
  // ---- per-cell arrays (device order)
  std::vector<int16_t> cellblk(N);
  std::vector<int> cellcase(N);
  hvec<int> nb27((size_t)N * 27);
  std::vector<double> sd(N);
  host_parallel(N, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const int g = order[i];
      cellblk[i] = (int16_t)hblk[g];
      cellcase[i] = hblk[g] / L;
      sd[i] = m->signdis[g];
      for (int q = 0; q < 27; ++q) {
        const int a = m->nb[(size_t)g * 27 + q];
        nb27[(size_t)i * 27 + q] = a < 0 ? -1 : h2d[a];
      }
    }
  });
  std::vector<uint8_t> flags(Nn, 0);
  host_parallel(Nn, [&](int lo, int hi) {
  for (int i = lo; i < hi; ++i) {
    const int g = order[i];
    const BlkInfo &B = hb[hblk[g]];
    const bool ib = m->ibidx[g] >= 0;
    const bool solid = m->signdis[g] < 0.0;
    uint8_t f = 0;
    if (!B.finest) {
      // every normal cell of a coarser level (NS_Solver.cpp:856-937, 19-113)
      f = F_UPDATE | F_RESID | ((m->signdis[g] > 0.0) ? F_PRIM : F_FREE);
    } else {
      // finest level with bodies (NS_Solver.cpp:938-983, 114-179)
      f = F_FINEST;
      if (!solid && !ib) f |= F_UPDATE;
      if (m->signdis[g] > 0.0 && !ib) f |= F_PRIM | F_RESID;
      else if (!ib) f |= F_FREE;
    }
    if (m->signdis[g] >= 0.0) f |= F_SA;
    flags[i] = f;
  }
  });
  hvec<int> cfaces((size_t)Nn * 2 * D);
  host_parallel(Nn, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const int g = order[i];
      for (int d = 0; d < D; ++d)
        for (int s2 = 0; s2 < 2; ++s2)
          cfaces[(size_t)i * 2 * D + 2 * d + s2] = c->fh2d[m->cell_faces[(size_t)g * 6 + d * 2 + s2]];
    }
  });
  // ---- faces: stencils
  auto nbo = [&](int dc, int dx, int dy, int dz) {
    return nb27[(size_t)dc * 27 + (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)];
  };
  hvec<int4> face_st(F), face_vst((size_t)F * (D - 1));
  std::vector<int8_t> face_dir(F);
  std::vector<int16_t> face_blk(F);
  host_parallel(F, [&](int ilo, int ihi) {
  for (int i = ilo; i < ihi; ++i) {
    const int f = forder[i];
    const int lo = h2d[m->face_cells[2 * f]], hi = h2d[m->face_cells[2 * f + 1]];
    const int d = m->face_dir[f];
    face_dir[i] = (int8_t)d;
    face_blk[i] = (int16_t)hblk[m->face_cells[2 * f + 1]];
    int o[3] = {0, 0, 0};
    o[d] = -1;
    int LL = nbo(lo, o[0], o[1], o[2]);
    o[d] = 1;
    int RR = nbo(hi, o[0], o[1], o[2]);
    // stencil fallback of ComputeFaceFlux NS_Solver.H:743-750
    const int LL0 = LL;
    if (LL < 0) LL = RR;
    if (RR < 0) RR = LL0;
    face_st[i] = make_int4(LL, lo, hi, RR);
    // viscous: L+t, R+t, L-t, R-t for the tangential axes t1 < t2
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    for (int k = 0; k < D - 1; ++k) {
      const int t = (k == 0) ? t1 : t2;
      int e[3] = {0, 0, 0};
      e[t] = 1;
      face_vst[(size_t)i * (D - 1) + k] =
          make_int4(nbo(lo, e[0], e[1], e[2]), nbo(hi, e[0], e[1], e[2]),
                    nbo(lo, -e[0], -e[1], -e[2]), nbo(hi, -e[0], -e[1], -e[2]));
    }
  }
  });
  // ---- tiles over each (level, case) run of normal cells.  A tile's face
  // list is the faces of its cells in first-occurrence order (cells in
  // order, each cell's faces low/high per direction); the lists are formed
  // per tile in parallel, placed by a prefix sum, then filled in parallel.
  std::vector<TileInfo> tiles;
  std::vector<int> case_tiles((size_t)nc * L * 2, 0);   // [case][level] = {first, end}
  c->tile_end.assign(L, 0);
  {
    int pos = 0;
    for (int l = L - 1; l >= 0; --l) {
      for (int cs = 0; cs < nc; ++cs) {
        const int b = cs * L + l;
        const int n = m->blk_nnormal[b];
        case_tiles[(size_t)b * 2] = (int)tiles.size();
        for (int t0 = 0; t0 < n; t0 += c->tile) {
          TileInfo ti{};
          ti.cell0 = pos + t0;
          ti.ncell = std::min(c->tile, n - t0);
          ti.blk = b;
          tiles.push_back(ti);
        }
        case_tiles[(size_t)b * 2 + 1] = (int)tiles.size();
        pos += n;
      }
      c->tile_end[l] = (int)tiles.size();
    }
  }
  const int ntl = (int)tiles.size();
  const int q2 = 2 * D;
  hvec<uint16_t> slots((size_t)Nn * q2);   // every entry set below
  // per tile: its faces in first-occurrence order (a small open-addressed
  // table per thread removes the repeats) and each cell face's slot
  std::vector<std::vector<int>> tface(ntl);
  host_parallel(ntl, [&](int lo, int hi) {
    constexpr int kH = 4096;   // > the 6 x 256 faces a tile can name
    std::vector<int> hkey(kH, -1), hval(kH, 0);
    for (int t = lo; t < hi; ++t) {
      const TileInfo &ti = tiles[t];
      std::vector<int> &used = tface[t];
      used.clear();
      std::vector<int> touched;
      for (int k = 0; k < ti.ncell; ++k) {
        const int dc = ti.cell0 + k;
        for (int q = 0; q < q2; ++q) {
          const int f = cfaces[(size_t)dc * q2 + q];
          unsigned h = ((unsigned)f * 2654435761u) >> 20;   // 12 bits
          while (hkey[h] >= 0 && hkey[h] != f) h = (h + 1) & (kH - 1);
          if (hkey[h] < 0) {
            hkey[h] = f;
            hval[h] = (int)used.size();
            used.push_back(f);
            touched.push_back((int)h);
          }
          slots[(size_t)dc * q2 + q] = (uint16_t)hval[h];
        }
      }
      for (int h : touched) hkey[h] = -1;
    }
  }, 64);
  int maxf = 1;
  {
    int off = 0;
    for (int t = 0; t < ntl; ++t) {
      tiles[t].face0 = off;
      tiles[t].nface = (int)tface[t].size();
      if (tiles[t].nface > 65535) return IBG_ERR_ARG;
      maxf = std::max(maxf, tiles[t].nface);
      off += tiles[t].nface;
    }
  }
  const size_t nts = ntl ? (size_t)tiles[ntl - 1].face0 + tiles[ntl - 1].nface : 0;
  hvec<int4> tst(nts), tvs(nts * (D - 1));
  hvec<int8_t> tdir(nts);
  // per device face: its first tile face (the lowest slot), the number of
  // tiles naming it; set with atomics while the tile faces are filled
  std::vector<std::atomic<int>> fslot(F), fcount(F);
  host_parallel(F, [&](int lo, int hi) {
    for (int f = lo; f < hi; ++f) {
      fslot[f].store(INT_MAX, std::memory_order_relaxed);
      fcount[f].store(0, std::memory_order_relaxed);
    }
  });
  hvec<int> slot_tile(nts);
  host_parallel(ntl, [&](int lo, int hi) {
    for (int t = lo; t < hi; ++t) {
      const size_t f0 = (size_t)tiles[t].face0;
      const std::vector<int> &used = tface[t];
      for (size_t k = 0; k < used.size(); ++k) {
        const int f = used[k];
        tst[f0 + k] = face_st[f];
        tdir[f0 + k] = face_dir[f];
        for (int k2 = 0; k2 < D - 1; ++k2) tvs[(f0 + k) * (D - 1) + k2] = face_vst[(size_t)f * (D - 1) + k2];
        slot_tile[f0 + k] = t;
        const int sl = (int)(f0 + k);
        int cur = fslot[f].load(std::memory_order_relaxed);
        while (sl < cur && !fslot[f].compare_exchange_weak(cur, sl, std::memory_order_relaxed)) {
        }
        fcount[f].fetch_add(1, std::memory_order_relaxed);
      }
    }
  }, 64);
  tface.clear();
  tface.shrink_to_fit();
  // that first tile face, its tile's block, the tile count (-1 / 0 for a
  // face no tile names)
  std::vector<int> face_slot0(F), face_tblk(F), face_ntile(F);
  host_parallel(F, [&](int lo, int hi) {
    for (int f = lo; f < hi; ++f) {
      const int sl = fslot[f].load(std::memory_order_relaxed);
      const bool named = sl != INT_MAX;
      face_slot0[f] = named ? sl : -1;
      face_tblk[f] = named ? tiles[slot_tile[sl]].blk : -1;
      face_ntile[f] = fcount[f].load(std::memory_order_relaxed);
    }
  });
  c->ntiles = (int)tiles.size();
  c->max_tile_faces = maxf;
  // ---- optional coarse/fine balance: the pairs (tile faces and blocks of
  // the coarse face and its fine faces) and the tile face -> entry map
  std::vector<RfPair> rfp;
  std::vector<int> tmod(tst.size(), -1);
  for (int i = 0; i < m->nmdy; ++i) {
    const int32_t *r = m->mdy + (size_t)i * 6;
    RfPair p{};
    p.mode = r[0];
    const int nf = rf_nfine(p.mode);
    const int j = (int)rfp.size();
    for (int k = 0; k < 4; ++k) p.fslot[k] = -1;
    for (int k = 0; k <= nf; ++k) {
      const int df = c->fh2d[r[1 + k]];
      // each balanced face lies in exactly one tile (its other cell is a
      // ghost) and in one pair
      if (df < 0 || face_ntile[df] != 1 || tmod[face_slot0[df]] >= 0) return IBG_ERR_ARG;
      const int sl = face_slot0[df];
      tmod[sl] = j * 5 + k;
      if (k == 0) {
        p.cslot = sl;
        p.cblk = face_tblk[df];
      } else {
        if (k > 1 && face_tblk[df] != p.fblk) return IBG_ERR_ARG;
        p.fslot[k - 1] = sl;
        p.fblk = face_tblk[df];
      }
    }
    rfp.push_back(p);
  }
  c->nrf = (int)rfp.size();
  // ---- domain ghosts: edge entries first, then corners (whose refcell is an
  // edge ghost)
  std::vector<int4> dmg;
  std::vector<DmgWall> dmgw;
  {
    std::vector<int4> edge, corner;
    std::vector<DmgWall> wedge, wcorner;
    for (int i = 0; i < m->ndmg; ++i) {
      const int g = m->dmg_cell[i], r = m->dmg_ref[i];
      const int4 e = make_int4(h2d[g], h2d[r], m->dmg_type[i], hblk[g] / L);
      DmgWall w{};
      w.nbr = -1;
      if (m->dmg_nbr && m->dmg_type[i] == -81) {
        const int a = m->dmg_nbr[i];
        w.nbr = h2d[a];
        for (int k = 0; k < 3; ++k) w.nmv[k] = m->dmg_nmv[(size_t)i * 3 + k];
        w.dref = m->dist_dm[r];
        w.dnbr = m->dist_dm[a];
      }
      if (m->kind[r] == IBG_DMGHOST) { corner.push_back(e); wcorner.push_back(w); }
      else { edge.push_back(e); wedge.push_back(w); }
    }
    dmg = edge;
    dmg.insert(dmg.end(), corner.begin(), corner.end());
    dmgw = wedge;
    dmgw.insert(dmgw.end(), wcorner.begin(), wcorner.end());
    c->ndmg_edge = (int)edge.size();
    c->ndmg_corner = (int)corner.size();
  }
  // ---- pairs grouped by coarse level
  std::vector<int> res_ci, res_fi, pro_ci, pro_fi;
  c->res_off.assign(L + 1, 0);
  c->pro_off.assign(L + 1, 0);
  for (int l = 0; l < L; ++l) {
    c->res_off[l] = (int)res_ci.size();
    for (int i = 0; i < m->nres; ++i) {
      if (m->res_blk[i] % L != l) continue;
      res_ci.push_back(h2d[m->res_ci[i]]);
      for (int q = 0; q < 8; ++q) {
        const int a = m->res_fi[(size_t)i * 8 + q];
        res_fi.push_back(a < 0 ? -1 : h2d[a]);
      }
    }
    c->pro_off[l] = (int)pro_ci.size();
    for (int i = 0; i < m->npro; ++i) {
      if (m->pro_blk[i] % L != l) continue;
      pro_ci.push_back(h2d[m->pro_ci[i]]);
      for (int q = 0; q < 8; ++q) {
        const int a = m->pro_fi[(size_t)i * 8 + q];
        pro_fi.push_back(a < 0 ? -1 : h2d[a]);
      }
    }
  }
  c->res_off[L] = (int)res_ci.size();
  c->pro_off[L] = (int)pro_ci.size();
  // ---- IB
  const int nib = m->nib;
  c->nib = nib;
  std::vector<int> ib_cell(nib), ib_intp((size_t)nib * 8), ib_body(nib), ib_case(nib);
  std::vector<double> ib_coef((size_t)nib * 8), ib_hgdis(nib), ib_att((size_t)nib * 3),
      ib_nvh((size_t)nib * 3, 0.0);
  for (int i = 0; i < nib; ++i) {
    const int g = m->ib_cell[i];
    ib_cell[i] = h2d[g];
    ib_body[i] = m->ib_body[i];
    ib_case[i] = hblk[g] / L;
    ib_hgdis[i] = m->ib_hgdis[i];
    for (int k = 0; k < 3; ++k) ib_att[(size_t)i * 3 + k] = m->ib_attach[(size_t)i * 3 + k];
    if (m->ib_nv)
      for (int k = 0; k < 3; ++k) ib_nvh[(size_t)i * 3 + k] = m->ib_nv[(size_t)i * 3 + k];
    for (int q = 0; q < 8; ++q) {
      int a = m->ib_intp[(size_t)i * 8 + q];
      double w = m->ib_coef[(size_t)i * 8 + q];
      if (D == 2 && (q & 1) == 0) { a = -1; w = 0.0; }   // zero-weight z layer
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
    for (int k = 0; k < 3; ++k)
      case_center[(size_t)m->body_case[bi] * 3 + k] = m->body_center[(size_t)bi * 3 + k];
  c->body_center.assign(m->body_center, m->body_center + (size_t)m->nbody * 3);
  c->hbody_case.assign(m->body_case, m->body_case + m->nbody);
  // ---- device allocation and upload
  ibg_status st;
#define UP(dst, vec) if ((st = dupload(c, &dst, vec))) return st
#define AL(dst, n) if ((st = dalloc(c, &dst, n))) return st
  const size_t rs = c->mixed ? sizeof(float) : sizeof(double);
  if ((st = dalloc_bytes(c, &c->W64a, sizeof(D4) * (size_t)N))) return st;
  if ((st = dalloc_bytes(c, &c->Rva, 4 * rs * (size_t)Nn))) return st;
  if (D == 3) {
    AL(c->W64e, N);
    if ((st = dalloc_bytes(c, &c->Rve, rs * (size_t)Nn))) return st;
  }
  // stencil buffers: S0 is the master in the plain mode
  for (int k = 0; k < 3; ++k) { c->Sa[k] = nullptr; c->Se[k] = nullptr; c->NUs[k] = nullptr; }
  c->nu64 = nullptr; c->mutf = nullptr; c->nuprobe = nullptr; c->mprobe = nullptr;
  if (c->turb) {
    AL(c->nu64, N);
    if ((st = dalloc_bytes(c, &c->mutf, rs * (size_t)Nn))) return st;
    AL(c->nuprobe, N);
    AL(c->mprobe, (size_t)std::max(F, 1));
  }
  for (int k = 0; k < c->nbuf; ++k) {
    if (k == 0 && !c->mixed) {
      c->Sa[0] = c->W64a; c->Se[0] = c->W64e; c->NUs[0] = c->nu64;
      continue;
    }
    if ((st = dalloc_bytes(c, &c->Sa[k], 4 * rs * (size_t)N))) return st;
    if (D == 3 && (st = dalloc_bytes(c, &c->Se[k], rs * (size_t)N))) return st;
    if (c->turb && (st = dalloc_bytes(c, &c->NUs[k], rs * (size_t)N))) return st;
  }
  UP(c->flags, flags);
  UP(c->cellcase, cellcase);
  UP(c->nb27, nb27);
  c->hcellblk = cellblk;
  UP(c->tiles, tiles);
  UP(c->tst, tst);
  UP(c->tvs, tvs);
  UP(c->rfp, rfp);
  UP(c->tmod, tmod);
  if ((st = dalloc_bytes(c, &c->rbal, (size_t)std::max(c->nrf, 1) * 5 * 5 * sizeof(double)))) return st;
  if ((st = dalloc_bytes(c, &c->rbalv, (size_t)std::max(c->nrf, 1) * 5 * 5 * sizeof(double)))) return st;
  UP(c->tdir, tdir);
  UP(c->slots, slots);
  UP(c->face_st, face_st);
  UP(c->face_vst, face_vst);
  UP(c->face_dir, face_dir);
  UP(c->face_blk, face_blk);
  UP(c->blk, hb);
  UP(c->phys, c->hphys);
  UP(c->dmg, dmg);
  c->dmgw = nullptr;
  c->dist_dm = nullptr;
  if (m->dmg_nbr) {
    UP(c->dmgw, dmgw);
    std::vector<double> ddm(N);
    for (int i = 0; i < N; ++i) ddm[i] = m->dist_dm[order[i]];
    UP(c->dist_dm, ddm);
  }
  UP(c->res_ci, res_ci); UP(c->res_fi, res_fi); UP(c->pro_ci, pro_ci); UP(c->pro_fi, pro_fi);
  UP(c->ib_cell, ib_cell); UP(c->ib_intp, ib_intp); UP(c->ib_body, ib_body); UP(c->ib_case, ib_case);
  UP(c->ib_coef, ib_coef); UP(c->ib_hgdis, ib_hgdis); UP(c->ib_attach, ib_att); UP(c->ib_nv, ib_nvh);
  UP(c->signdis, sd);
  {
    std::vector<float> sd32(sd.begin(), sd.end());
    UP(c->sd32, sd32);
  }
  UP(c->pat_intp, pat_intp); UP(c->pat_case, pat_case); UP(c->pat_coef, pat_coef);
  UP(c->pat_hgdis, pat_hgdis); UP(c->pat_nv, pat_nv); UP(c->pat_pc, pat_pc); UP(c->pat_area, pat_area);
  UP(c->case_center, case_center);
  AL(c->motion, (size_t)std::max(m->nbody, nc) + 1);
  std::vector<double> dt0(nc);
  for (int cs = 0; cs < nc; ++cs) dt0[cs] = ph[cs].dt_init;
  UP(c->dt, dt0);
  AL(c->cflpart, (size_t)std::max(c->ntiles, 1));
  UP(c->case_tiles, case_tiles);
  AL(c->err, 1);
  AL(c->err2, 1);
  AL(c->rpart, (size_t)std::max(c->ntiles, 1) * kRpart);
  AL(c->resid, (size_t)nc * kRpart);
  AL(c->probe, (size_t)std::max(N, F) * 5);
#undef UP
#undef AL
  // cell-stage update count per step, for each lmin (UPDATE cells marched)
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

}  // extern "C"

namespace {
// pinned staging for the state boundary (grown on demand, kept until the
// context is destroyed): page-locked copies run at the link rate
void *host_stage(ibg_ctx *c, size_t bytes) {
  if (c->hstage_bytes < bytes) {
    if (c->hstage) (void)hipHostFree(c->hstage);
    c->hstage = nullptr;
    c->hstage_bytes = 0;
    if (hipHostMalloc(&c->hstage, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    c->hstage_bytes = bytes;
  }
  return c->hstage;
}

// the stencil copies from the FP64 master on the device
template <typename R, int D>
ibg_status fill_stencil_from_master(ibg_ctx *c) {
  for (int k = 0; k < c->nbuf; ++k) {
    if (c->Sa[k] == c->W64a) continue;
    hipLaunchKernelGGL((k_fill_from_master<R, D>), dim3(grid(c->N, 256)), dim3(256), 0, c->s, c->N,
                       c->w64<D>(), c->sbuf<R, D>(k));
    HIPCHK(hipGetLastError());
  }
  return IBG_OK;
}
}  // namespace

extern "C" {

ibg_status ibg_set_state(ibg_ctx *c, const double *fv, int64_t n) {
  if (!c || !fv || !c->have_mesh || n != c->N) return IBG_ERR_ARG;
  const int N = c->N;
  HIPCHK(hipStreamSynchronize(c->s));
  D4 *a = (D4 *)host_stage(c, (sizeof(D4) + (c->D == 3 ? sizeof(double) : 0)) * (size_t)N);
  if (!a) return IBG_ERR_HIP;
  double *e = (double *)(a + N);
  host_parallel(N, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      const double *q = fv + (size_t)c->d2h[i] * IBG_NFV;   // roe u v w p T e
      if (c->D == 2) a[i] = D4{q[0], q[1], q[2], q[4]};
      else { a[i] = D4{q[0], q[1], q[2], q[3]}; e[i] = q[4]; }
    }
  });
  HIPCHK(hipMemcpy(c->W64a, a, sizeof(D4) * N, hipMemcpyHostToDevice));
  if (c->D == 3) HIPCHK(hipMemcpy(c->W64e, e, sizeof(double) * N, hipMemcpyHostToDevice));
  ibg_status st;
  if ((st = reset_err(c))) return st;   // a new state clears an earlier error
  if (c->D == 2) st = c->mixed ? fill_stencil_from_master<float, 2>(c) : fill_stencil_from_master<double, 2>(c);
  else st = c->mixed ? fill_stencil_from_master<float, 3>(c) : fill_stencil_from_master<double, 3>(c);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_set_nut(ibg_ctx *c, const double *nut, int64_t n) {
  if (!c || !nut || !c->have_mesh || n != c->N) return IBG_ERR_ARG;
  if (!c->turb) return IBG_ERR_STATE;
  const int N = c->N;
  HIPCHK(hipStreamSynchronize(c->s));
  double *a = (double *)host_stage(c, sizeof(double) * (size_t)N);
  if (!a) return IBG_ERR_HIP;
  host_parallel(N, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) a[i] = nut[c->d2h[i]];
  });
  HIPCHK(hipMemcpy(c->nu64, a, sizeof(double) * N, hipMemcpyHostToDevice));
  for (int k = 0; k < c->nbuf; ++k) {
    if (c->NUs[k] == (void *)c->nu64) continue;
    if (c->mixed)
      hipLaunchKernelGGL((k_fill_nu_from_master<float>), dim3(grid(N, 256)), dim3(256), 0, c->s, N,
                         (const double *)c->nu64, (float *)c->NUs[k]);
    else
      hipLaunchKernelGGL((k_fill_nu_from_master<double>), dim3(grid(N, 256)), dim3(256), 0, c->s, N,
                         (const double *)c->nu64, (double *)c->NUs[k]);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_get_nut(ibg_ctx *c, double *nut, int64_t n) {
  if (!c || !nut || !c->have_mesh || n != c->N) return IBG_ERR_ARG;
  if (!c->turb) return IBG_ERR_STATE;
  const int N = c->N;
  HIPCHK(hipStreamSynchronize(c->s));
  double *a = (double *)host_stage(c, sizeof(double) * (size_t)N);
  if (!a) return IBG_ERR_HIP;
  HIPCHK(hipMemcpy(a, c->nu64, sizeof(double) * N, hipMemcpyDeviceToHost));
  host_parallel(N, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) nut[c->d2h[i]] = a[i];
  });
  return IBG_OK;
}

ibg_status ibg_get_state(ibg_ctx *c, double *fv, int64_t n) {
  if (!c || !fv || !c->have_mesh || n != c->N) return IBG_ERR_ARG;
  const int N = c->N;
  HIPCHK(hipStreamSynchronize(c->s));
  D4 *a = (D4 *)host_stage(c, (sizeof(D4) + (c->D == 3 ? sizeof(double) : 0)) * (size_t)N);
  if (!a) return IBG_ERR_HIP;
  double *e = (double *)(a + N);
  HIPCHK(hipMemcpy(a, c->W64a, sizeof(D4) * N, hipMemcpyDeviceToHost));
  if (c->D == 3) HIPCHK(hipMemcpy(e, c->W64e, sizeof(double) * N, hipMemcpyDeviceToHost));
  host_parallel(N, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      double *q = fv + (size_t)c->d2h[i] * IBG_NFV;
      const DevPhys &P = c->hphys[c->hcellblk[i] / c->L];
      const double r = a[i].x, u = a[i].y, v = a[i].z;
      const double w = (c->D == 3) ? a[i].w : 0.0;
      const double p = (c->D == 3) ? e[i] : a[i].w;
      q[0] = r; q[1] = u; q[2] = v; q[3] = w; q[4] = p;
      // derived: Get_T_ideal_gas and Get_E (FlowVariables.H:23,33)
      q[5] = p / (r * P.Rg);
      q[6] = P.Cv * q[5] + 0.5 * (u * u + v * v + w * w);
    }
  });
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
    for (int b = 0; b < c->nbody; ++b)
      for (int k = 0; k < 3; ++k) {
        hm[b].tra[k] = mo[b].tra_vel[k];
        hm[b].rot[k] = mo[b].rot_vel[k];
        hm[b].center[k] = c->body_center[(size_t)b * 3 + k];
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
  return run_visc_prep(c, ts);
}

ibg_status ibg_viscous_flux(ibg_ctx *c, int32_t ts) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  return run_visc_prep(c, ts);
}

ibg_status ibg_stage(ibg_ctx *c, int32_t ts, int32_t subt) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (subt < 0 || subt >= c->cfg.nstage) return IBG_ERR_ARG;
  const bool last = (subt == c->cfg.nstage - 1);
  return run_stage(c, ts, subt, last ? c->probe : nullptr, last ? c->nuprobe : nullptr);
}

}  // extern "C"

namespace {

// A run of steps that one captured graph replays: from ts up to the next
// boundary of the level schedule's period (2^(L-1) steps under temporal
// sub-cycling), within [ts, end).  Its launch sequence depends only on the
// phase of ts in the period, its length and the dt updates it holds; a
// range is captured when those are fixed by its first step (dt_var_step a
// multiple of the period: only a range's first step can update dt; or
// dt_var_step 1: every step does).
struct MarchRange {
  int ts, n, dtflag;
  bool capturable;
};

MarchRange next_range(const ibg_ctx *c, int ts, int end) {
  const int P = c->cfg.temporal_refine ? (1 << (c->L - 1)) : 1;
  MarchRange r;
  r.ts = ts;
  r.n = std::min(end - ts, P - ts % P);
  const int dv = c->cfg.dt_var_step;
  if (dv <= 0) { r.dtflag = 0; r.capturable = true; }
  else if (dv == 1) { r.dtflag = 2; r.capturable = true; }
  else if (dv % P == 0) { r.dtflag = (ts % dv == 0) ? 1 : 0; r.capturable = true; }
  else { r.dtflag = 0; r.capturable = false; }
  return r;
}

long long graph_key(const ibg_ctx *c, const MarchRange &r) {
  const int P = c->cfg.temporal_refine ? (1 << (c->L - 1)) : 1;
  return ((long long)(r.ts % P) * (P + 1) + r.n) * 4 + r.dtflag;
}

// the graph of a range from the cache (captured on first use, without
// executing anything); launched on the stream when launch is set
ibg_status range_graph(ibg_ctx *c, const MarchRange &r, bool launch) {
  const long long key = graph_key(c, r);
  auto it = c->gexec.find(key);
  if (it == c->gexec.end()) {
    const int keep_ntiles = c->last_ntiles;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < r.n; ++k) {
      const ibg_status st = run_step(c, r.ts + k, k > 0, k + 1 < r.n);
      if (st) { hipGraph_t tmp; (void)hipStreamEndCapture(c->s, &tmp); if (tmp) (void)hipGraphDestroy(tmp); return st; }
    }
    HIPCHK(hipStreamEndCapture(c->s, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      fprintf(stderr, "ibgpu: hipGraphInstantiate failed: %s\n", hipGetErrorString(e));
      return IBG_ERR_HIP;
    }
    it = c->gexec.emplace(key, ge).first;
    c->last_ntiles = keep_ntiles;   // nothing ran
  }
  if (launch) {
    HIPCHK(hipGraphLaunch(it->second, c->s));
    c->last_ntiles = c->tile_end[lmin_of(c, r.ts + r.n - 1)];
  }
  return IBG_OK;
}

}  // namespace

extern "C" {

ibg_status ibg_prepare_march(ibg_ctx *c, int32_t ts0, int32_t nsteps) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (nsteps < 0) return IBG_ERR_ARG;
  const int keep_motion = c->have_motion;
  c->have_motion = c->mv.law ? 1 : 0;   // as ibg_time_march sets it
  int ts = ts0;
  const int end = ts0 + nsteps;
  ibg_status st = IBG_OK;
  while (ts < end && !st) {
    const MarchRange r = next_range(c, ts, end);
    if (r.capturable) st = range_graph(c, r, false);
    ts += r.n;
  }
  c->have_motion = keep_motion;
  return st;
}

ibg_status ibg_time_march(ibg_ctx *c, int32_t ts0, int32_t nsteps, int32_t use_graph) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (nsteps < 0) return IBG_ERR_ARG;
  ibg_status st;
  if ((st = reset_err(c))) return st;
  c->have_motion = c->mv.law ? 1 : 0;
  const int P = c->cfg.temporal_refine ? (1 << (c->L - 1)) : 1;
  int64_t upd = 0;
  for (int k = 0; k < nsteps; ++k) upd += c->upd_per_lmin[lmin_of(c, ts0 + k)] * c->cfg.nstage;
  c->last_updates = upd;
  if (c->mv.law) {
    // the moving-body masks change every step: counted on the device
    hipLaunchKernelGGL(k_fill_u64, dim3(grid(kUpdSlots * kUpdSlotStride, 256)), dim3(256), 0, c->s,
                       c->mv.upd, kUpdSlots * kUpdSlotStride, 0ull);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(c->ev0, c->s));
  int ts = ts0;
  const int end = ts0 + nsteps;
  while (ts < end) {
    const MarchRange r = next_range(c, ts, end);
    if (use_graph && r.capturable) {
      if ((st = range_graph(c, r, true))) return st;
    } else {
      for (int k = 0; k < r.n; ++k)
        if ((st = run_step(c, ts + k, k > 0, k + 1 < r.n))) return st;
    }
    ts += r.n;
  }
  HIPCHK(hipEventRecord(c->ev1, c->s));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_total_ms = ms;
  if (c->mv.law) {
    std::vector<unsigned long long> u((size_t)kUpdSlots * kUpdSlotStride);
    HIPCHK(hipMemcpyAsync(u.data(), c->mv.upd, sizeof(unsigned long long) * u.size(),
                          hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    unsigned long long sum = 0;
    for (int k = 0; k < kUpdSlots; ++k) sum += u[(size_t)k * kUpdSlotStride];
    c->last_updates = (int64_t)sum;
  }
  int64_t cell;
  int32_t code;
  if ((st = read_err(c, &cell, &code))) return st;
  if (code) return (ibg_status)code;
  return IBG_OK;
}

ibg_status ibg_residual(ibg_ctx *c, double *linf, double *l2) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  hipLaunchKernelGGL(k_resid_final, dim3(c->ncases), dim3(256), 0, c->s, c->last_ntiles, c->L,
                     c->case_tiles, c->rpart, c->resid, c->ncases);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)c->ncases * kRpart);
  HIPCHK(hipMemcpyAsync(h.data(), c->resid, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int cs = 0; cs < c->ncases; ++cs)
    for (int k = 0; k < 5; ++k) {
      if (linf) linf[cs * 5 + k] = h[cs * kRpart + k];
      if (l2) l2[cs * 5 + k] = h[cs * kRpart + 5 + k];
    }
  return IBG_OK;
}

ibg_status ibg_residual_nut(ibg_ctx *c, double *linf, double *l2) {
  if (!c || !c->have_mesh) return IBG_ERR_STATE;
  if (!c->turb) return IBG_ERR_STATE;
  hipLaunchKernelGGL(k_resid_final, dim3(c->ncases), dim3(256), 0, c->s, c->last_ntiles, c->L,
                     c->case_tiles, c->rpart, c->resid, c->ncases);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)c->ncases * kRpart);
  HIPCHK(hipMemcpyAsync(h.data(), c->resid, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int cs = 0; cs < c->ncases; ++cs) {
    if (linf) linf[cs] = h[cs * kRpart + 10];
    if (l2) l2[cs] = h[cs * kRpart + 11];
  }
  return IBG_OK;
}

ibg_status ibg_forces(ibg_ctx *c, const ibg_body_motion *mo, double *out) {
  if (!c || !c->have_mesh || !out) return IBG_ERR_ARG;
  int hm = 0;
  if (mo) {
    // the bodies of one airfoil move together: the motion of the first body
    // of each case about its center (Surface_Point_Velocity Body.H:709-721)
    std::vector<Motion> h(c->ncases);
    std::vector<int> first(c->ncases, -1);
    for (int b = 0; b < c->nbody; ++b)
      if (first[c->hbody_case[b]] < 0) first[c->hbody_case[b]] = b;
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
  const Motion *mp = c->motion;
  if (!mo && c->mv.law) { mp = c->mv.motc; hm = 1; }   // the prescribed law's motion
  if (c->D == 2)
    hipLaunchKernelGGL((k_forces<2>), dim3(c->ncases), dim3(256), 0, c->s, c->npatch, c->pat_case,
                       c->pat_intp, c->pat_coef, c->pat_hgdis, c->pat_nv, c->pat_pc, c->pat_area,
                       c->case_center, mp, hm, c->w64<2>(), c->phys, (int)c->turb, c->resid);
  else
    hipLaunchKernelGGL((k_forces<3>), dim3(c->ncases), dim3(256), 0, c->s, c->npatch, c->pat_case,
                       c->pat_intp, c->pat_coef, c->pat_hgdis, c->pat_nv, c->pat_pc, c->pat_area,
                       c->case_center, mp, hm, c->w64<3>(), c->phys, (int)c->turb, c->resid);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->resid, sizeof(double) * 6 * c->ncases, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

// ------------------------------------------------------------ moving bodies
ibg_status ibg_upload_surface(ibg_ctx *c, const ibg_mesh_soa *m, const ibg_surface *sf) {
  if (!c || !m || !sf) return IBG_ERR_ARG;
  if (!c->have_mesh) return IBG_ERR_STATE;
  auto fail = [&](const char *msg) {
    c->errmsg = msg;
    fprintf(stderr, "ibgpu: surface rejected: %s\n", msg);
    return IBG_ERR_ARG;
  };
  if (m->ncells != c->N || m->npatch != c->npatch || m->nbody != c->nbody || m->dim != c->D)
    return fail("surface and uploaded mesh differ");
  if (c->D != 2) return fail("moving bodies: 2D cases only");
  if (!sf->pt_off || !sf->pat_pt0 || !sf->pat_body || !sf->pat_flip || !sf->pat_dz ||
      !sf->body_p0 || !sf->case_origin || !sf->cell_patch || !sf->pat_close || sf->npoint < 2)
    return fail("missing surface array");
  const int N = c->N, np = c->npatch, nbody = c->nbody, L = c->L;
  if (sf->body_p0[0] != 0 || sf->body_p0[nbody] != np) return fail("body patch ranges");
  for (int b = 0; b < nbody; ++b)
    if (sf->body_p0[b + 1] < sf->body_p0[b]) return fail("body patch ranges");
  std::vector<int> pbody(np), pcase(np);
  HIPCHK(hipMemcpy(pcase.data(), c->pat_case, sizeof(int) * (size_t)np, hipMemcpyDeviceToHost));
  for (int p = 0; p < np; ++p) {
    const int b = sf->pat_body[p];
    if (b < 0 || b >= nbody || p < sf->body_p0[b] || p >= sf->body_p0[b + 1])
      return fail("patch body out of range");
    if (c->hbody_case[b] != pcase[p]) return fail("patch and body cases differ");
    if (sf->pat_pt0[p] < 0 || sf->pat_pt0[p] + 1 >= sf->npoint) return fail("patch point out of range");
  }
  auto finest_cell = [&](int g) {
    const int b = c->hcellblk[c->h2d[g]];
    return (b % L) == L - 1;
  };
  {
    const char *msg = nullptr;
    if (first_failure(N, [&](int g) -> const char * {
          const int p = sf->cell_patch[g];
          if (p < -1 || p >= np) return "cell patch out of range";
          if (p >= 0 && pcase[p] != (int)c->hcellblk[c->h2d[g]] / L) return "cell patch of another case";
          return nullptr;
        }, &msg) >= 0)
      return fail(msg);
  }
  for (int p = 0; p < np; ++p) {
    const int g = sf->pat_close[p];
    if (g < 0 || g >= N || !finest_cell(g)) return fail("patch close cell not on the finest level");
  }
  (void)hipStreamSynchronize(c->s);
  // device-order per-cell arrays
  std::vector<int> key((size_t)N * 3), cp(N);
  std::vector<int8_t> kind(N), fin(N);
  host_parallel(N, [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const int g = c->d2h[i];
      for (int d = 0; d < 3; ++d) key[(size_t)i * 3 + d] = m->key[(size_t)g * 3 + d];
      kind[i] = m->kind[g];
      const int b = c->hcellblk[i];
      fin[i] = ((b % L) == L - 1 && nbody > 0) ? 1 : 0;
      cp[i] = sf->cell_patch[g];
    }
  });
  std::vector<int> pt0(sf->pat_pt0, sf->pat_pt0 + np), pb(sf->pat_body, sf->pat_body + np);
  std::vector<int8_t> flip(sf->pat_flip, sf->pat_flip + np);
  std::vector<double> dz(sf->pat_dz, sf->pat_dz + np);
  std::vector<int> bp0(sf->body_p0, sf->body_p0 + nbody + 1);
  std::vector<double> off(sf->pt_off, sf->pt_off + 2 * (size_t)sf->npoint);
  std::vector<double> org(sf->case_origin, sf->case_origin + 3 * (size_t)c->ncases);
  std::vector<int> cc(np), bcase(c->hbody_case);
  for (int p = 0; p < np; ++p) cc[p] = c->h2d[sf->pat_close[p]];
  std::vector<int16_t> cblk(c->hcellblk);
  ibg_ctx::MoveState &v = c->mv;
  v = ibg_ctx::MoveState{};
  ibg_status st;
#define UP(dst, vec) if ((st = dupload(c, &dst, vec))) return st
#define AL(dst, n) if ((st = dalloc(c, &dst, n))) return st
  UP(v.key, key); UP(v.kind, kind); UP(v.fin, fin); UP(v.cblk, cblk); UP(v.cpatch, cp);
  UP(v.origin, org); UP(v.pt_off, off);
  {
    std::vector<double> borg((size_t)c->nblk * 3);
    for (int b = 0; b < c->nblk; ++b)
      for (int d = 0; d < 3; ++d) borg[(size_t)b * 3 + d] = org[(size_t)(b / c->L) * 3 + d];
    UP(v.borg, borg);
  }
  UP(v.pat_pt0, pt0); UP(v.pat_body, pb);
  UP(v.pat_flip, flip); UP(v.pat_dz, dz); UP(v.body_p0, bp0); UP(v.pat_cc, cc);
  UP(v.bcase_d, bcase);
  {
    std::vector<int2> lh(np);
    for (int p = 0; p < np; ++p) {
      const int b = sf->pat_body[p];
      lh[p] = make_int2(sf->body_p0[b], sf->body_p0[b + 1]);
    }
    UP(v.pat_lh, lh);
  }
  {
    // finest-level key grids: the cell (normal or ghost) of every in-plane
    // key of the level's bounding box, the table the image searches read
    // instead of walking the neighbour table one cell at a time
    std::vector<int4> box(c->ncases, make_int4(0, 0, 0, 0));
    std::vector<long long> foff(c->ncases + 1, 0);
    for (int cs = 0; cs < c->ncases; ++cs) {
      const int b = cs * L + L - 1;
      int x0 = 1 << 30, y0 = 1 << 30, x1 = -(1 << 30), y1 = -(1 << 30);
      for (int g = m->blk_start[b]; g < m->blk_start[b + 1]; ++g) {
        x0 = std::min(x0, m->key[(size_t)g * 3]); x1 = std::max(x1, m->key[(size_t)g * 3]);
        y0 = std::min(y0, m->key[(size_t)g * 3 + 1]); y1 = std::max(y1, m->key[(size_t)g * 3 + 1]);
      }
      if (x1 >= x0) box[cs] = make_int4(x0, y0, x1 - x0 + 1, y1 - y0 + 1);
      foff[cs + 1] = foff[cs] + (long long)box[cs].z * box[cs].w;
    }
    std::vector<int> grid(std::max<long long>(foff[c->ncases], 1), -1);
    for (int cs = 0; cs < c->ncases; ++cs) {
      const int b = cs * L + L - 1;
      for (int g = m->blk_start[b]; g < m->blk_start[b + 1]; ++g) {
        const long long rx = m->key[(size_t)g * 3] - box[cs].x, ry = m->key[(size_t)g * 3 + 1] - box[cs].y;
        grid[(size_t)(foff[cs] + ry * box[cs].z + rx)] = c->h2d[g];
      }
    }
    UP(v.fbox, box); UP(v.foff, foff); UP(v.fgrid, grid);
  }
  {
    // ghosts by level, finest first: the ghosts of the levels >= l are a
    // prefix of this list
    std::vector<std::vector<int>> byl(L);
    for (int i = c->Nn; i < N; ++i) byl[c->hcellblk[i] % L].push_back(i);
    std::vector<int> gl;
    v.gend.assign(L, 0);
    for (int l = L - 1; l >= 0; --l) {
      gl.insert(gl.end(), byl[l].begin(), byl[l].end());
      v.gend[l] = (int)gl.size();
    }
    if (gl.empty()) gl.push_back(0);
    UP(v.gidx, gl);
  }
  {
    // in-plane neighbours in SoA (Point_iterator_2d order without the cell)
    static const int q8[kNb8] = {1, 4, 7, 10, 16, 19, 22, 25};
    // (from the host mesh, in device order)
    std::vector<int> nb8((size_t)kNb8 * N);
    host_parallel(N, [&](int lo, int hi) {
      for (int i = lo; i < hi; ++i) {
        const int g = c->d2h[i];
        for (int k = 0; k < kNb8; ++k) {
          const int a = m->nb[(size_t)g * 27 + q8[k]];
          nb8[(size_t)k * N + i] = a < 0 ? -1 : c->h2d[a];
        }
      }
    });
    UP(v.nb8, nb8);
    // mixed-body neighbourhoods of the uploaded closest patches
    std::vector<int8_t> cmix(N, 0);
    host_parallel(N, [&](int lo, int hi) {
      for (int i = lo; i < hi; ++i) {
        if (cp[i] < 0) continue;
        const int bi = sf->pat_body[cp[i]];
        for (int k = 0; k < kNb8; ++k) {
          const int y = nb8[(size_t)k * N + i];
          if (y >= 0 && cp[y] >= 0 && sf->pat_body[cp[y]] != bi) { cmix[i] = 1; break; }
        }
      }
    });
    UP(v.cmix, cmix);
    // move stamps: the uploaded IB set is move 1
    std::vector<int> mkib(N, 0), one(1, 1);
    for (int j = 0; j < m->nib; ++j) mkib[c->h2d[m->ib_cell[j]]] = 1;
    UP(v.mkib, mkib);
    UP(v.epoch, one);
    // list capacities from the uploaded geometry: finest-level solid cells
    int nsolid = 0;
    for (int i = 0; i < N; ++i)
      if (fin[i] && kind[i] != IBG_DMGHOST && m->signdis[c->d2h[i]] < 0.0) ++nsolid;
    v.cap_s = 2 * nsolid + 4096;
    v.cap_io = c->nib + 4096;
    v.cap_chg = 8192;
  }
  AL(v.cp1, N); AL(v.sd1, N); AL(v.mk1, N);
  AL(v.pose, c->ncases); AL(v.motc, c->ncases); AL(v.lawd, c->ncases); AL(v.t, c->ncases);
  AL(v.upd, (size_t)kUpdSlots * kUpdSlotStride); AL(v.lvlcnt, IBG_MAX_LEVELS);
  // IB arrays at capacity, holding the current list
  const int cap = sf->ib_capacity > 0 ? sf->ib_capacity : 2 * c->nib + 4096;
  if (cap < c->nib) return fail("IB capacity below the current list");
  int *nc_ = nullptr, *ni = nullptr, *nbd = nullptr, *ncs = nullptr;
  double *nco = nullptr, *nh = nullptr, *nat = nullptr, *nnv = nullptr;
  AL(nc_, cap); AL(ni, (size_t)cap * 8); AL(nbd, cap); AL(ncs, cap);
  AL(nco, (size_t)cap * 8); AL(nh, cap); AL(nat, (size_t)cap * 3); AL(nnv, (size_t)cap * 3);
#undef UP
#undef AL
  const size_t n0 = (size_t)c->nib;
  if (n0) {
    HIPCHK(hipMemcpy(nc_, c->ib_cell, n0 * sizeof(int), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(ni, c->ib_intp, n0 * 8 * sizeof(int), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(nbd, c->ib_body, n0 * sizeof(int), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(ncs, c->ib_case, n0 * sizeof(int), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(nco, c->ib_coef, n0 * 8 * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(nh, c->ib_hgdis, n0 * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(nat, c->ib_attach, n0 * 3 * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(nnv, c->ib_nv, n0 * 3 * sizeof(double), hipMemcpyDeviceToDevice));
  }
  c->ib_cell = nc_; c->ib_intp = ni; c->ib_body = nbd; c->ib_case = ncs;
  c->ib_coef = nco; c->ib_hgdis = nh; c->ib_attach = nat; c->ib_nv = nnv;
  v.cap = cap;
  {
    ibg_status st2;
    std::vector<int> cnt0 = {0, 0, 0, c->nib, 0, 0};
    if ((st2 = dupload(c, &v.cnt, cnt0))) return st2;
    if ((st2 = dalloc(c, &v.mixlist, (size_t)N))) return st2;
    if ((st2 = dalloc(c, &v.chg, (size_t)v.cap_chg))) return st2;
    if ((st2 = dalloc(c, &v.slist, (size_t)v.cap_s))) return st2;
    if ((st2 = dalloc(c, &v.iolist, (size_t)v.cap_io))) return st2;
  }
  v.npoint = sf->npoint;
  v.on = true;
  // captured graphs hold the old IB launch
  drop_graphs(c);
  return IBG_OK;
}

namespace {
// per-level update counts after a move: the cell-stage updates of a step
ibg_status recount_updates(ibg_ctx *c) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->s, c->mv.lvlcnt, IBG_MAX_LEVELS, 0ull);
  const MoveArgs a = move_args(c, nullptr, 0);
  hipLaunchKernelGGL(k_move_levelcount, dim3(grid(c->Nn, 256)), dim3(256), 0, c->s, a, c->L);
  HIPCHK(hipGetLastError());
  std::vector<unsigned long long> h(IBG_MAX_LEVELS);
  HIPCHK(hipMemcpyAsync(h.data(), c->mv.lvlcnt, sizeof(unsigned long long) * IBG_MAX_LEVELS,
                        hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int lm = 0; lm < c->L; ++lm) {
    int64_t s = 0;
    for (int l = lm; l < c->L; ++l) s += (int64_t)h[l];
    c->upd_per_lmin[lm] = s;
  }
  return IBG_OK;
}
}  // namespace

ibg_status ibg_move_bodies(ibg_ctx *c, const ibg_body_pose *pose) {
  if (!c || !pose) return IBG_ERR_ARG;
  if (!c->have_mesh || !c->mv.on) return IBG_ERR_STATE;
  std::vector<Pose> hp(c->ncases);
  for (int k = 0; k < c->ncases; ++k) {
    hp[k].px = pose[k].pivot[0]; hp[k].py = pose[k].pivot[1];
    hp[k].c = pose[k].cosa; hp[k].s = pose[k].sina;
  }
  for (int b = 0; b < c->nbody; ++b) {
    c->body_center[(size_t)b * 3] = hp[c->hbody_case[b]].px;
    c->body_center[(size_t)b * 3 + 1] = hp[c->hbody_case[b]].py;
  }
  ibg_status st;
  if ((st = reset_err(c))) return st;
  HIPCHK(hipMemcpyAsync(c->mv.pose, hp.data(), sizeof(Pose) * hp.size(), hipMemcpyHostToDevice, c->s));
  if ((st = run_move(c, nullptr, 0))) return st;
  if ((st = recount_updates(c))) return st;
  int64_t cell;
  int32_t code;
  if ((st = read_err(c, &cell, &code))) return st;
  return code ? (ibg_status)code : IBG_OK;
}

ibg_status ibg_set_motion(ibg_ctx *c, const ibg_pitch_heave *law, double t0) {
  if (!c) return IBG_ERR_ARG;
  if (!c->have_mesh || !c->mv.on) return IBG_ERR_STATE;
  drop_graphs(c);
  if (!law) { c->mv.law = false; return IBG_OK; }
  const double d2r = 3.14159265358979323846 / 180.0;
  std::vector<PitchHeave> h(c->ncases);
  std::vector<double> t(c->ncases, t0);
  for (int k = 0; k < c->ncases; ++k) {
    h[k].x0 = law[k].x0; h[k].y0 = law[k].y0;
    h[k].aoa0 = law[k].aoa0_deg * d2r; h[k].amp = law[k].amp_deg * d2r;
    h[k].omega = law[k].omega; h[k].heave = law[k].heave;
  }
  HIPCHK(hipMemcpyAsync(c->mv.lawd, h.data(), sizeof(PitchHeave) * h.size(), hipMemcpyHostToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->mv.t, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int b = 0; b < c->nbody; ++b) {
    c->body_center[(size_t)b * 3] = h[c->hbody_case[b]].x0;
    c->body_center[(size_t)b * 3 + 1] = h[c->hbody_case[b]].y0;
  }
  c->mv.law = true;
  return IBG_OK;
}

ibg_status ibg_get_time(ibg_ctx *c, double *t) {
  if (!c || !t) return IBG_ERR_ARG;
  if (!c->have_mesh || !c->mv.on) return IBG_ERR_STATE;
  HIPCHK(hipMemcpyAsync(t, c->mv.t, sizeof(double) * c->ncases, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_set_time(ibg_ctx *c, const double *t) {
  if (!c || !t) return IBG_ERR_ARG;
  if (!c->have_mesh || !c->mv.on) return IBG_ERR_STATE;
  HIPCHK(hipMemcpyAsync(c->mv.t, t, sizeof(double) * c->ncases, hipMemcpyHostToDevice, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_get_geometry(ibg_ctx *c, double *signdis, int32_t *cell_patch, int8_t *ib,
                            int32_t *nib) {
  if (!c) return IBG_ERR_ARG;
  if (!c->have_mesh || !c->mv.on) return IBG_ERR_STATE;
  const int N = c->N;
  std::vector<double> sd(N);
  std::vector<int> cp(N), mk(N);
  int n = 0, ep = 0;
  HIPCHK(hipMemcpyAsync(sd.data(), c->signdis, sizeof(double) * N, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(cp.data(), c->mv.cpatch, sizeof(int) * N, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(mk.data(), c->mv.mkib, sizeof(int) * N, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(&n, c->mv.cnt + 3, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(&ep, c->mv.epoch, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int i = 0; i < N; ++i) {
    const int g = c->d2h[i];
    if (signdis) signdis[g] = sd[i];
    if (cell_patch) cell_patch[g] = cp[i];
    if (ib) ib[g] = mk[i] == ep ? 1 : 0;
  }
  if (nib) *nib = std::min(n, c->mv.cap);
  return IBG_OK;
}

ibg_status ibg_get_ib_list(ibg_ctx *c, int32_t cap, int32_t *cell, int32_t *intp, double *coef,
                           double *hgdis, double *nv, double *attach, int32_t *body) {
  if (!c || cap < 0) return IBG_ERR_ARG;
  if (!c->have_mesh) return IBG_ERR_STATE;
  int n = c->nib;
  if (c->mv.on) {
    HIPCHK(hipMemcpyAsync(&n, c->mv.cnt + 3, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    n = std::min(n, c->mv.cap);
  }
  if (n > cap) return IBG_ERR_ARG;
  std::vector<int> hc(n), hi((size_t)n * 8), hb(n);
  std::vector<double> hco((size_t)n * 8), hh(n), hn((size_t)n * 3), ha((size_t)n * 3);
  if (n) {
    HIPCHK(hipMemcpyAsync(hc.data(), c->ib_cell, sizeof(int) * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(hi.data(), c->ib_intp, sizeof(int) * 8 * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(hb.data(), c->ib_body, sizeof(int) * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(hco.data(), c->ib_coef, sizeof(double) * 8 * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(hh.data(), c->ib_hgdis, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(hn.data(), c->ib_nv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(ha.data(), c->ib_attach, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
  }
  for (int j = 0; j < n; ++j) {
    if (cell) cell[j] = c->d2h[hc[j]];
    for (int q = 0; q < 8; ++q) {
      const int a = hi[(size_t)j * 8 + q];
      if (intp) intp[(size_t)j * 8 + q] = a < 0 ? -1 : c->d2h[a];
      if (coef) coef[(size_t)j * 8 + q] = hco[(size_t)j * 8 + q];
    }
    if (hgdis) hgdis[j] = hh[j];
    for (int d = 0; d < 3; ++d) {
      if (nv) nv[(size_t)j * 3 + d] = hn[(size_t)j * 3 + d];
      if (attach) attach[(size_t)j * 3 + d] = ha[(size_t)j * 3 + d];
    }
    if (body) body[j] = hb[j];
  }
  return IBG_OK;
}

ibg_status ibg_get_patches(ibg_ctx *c, double *pc, double *nv, double *area, int32_t *intp,
                           double *coef, double *hgdis) {
  if (!c) return IBG_ERR_ARG;
  if (!c->have_mesh) return IBG_ERR_STATE;
  const int np = c->npatch;
  std::vector<int> hi((size_t)np * 8);
  HIPCHK(hipStreamSynchronize(c->s));
  if (pc) HIPCHK(hipMemcpy(pc, c->pat_pc, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
  if (nv) HIPCHK(hipMemcpy(nv, c->pat_nv, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
  if (area) HIPCHK(hipMemcpy(area, c->pat_area, sizeof(double) * np, hipMemcpyDeviceToHost));
  if (coef) HIPCHK(hipMemcpy(coef, c->pat_coef, sizeof(double) * 8 * np, hipMemcpyDeviceToHost));
  if (hgdis) HIPCHK(hipMemcpy(hgdis, c->pat_hgdis, sizeof(double) * np, hipMemcpyDeviceToHost));
  if (intp) {
    HIPCHK(hipMemcpy(hi.data(), c->pat_intp, sizeof(int) * 8 * np, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < hi.size(); ++k) intp[k] = hi[k] < 0 ? -1 : c->d2h[hi[k]];
  }
  return IBG_OK;
}

}  // extern "C"

namespace {
template <typename R, int D>
void launch_probes(ibg_ctx *c, bool visc) {
  const int F = c->F;
  if (F <= 0) return;
  if (visc)
    hipLaunchKernelGGL((k_visc_probe<R, D>), dim3(grid(F, 256)), dim3(256), 0, c->s, F, c->face_st,
                       c->face_vst, c->face_dir, c->face_blk, c->wstencil<R, D>(), c->blk, c->phys,
                       c->turb ? c->nubuf<R>(0) : (const R *)nullptr, c->probe, c->mprobe);
  else
    hipLaunchKernelGGL((k_face_probe<R, D>), dim3(grid(F, 256)), dim3(256), 0, c->s, F, c->face_st,
                       c->face_dir, c->face_blk, c->wstencil<R, D>(), c->blk, c->phys,
                       c->cfg.limiter, c->probe);
}

ibg_status face_probe(ibg_ctx *c, double *out, bool visc) {
  if (c->D == 2) { if (c->mixed) launch_probes<float, 2>(c, visc); else launch_probes<double, 2>(c, visc); }
  else { if (c->mixed) launch_probes<float, 3>(c, visc); else launch_probes<double, 3>(c, visc); }
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)c->F * 5);
  HIPCHK(hipMemcpyAsync(h.data(), c->probe, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  for (int hf = 0; hf < c->F; ++hf)
    for (int k = 0; k < 5; ++k) out[(size_t)hf * 5 + k] = h[(size_t)c->fh2d[hf] * 5 + k];
  return IBG_OK;
}
}  // namespace

extern "C" {

ibg_status ibg_get_faceflux(ibg_ctx *c, double *out, int64_t nf) {
  if (!c || !c->have_mesh || !out || nf != c->F) return IBG_ERR_ARG;
  return face_probe(c, out, false);
}

ibg_status ibg_get_vsflux(ibg_ctx *c, double *out, int64_t nf) {
  if (!c || !c->have_mesh || !out || nf != c->F) return IBG_ERR_ARG;
  return face_probe(c, out, true);
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

ibg_status ibg_get_mutflux(ibg_ctx *c, double *out, int64_t nf) {
  if (!c || !c->have_mesh || !out || nf != c->F) return IBG_ERR_ARG;
  if (!c->turb) return IBG_ERR_STATE;
  std::vector<double> tmp((size_t)c->F * 5);
  ibg_status st = face_probe(c, tmp.data(), true);
  if (st) return st;
  std::vector<double> h((size_t)c->F);
  HIPCHK(hipMemcpy(h.data(), c->mprobe, h.size() * 8, hipMemcpyDeviceToHost));
  for (int hf = 0; hf < c->F; ++hf) out[hf] = h[(size_t)c->fh2d[hf]];
  return IBG_OK;
}

ibg_status ibg_get_nut_source(ibg_ctx *c, double *out, int64_t n) {
  if (!c || !c->have_mesh || !out || n != c->N) return IBG_ERR_ARG;
  if (!c->turb) return IBG_ERR_STATE;
  HIPCHK(hipStreamSynchronize(c->s));
  std::vector<double> h((size_t)c->N);
  HIPCHK(hipMemcpy(h.data(), c->nuprobe, h.size() * 8, hipMemcpyDeviceToHost));
  for (int i = 0; i < c->N; ++i) out[(size_t)c->d2h[i]] = (i < c->Nn) ? h[(size_t)i] : 0.0;
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

ibg_status ibg_bench_stage(ibg_ctx *c, int32_t ts, int32_t subt, int32_t reps, double *avg_ms,
                           int64_t *cells, int64_t *faces) {
  if (!c || !c->have_mesh || reps < 1 || subt < 0 || subt >= c->cfg.nstage) return IBG_ERR_ARG;
  const int nt = c->tile_end[lmin_of(c, ts)];
  int64_t nc = 0, nf = 0;
  {
    std::vector<TileInfo> h(nt);
    if (nt > 0) HIPCHK(hipMemcpy(h.data(), c->tiles, sizeof(TileInfo) * nt, hipMemcpyDeviceToHost));
    for (const TileInfo &t : h) { nc += t.ncell; nf += t.nface; }
  }
  ibg_status st;
  // a measurement, not a march: the repeated stages must not stop at an
  // error they produce themselves (the state is not advanced consistently)
  if ((st = reset_err(c))) return st;
  HIPCHK(hipEventRecord(c->ev0, c->s));
  for (int r = 0; r < reps; ++r)
    if ((st = run_stage(c, ts, subt, nullptr, nullptr, fused_visc(c)))) return st;
  HIPCHK(hipEventRecord(c->ev1, c->s));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_stage_ms = ms / reps;
  c->last_stage_launches = reps;
  if (avg_ms) *avg_ms = ms / reps;
  if (cells) *cells = nc;
  if (faces) *faces = nf;
  return IBG_OK;
}

ibg_status ibg_bench_copy(ibg_ctx *c, int64_t nbytes, int32_t reps, double *avg_ms) {
  if (!c || nbytes < 16 || reps < 1) return IBG_ERR_ARG;
  const long n = (long)(nbytes / 16);
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, (size_t)n * 16));
  if (hipMalloc(&b, (size_t)n * 16) != hipSuccess) { (void)hipFree(a); return IBG_ERR_HIP; }
  (void)hipMemsetAsync(a, 0, (size_t)n * 16, c->s);
  // one 16-byte element per thread over a full grid: the fastest plain copy
  // measured on MI355X (6.2 TB/s read+write for 1 GiB; a grid-stride loop
  // over 16K workgroups reaches 4.8)
  const long nbl = (n + 255) / 256;
  if (nbl > 0x7fffffffL) { (void)hipFree(a); (void)hipFree(b); return IBG_ERR_ARG; }
  const unsigned nb = (unsigned)nbl;
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c->s, (const F4 *)a, (F4 *)b, n);
  (void)hipEventRecord(c->ev0, c->s);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c->s, (const F4 *)a, (F4 *)b, n);
  (void)hipEventRecord(c->ev1, c->s);
  hipError_t e = hipEventSynchronize(c->ev1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return IBG_ERR_HIP;
  if (avg_ms) *avg_ms = ms / reps;
  return IBG_OK;
}

ibg_status ibg_autotune_tile(ibg_ctx *c, const ibg_mesh_soa *m, const ibg_physics *ph,
                             const int32_t *cand, int32_t ncand, int32_t reps, int32_t *best,
                             double *ms) {
  if (!c || !m || !ph || !cand || ncand < 1 || reps < 1) return IBG_ERR_ARG;
  for (int k = 0; k < ncand; ++k)
    if (cand[k] < 64 || cand[k] > kTile || cand[k] % 64 != 0) return IBG_ERR_ARG;
  // freestream state of every case (InitFlow NS_Solver.H:432-446)
  std::vector<double> fv((size_t)m->ncells * IBG_NFV), nu(m->ncells);
  const int L = m->nlevels;
  for (int cs = 0; cs < m->ncases; ++cs)
    for (int g = m->blk_start[cs * L]; g < m->blk_start[(cs + 1) * L]; ++g) {
      for (int v = 0; v < IBG_NFV; ++v) fv[(size_t)g * IBG_NFV + v] = ph[cs].initvar[v];
      nu[g] = ph[cs].nut_inf;
    }
  double tbest = 1e300;
  int kbest = 0, sbest = 0;
  ibg_status st;
  for (int k = 0; k < ncand; ++k) {
    c->cfg.tile = cand[k];
    c->tile = cand[k];
    c->cfg.schedule = 0;
    if ((st = ibg_upload_mesh(c, m, ph))) return st;
    if ((st = ibg_set_state(c, fv.data(), m->ncells))) return st;
    if (c->turb && (st = ibg_set_nut(c, nu.data(), m->ncells))) return st;
    if ((st = run_dt(c))) return st;
    for (int sch = 0; sch < 2; ++sch) {
      c->cfg.schedule = sch;
      double t = 0.0;
      if ((st = ibg_bench_stage(c, 0, 0, 2, &t, nullptr, nullptr))) return st;   // warm
      if ((st = ibg_bench_stage(c, 0, 0, reps, &t, nullptr, nullptr))) return st;
      if (ms) ms[2 * k + sch] = t;
      if (t < tbest) { tbest = t; kbest = k; sbest = sch; }
    }
  }
  c->cfg.tile = cand[kbest];
  c->tile = cand[kbest];
  c->cfg.schedule = sbest;
  if ((st = ibg_upload_mesh(c, m, ph))) return st;
  if (best) *best = cand[kbest];
  return IBG_OK;
}

ibg_status ibg_cell_stage_updates(ibg_ctx *c, int64_t *count) {
  if (!c || !count) return IBG_ERR_ARG;
  *count = c->last_updates;
  return IBG_OK;
}

}  // extern "C"

// ======================================================= learned closure
namespace {

template <typename T>
ibg_status cl_alloc(ibg_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n * sizeof(T), 64)));
  HIPCHK(hipMemset(q, 0, std::max<size_t>(n * sizeof(T), 64)));
  c->cl.allocs.push_back(q);
  *p = reinterpret_cast<T *>(q);
  return IBG_OK;
}

uint64_t host_mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

ibg_status cl_pack(ibg_ctx *c) {
  hipLaunchKernelGGL(k_closure_pack, dim3(grid(64 * kH, 256)), dim3(256), 0, c->s, c->cl.p, c->cl.W1t,
                     c->cl.W2t, c->cl.W2, c->cl.W2p, c->cl.w3, c->cl.b1, c->cl.b2, c->cl.b3);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

constexpr size_t kTrainSmem = (size_t)64 * kX * 2 * 2 + (size_t)64 * kH * 2 * 5 + (5 * 64 + 256) * 4;

ibg_status cl_train_launch(ibg_ctx *c, const float *X, const float *Z, long long n, int nwg,
                           int batch, int explicit_rows, uint64_t seed) {
  TrainArgs a;
  a.M = c->mlpdev();
  a.X = X; a.Z = Z; a.n = n; a.batch = batch; a.explicit_rows = explicit_rows; a.seed = seed;
  a.grad = c->cl.grad; a.lossp = c->cl.lossp;
  hipLaunchKernelGGL(k_closure_train, dim3(nwg), dim3(256), kTrainSmem, c->s, a);
  HIPCHK(hipGetLastError());
  return IBG_OK;
}

}  // namespace

extern "C" {

// the 34 statistics plus the reciprocal standard deviations (IEEE FP32
// division, the value the standardisation multiplies by)
static ibg_status cl_put_stats(ibg_ctx *c, const float *st34) {
  float h[42];
  for (int k = 0; k < 34; ++k) h[k] = st34[k];
  for (int k = 0; k < 8; ++k) h[34 + k] = 1.0f / st34[8 + k];
  HIPCHK(hipMemcpy(c->cl.stat, h, sizeof(h), hipMemcpyHostToDevice));
  return IBG_OK;
}

ibg_status ibg_closure_init(ibg_ctx *c, int64_t capacity, uint64_t seed) {
  if (!c || capacity < 64) return IBG_ERR_ARG;
  // captured time-march graphs may hold the closure buffers about to be
  // freed (the network variant of the stage kernel): drop them
  HIPCHK(hipStreamSynchronize(c->s));
  drop_graphs(c);
  for (void *p : c->cl.allocs) (void)hipFree(p);
  c->cl = ibg_ctx::Closure();
  ibg_ctx::Closure &L = c->cl;
  L.cap = capacity;
  L.max_wg = 1024;   // batch up to 65536 rows
  ibg_status st;
  if ((st = cl_alloc(c, &L.X, (size_t)capacity * kMlpIn))) return st;
  if ((st = cl_alloc(c, &L.Z, (size_t)capacity))) return st;
  if ((st = cl_alloc(c, &L.cell, (size_t)capacity))) return st;
  if ((st = cl_alloc(c, &L.count, 2))) return st;
  if ((st = cl_alloc(c, &L.mlcount, (size_t)kMlSlots * kMlSlotStride))) return st;
  if ((st = cl_alloc(c, &L.p, kNParam))) return st;
  if ((st = cl_alloc(c, &L.m, kNParam))) return st;
  if ((st = cl_alloc(c, &L.v, kNParam))) return st;
  if ((st = cl_alloc(c, &L.gsum, kNParam))) return st;
  if ((st = cl_alloc(c, &L.grad, (size_t)L.max_wg * kNParam))) return st;
  if ((st = cl_alloc(c, &L.lossp, (size_t)L.max_wg))) return st;
  if ((st = cl_alloc(c, &L.W1t, (size_t)64 * kX))) return st;
  if ((st = cl_alloc(c, &L.W2t, (size_t)64 * kH))) return st;
  if ((st = cl_alloc(c, &L.W2, (size_t)64 * kH))) return st;
  if ((st = cl_alloc(c, &L.W2p, (size_t)64 * 64))) return st;
  if ((st = cl_alloc(c, &L.w3, 64))) return st;
  if ((st = cl_alloc(c, &L.b1, 64))) return st;
  if ((st = cl_alloc(c, &L.b2, 64))) return st;
  if ((st = cl_alloc(c, &L.b3, 1))) return st;
  if ((st = cl_alloc(c, &L.stat, 42))) return st;
  if ((st = cl_alloc(c, &L.mom, (size_t)256 * (kMlpIn + 1) * 4))) return st;
  // weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0 (seeded)
  std::vector<float> p(kNParam, 0.0f);
  uint64_t s = seed;
  auto uni = [&](float a) {
    s = host_mix64(s);
    return (float)(((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * a);
  };
  for (int i = 0; i < 8 * 64; ++i) p[kOffW1 + i] = uni(1.0f / std::sqrt(8.0f));
  for (int i = 0; i < 64 * 64; ++i) p[kOffW2 + i] = uni(1.0f / 8.0f);
  for (int i = 0; i < 64; ++i) p[kOffW3 + i] = uni(1.0f / 8.0f);
  HIPCHK(hipMemcpy(L.p, p.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice));
  std::vector<float> st0(34, 0.0f);
  for (int k = 0; k < 8; ++k) { st0[8 + k] = 1.0f; st0[16 + k] = -1e30f; st0[24 + k] = 1e30f; }
  st0[33] = 1.0f;
  if ((st = cl_put_stats(c, st0.data()))) return st;
  if ((st = cl_pack(c))) return st;
  HIPCHK(hipStreamSynchronize(c->s));
  L.ready = true;
  return IBG_OK;
}

static int *cl_cells(ibg_ctx *c) { return c->cl.cell; }

ibg_status ibg_closure_sample(ibg_ctx *c, double frac, uint64_t seed, int64_t *total) {
  if (!c || !c->have_mesh || !c->cl.ready) return IBG_ERR_STATE;
  if (!c->turb || frac < 0.0 || frac > 1.0) return IBG_ERR_ARG;
  if (c->ntiles > 0) {
    if (c->cl.ntc < c->ntiles) {
      // per-tile counts and row offsets of the two-pass sampler
      ibg_status st;
      if ((st = cl_alloc(c, &c->cl.tcount, (size_t)c->ntiles))) return st;
      if ((st = cl_alloc(c, &c->cl.toff, (size_t)c->ntiles))) return st;
      c->cl.ntc = c->ntiles;
    }
    const long long cap = (long long)c->cl.cap;
    int *cells = cl_cells(c);
    for (int pass = 0; pass < 2; ++pass) {
      if (c->D == 2) {
        if (pass == 0)
          hipLaunchKernelGGL((k_closure_sample<2, false>), dim3(c->ntiles), dim3(c->tile), 0, c->s,
                             c->tiles, c->blk, c->phys, c->flags, c->w64<2>(), c->nu64, c->signdis,
                             c->nb27, frac, seed, c->cl.X, c->cl.Z, cells, c->cl.tcount,
                             (const long long *)c->cl.toff, cap);
        else
          hipLaunchKernelGGL((k_closure_sample<2, true>), dim3(c->ntiles), dim3(c->tile), 0, c->s,
                             c->tiles, c->blk, c->phys, c->flags, c->w64<2>(), c->nu64, c->signdis,
                             c->nb27, frac, seed, c->cl.X, c->cl.Z, cells, c->cl.tcount,
                             (const long long *)c->cl.toff, cap);
      } else {
        if (pass == 0)
          hipLaunchKernelGGL((k_closure_sample<3, false>), dim3(c->ntiles), dim3(c->tile), 0, c->s,
                             c->tiles, c->blk, c->phys, c->flags, c->w64<3>(), c->nu64, c->signdis,
                             c->nb27, frac, seed, c->cl.X, c->cl.Z, cells, c->cl.tcount,
                             (const long long *)c->cl.toff, cap);
        else
          hipLaunchKernelGGL((k_closure_sample<3, true>), dim3(c->ntiles), dim3(c->tile), 0, c->s,
                             c->tiles, c->blk, c->phys, c->flags, c->w64<3>(), c->nu64, c->signdis,
                             c->nb27, frac, seed, c->cl.X, c->cl.Z, cells, c->cl.tcount,
                             (const long long *)c->cl.toff, cap);
      }
      HIPCHK(hipGetLastError());
      if (pass == 0) {
        hipLaunchKernelGGL(k_closure_offsets, dim3(1), dim3(1024), 0, c->s, c->cl.tcount, c->ntiles,
                           c->cl.toff, c->cl.count);
        HIPCHK(hipGetLastError());
      }
    }
  }
  unsigned long long n = 0;
  HIPCHK(hipMemcpyAsync(&n, c->cl.count, sizeof(n), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (total) *total = (int64_t)std::min<unsigned long long>(n, (unsigned long long)c->cl.cap);
  return IBG_OK;
}

static long long cl_n(ibg_ctx *c) {
  unsigned long long n = 0;
  if (hipMemcpy(&n, c->cl.count, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return (long long)std::min<unsigned long long>(n, (unsigned long long)c->cl.cap);
}

ibg_status ibg_closure_fit_stats(ibg_ctx *c, double margin) {
  if (!c || !c->cl.ready) return IBG_ERR_STATE;
  HIPCHK(hipStreamSynchronize(c->s));
  const long long n = cl_n(c);
  if (n < 2) return IBG_ERR_STATE;
  const int nb = 256;
  hipLaunchKernelGGL(k_closure_moments, dim3(nb), dim3(256), 0, c->s, c->cl.X, c->cl.Z, n, c->cl.mom);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)nb * (kMlpIn + 1) * 4);
  HIPCHK(hipMemcpyAsync(h.data(), c->cl.mom, h.size() * 8, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  std::vector<float> st(34);
  for (int j = 0; j <= kMlpIn; ++j) {
    double s = 0, s2 = 0, mn = 1e300, mx = -1e300;
    for (int b = 0; b < nb; ++b) {
      const double *o = &h[((size_t)b * (kMlpIn + 1) + j) * 4];
      s += o[0]; s2 += o[1]; mn = std::min(mn, o[2]); mx = std::max(mx, o[3]);
    }
    const double mean = s / n;
    const double sd = std::sqrt(std::max(s2 / n - mean * mean, 1e-12));
    if (j < kMlpIn) {
      st[j] = (float)mean;
      st[8 + j] = (float)sd;
      const double pad = margin * (mx - mn);
      st[16 + j] = (float)(mn - pad);
      st[24 + j] = (float)(mx + pad);
    } else {
      st[32] = (float)mean;
      st[33] = (float)sd;
    }
  }
  ibg_status rs = cl_put_stats(c, st.data());
  if (rs) return rs;
  c->cl.stats = true;
  return IBG_OK;
}

ibg_status ibg_closure_train(ibg_ctx *c, int32_t steps, int32_t batch, double lr, uint64_t seed,
                             double *loss) {
  if (!c || !c->cl.ready || !c->cl.stats) return IBG_ERR_STATE;
  if (steps < 0 || batch < 64 || batch % 64 != 0 || batch / 64 > c->cl.max_wg) return IBG_ERR_ARG;
  const long long n = cl_n(c);
  if (n < 1) return IBG_ERR_STATE;
  const int nwg = batch / 64;
  ibg_status st;
  for (int k = 0; k < steps; ++k) {
    const uint64_t sk = host_mix64(seed) + (uint64_t)(c->cl.step) * (uint64_t)batch;
    if ((st = cl_train_launch(c, c->cl.X, c->cl.Z, n, nwg, batch, 0, sk))) return st;
    ++c->cl.step;
    // cosine decay of the learning rate over the call, down to lr/100
    const double lrk = lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.141592653589793 * k / std::max(steps, 1))));
    hipLaunchKernelGGL(k_closure_adam, dim3(grid(kNParam, 256)), dim3(256), 0, c->s, c->cl.p, c->cl.m,
                       c->cl.v, c->cl.grad, nwg, (float)lrk, c->cl.step, (float *)nullptr);
    HIPCHK(hipGetLastError());
    if ((st = cl_pack(c))) return st;
  }
  if (loss) {
    std::vector<float> h(nwg);
    HIPCHK(hipMemcpyAsync(h.data(), c->cl.lossp, sizeof(float) * nwg, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    double s = 0;
    for (float v : h) s += v;
    *loss = s / batch;
  }
  return IBG_OK;
}

ibg_status ibg_closure_enable(ibg_ctx *c, int32_t on) {
  if (!c) return IBG_ERR_ARG;
  if (on) {
    if (!c->cl.ready || !c->cl.stats) return IBG_ERR_STATE;
    if (!c->mixed || !c->turb) return IBG_ERR_ARG;
  }
  if (c->cl.on != (on != 0)) {
    drop_graphs(c);
  }
  c->cl.on = (on != 0);
  return IBG_OK;
}

ibg_status ibg_closure_params(ibg_ctx *c, float *out, int64_t n) {
  if (!c || !out || !c->cl.ready) return IBG_ERR_STATE;
  if (n != kNParam + 34) return IBG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(out, c->cl.p, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(out + kNParam, c->cl.stat, sizeof(float) * 34, hipMemcpyDeviceToHost));
  return IBG_OK;
}

ibg_status ibg_closure_set_params(ibg_ctx *c, const float *in, int64_t n) {
  if (!c || !in || !c->cl.ready) return IBG_ERR_STATE;
  if (n != kNParam + 34) return IBG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(c->cl.p, in, sizeof(float) * kNParam, hipMemcpyHostToDevice));
  ibg_status rs = cl_put_stats(c, in + kNParam);
  if (rs) return rs;
  c->cl.stats = true;
  return cl_pack(c);
}

ibg_status ibg_closure_infer(ibg_ctx *c, const float *x, float *z, int64_t n) {
  if (!c || !x || !z || !c->cl.ready || n < 1) return IBG_ERR_ARG;
  float *dx = nullptr, *dz = nullptr;
  HIPCHK(hipMalloc(&dx, sizeof(float) * n * kMlpIn));
  if (hipMalloc(&dz, sizeof(float) * n) != hipSuccess) { (void)hipFree(dx); return IBG_ERR_HIP; }
  hipError_t e = hipMemcpy(dx, x, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_closure_infer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s, c->mlpdev(),
                       dx, dz, (long long)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e == hipSuccess) e = hipMemcpy(z, dz, sizeof(float) * n, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dz);
  return e == hipSuccess ? IBG_OK : IBG_ERR_HIP;
}

ibg_status ibg_closure_infer_noalloc(ibg_ctx *c, const float *x, float *z, int64_t n) {
  if (!c || !x || !z || !c->cl.ready || n < 1) return IBG_ERR_ARG;
  if (n > kScratchRows) {
    fprintf(stderr, "ibgpu: ibg_closure_infer_noalloc: %lld rows exceed the staging limit of %lld\n",
            (long long)n, kScratchRows);
    return IBG_ERR_ARG;
  }
  float *dx = g_scratch_dev[c->dev];   // set by ibg_create on this device
  if (!dx) return IBG_ERR_STATE;
  float *dz = dx + (size_t)kScratchRows * kMlpIn;
  // rows [0, n) of x are copied in and rows [0, n) of z written: nothing
  // left from an earlier call is read
  HIPCHK(hipMemcpyAsync(dx, x, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, c->s));
  hipLaunchKernelGGL(k_closure_infer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s, c->mlpdev(),
                     (const float *)dx, dz, (long long)n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(z, dz, sizeof(float) * n, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return IBG_OK;
}

ibg_status ibg_closure_grad(ibg_ctx *c, const float *x, const float *z, int64_t n, float *grad,
                            double *loss) {
  if (!c || !x || !z || !grad || !c->cl.ready || n < 1) return IBG_ERR_ARG;
  const int nwg = (int)((n + 63) / 64);
  if (nwg > c->cl.max_wg) return IBG_ERR_ARG;
  float *dx = nullptr, *dz = nullptr;
  HIPCHK(hipMalloc(&dx, sizeof(float) * n * kMlpIn));
  if (hipMalloc(&dz, sizeof(float) * n) != hipSuccess) { (void)hipFree(dx); return IBG_ERR_HIP; }
  hipError_t e = hipMemcpy(dx, x, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dz, z, sizeof(float) * n, hipMemcpyHostToDevice);
  ibg_status st = IBG_OK;
  if (e == hipSuccess) st = cl_train_launch(c, dx, dz, n, nwg, (int)n, 1, 0);
  if (e == hipSuccess && !st) {
    // the Adam kernel in its summing mode: the gradient, no update
    hipLaunchKernelGGL(k_closure_adam, dim3(grid(kNParam, 256)), dim3(256), 0, c->s, c->cl.p, c->cl.m,
                       c->cl.v, c->cl.grad, nwg, 0.0f, 1, c->cl.gsum);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    if (e == hipSuccess) e = hipMemcpy(grad, c->cl.gsum, sizeof(float) * kNParam, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) {
      std::vector<float> h(nwg);
      e = hipMemcpy(h.data(), c->cl.lossp, sizeof(float) * nwg, hipMemcpyDeviceToHost);
      double s = 0;
      for (float q : h) s += q;
      *loss = s / (double)n;
    }
  }
  (void)hipFree(dx);
  (void)hipFree(dz);
  if (st) return st;
  return e == hipSuccess ? IBG_OK : IBG_ERR_HIP;
}

ibg_status ibg_closure_dataset(ibg_ctx *c, float *x, float *z, int64_t *cell, int64_t n) {
  if (!c || !c->cl.ready || n < 0) return IBG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  const long long have = cl_n(c);
  if (n > have) return IBG_ERR_ARG;
  if (n == 0) return IBG_OK;
  if (x) HIPCHK(hipMemcpy(x, c->cl.X, sizeof(float) * n * kMlpIn, hipMemcpyDeviceToHost));
  if (z) HIPCHK(hipMemcpy(z, c->cl.Z, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (cell) {
    std::vector<int> h(n);
    HIPCHK(hipMemcpy(h.data(), cl_cells(c), sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) cell[i] = (h[i] >= 0 && h[i] < c->N) ? c->d2h[h[i]] : -1;
  }
  return IBG_OK;
}

ibg_status ibg_closure_counts(ibg_ctx *c, int64_t *on_net, int64_t *fallback) {
  if (!c || !c->cl.ready) return IBG_ERR_STATE;
  std::vector<unsigned long long> slots((size_t)kMlSlots * kMlSlotStride);
  const size_t bytes = slots.size() * sizeof(unsigned long long);
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(slots.data(), c->cl.mlcount, bytes, hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(c->cl.mlcount, 0, bytes));
  unsigned long long h[2] = {0, 0};
  for (int k = 0; k < kMlSlots; ++k) {
    h[0] += slots[(size_t)k * kMlSlotStride];
    h[1] += slots[(size_t)k * kMlSlotStride + 1];
  }
  if (on_net) *on_net = (int64_t)h[0];
  if (fallback) *fallback = (int64_t)h[1];
  return IBG_OK;
}

}  // extern "C"
