// ns_solver.hpp — C++ host class with the reference solver's call surface.
//
// NS_Solver (NS_Solver.H:190-872) is the class the reference's driver calls
// once per step (main.cpp:87-97).  This class keeps its method names and
// call order and forwards each one to the C ABI of libibgpu.so, so a driver
// written against the reference's NS_Solver / AMR / Body / BCValues surface
// runs the residual path on the GPU.  The mesh is passed once (and after every
// regrid) as a flattened ibg_mesh_soa.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "ibgpu.h"

namespace ibgpu {

struct Error : std::runtime_error {
  int code;
  long long cell;
  Error(const std::string &w, int c, long long cell_)
      : std::runtime_error(w + " returned status " + std::to_string(c) + " (cell " +
                           std::to_string(cell_) + ")"),
        code(c), cell(cell_) {}
};

class NS_Solver {
 public:
  // NS_Solver(Mesh&, Domain&, int) NS_Solver.H:195-237: the stage table
  // (tp[]) is chosen by the stage count, as sub_time_step selects it.
  // turbulence selects the -D TURBULENCE build (the SA closure).
  NS_Solver(const ibg_mesh_soa &mesh, const std::vector<ibg_physics> &phys, int sub_time_step = 3,
            int mode = IBG_MODE_FP64, bool temporal_refine = true, bool turbulence = false)
      : ncells_(mesh.ncells), ncases_(mesh.ncases), mesh_(&mesh), phys_(phys) {
    ibg_config cfg{};
    cfg.mode = mode;
    cfg.nstage = sub_time_step;
    const double a3[3] = {0.1918, 0.4929, 1.0};
    const double a4[4] = {0.1084, 0.2602, 0.5052, 1.0};
    for (int i = 0; i < sub_time_step; ++i)
      cfg.alpha[i] = sub_time_step == 4 ? a4[i] : (sub_time_step == 3 ? a3[i] : 1.0);
    cfg.temporal_refine = temporal_refine ? 1 : 0;
    cfg.dt_var_step = temporal_refine ? (1 << (mesh.nlevels - 1)) : 1;
    cfg.viscous = 1;
    cfg.turbulence = turbulence ? 1 : 0;
    check(ibg_create(&cfg, &ctx_), "ibg_create");
    check(ibg_upload_mesh(ctx_, &mesh, phys.data()), "ibg_upload_mesh");
  }
  ~NS_Solver() { ibg_destroy(ctx_); }
  NS_Solver(const NS_Solver &) = delete;
  NS_Solver &operator=(const NS_Solver &) = delete;

  // InitFlow / direct m_data access (NS_Solver.H:432-446, Meshlevel.H:233)
  void SetState(const std::vector<double> &fv) { check(ibg_set_state(ctx_, fv.data(), ncells_), "ibg_set_state"); }
  std::vector<double> GetState() {
    std::vector<double> fv((size_t)ncells_ * IBG_NFV);
    check(ibg_get_state(ctx_, fv.data(), ncells_), "ibg_get_state");
    return fv;
  }
  // FlowVariables::viseddy of the TURBULENCE build (FlowVariables.H:327-329)
  void SetViseddy(const std::vector<double> &nut) { check(ibg_set_nut(ctx_, nut.data(), ncells_), "ibg_set_nut"); }
  std::vector<double> GetViseddy() {
    std::vector<double> nut(ncells_);
    check(ibg_get_nut(ctx_, nut.data(), ncells_), "ibg_get_nut");
    return nut;
  }

  // the per-step sequence of main.cpp:87-97
  void InterfaceExchange(int ts) { check(ibg_interface_exchange(ctx_, ts), "InterfaceExchange"); }
  void IBConditions_Image(const ibg_body_motion *motion = nullptr) { check(ibg_ib_fill(ctx_, motion), "IBConditions_Image"); }
  std::vector<double> Comptdt(int ts) {
    std::vector<double> dt(ncases_);
    check(ibg_compute_dt(ctx_, ts, dt.data()), "Comptdt");
    return dt;
  }
  void ComptOldConsVar(int ts) { check(ibg_old_cons(ctx_, ts), "ComptOldConsVar"); }
  void BCBoxTreat(int ts) { check(ibg_bc_fill(ctx_, ts), "BCBoxTreat"); }
  void ComptViscousStress(int ts) { check(ibg_viscous_flux(ctx_, ts), "ComptViscousStress"); }
  void Stage(int ts, int subt) { check(ibg_stage(ctx_, ts, subt), "Stage"); }
  // TimeMarching(AMR&) NS_Solver.H:657-702 (with the exchange and IB fill
  // of main.cpp in front of every step), nsteps steps from ts
  void TimeMarching(int ts, int nsteps = 1, bool use_graph = true) {
    check(ibg_time_march(ctx_, ts, nsteps, use_graph ? 1 : 0), "TimeMarching");
  }
  // ResinDomain NS_Solver.H:264-378 (unnormalised L-inf and L2 per case)
  void ResinDomain(std::vector<double> &linf, std::vector<double> &l2) {
    linf.assign((size_t)ncases_ * 5, 0.0);
    l2.assign((size_t)ncases_ * 5, 0.0);
    check(ibg_residual(ctx_, linf.data(), l2.data()), "ResinDomain");
  }
  // Body::ComptWallForce Body.cpp:439-558 summed per case
  std::vector<double> ComptWallForce(const ibg_body_motion *motion = nullptr) {
    std::vector<double> f((size_t)ncases_ * 6);
    check(ibg_forces(ctx_, motion, f.data()), "ComptWallForce");
    return f;
  }
  // SA residual (res_vis_sum, NS_Solver.H:297, 321)
  void ResinDomainViseddy(std::vector<double> &linf, std::vector<double> &l2) {
    linf.assign(ncases_, 0.0);
    l2.assign(ncases_, 0.0);
    check(ibg_residual_nut(ctx_, linf.data(), l2.data()), "ibg_residual_nut");
  }

  // Body::MovingBody_Image (Bodymove.cpp:37-125) between regrids: the
  // surface once, then a move per step to the structural solver's pose, or
  // a prescribed pitch/heave law applied by TimeMarching itself
  void UploadSurface(const ibg_mesh_soa &mesh, const ibg_surface &surf) {
    check(ibg_upload_surface(ctx_, &mesh, &surf), "ibg_upload_surface");
  }
  void MovingBody_Image(const std::vector<ibg_body_pose> &pose) {
    check(ibg_move_bodies(ctx_, pose.data()), "MovingBody_Image");
  }
  void SetMotion(const std::vector<ibg_pitch_heave> &law, double t0) {
    check(ibg_set_motion(ctx_, law.data(), t0), "ibg_set_motion");
  }
  std::vector<double> GetTime() {
    std::vector<double> t(ncases_);
    check(ibg_get_time(ctx_, t.data()), "ibg_get_time");
    return t;
  }
  std::vector<double> GetDt() {
    std::vector<double> dt(ncases_);
    check(ibg_get_dt(ctx_, dt.data()), "ibg_get_dt");
    return dt;
  }

  // Learned closure (no reference counterpart): sample the current state,
  // fit the statistics, train, and switch the stage to the network.
  void ClosureInit(long long capacity, unsigned long long seed) {
    check(ibg_closure_init(ctx_, capacity, seed), "ibg_closure_init");
  }
  long long ClosureSample(double frac, unsigned long long seed) {
    int64_t n = 0;
    check(ibg_closure_sample(ctx_, frac, seed, &n), "ibg_closure_sample");
    return (long long)n;
  }
  double ClosureTrain(int steps, int batch, double lr, unsigned long long seed) {
    double loss = 0.0;
    check(ibg_closure_fit_stats(ctx_, 0.0), "ibg_closure_fit_stats");
    check(ibg_closure_train(ctx_, steps, batch, lr, seed, &loss), "ibg_closure_train");
    return loss;
  }
  std::vector<float> ClosureParams() {
    std::vector<float> p(IBG_CLOSURE_NPARAM + 34);
    check(ibg_closure_params(ctx_, p.data(), (int64_t)p.size()), "ibg_closure_params");
    return p;
  }
  void ClosureSetParams(const std::vector<float> &p) {
    check(ibg_closure_set_params(ctx_, p.data(), (int64_t)p.size()), "ibg_closure_set_params");
  }
  void ClosureEnable(bool on) { check(ibg_closure_enable(ctx_, on ? 1 : 0), "ibg_closure_enable"); }
  // Stage-kernel tile size by timing (resets the state: call before SetState)
  int AutotuneTile(const std::vector<int> &cand, int reps = 20) {
    int32_t best = 0;
    std::vector<int32_t> c(cand.begin(), cand.end());
    check(ibg_autotune_tile(ctx_, mesh_, phys_.data(), c.data(), (int32_t)c.size(), reps, &best, nullptr),
          "ibg_autotune_tile");
    return best;
  }

  long long CellStageUpdates() {
    int64_t n = 0;
    ibg_cell_stage_updates(ctx_, &n);
    return (long long)n;
  }
  ibg_ctx *ctx() { return ctx_; }

 private:
  void check(ibg_status st, const char *where) {
    if (st == IBG_OK) return;
    int64_t cell = -1;
    int32_t code = 0;
    if (ctx_) ibg_last_error(ctx_, &cell, &code);
    throw Error(where, (int)st, (long long)cell);
  }
  ibg_ctx *ctx_ = nullptr;
  int ncells_, ncases_;
  const ibg_mesh_soa *mesh_;
  std::vector<ibg_physics> phys_;
};

}  // namespace ibgpu
