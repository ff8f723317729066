// ns_solver.cpp — C++ driver in the shape of the reference's main.cpp.
//
//   ibg_run [steps] [fp64|mixed] [stages] [laminar|sa|ml] [static|move] [solution.plt]
//
// move pitches the section 5 deg about its quarter chord and heaves it 0.05
// chord at reduced frequency 0.1 (the IB re-tagged every step,
// Bodymove.cpp:37-125); a last argument writes the final Tecplot solution
// (Pltoutput::Output_Soln, a restart file).
// sa runs the TURBULENCE build (SA closure, working variable from its
// freestream, Userfunction_cy_naca15.cpp:368-370); ml trains the learned
// closure on cells of an FP64 SA run of the same case first (mixed mode).
// Builds the small 2D case (NACA0012, Mach 0.5, AoA 2 deg, Re 5000, 64x64
// base grid + 3 AMR levels), initialises the freestream (InitFlow
// NS_Solver.H:432-446) and runs the time loop of main.cpp:84-99 through the
// NS_Solver class of ns_solver.hpp, printing the residual line of
// ResinDomain (NS_Solver.H:353-354, normalised by its step-0 value) and the
// force coefficients of the case file (Userfunction_cy_naca15.cpp:437-455).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ib_mesh.h"
#include "ns_solver.hpp"

int main(int argc, char **argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 64;
  const int mode = (argc > 2 && std::string(argv[2]) == "mixed") ? IBG_MODE_MIXED : IBG_MODE_FP64;
  const int stages = argc > 3 ? std::atoi(argv[3]) : 3;
  const std::string closure = argc > 4 ? std::string(argv[4]) : std::string("laminar");
  const bool turb = closure != "laminar";
  const bool move = argc > 5 && std::string(argv[5]) == "move";
  const char *soln = argc > 6 ? argv[6] : nullptr;

  const int npts = 100;
  std::vector<double> xy(4 * npts);
  ibm_naca4(0.12, npts, 0.0, 0.0, 2.0, xy.data());
  ibm_case cs;
  std::memset(&cs, 0, sizeof(cs));
  cs.dim = 2;
  cs.nx = cs.ny = 64;
  cs.nz = 1;
  cs.dh0[0] = cs.dh0[1] = 20.0 / 64;
  cs.dh0[2] = 0.1;
  cs.origin[0] = cs.origin[1] = -10.0;
  cs.nlevels = 4;
  cs.twod_from = 1;
  const double blo[3] = {-0.6, -0.3, -0.15}, bhi[3] = {2.0, 1.0, 0.5};
  for (int l = 0; l < 3; ++l) { cs.band_lo[l] = blo[l]; cs.band_hi[l] = bhi[l]; }
  cs.nbody = 2;
  cs.body_npts[0] = cs.body_npts[1] = npts;
  cs.body_xy = xy.data();
  cs.body_center[2] = 0.05;
  cs.span = 0.1;
  cs.span_patches = 1;
  cs.gama = 1.4; cs.Ma = 0.5; cs.Re = 5000.0; cs.Pr = 0.72; cs.cfl = 0.5;
  cs.backpressure = 0.0;
  cs.dt_init = 1.0e30;

  ibm_handle *h = ibm_build(&cs, 1);
  if (!h) { std::fprintf(stderr, "mesh build failed: %s\n", ibm_last_error()); return 1; }
  const ibg_mesh_soa *m = ibm_mesh(h);
  std::vector<ibg_physics> phys(1);
  ibm_physics(&cs, &phys[0]);
  std::printf("mesh: %d cells, %d faces, %d IB cells, %d patches\n", m->ncells, m->nfaces, m->nib, m->npatch);

  try {
    std::vector<float> net;
    if (closure == "ml") {
      // training data from the plain closure (FP64), as the reference runs it
      ibgpu::NS_Solver teacher(*m, phys, stages, IBG_MODE_FP64, true, true);
      std::vector<double> fv((size_t)m->ncells * IBG_NFV);
      ibm_init_state(h, phys.data(), fv.data());
      teacher.SetState(fv);
      teacher.SetViseddy(std::vector<double>(m->ncells, phys[0].nut_inf));
      teacher.ClosureInit(1 << 20, 20261202);
      long long n = 0;
      for (int k = 0; k < 16; ++k) {
        teacher.TimeMarching(8 * k, 8, true);
        n = teacher.ClosureSample(0.5, 100 + k);
      }
      const double loss = teacher.ClosureTrain(2000, 1024, 1e-3, 7);
      std::printf("closure: %lld samples, training loss %.3e\n", n, loss);
      net = teacher.ClosureParams();
    }
    ibgpu::NS_Solver solver(*m, phys, stages, mode, true, turb);
    std::vector<double> fv((size_t)m->ncells * IBG_NFV);
    ibm_init_state(h, phys.data(), fv.data());
    solver.SetState(fv);
    if (turb) solver.SetViseddy(std::vector<double>(m->ncells, phys[0].nut_inf));
    if (!net.empty()) {
      solver.ClosureInit(64, 1);
      solver.ClosureSetParams(net);
      solver.ClosureEnable(true);
    }
    std::vector<double> pt_off;
    if (move) {
      pt_off.resize(4 * npts);
      ibm_naca4_offsets(0.12, npts, pt_off.data());
      ibg_surface surf;
      if (ibm_surface(h, pt_off.data(), 2 * npts, &surf) != 0) {
        std::fprintf(stderr, "surface: %s\n", ibm_last_error());
        ibm_free(h);
        return 1;
      }
      solver.UploadSurface(*m, surf);
      const ibg_pitch_heave law = {0.0, 0.0, 2.0, 5.0, 0.2, 0.05};
      solver.SetMotion(std::vector<ibg_pitch_heave>(1, law), 0.0);
    }
    std::vector<double> linf, l2, init_res;
    for (int ts = 0; ts < steps; ++ts) {
      solver.TimeMarching(ts, 1, false);
      solver.ResinDomain(linf, l2);
      if (ts == 0) init_res = linf;
      double r[5];
      for (int k = 0; k < 5; ++k) r[k] = init_res[k] > 0 ? linf[k] / init_res[k] : 0.0;
      r[3] = 0.0;   // w residual is zeroed in 2D (NS_Solver.H:343-345)
      if (ts % 8 == 0 || ts == steps - 1)
        std::printf("Time step %d roe %e u %e v %e w %e e %e\n", ts, r[0], r[1], r[2], r[3], r[4]);
    }
    const std::vector<double> F = solver.ComptWallForce();
    std::printf("C_D %.8f C_L %.8f C_M %.8f\n", F[0] / phys[0].force_ref, F[1] / phys[0].force_ref,
                F[5] / phys[0].moment_ref);
    if (move) std::printf("body time %.8e\n", solver.GetTime()[0]);
    if (soln) {
      const std::vector<double> st = solver.GetState();
      const std::vector<double> nut = turb ? solver.GetViseddy() : std::vector<double>();
      const std::vector<double> dt = solver.GetDt();
      const std::vector<double> t = move ? solver.GetTime() : std::vector<double>();
      if (ibm_write_solution(h, phys.data(), st.data(), turb ? nut.data() : nullptr, steps, dt.data(),
                             move ? t.data() : nullptr, soln) != 0) {
        std::fprintf(stderr, "solution: %s\n", ibm_last_error());
        ibm_free(h);
        return 1;
      }
      std::printf("solution written to %s\n", soln);
    }
  } catch (const ibgpu::Error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    ibm_free(h);
    return 2;
  }
  ibm_free(h);
  return 0;
}
