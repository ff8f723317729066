"""Python mirror of the NS_Solver call surface over libibgpu.so (ctypes).

Method names follow the reference's solver (NS_Solver.H:195-703,
Body.H:723, AMR.H:1012, BCValues.H:51): TimeMarching, ComputeFaceFlux,
ComptViscousStress, Comptdt, ResinDomain, IBConditions_Image,
InterfaceExchange, BCBoxTreat, ComptWallForce.  Every call goes to the HIP
engine; a missing library raises (there is no CPU path here).
"""
import ctypes as C

import numpy as np

from . import _ffi

ALPHA = {1: [1.0], 3: [0.1918, 0.4929, 1.0], 4: [0.1084, 0.2602, 0.5052, 1.0]}   # NS_Solver.H:202-231
MODES = {"fp64": 0, "mixed": 1}


class IBGError(RuntimeError):
    def __init__(self, code, where, cell=-1):
        super().__init__(f"{where} returned ibg_status {code} (first offending cell {cell})")
        self.code = code
        self.cell = cell


def make_config(mode="fp64", nstage=3, dt_var_step=None, temporal_refine=True, limiter=0,
                viscous=True, tile=0, nlevels=1, turbulence=False, schedule=0, reflux=False):
    cfg = _ffi.Config()
    cfg.mode = MODES[mode]
    cfg.nstage = nstage
    for i, a in enumerate(ALPHA[nstage]):
        cfg.alpha[i] = a
    period = (1 << (nlevels - 1)) if temporal_refine else 1
    cfg.dt_var_step = period if dt_var_step is None else dt_var_step
    cfg.temporal_refine = 1 if temporal_refine else 0
    cfg.limiter = limiter
    cfg.viscous = 1 if viscous else 0
    cfg.tile = tile
    cfg.turbulence = 1 if turbulence else 0
    cfg.schedule = schedule
    cfg.reflux = 1 if reflux else 0
    return cfg


class Engine:
    def __init__(self, mesh, mode="fp64", nstage=3, temporal_refine=True, limiter=0,
                 viscous=True, tile=0, dt_var_step=None, turbulence=False, schedule=0,
                 reflux=False):
        """reflux: the optional coarse/fine flux balance (BalanceFaceFlux
        Fluxmodification.cpp:291-358, inactive in the reference), which
        needs temporal_refine=False."""
        self.lib = _ffi.gpu_lib()
        self.mesh = mesh
        self.turbulence = bool(turbulence)
        self.cfg = make_config(mode, nstage, dt_var_step, temporal_refine, limiter, viscous, tile,
                               mesh.nlevels, turbulence, schedule, reflux)
        self.ctx = C.c_void_p()
        self._chk(self.lib.ibg_create(C.byref(self.cfg), C.byref(self.ctx)), "ibg_create")
        self._chk(self.lib.ibg_upload_mesh(self.ctx, mesh.soa_ptr, mesh.phys), "ibg_upload_mesh")

    def _chk(self, code, where):
        if code != 0:
            cell = C.c_int64(-1)
            c2 = C.c_int32(0)
            try:
                if self.ctx:
                    self.lib.ibg_last_error(self.ctx, C.byref(cell), C.byref(c2))
            except Exception:
                pass
            raise IBGError(code, where, cell.value)

    # state ------------------------------------------------------------------
    def set_state(self, fv):
        fv = np.ascontiguousarray(fv, dtype=np.float64)
        self._chk(self.lib.ibg_set_state(self.ctx, fv.ctypes.data_as(_ffi.f64p), fv.shape[0]),
                  "ibg_set_state")

    def get_state(self):
        fv = np.zeros((self.mesh.ncells, _ffi.IBG_NFV))
        self._chk(self.lib.ibg_get_state(self.ctx, fv.ctypes.data_as(_ffi.f64p), fv.shape[0]),
                  "ibg_get_state")
        return fv

    def set_nut(self, nut):
        nut = np.ascontiguousarray(nut, dtype=np.float64)
        self._chk(self.lib.ibg_set_nut(self.ctx, nut.ctypes.data_as(_ffi.f64p), nut.shape[0]),
                  "ibg_set_nut")

    def get_nut(self):
        nut = np.zeros(self.mesh.ncells)
        self._chk(self.lib.ibg_get_nut(self.ctx, nut.ctypes.data_as(_ffi.f64p), nut.shape[0]),
                  "ibg_get_nut")
        return nut

    def set_dt(self, dt):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        self._chk(self.lib.ibg_set_dt(self.ctx, dt.ctypes.data_as(_ffi.f64p)), "ibg_set_dt")

    def get_dt(self):
        dt = np.zeros(self.mesh.ncases)
        self._chk(self.lib.ibg_get_dt(self.ctx, dt.ctypes.data_as(_ffi.f64p)), "ibg_get_dt")
        return dt

    # reference call surface -------------------------------------------------
    def InterfaceExchange(self, ts):
        self._chk(self.lib.ibg_interface_exchange(self.ctx, ts), "ibg_interface_exchange")

    def IBConditions_Image(self, motion=None):
        self._chk(self.lib.ibg_ib_fill(self.ctx, motion), "ibg_ib_fill")

    def BCBoxTreat(self, ts):
        self._chk(self.lib.ibg_bc_fill(self.ctx, ts), "ibg_bc_fill")

    def Comptdt(self, ts=0):
        dt = np.zeros(self.mesh.ncases)
        self._chk(self.lib.ibg_compute_dt(self.ctx, ts, dt.ctypes.data_as(_ffi.f64p)), "ibg_compute_dt")
        return dt

    def ComptViscousStress(self, ts):
        self._chk(self.lib.ibg_viscous_flux(self.ctx, ts), "ibg_viscous_flux")

    def Stage(self, ts, subt):
        self._chk(self.lib.ibg_stage(self.ctx, ts, subt), "ibg_stage")

    def prepare_march(self, ts0, nsteps):
        """Capture the graphs TimeMarching(ts0, nsteps, use_graph=True)
        replays, without running anything."""
        self._chk(self.lib.ibg_prepare_march(self.ctx, ts0, nsteps), "ibg_prepare_march")

    def TimeMarching(self, ts0, nsteps=1, use_graph=False):
        self._chk(self.lib.ibg_time_march(self.ctx, ts0, nsteps, 1 if use_graph else 0),
                  "ibg_time_march")

    def ResinDomain(self):
        n = self.mesh.ncases * 5
        linf, l2 = np.zeros(n), np.zeros(n)
        self._chk(self.lib.ibg_residual(self.ctx, linf.ctypes.data_as(_ffi.f64p),
                                        l2.ctypes.data_as(_ffi.f64p)), "ibg_residual")
        return linf, l2

    def ResinDomain_nut(self):
        n = self.mesh.ncases
        linf, l2 = np.zeros(n), np.zeros(n)
        self._chk(self.lib.ibg_residual_nut(self.ctx, linf.ctypes.data_as(_ffi.f64p),
                                            l2.ctypes.data_as(_ffi.f64p)), "ibg_residual_nut")
        return linf, l2

    def ComptWallForce(self, motion=None):
        out = np.zeros(self.mesh.ncases * 6)
        self._chk(self.lib.ibg_forces(self.ctx, motion, out.ctypes.data_as(_ffi.f64p)), "ibg_forces")
        return out

    # probes -----------------------------------------------------------------
    def faceflux(self):
        out = np.zeros((self.mesh.nfaces, 5))
        self._chk(self.lib.ibg_get_faceflux(self.ctx, out.ctypes.data_as(_ffi.f64p), out.shape[0]),
                  "ibg_get_faceflux")
        return out

    def vsflux(self):
        out = np.zeros((self.mesh.nfaces, 5))
        self._chk(self.lib.ibg_get_vsflux(self.ctx, out.ctypes.data_as(_ffi.f64p), out.shape[0]),
                  "ibg_get_vsflux")
        return out

    def source(self):
        out = np.zeros((self.mesh.ncells, 5))
        self._chk(self.lib.ibg_get_source(self.ctx, out.ctypes.data_as(_ffi.f64p), out.shape[0]),
                  "ibg_get_source")
        return out

    def mutflux(self):
        out = np.zeros(self.mesh.nfaces)
        self._chk(self.lib.ibg_get_mutflux(self.ctx, out.ctypes.data_as(_ffi.f64p), out.shape[0]),
                  "ibg_get_mutflux")
        return out

    def nut_source(self):
        out = np.zeros(self.mesh.ncells)
        self._chk(self.lib.ibg_get_nut_source(self.ctx, out.ctypes.data_as(_ffi.f64p),
                                              out.shape[0]), "ibg_get_nut_source")
        return out

    def autotune_tile(self, candidates=(64, 128, 192, 256), reps=20):
        """Pick the stage-kernel tile size and schedule by timing (resets the
        state).  Returns (best tile, {(tile, schedule): ms}); the schedule
        stays in the context (self.schedule)."""
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(2 * cand.size)
        best = C.c_int32()
        self._chk(self.lib.ibg_autotune_tile(self.ctx, self.mesh.soa_ptr, self.mesh.phys,
                                             cand.ctypes.data_as(_ffi.i32p), cand.size, reps,
                                             C.byref(best), ms.ctypes.data_as(_ffi.f64p)),
                  "ibg_autotune_tile")
        table = {(int(t), s): float(ms[2 * k + s]) for k, t in enumerate(cand.tolist()) for s in (0, 1)}
        self.schedule = min(table, key=table.get)[1]
        return best.value, table

    # learned closure ----------------------------------------------------------
    def closure_init(self, capacity, seed=20261202):
        self._chk(self.lib.ibg_closure_init(self.ctx, int(capacity), seed), "ibg_closure_init")

    def closure_sample(self, frac, seed):
        n = C.c_int64()
        self._chk(self.lib.ibg_closure_sample(self.ctx, float(frac), seed, C.byref(n)),
                  "ibg_closure_sample")
        return n.value

    def closure_fit_stats(self, margin=0.0):
        self._chk(self.lib.ibg_closure_fit_stats(self.ctx, float(margin)), "ibg_closure_fit_stats")

    def closure_train(self, steps, batch=8192, lr=1e-3, seed=1):
        loss = C.c_double()
        self._chk(self.lib.ibg_closure_train(self.ctx, steps, batch, float(lr), seed, C.byref(loss)),
                  "ibg_closure_train")
        return loss.value

    def closure_enable(self, on=True):
        self._chk(self.lib.ibg_closure_enable(self.ctx, 1 if on else 0), "ibg_closure_enable")

    def closure_params(self):
        out = np.zeros(_ffi.IBG_CLOSURE_NPARAM + 34, dtype=np.float32)
        self._chk(self.lib.ibg_closure_params(self.ctx, out.ctypes.data_as(_ffi.f32p), out.size),
                  "ibg_closure_params")
        return out

    def closure_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self.lib.ibg_closure_set_params(self.ctx, p.ctypes.data_as(_ffi.f32p), p.size),
                  "ibg_closure_set_params")

    def closure_infer(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        z = np.zeros(x.shape[0], dtype=np.float32)
        self._chk(self.lib.ibg_closure_infer(self.ctx, x.ctypes.data_as(_ffi.f32p),
                                             z.ctypes.data_as(_ffi.f32p), x.shape[0]),
                  "ibg_closure_infer")
        return z

    def closure_infer_noalloc(self, x):
        """closure_infer staged through the library's fixed per-device
        array (no device allocation; at most 65536 rows)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        z = np.zeros(x.shape[0], dtype=np.float32)
        self._chk(self.lib.ibg_closure_infer_noalloc(self.ctx, x.ctypes.data_as(_ffi.f32p),
                                                     z.ctypes.data_as(_ffi.f32p), x.shape[0]),
                  "ibg_closure_infer_noalloc")
        return z

    def closure_grad(self, x, z):
        x = np.ascontiguousarray(x, dtype=np.float32)
        z = np.ascontiguousarray(z, dtype=np.float32)
        g = np.zeros(_ffi.IBG_CLOSURE_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._chk(self.lib.ibg_closure_grad(self.ctx, x.ctypes.data_as(_ffi.f32p),
                                            z.ctypes.data_as(_ffi.f32p), x.shape[0],
                                            g.ctypes.data_as(_ffi.f32p), C.byref(loss)),
                  "ibg_closure_grad")
        return g, loss.value

    def closure_dataset(self, n):
        x = np.zeros((n, 8), dtype=np.float32)
        z = np.zeros(n, dtype=np.float32)
        cell = np.zeros(n, dtype=np.int64)
        self._chk(self.lib.ibg_closure_dataset(self.ctx, x.ctypes.data_as(_ffi.f32p),
                                               z.ctypes.data_as(_ffi.f32p),
                                               cell.ctypes.data_as(_ffi.i64p), n),
                  "ibg_closure_dataset")
        return x, z, cell

    def closure_counts(self):
        a, b = C.c_int64(), C.c_int64()
        self._chk(self.lib.ibg_closure_counts(self.ctx, C.byref(a), C.byref(b)), "ibg_closure_counts")
        return a.value, b.value

    def timing(self):
        tot, st = C.c_double(), C.c_double()
        n = C.c_int64()
        self.lib.ibg_last_timing(self.ctx, C.byref(tot), C.byref(st), C.byref(n))
        return tot.value, st.value, n.value

    def bench_stage(self, ts, subt, reps):
        ms = C.c_double()
        nc, nf = C.c_int64(), C.c_int64()
        self._chk(self.lib.ibg_bench_stage(self.ctx, ts, subt, reps, C.byref(ms), C.byref(nc),
                                           C.byref(nf)), "ibg_bench_stage")
        return ms.value, nc.value, nf.value

    def bench_copy(self, nbytes, reps=20):
        ms = C.c_double()
        self._chk(self.lib.ibg_bench_copy(self.ctx, int(nbytes), reps, C.byref(ms)), "ibg_bench_copy")
        return ms.value

    def cell_stage_updates(self):
        n = C.c_int64()
        self.lib.ibg_cell_stage_updates(self.ctx, C.byref(n))
        return n.value

    # moving bodies ------------------------------------------------------------
    def upload_surface(self, ib_capacity=0):
        """Enable moving bodies (Body::MovingBody_Image between regrids)."""
        self._surface = self.mesh.surface(ib_capacity)
        self._chk(self.lib.ibg_upload_surface(self.ctx, self.mesh.soa_ptr,
                                              C.byref(self._surface[0])), "ibg_upload_surface")

    def move_bodies(self, poses):
        """poses: per case {pivot x, pivot y, cos a, sin a} (Mesh.pose)."""
        arr = (_ffi.BodyPose * self.mesh.ncases)()
        for k, p in enumerate(poses):
            arr[k].pivot[0], arr[k].pivot[1] = float(p[0]), float(p[1])
            arr[k].cosa, arr[k].sina = float(p[2]), float(p[3])
        self._chk(self.lib.ibg_move_bodies(self.ctx, arr), "ibg_move_bodies")

    def set_motion(self, laws, t0=0.0):
        """laws: per case (x0, y0, aoa0_deg, amp_deg, omega, heave) or None."""
        if laws is None:
            self._chk(self.lib.ibg_set_motion(self.ctx, None, 0.0), "ibg_set_motion")
            self._law = None
            return
        self._law = [tuple(lw) for lw in laws]
        arr = (_ffi.PitchHeave * self.mesh.ncases)()
        for k, lw in enumerate(laws):
            (arr[k].x0, arr[k].y0, arr[k].aoa0_deg, arr[k].amp_deg, arr[k].omega,
             arr[k].heave) = [float(x) for x in lw]
        self._chk(self.lib.ibg_set_motion(self.ctx, arr, float(t0)), "ibg_set_motion")

    def regrid(self, new_mesh):
        """Regrid (AMR::PerformAdaptive AMR.cpp:6-209) to `new_mesh`, a rebuild
        of the same cases (for a moving body: at its current pose): the state
        moves over by canonical keys (Mesh.transfer_to) and replaces the
        uploaded mesh; dt, the time and a prescribed motion carry on."""
        import time
        tick = [time.perf_counter()]
        timing = {}

        def lap(name):
            now = time.perf_counter()
            timing[name] = now - tick[0]
            tick[0] = now

        fv = self.get_state()
        nut = self.get_nut() if self.turbulence else None
        dt = self.get_dt()
        moving = getattr(self, "_surface", None) is not None
        t = self.get_time() if moving else None
        lap("state_out_s")
        fv2, nut2 = self.mesh.transfer_to(new_mesh, fv, nut)
        lap("transfer_s")
        self._chk(self.lib.ibg_upload_mesh(self.ctx, new_mesh.soa_ptr, new_mesh.phys),
                  "ibg_upload_mesh")
        lap("upload_mesh_s")
        self.mesh = new_mesh
        self._surface = None
        self.set_state(fv2)
        if nut2 is not None:
            self.set_nut(nut2)
        self.set_dt(dt)
        lap("state_in_s")
        if moving:
            self.upload_surface()
            if getattr(self, "_law", None) is not None:
                self.set_motion(self._law)
                self._chk(self.lib.ibg_set_time(self.ctx, np.ascontiguousarray(t).ctypes.data_as(_ffi.f64p)),
                          "ibg_set_time")
            lap("surface_s")
        # wall-clock seconds of each part of the last regrid
        self.last_regrid_timing = timing

    def write_solution(self, path, ts):
        """Tecplot solution of the current state (Pltoutput::Output_Soln) with
        ts, dt and the motion time: a restart point."""
        moving = getattr(self, "_surface", None) is not None
        self.mesh.write_solution(path, self.get_state(),
                                 self.get_nut() if self.turbulence else None, ts,
                                 self.get_dt(), self.get_time() if moving else None)

    def read_solution(self, path):
        """Restart from a solution file of this level set (InputSoln): state,
        SA variable, dt and, for a prescribed motion, its time; returns ts."""
        fv, nut, ts, dt, t = self.mesh.read_solution(path, self.turbulence)
        self.set_state(fv)
        if nut is not None:
            self.set_nut(nut)
        if np.all(np.isfinite(dt)):
            self.set_dt(dt)
        if getattr(self, "_surface", None) is not None and np.all(np.isfinite(t)):
            self._chk(self.lib.ibg_set_time(self.ctx, np.ascontiguousarray(t).ctypes.data_as(_ffi.f64p)),
                      "ibg_set_time")
        return ts

    def get_time(self):
        t = np.zeros(self.mesh.ncases)
        self._chk(self.lib.ibg_get_time(self.ctx, t.ctypes.data_as(_ffi.f64p)), "ibg_get_time")
        return t

    def geometry(self):
        """(signdis, closest patch, IB flag) per cell and the IB count."""
        n = self.mesh.ncells
        sd = np.zeros(n)
        cp = np.zeros(n, np.int32)
        ib = np.zeros(n, np.int8)
        nib = C.c_int32()
        self._chk(self.lib.ibg_get_geometry(self.ctx, sd.ctypes.data_as(_ffi.f64p),
                                            cp.ctypes.data_as(_ffi.i32p),
                                            ib.ctypes.data_as(_ffi.i8p), C.byref(nib)),
                  "ibg_get_geometry")
        return sd, cp, ib, nib.value

    def ib_list(self, cap):
        o = {"cell": np.zeros(cap, np.int32), "intp": np.zeros((cap, 8), np.int32),
             "coef": np.zeros((cap, 8)), "hgdis": np.zeros(cap), "nv": np.zeros((cap, 3)),
             "attach": np.zeros((cap, 3)), "body": np.zeros(cap, np.int32)}
        self._chk(self.lib.ibg_get_ib_list(
            self.ctx, cap, o["cell"].ctypes.data_as(_ffi.i32p), o["intp"].ctypes.data_as(_ffi.i32p),
            o["coef"].ctypes.data_as(_ffi.f64p), o["hgdis"].ctypes.data_as(_ffi.f64p),
            o["nv"].ctypes.data_as(_ffi.f64p), o["attach"].ctypes.data_as(_ffi.f64p),
            o["body"].ctypes.data_as(_ffi.i32p)), "ibg_get_ib_list")
        return o

    def patches(self):
        np_ = self.mesh.npatch
        o = {"pc": np.zeros((np_, 3)), "nv": np.zeros((np_, 3)), "area": np.zeros(np_),
             "intp": np.zeros((np_, 8), np.int32), "coef": np.zeros((np_, 8)),
             "hgdis": np.zeros(np_)}
        self._chk(self.lib.ibg_get_patches(
            self.ctx, o["pc"].ctypes.data_as(_ffi.f64p), o["nv"].ctypes.data_as(_ffi.f64p),
            o["area"].ctypes.data_as(_ffi.f64p), o["intp"].ctypes.data_as(_ffi.i32p),
            o["coef"].ctypes.data_as(_ffi.f64p), o["hgdis"].ctypes.data_as(_ffi.f64p)),
            "ibg_get_patches")
        return o

    def synchronize(self):
        self._chk(self.lib.ibg_synchronize(self.ctx), "ibg_synchronize")

    def close(self):
        if self.ctx:
            self.lib.ibg_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
