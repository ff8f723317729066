"""Throughput of the per-stage residual path on one MI355X.

Workload: the mid-size 2D FSI case with the SA closure — NACA0015 at Mach
0.3, Re 10000, in [-10,10]^2 with a 256x256 base grid and 5 AMR levels (6
levels, ~2.6M leaf cells), 4-stage low-storage scheme (NS_Solver.H:204-208)
with temporal sub-cycling, the section pitching 5 deg about its quarter
chord and heaving 0.05 chord at reduced frequency 0.1 (the immersed
boundary re-tagged every step on the fixed level set, Bodymove.cpp:37-125;
--static holds it at its initial pose), Spalart-Allmaras working variable
from a freestream of 5 (the reference's TURBULENCE build).  A "step" is one
InterfaceExchange + IBConditions_Image (with the body move) + TimeMarching
(main.cpp:87-97).

Three forms run in the same process on the same mesh:
  plain     every operation FP64, plain SA closure (the reference's form);
  mixed     FP32 limiter/Roe/viscous/SA arithmetic, FP64 state and
            accumulation, plain closure;
  mixed+ML  as mixed, with the point-wise SA term from the learned closure
            (BF16 MFMA), trained in this run on cells sampled from the plain
            run (the plain closure is the fallback outside its range).
The metric is cell-stage updates/s (one UpdateConsVars cell body per stage,
NS_Solver.cpp:868-872 / 948-953, only marched levels); `value` is mixed+ML.
Accuracy: C_L/C_D of each form against the plain form after the same
steps.  The CPU baseline is the FP64 oracle (this project's C restatement
of the reference's path, OpenMP) on the host cores, timed on a bounded
sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "immsered-boundary_amd"))

import numpy as np  # noqa: E402

import ibgpu  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

# Measured HBM traffic of the learned-closure stage kernel, bytes per cell of
# a launch: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch) over
# every dispatch of k_stage_ml<2> in this script's profiled run,
# scaled by what the counters read for a plain 1 GiB copy (FETCH_SIZE x2.0,
# WRITE_SIZE x1.0), divided by the cells of the launch (the launch mix of the
# timed steps), as recorded with the traffic profile of this workload; update it
# with the kernel.  FETCH_SIZE counts what the L2s fetch from the fabric,
# last-level-cache (MALL) hits included, so it bounds the HBM bytes from above.
TRAFFIC_BYTES_PER_CELL = 198.3


def stage_bytes(mode, dim, turb, cells, faces, last=False, first=False):
    """Algorithmic HBM bytes of one stage-kernel launch.

    Per normal cell: its stencil state read once (neighbour reads hit in
    cache) and written once, the frozen viscous residual, the FP64 master
    read for U_old, the face slots (2 B per cell face) and the flag byte;
    with the SA closure the working variable read and written, its FP64
    master (the old value), the frozen face term and the wall distance
    (FP32 in the mixed mode).  The last stage of a step in mixed mode also
    writes the FP64 masters.  Per tile face the 4 stencil indices and the
    direction byte.  The first stage of a 2D FP64 step also computes the
    frozen viscous residual and face term: it writes them instead of reading
    them (same bytes) and reads the tangential viscous stencil (16 B per
    face).
    """
    nv = dim + 2
    sw = 8 if mode == "fp64" else 4
    per_cell = 3 * nv * sw + nv * 8 + 2 * dim * 2 + 1
    if turb:
        per_cell += 3 * sw + 8 + (8 if mode == "fp64" else 4)
    if last and mode != "fp64":
        per_cell += nv * 8 + (8 if turb else 0)
    per_face = 17 + (16 if first and dim == 2 and mode == "fp64" else 0)
    return cells * per_cell + faces * per_face


def stage_mix(eng, mesh, mode, ts0, nsteps, nstage, reps):
    """Stage-kernel launches of the timed steps ts0 .. ts0 + nsteps - 1, the
    launch mix of the timed region: a step's launches depend only on its
    phase in the level schedule (ts mod 2^(L-1)), so each phase present is
    timed once (every stage, HIP events on the engine's stream) and
    weighted by how often it occurs.  Returns the average launch duration,
    algorithmic bytes and cells per launch."""
    period = 1 << (mesh.nlevels - 1)
    count = {}
    for ts in range(ts0, ts0 + nsteps):
        count.setdefault(ts % period, [ts, 0])[1] += 1
    ms_sum, by_sum, c_sum, n = 0.0, 0.0, 0.0, 0
    for ts, w in count.values():
        for k in range(nstage):
            ms, nc, nf = eng.bench_stage(ts, k, reps)
            ms_sum += w * ms
            by_sum += w * stage_bytes(mode, mesh.dim, True, nc, nf, last=(k == nstage - 1), first=(k == 0))
            c_sum += w * nc
            n += w
    return ms_sum / n, by_sum / n, c_sum / n


def updates_per_step(mesh, nstage, ts):
    """cell-stage updates of step ts under temporal sub-cycling"""
    L = mesh.nlevels
    n = 0
    for c in range(mesh.ncases):
        for l in range(L):
            if ts % (1 << (L - 1 - l)) != 0:
                continue
            b = mesh.block(c, l)
            s, e = mesh.blk_start[b], mesh.blk_start[b] + mesh.blk_nnormal[b]
            if l == L - 1:
                sd = mesh.signdis[s:e]
                ib = mesh.ibidx[s:e]
                n += int(np.count_nonzero((sd >= 0.0) & (ib < 0)))
            else:
                n += int(e - s)
    return n * nstage


# Pitch and heave of the FSI case: aoa(t) = 5 deg sin(omega t) about the
# quarter chord, y(t) = 0.05 sin(omega t), reduced frequency
# k = omega c / (2 U) = 0.1 with c = U = 1.
MOTION = (0.0, 0.0, 0.0, 5.0, 0.2, 0.05)


def new_engine(mesh, mode, nstage, motion=True):
    eng = ibgpu.Engine(mesh, mode=mode, nstage=nstage, turbulence=True)
    eng.set_state(mesh.init_state())
    eng.set_nut(np.full(mesh.ncells, mesh.phys[0].nut_inf))
    if motion:
        eng.upload_surface()
        eng.set_motion([MOTION] * mesh.ncases, t0=0.0)
    return eng


def timed_run(eng, mesh, mode, warmup, steps, reps, ml=False):
    """W warmup steps, then exactly K timed steps replayed from HIP graphs
    captured beforehand (ibg_prepare_march: the timed region holds no
    capture, instantiation or direct launch), stream synchronised on both
    sides; then the stage kernel alone over the same steps' launch mix."""
    eng.TimeMarching(0, warmup, use_graph=True)
    eng.prepare_march(warmup, steps)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.TimeMarching(warmup, steps, use_graph=True)
    eng.synchronize()
    t1 = time.perf_counter()
    upd = eng.cell_stage_updates()
    P = mesh.phys[0]
    F = eng.ComptWallForce()
    cd, cl = F[0] / P.force_ref, F[1] / P.force_ref
    forces = F.reshape(mesh.ncases, 6)[:, :2].copy()   # every case, before the timing launches
    state, nut = eng.get_state(), eng.get_nut()
    counts = eng.closure_counts() if ml else (0, 0)   # cells of the warmup + timed steps
    t_end = float(eng.get_time()[0]) if getattr(eng, "_surface", None) else None
    # the stage kernel alone, launch by launch over the timed steps' mix
    ms, nbytes, ncell = stage_mix(eng, mesh, mode, warmup, steps, eng.cfg.nstage, reps)
    return {"seconds": t1 - t0, "updates": upd, "rate": upd / (t1 - t0), "cd": cd, "cl": cl,
            "stage_ms": ms, "stage_bytes": nbytes, "stage_cells": ncell, "state": state, "nut": nut,
            "counts": counts, "forces": forces, "t_end": t_end}


def accuracy_run(eng, mesh, nsteps):
    """The form's run from the start to nsteps (graph replay, untimed):
    C_L, C_D, state and SA variable there, for the comparison between forms
    at the configuration's horizon."""
    eng.TimeMarching(0, nsteps, use_graph=True)
    P = mesh.phys[0]
    F = eng.ComptWallForce()
    t_end = float(eng.get_time()[0]) if getattr(eng, "_surface", None) else None
    return {"cd": F[0] / P.force_ref, "cl": F[1] / P.force_ref, "state": eng.get_state(),
            "nut": eng.get_nut(), "t_end": t_end}


def regrid_cost(eng, spec, every=20):
    """One regrid (AMR::PerformAdaptive AMR.cpp:6-209 as a rebuild of the
    level set at the section's current pose, Engine.regrid) timed from the
    device time to the context marching on the new mesh, and the cell
    counts before and after."""
    import dataclasses
    eng.synchronize()
    t0 = time.perf_counter()
    t = float(eng.get_time()[0])
    x0, y0, aoa0, amp, omega, heave = MOTION
    sw = np.sin(omega * t)
    new = ibgpu.Mesh(dataclasses.replace(spec, aoa=aoa0 + amp * sw, pivot_y=y0 + heave * sw))
    t1 = time.perf_counter()
    n0 = eng.mesh.ncells
    eng.regrid(new)
    eng.synchronize()
    t2 = time.perf_counter()
    return {"every_steps": every, "seconds": t2 - t0, "build_s": t1 - t0, "transfer_upload_s": t2 - t1,
            "parts_s": getattr(eng, "last_regrid_timing", None),
            "cells_before": int(n0), "cells_after": int(new.ncells), "t": t,
            "note": "host rebuild of the level set at the pose (not incremental); not in value"}


def train_closure(mesh, nstage, args):
    """Network trained on cells sampled from the plain FP64 run
    (ibgpu.train_closure: the configuration's sampling and Adam recipe)."""
    return ibgpu.train_closure(mesh, nstage=nstage, samples=args.samples,
                               sample_rounds=args.sample_rounds, batch=args.batch,
                               epochs=args.epochs, motion=None if args.static else MOTION)


def cpu_baseline(mesh, nstage, budget_s, motion=True):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    ocfg = oracle_ffi.cfg(nstage=nstage, dt_var_step=1 << (mesh.nlevels - 1))
    orc = oracle_ffi.Oracle(mesh, ocfg)
    mb = oracle_ffi.MovingBodies(mesh) if motion else None
    fv = mesh.init_state()
    nut = np.full(mesh.ncells, mesh.phys[0].nut_inf)
    dt = np.array([mesh.phys[c].dt_init for c in range(mesh.ncases)])
    upd, ts, t0, t = 0, 0, time.perf_counter(), 0.0
    while True:
        if mb is not None:
            # the body move of the step (oracle/retag.c), then the step on
            # the moved geometry
            x0, y0, aoa0, amp, omega, heave = MOTION
            a = np.deg2rad(aoa0 + amp * np.sin(omega * t))
            r = mb.move([[x0, y0 + heave * np.sin(omega * t), np.cos(-a), np.sin(-a)]] * mesh.ncases)
            if r["code"]:
                raise RuntimeError(f"oracle move abort {r['code']}")
            orc = oracle_ffi.Oracle(oracle_ffi.MovedMesh(mb, r), ocfg)
        code, _, _, _ = orc.step_sa(fv, nut, dt, ts)
        t += float(dt[0])
        if code:
            raise RuntimeError(f"oracle abort {code}")
        upd += updates_per_step(mesh, nstage, ts)
        ts += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": upd / el, "unit": "cell-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle with the SA closure (C restatement of the reference path, "
                      f"OpenMP{', body move per step single-threaded' if motion else ''}), "
                      f"same mesh, steps 0..{ts - 1} from freestream "
                      f"({upd} cell-stage updates in {el:.1f} s)"}


def secondary(args):
    """The 3D quasi-span and the batched-ensemble configurations: plain FP64
    and mixed (FP32 flux, FP64 state) with the SA closure, one line each.
    The ensemble also picks its tile size with the autotuner and serves as
    the closure's training-data generator (one sampling pass, timed)."""
    if args.workload == "naca3d":
        mesh = ibgpu.Mesh(ibgpu.config3())
        desc = ("NACA0012 3D quasi-span, M0.6 Re5e4 AoA 4, periodic span 0.1 chord, 128x128x8 base "
                "+ 4 AMR levels, 4-stage, temporal sub-cycling, Spalart-Allmaras closure")
    else:
        mesh = ibgpu.Mesh(ibgpu.ensemble_cases())
        desc = ("32 NACA 4-digit sections (t U[0.09,0.18], AoA U[-4,8] deg, M U[0.3,0.7], seed "
                "20261202), Re1e4, 256^2 base + 5 AMR levels each with uneven refinement, one launch "
                "per stage for all cases, Spalart-Allmaras closure")
    per_case = [int(sum(r)) for r in mesh.counts()]
    print(f"[bench] {args.workload}: {mesh.ncases} case(s), {mesh.ncells} cells, normal cells per case "
          f"{min(per_case)}..{max(per_case)} (total {sum(per_case)})", file=sys.stderr, flush=True)
    res, extra = {}, {}
    for mode in ("fp64", "mixed"):
        eng = new_engine(mesh, mode, args.stages, False)
        if args.workload == "ensemble" and mode == "mixed":
            best, ms = eng.autotune_tile()
            eng.set_state(mesh.init_state())
            eng.set_nut(np.full(mesh.ncells, mesh.phys[0].nut_inf))
            extra[f"tile_{mode}"] = {
                "chosen": best, "schedule": eng.schedule,
                "stage_ms_by_tile_and_schedule": {f"tile {t}, "
                                                  f"{'persistent' if sch else 'one workgroup per tile'}": v
                                                  for (t, sch), v in ms.items()}}
        r = timed_run(eng, mesh, mode, args.warmup, args.steps, args.reps)
        if args.workload == "ensemble" and mode == "fp64":
            eng.closure_init(1 << 23)
            t0 = time.perf_counter()
            n = eng.closure_sample(0.1, 1)
            eng.synchronize()
            extra["training_data"] = {"rows": n, "fraction": 0.1, "seconds": time.perf_counter() - t0}
        res[mode] = r
        eng.close()
        print(f"[bench] {mode}+SA: {r['rate']:.4e} cell-stage/s, {1e3 * r['seconds'] / args.steps:.3f} "
              f"ms/step, stage kernel {r['stage_ms'] * 1e3:.1f} us "
              f"({r['stage_bytes'] / (r['stage_ms'] * 1e-3) / 1e9:.0f} GB/s algorithmic)",
              file=sys.stderr, flush=True)
    # plain copy of 1 GiB (the achievable HBM bandwidth; the context needs no large mesh)
    probe = ibgpu.Engine(ibgpu.Mesh(ibgpu.tiny_case()), mode="fp64")
    copy_ms = probe.bench_copy(1 << 30, 20)
    probe.close()
    copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    fp, mx = res["fp64"], res["mixed"]
    cf_fp = fp["forces"] / np.array([[mesh.phys[c].force_ref] for c in range(mesh.ncases)])
    cf_mx = mx["forces"] / np.array([[mesh.phys[c].force_ref] for c in range(mesh.ncases)])
    dF = np.abs(cf_mx - cf_fp) / np.maximum(np.abs(cf_fp), 1e-30)
    out = {
        "metric": "cell-stage updates/s, mixed vs plain FP64 (SA closure)",
        "value": mx["rate"], "unit": "cell-stage updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * mx["seconds"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 flux + f64 state", "data": "synthetic (freestream start, analytic NACA geometry)",
        "config": {"workload": desc, "cases": mesh.ncases, "cells": int(mesh.ncells),
                   "normal_cells": int(sum(per_case)), "levels": mesh.nlevels, "stages": args.stages},
        "plain_fp64": {"value": fp["rate"], "ms_per_step": 1e3 * fp["seconds"] / args.steps,
                       "stage_kernel_ms": fp["stage_ms"],
                       "stage_kernel_GBs": fp["stage_bytes"] / (fp["stage_ms"] * 1e-3) / 1e9},
        "mixed_over_fp64": mx["rate"] / fp["rate"],
        "accuracy": {"vs": "plain FP64 after the same steps",
                     "max_dCD_rel": float(dF[:, 0].max()), "max_dCL_rel": float(dF[:, 1].max()),
                     "max_dCL_abs": float(np.abs(cf_mx[:, 1] - cf_fp[:, 1]).max()),
                     "max_dCD_abs": float(np.abs(cf_mx[:, 0] - cf_fp[:, 0]).max()),
                     "CD_fp64": cf_fp[:, 0].tolist(), "CL_fp64": cf_fp[:, 1].tolist(),
                     "CD_mixed": cf_mx[:, 0].tolist(), "CL_mixed": cf_mx[:, 1].tolist(),
                     "state_rel_L2": float(np.linalg.norm(mx["state"][:, :5] - fp["state"][:, :5]) /
                                           np.linalg.norm(fp["state"][:, :5]))},
        "roofline": {"bound": "hbm", "achieved": mx["stage_bytes"] / (mx["stage_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": mx["stage_bytes"] / (mx["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "kernel": f"k_stage<float,{mesh.dim},false>",
                     "bytes_per_launch": mx["stage_bytes"], "avg_launch_ms": mx["stage_ms"],
                     "cells_per_launch": mx["stage_cells"], "copy_probe_GBs": copy_gbs},
        "cpu_baseline": None,
    }
    out.update(extra)
    print(json.dumps(out))


def _vs(r, fp):
    return {"CL": r["cl"], "CD": r["cd"],
            "dCL_rel": abs(r["cl"] - fp["cl"]) / max(abs(fp["cl"]), 1e-30),
            "dCD_rel": abs(r["cd"] - fp["cd"]) / max(abs(fp["cd"]), 1e-30),
            "state_rel_L2": float(np.linalg.norm(r["state"][:, :5] - fp["state"][:, :5]) /
                                  np.linalg.norm(fp["state"][:, :5])),
            "nut_rel_L2": float(np.linalg.norm(r["nut"] - fp["nut"]) / np.linalg.norm(fp["nut"]))}


def accuracy_block(acc, fp, mx, ml, args):
    """C_L / C_D and state of the mixed forms against plain FP64: at the
    configuration's horizon (acc_steps, own untimed runs) and after the
    timed region's steps."""
    tr = {"after_steps": args.warmup + args.steps, "CL_fp64": fp["cl"], "CD_fp64": fp["cd"],
          "mixed": _vs(mx, fp), "mixed_ml": _vs(ml, fp)}
    if not acc:
        return dict(vs="plain FP64 after the same steps", **tr)
    a = acc["fp64"]
    out = {"vs": "plain FP64 after the same steps", "after_steps": args.acc_steps,
           "t_end": a["t_end"], "CL_fp64": a["cl"], "CD_fp64": a["cd"],
           "mixed": _vs(acc["mixed"], a), "mixed_ml": _vs(acc["ml"], a),
           "mixed_ml_counts": {"network_cells": acc["ml"]["counts"][0],
                               "fallback_cells": acc["ml"]["counts"][1]},
           "timed_region": tr}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--stages", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--acc-steps", type=int, default=2000,
                    help="horizon of the C_L/C_D comparison between the forms (0: the timed steps)")
    ap.add_argument("--samples", type=int, default=5_000_000)
    ap.add_argument("--sample-rounds", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--epochs", type=float, default=20.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--static", action="store_true",
                    help="hold the body at its initial pose (no per-step re-tagging)")
    ap.add_argument("--workload", choices=("naca2d", "naca3d", "ensemble"), default="naca2d",
                    help="naca2d (default): the measured configuration with the learned closure; "
                         "naca3d, ensemble: plain FP64 vs mixed on the 3D and batched cases")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (the whole mesh lives in one device's HBM); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload != "naca2d":
        secondary(args)
        return
    mesh = ibgpu.Mesh(ibgpu.config2())
    leaf = int(sum(mesh.counts()[0]))
    print(f"[bench] mesh: {mesh.ncells} cells, normal per level {mesh.counts()[0]}", file=sys.stderr, flush=True)
    res = {}
    motion = not args.static
    acc = {}
    for mode in ("fp64", "mixed"):
        eng = new_engine(mesh, mode, args.stages, motion)
        res[mode] = timed_run(eng, mesh, mode, args.warmup, args.steps, args.reps)
        eng.close()
        r = res[mode]
        print(f"[bench] {mode}+SA: {r['rate']:.4e} cell-stage/s, {1e3 * r['seconds'] / args.steps:.3f} ms/step, "
              f"CL {r['cl']:.6f} CD {r['cd']:.6f}, stage kernel {r['stage_ms'] * 1e3:.1f} us",
              file=sys.stderr, flush=True)
        if args.acc_steps:
            eng = new_engine(mesh, mode, args.stages, motion)
            acc[mode] = accuracy_run(eng, mesh, args.acc_steps)
            eng.close()
    prm, tinfo = train_closure(mesh, args.stages, args)
    print(f"[bench] closure: {tinfo['samples']} samples, {tinfo['train_steps']} Adam steps "
          f"(batch {args.batch}), loss {tinfo['final_loss']:.3e}, sampling {tinfo['sample_s']:.1f} s, "
          f"training {tinfo['train_s']:.2f} s", file=sys.stderr, flush=True)
    eng = new_engine(mesh, "mixed", args.stages, motion)
    eng.closure_init(1 << 10)
    eng.closure_set_params(prm)
    eng.closure_enable(True)
    res["ml"] = timed_run(eng, mesh, "mixed", args.warmup, args.steps, args.reps, ml=True)
    on_net, fallback = res["ml"]["counts"]
    if args.acc_steps:
        e2 = new_engine(mesh, "mixed", args.stages, motion)
        e2.closure_init(1 << 10)
        e2.closure_set_params(prm)
        e2.closure_enable(True)
        acc["ml"] = accuracy_run(e2, mesh, args.acc_steps)
        acc["ml"]["counts"] = e2.closure_counts()
        e2.close()
    # host boundary: the AoS state in and out through the C ABI (PCIe + layout
    # conversion), which a caller pays once around a run of steps
    fv_host = mesh.init_state()
    eng.synchronize()
    t0 = time.perf_counter()
    eng.set_state(fv_host)
    eng.set_nut(np.full(mesh.ncells, mesh.phys[0].nut_inf))
    _ = eng.get_state()
    _ = eng.get_nut()
    host_io_s = time.perf_counter() - t0
    regrid = None
    if motion:
        # the configuration regrids every 20 steps: one regrid of the moving
        # section at the law's pose of the device time (host rebuild of the
        # level set, state moved over by canonical keys, re-upload), timed
        eng.TimeMarching(0, 32, use_graph=True)
        regrid = regrid_cost(eng, ibgpu.config2())
    eng.close()
    r = res["ml"]
    print(f"[bench] mixed+ML: {r['rate']:.4e} cell-stage/s, {1e3 * r['seconds'] / args.steps:.3f} ms/step, "
          f"CL {r['cl']:.6f} CD {r['cd']:.6f}, stage kernel {r['stage_ms'] * 1e3:.1f} us, "
          f"network cells {on_net}, fallback cells {fallback}", file=sys.stderr, flush=True)
    # plain copy of 1 GiB: the achievable HBM bandwidth on this box
    probe = ibgpu.Engine(mesh, mode="fp64", nstage=args.stages)
    copy_ms = probe.bench_copy(1 << 30, 20)
    probe.close()
    copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    print(f"[bench] copy probe: 1 GiB in {copy_ms:.3f} ms = {copy_gbs:.0f} GB/s (read+write)", file=sys.stderr, flush=True)
    fp, mx, ml = res["fp64"], res["mixed"], res["ml"]

    def kernel_gbs(r):
        return r["stage_bytes"] / (r["stage_ms"] * 1e-3) / 1e9

    ml_bytes, ml_gbs = ml["stage_bytes"], kernel_gbs(ml)
    mx_gbs, fp_gbs = kernel_gbs(mx), kernel_gbs(fp)

    out = {
        "metric": "cell-stage updates/s, mixed+ML vs plain FP64 at equal C_L/C_D error",
        "value": ml["rate"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * ml["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flux + f64 state, bf16 closure network",
        "data": "synthetic (freestream start, analytic NACA0015 geometry); closure trained in-run "
                "on cells of the plain FP64 run",
        "config": {
            "workload": "NACA0015 2D, M0.3 Re1e4, [-10,10]^2, 256^2 base + 5 AMR levels, "
                        f"{leaf} leaf cells, 4-stage, temporal sub-cycling, "
                        + ("body at initial pose, " if args.static else
                           "pitch 5 deg + heave 0.05c at k=0.1 re-tagged every step (no regrid), ")
                        + "Spalart-Allmaras closure (plain FP64 vs mixed + learned)",
            "motion": None if args.static else dict(zip(
                ("x0", "y0", "aoa0_deg", "amp_deg", "omega", "heave"), MOTION),
                t_end=ml["t_end"]),
            "cells": int(mesh.ncells),
            "leaf_cells": leaf,
            "levels": mesh.nlevels,
            "stages": args.stages,
            "closure_network": "8-64-64-1, ReLU hidden units (config's SiLU measured slower at "
                               "equal closure error), BF16 weights, FP32 accumulation",
            "parallelism": "1 GPU",
        },
        "plain_fp64": {"value": fp["rate"], "ms_per_step": 1e3 * fp["seconds"] / args.steps,
                       "stage_kernel_ms": fp["stage_ms"], "stage_kernel_GBs": fp_gbs},
        "mixed_plain_closure": {"value": mx["rate"], "ms_per_step": 1e3 * mx["seconds"] / args.steps,
                                "stage_kernel_ms": mx["stage_ms"], "stage_kernel_GBs": mx_gbs},
        "ml_over_fp64": ml["rate"] / fp["rate"],
        "mixed_over_fp64": mx["rate"] / fp["rate"],
        "accuracy": accuracy_block(acc, fp, mx, ml, args),
        "closure": dict(tinfo, network_cells=on_net, fallback_cells=fallback),
        "regrid": (dict(regrid, value_incl_regrid=ml["updates"] / (
            ml["seconds"] + regrid["seconds"] * args.steps / regrid["every_steps"]))
            if regrid else None),
        "host_boundary": {"state_in_out_ms": 1e3 * host_io_s,
                          "value_incl_state_in_out": ml["updates"] / (ml["seconds"] + host_io_s)},
        "roofline": {"bound": "hbm", "achieved": ml_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ml_gbs / HBM_PEAK_GBS,
                     "traffic": (TRAFFIC_BYTES_PER_CELL * ml["stage_cells"]
                                 if TRAFFIC_BYTES_PER_CELL else None),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this kernel in a "
                                       "profiled run of this script (profiles/k_stage_naca2d_sa_"
                                       "traffic.md), bytes per cell x this run's cells per launch; "
                                       "not counted in this run; L2 fetches including hits in "
                                       "the shared last-level cache, an upper bound on HBM bytes",
                     "kernel": "k_stage_ml<2>", "bytes_per_launch": ml_bytes,
                     "avg_launch_ms": ml["stage_ms"], "cells_per_launch": ml["stage_cells"],
                     "copy_probe_GBs": copy_gbs},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, args.stages, args.cpu_seconds, motion)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
