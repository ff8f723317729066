"""Throughput of the per-stage residual path on one MI355X.

Workload: the mid-size 2D case — NACA0015 at Mach 0.3, Re 10000, in
[-10,10]^2 with a 256x256 base grid and 5 AMR levels (6 levels, ~2.6M leaf
cells), laminar, 4-stage low-storage scheme (NS_Solver.H:204-208) with
temporal sub-cycling, body held at its initial pose.  A "step" is one
InterfaceExchange + IBConditions_Image + TimeMarching (main.cpp:87-97).

The metric is cell-stage updates/s (one UpdateConsVars cell body per stage,
NS_Solver.cpp:868-872 / 948-953, only marched levels).  `value` is the mixed
precision form (FP32 limiter/Roe/viscous arithmetic, FP64 state and
accumulation); the plain FP64 form is run in the same process and reported
beside it with the C_L/C_D difference.  The CPU baseline is the FP64 oracle
(this project's C restatement of the reference's path, OpenMP) on the host
cores, timed on a bounded sample of the same workload.

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


def stage_bytes(mode, dim, cells, faces):
    """Algorithmic HBM bytes of one stage-kernel launch.

    Per normal cell: the cell's W read once (its neighbours' reads hit in
    cache), U_old read (FP64), frozen viscous residual read, W written, the
    face slots (2 B per cell face) and the flag byte; per tile face the 4
    stencil indices and the direction byte.
    """
    nv = dim + 2
    sw = 8 if mode == "fp64" else 4
    per_cell = nv * sw + nv * 8 + nv * sw + nv * sw + 2 * dim * 2 + 1
    return cells * per_cell + faces * 17


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


def run_mode(mesh, mode, nstage, warmup, steps, reps):
    eng = ibgpu.Engine(mesh, mode=mode, nstage=nstage)
    eng.set_state(mesh.init_state())
    eng.TimeMarching(0, warmup, use_graph=True)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.TimeMarching(warmup, steps, use_graph=True)
    eng.synchronize()
    t1 = time.perf_counter()
    upd = eng.cell_stage_updates()
    P = mesh.phys[0]
    F = eng.ComptWallForce()
    cd, cl = F[0] / P.force_ref, F[1] / P.force_ref
    state = eng.get_state()
    # stage kernel alone, all levels marched (ts multiple of the period)
    period = 1 << (mesh.nlevels - 1)
    ts_full = ((warmup + steps) // period) * period
    ms, nc, nf = eng.bench_stage(ts_full, 0, reps)
    eng.close()
    return {"seconds": t1 - t0, "updates": upd, "rate": upd / (t1 - t0), "cd": cd, "cl": cl,
            "stage_ms": ms, "stage_cells": nc, "stage_faces": nf, "state": state}


def cpu_baseline(mesh, nstage, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    orc = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(nstage=nstage, dt_var_step=1 << (mesh.nlevels - 1)))
    fv = mesh.init_state()
    dt = np.array([mesh.phys[c].dt_init for c in range(mesh.ncases)])
    upd, ts, t0 = 0, 0, time.perf_counter()
    while True:
        code, _, _, _ = orc.step(fv, dt, ts)
        if code:
            raise RuntimeError(f"oracle abort {code}")
        upd += updates_per_step(mesh, nstage, ts)
        ts += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": upd / el, "unit": "cell-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (C restatement of the reference path, OpenMP), same mesh, "
                      f"steps 0..{ts - 1} from freestream ({upd} cell-stage updates in {el:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--stages", type=int, default=4)
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (the whole mesh lives in one device's HBM); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    spec = ibgpu.config2()
    mesh = ibgpu.Mesh(spec)
    print(f"[bench] mesh: {mesh.ncells} cells, normal per level {mesh.counts()[0]}", file=sys.stderr, flush=True)
    res = {}
    for mode in ("fp64", "mixed"):
        res[mode] = run_mode(mesh, mode, args.stages, args.warmup, args.steps, args.reps)
        r = res[mode]
        print(f"[bench] {mode}: {r['rate']:.4e} cell-stage/s, {1e3 * r['seconds'] / args.steps:.3f} ms/step, "
              f"CL {r['cl']:.6f} CD {r['cd']:.6f}, stage kernel {r['stage_ms'] * 1e3:.1f} us",
              file=sys.stderr, flush=True)
    # plain copy of 1 GiB: the achievable HBM bandwidth on this box
    probe = ibgpu.Engine(mesh, mode="fp64", nstage=args.stages)
    copy_ms = probe.bench_copy(1 << 30, 20)
    probe.close()
    copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    print(f"[bench] copy probe: 1 GiB in {copy_ms:.3f} ms = {copy_gbs:.0f} GB/s (read+write)", file=sys.stderr, flush=True)
    mx, fp = res["mixed"], res["fp64"]
    nbytes = stage_bytes("mixed", mesh.dim, mx["stage_cells"], mx["stage_faces"])
    achieved = nbytes / (mx["stage_ms"] * 1e-3) / 1e9
    fp_bytes = stage_bytes("fp64", mesh.dim, fp["stage_cells"], fp["stage_faces"])
    fp_achieved = fp_bytes / (fp["stage_ms"] * 1e-3) / 1e9
    w_rel = float(np.linalg.norm(mx["state"][:, :5] - fp["state"][:, :5]) / np.linalg.norm(fp["state"][:, :5]))
    out = {
        "metric": "cell-stage updates/s, mixed+ML vs plain FP64 at equal C_L/C_D error",
        "value": mx["rate"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mx["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flux + f64 state",
        "data": "synthetic (freestream start, analytic NACA0015 geometry)",
        "config": {
            "workload": "NACA0015 2D, M0.3 Re1e4, [-10,10]^2, 256^2 base + 5 AMR levels, "
                        f"{int(sum(mesh.counts()[0]))} leaf cells, 4-stage, temporal sub-cycling, "
                        "body at initial pose, laminar (no closure)",
            "cells": int(mesh.ncells),
            "leaf_cells": int(sum(mesh.counts()[0])),
            "levels": mesh.nlevels,
            "stages": args.stages,
            "parallelism": "1 GPU",
        },
        "plain_fp64": {"value": fp["rate"], "ms_per_step": 1e3 * fp["seconds"] / args.steps,
                       "stage_kernel_GBs": fp_achieved},
        "mixed_over_fp64": mx["rate"] / fp["rate"],
        "accuracy": {"CL_fp64": fp["cl"], "CD_fp64": fp["cd"], "CL_mixed": mx["cl"], "CD_mixed": mx["cd"],
                     "dCL_rel": abs(mx["cl"] - fp["cl"]) / max(abs(fp["cl"]), 1e-30),
                     "dCD_rel": abs(mx["cd"] - fp["cd"]) / max(abs(fp["cd"]), 1e-30),
                     "state_rel_L2": w_rel,
                     "after_steps": args.warmup + args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_stage<float,2,mixed>", "bytes_per_launch": nbytes,
                     "avg_launch_ms": mx["stage_ms"], "cells_per_launch": mx["stage_cells"],
                     "copy_probe_GBs": copy_gbs},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, args.stages, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
