"""Training recipe of the learned SA closure: cells sampled from this
solver's own plain-closure (FP64) run, then Adam on the GPU.

The network stands in for the point-wise part of the per-cell SA source
(Turbulence.cpp, NS_Solver.cpp:632-840); the plain closure stays the
fallback for cells outside the sampled feature range.
"""
import time

import numpy as np

from .engine import Engine


def plain_engine(mesh, nstage, motion=None, t0=0.0):
    """FP64 context with the plain SA closure from the freestream state,
    moving under the pitch/heave law `motion` (one tuple per case) if given."""
    eng = Engine(mesh, mode="fp64", nstage=nstage, turbulence=True)
    eng.set_state(mesh.init_state())
    eng.set_nut(np.full(mesh.ncells, mesh.phys[0].nut_inf))
    if motion is not None:
        eng.upload_surface()
        eng.set_motion([motion] * mesh.ncases, t0=t0)
    return eng


def train_closure(mesh, nstage=4, samples=5_000_000, sample_rounds=8, batch=8192, epochs=20.0,
                  lr=1e-3, motion=None, seed=20261202, train_seed=7):
    """Sample `samples` cells from the plain FP64 run (one sampling pass per
    sub-cycling period, each cell kept with a fixed probability, seeded),
    fit the feature statistics and range, and train for `epochs` passes over
    the samples at `batch` rows per Adam step.  Returns (parameter vector,
    info dict)."""
    t0 = time.perf_counter()
    src = plain_engine(mesh, nstage, motion)
    try:
        src.closure_init(samples, seed)
        period = 1 << (mesh.nlevels - 1)
        total, ts, k = 0, 0, 0
        frac = min(1.0, samples / (sample_rounds * 0.95 * mesh.ncells))
        while total < samples and k < 4 * sample_rounds:
            src.TimeMarching(ts, period, use_graph=True)
            ts += period
            total = src.closure_sample(frac, 1000 + k)
            k += 1
        src.closure_fit_stats(0.0)
        t1 = time.perf_counter()
        steps = max(1, int(epochs * total / batch))
        loss = src.closure_train(steps, batch=batch, lr=lr, seed=train_seed)
        src.synchronize()
        t2 = time.perf_counter()
        prm = src.closure_params()
    finally:
        src.close()
    return prm, {"samples": int(total), "sampled_steps": ts, "train_steps": steps, "batch": batch,
                 "epochs": epochs, "final_loss": loss, "sample_s": t1 - t0, "train_s": t2 - t1}
