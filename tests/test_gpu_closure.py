"""GPU tests of the learned SA closure: BF16 MFMA kernels against the CPU
mirror (oracle/mlp.c), the training-data sampler against the oracle's
features, training, and what the trained network gains.

Bars: inference and the training gradient of the GPU kernels within 1e-3
relative L2 of the CPU mirror (same BF16 rounding points; differences come
from FP32 accumulation order and the hardware exp, which can flip single
BF16 roundings); sampled features within FP32 rounding of the oracle's;
training lowers the loss; the trained network keeps the SA working
variable and C_L/C_D of the FP64 plain-closure run within 1 %, much closer
than an untrained network or the laminar solver without a closure.
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi as O
from test_gpu_parity import perturbed_state, rel

pytestmark = pytest.mark.gpu


def rand_prm(seed=3, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, O.NPARAM)
    stats = np.concatenate([rng.normal(size=8), rng.uniform(0.5, 2.0, 8), -1e30 * np.ones(8),
                            1e30 * np.ones(8), [0.2, 1.7]])
    return np.concatenate([p, stats]).astype(np.float32)


@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case(Re=2000.0))


def eng_sa(mesh, mode="mixed"):
    e = ibgpu.Engine(mesh, mode=mode, nstage=4, turbulence=True)
    e.set_state(mesh.init_state())
    e.set_nut(np.full(mesh.ncells, mesh.phys[0].nut_inf))
    return e


def test_infer_matches_cpu_mirror(tiny):
    e = eng_sa(tiny)
    e.closure_init(1 << 12)
    prm = rand_prm()
    e.closure_set_params(prm)
    rng = np.random.default_rng(9)
    x = rng.normal(size=(1000, 8)).astype(np.float32)
    g = e.closure_infer(x)
    c = O.mlp_forward(prm, x)
    assert rel(g.astype(np.float64), c.astype(np.float64)) < 1e-3


@pytest.mark.parametrize("n", [64, 200, 1024])
def test_grad_matches_cpu_mirror(tiny, n):
    e = eng_sa(tiny)
    e.closure_init(1 << 12)
    prm = rand_prm(seed=n)
    e.closure_set_params(prm)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 8)).astype(np.float32)
    z = rng.normal(size=n).astype(np.float32)
    gg, lg = e.closure_grad(x, z)
    gc, lc = O.mlp_grad(prm, x, z)
    assert lg == pytest.approx(lc, rel=1e-4)
    assert rel(gg.astype(np.float64), gc) < 1e-3
    for a, b in [(0, 512), (512, 576), (576, 4672), (4672, 4736), (4736, 4800), (4800, 4801)]:
        assert rel(gg[a:b].astype(np.float64), gc[a:b]) < 5e-3, (a, b)


def test_sampler_matches_oracle_features(tiny):
    e = eng_sa(tiny, mode="fp64")
    e.TimeMarching(0, 16, use_graph=False)
    e.closure_init(1 << 14)
    n = e.closure_sample(0.5, 77)
    assert n > 500
    x, z, cell = e.closure_dataset(n)
    fv, nut = e.get_state(), e.get_nut()
    for k in range(0, n, max(1, n // 200)):
        ok, xo, zo = O.closure_features(tiny, fv, nut, int(cell[k]))
        assert ok
        assert np.allclose(x[k], xo, rtol=2e-6, atol=2e-6), (k, x[k], xo)
        assert z[k] == pytest.approx(zo, rel=1e-5, abs=1e-6)


def test_sampler_and_training_deterministic(tiny):
    """Same state, seeds and calls give the same dataset rows in the same
    order and the same trained parameters (the sampler places rows by a
    prefix sum over the tiles, not by atomics)."""
    def run():
        e = eng_sa(tiny, mode="fp64")
        e.TimeMarching(0, 8, use_graph=False)
        e.closure_init(1 << 14, 5)
        n1 = e.closure_sample(0.5, 11)
        e.TimeMarching(8, 8, use_graph=False)
        n2 = e.closure_sample(0.5, 12)
        x, z, cell = e.closure_dataset(n2)
        e.closure_fit_stats(0.0)
        loss = e.closure_train(50, batch=256, lr=1e-3, seed=3)
        return n1, n2, x, z, cell, e.closure_params(), loss

    a, b = run(), run()
    assert a[1] > a[0] > 0 and (a[0], a[1]) == (b[0], b[1])
    for u, v in zip(a[2:6], b[2:6]):
        assert np.array_equal(u, v)
    assert a[6] == b[6]
    assert len(np.unique(a[4][:a[0]])) == a[0]   # one row per kept cell in a round


def test_reinit_drops_captured_graphs(tiny):
    """A time-march graph captured with the network on holds the closure's
    buffers; re-initialising the closure (which frees them and turns the
    network off) must not replay it: the march continues as with the
    closure switched off."""
    P = 1 << (tiny.nlevels - 1)
    prm = rand_prm(seed=5, scale=0.05)
    prm[O.NPARAM + 16:O.NPARAM + 24] = -1e30     # training range: everything
    prm[O.NPARAM + 24:O.NPARAM + 32] = 1e30

    def start():
        e = eng_sa(tiny)
        e.closure_init(1 << 10)
        e.closure_set_params(prm)
        e.closure_enable(True)
        e.TimeMarching(0, P, use_graph=True)
        return e

    a = start()
    a.closure_init(1 << 10)
    a.TimeMarching(P, P, use_graph=True)
    b = start()
    b.closure_enable(False)
    b.TimeMarching(P, P, use_graph=True)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_nut(), b.get_nut())


def trained(mesh, steps_run=64, train_steps=1500, seed=20261202):
    """Network trained on cells of the plain-closure run (FP64 SA)."""
    src = eng_sa(mesh, mode="fp64")
    src.closure_init(1 << 18, seed)
    for k in range(steps_run // 8):
        src.TimeMarching(8 * k, 8, use_graph=False)
        src.closure_sample(0.6, seed + k)
    src.closure_fit_stats(0.0)
    loss0 = src.closure_train(1, batch=1024, lr=0.0 + 1e-12, seed=seed)
    loss = src.closure_train(train_steps, batch=1024, lr=1e-3, seed=seed + 1)
    return src.closure_params(), loss0, loss


@pytest.fixture(scope="module")
def net(tiny):
    return trained(tiny)


def test_training_reduces_loss(net):
    prm, loss0, loss = net
    assert loss < 0.2 * loss0, (loss0, loss)
    assert np.all(np.isfinite(prm))


def test_network_gain_over_no_closure(tiny, net):
    """The trained network against the FP64 plain closure: it tracks the SA
    working variable and the wall force; an untrained network and the
    solver without any closure (laminar) do not."""
    prm, _, _ = net
    steps = 96
    ref = eng_sa(tiny, mode="fp64")
    ref.TimeMarching(0, steps, use_graph=False)
    nut_ref = ref.get_nut()
    F_ref = ref.ComptWallForce()

    def run_ml(params):
        e = eng_sa(tiny, mode="mixed")
        e.closure_init(1 << 10)
        e.closure_set_params(params)
        e.closure_enable(True)
        e.TimeMarching(0, steps, use_graph=True)
        on, fb = e.closure_counts()
        return e.get_nut(), e.ComptWallForce(), on, fb

    nut_ml, F_ml, on, fb = run_ml(prm)
    assert on > fb > 0 or on > 0
    untrained = rand_prm(seed=1)
    untrained[O.NPARAM:] = prm[O.NPARAM:]        # same statistics and range
    nut_un, F_un, _, _ = run_ml(untrained)
    lam = ibgpu.Engine(tiny, mode="mixed", nstage=4)
    lam.set_state(tiny.init_state())
    lam.TimeMarching(0, steps, use_graph=True)
    F_lam = lam.ComptWallForce()
    e_ml = rel(nut_ml, nut_ref)
    e_un = rel(nut_un, nut_ref)
    assert e_ml < 1e-2, e_ml
    assert e_ml < 0.2 * e_un, (e_ml, e_un)
    dd = lambda F: abs(F[0] - F_ref[0]) / abs(F_ref[0])
    assert dd(F_ml) < 0.01, dd(F_ml)
    assert dd(F_ml) < dd(F_lam), (dd(F_ml), dd(F_lam))


# ------------------------------------------------ the network inside the stage


def fp32_state(mesh, fv, nut):
    """The state the mixed-mode stencil buffers hold: roe, u, v, w, p and the
    SA variable rounded to FP32 (T and e derived again in FP64)."""
    f = fv.copy()
    f[:, :5] = fv[:, :5].astype(np.float32).astype(np.float64)
    for c in range(mesh.ncases):
        s, e = mesh.blk_start[c * mesh.nlevels], mesh.blk_start[(c + 1) * mesh.nlevels]
        P = mesh.phys[c]
        f[s:e, 5] = f[s:e, 4] / (f[s:e, 0] * P.Rg)
        f[s:e, 6] = P.Cv * f[s:e, 5] + 0.5 * (f[s:e, 1] ** 2 + f[s:e, 2] ** 2 + f[s:e, 3] ** 2)
    return f, nut.astype(np.float32).astype(np.float64)


def range_params(x, z, seed=3, scale=0.3, q=(2.0, 98.0)):
    """Random weights with statistics fitted to feature rows x and targets z;
    the network's range is the q-percentile box of the features, so that a
    part of the cells takes the plain closure."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(-scale, scale, O.NPARAM)
    lo, hi = np.percentile(x, q[0], axis=0), np.percentile(x, q[1], axis=0)
    stats = np.concatenate([x.mean(0), x.std(0) + 1e-3, lo, hi, [z.mean(), z.std() + 1e-12]])
    return np.concatenate([w, stats]).astype(np.float32)


def network_stage_check(mesh, fv, nut, prm=None, sample=None, seed=7, schedule=0):
    """One stage of the learned-closure kernel (k_stage_ml) against the CPU
    mirror, cell by cell: the SA source of every sampled SA cell is either
    the plain closure (cells outside the network's range) or the plain
    closure with its point-wise term replaced by network output x production
    scale; the network/fallback counts equal the oracle's range test.
    Returns (params, statistics)."""
    e = ibgpu.Engine(mesh, mode="mixed", nstage=4, turbulence=True, schedule=schedule)
    e.set_state(fv)
    e.set_nut(nut)
    ts = 0
    e.InterfaceExchange(ts)
    e.IBConditions_Image()
    e.Comptdt(ts)
    e.BCBoxTreat(ts)
    e.ComptViscousStress(ts)
    f32, n32 = fp32_state(mesh, e.get_state(), e.get_nut())   # what the stage reads
    normal = np.asarray(mesh.kind) == 0
    sa = normal & (np.asarray(mesh.signdis) >= 0.0) & (np.asarray(mesh.ibidx) < 0)
    cells = np.flatnonzero(sa).astype(np.int32)
    x, zt, sc, ok = O.closure_cells(mesh, f32, n32, cells)
    if prm is None:
        prm = range_params(x[ok], zt[ok], seed=seed)
    e.closure_init(1 << 10)
    e.closure_set_params(prm)
    e.closure_enable(True)
    e.closure_counts()
    e.Stage(ts, 3)
    g = e.nut_source()
    on_net, fallback = e.closure_counts()
    e.close()
    # plain closure of the oracle on the same FP32-rounded state
    orc = O.Oracle(mesh, O.cfg(nstage=4))
    _, mutflux = orc.viscous_flux_sa(f32, n32, ts)
    plain = orc.mut_source(f32, n32, mutflux, ts)[cells]
    flo, fhi = prm[O.NPARAM + 16:O.NPARAM + 24], prm[O.NPARAM + 24:O.NPARAM + 32]
    inside = ok & np.all((x >= flo) & (x <= fhi), axis=1)
    # a cell whose FP64 feature lies within FP32 rounding of a bound may
    # fall on either side on the device
    tol = 1e-5 * np.maximum(1.0, np.abs(x))
    near = ok & np.any((np.abs(x - flo) <= tol) | (np.abs(x - fhi) <= tol), axis=1)
    assert on_net + fallback == cells.size, (on_net, fallback, cells.size)
    assert abs(on_net - int(inside.sum())) <= int(near.sum()), (on_net, int(inside.sum()), int(near.sum()))
    rng = np.random.default_rng(seed)
    net_cells = np.flatnonzero(inside & ~near)
    fb_cells = np.flatnonzero(~inside & ~near)
    if sample is not None:
        net_cells = rng.choice(net_cells, min(sample, net_cells.size), replace=False)
        fb_cells = rng.choice(fb_cells, min(sample, fb_cells.size), replace=False)
    gs = g[cells]
    # fallback cells: the FP32 plain closure against the FP64 one
    ref = 1e-3 * np.median(np.abs(plain)) + 1e-30
    err_fb = np.abs(gs[fb_cells] - plain[fb_cells]) / (np.abs(plain[fb_cells]) + ref)
    # network cells: plain - target * scale + network * scale
    znet = O.mlp_forward(prm, x[net_cells].astype(np.float32)).astype(np.float64)
    want = plain[net_cells] - zt[net_cells] * sc[net_cells] + znet * sc[net_cells]
    scale_term = np.abs(znet * sc[net_cells]) + np.abs(plain[net_cells]) + ref
    err_net = np.abs(gs[net_cells] - want) / scale_term
    stats = {"cells": int(cells.size), "on_net": on_net, "fallback": fallback,
             "near_bound": int(near.sum()), "net_checked": int(net_cells.size),
             "fb_checked": int(fb_cells.size),
             "fb_max": float(err_fb.max()) if err_fb.size else 0.0,
             "net_p99": float(np.percentile(err_net, 99)) if err_net.size else 0.0,
             "net_max": float(err_net.max()) if err_net.size else 0.0,
             "net_rel_l2": rel(gs[net_cells], want)}
    print("network stage check:", stats, flush=True)
    assert net_cells.size > 0 and fb_cells.size > 0
    assert stats["fb_max"] < 1e-3, stats
    # BF16 inputs: a feature within FP32 rounding of a BF16 rounding
    # boundary can round the other way on the device (rare); the bulk
    # agrees to FP32 accumulation order
    assert stats["net_p99"] < 1e-3, stats
    assert np.mean(err_net > 2e-2) < 2e-3, stats
    assert stats["net_rel_l2"] < 2e-3, stats
    return prm, stats


@pytest.mark.parametrize("schedule", [0, 1])
def test_network_in_stage_matches_cpu_mirror(tiny, schedule):
    print(f"network inside the stage kernel: tiny case, schedule {schedule}", flush=True)
    fv = perturbed_state(tiny)
    rng = np.random.default_rng(4)
    nut = tiny.phys[0].nut_inf * (1.0 + 0.5 * rng.uniform(-1.0, 1.0, tiny.ncells))
    network_stage_check(tiny, fv, nut, schedule=schedule)


def test_network_march_same_bits_either_schedule(tiny):
    """The learned-closure march under the persistent schedule (weights
    staged once per workgroup) gives the bits and the network / fallback
    counts of the one-workgroup-per-tile schedule."""
    prm = rand_prm(seed=5, scale=0.05)
    prm[O.NPARAM + 16:O.NPARAM + 24] = -1e30
    prm[O.NPARAM + 24:O.NPARAM + 32] = 1e30
    prm[O.NPARAM + 24] = 0.0          # feature 0 above 0: plain closure
    out = []
    for schedule in (0, 1):
        e = ibgpu.Engine(tiny, mode="mixed", nstage=4, turbulence=True, schedule=schedule)
        e.set_state(perturbed_state(tiny))
        e.set_nut(np.full(tiny.ncells, tiny.phys[0].nut_inf))
        e.closure_init(1 << 10)
        e.closure_set_params(prm)
        e.closure_enable(True)
        e.TimeMarching(0, 8, use_graph=True)
        out.append((e.get_state(), e.get_nut(), e.closure_counts()))
        e.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2] and out[0][2][0] > 0 and out[0][2][1] > 0
