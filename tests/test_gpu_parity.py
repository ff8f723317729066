"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Bar: the plain FP64 mode within 1e-12 relative L2 of the oracle on the
state, the residual source[0], the face fluxes, the time step, the residual
norms and the wall force, on the same seeded inputs; per-phase probes (IB
fill, domain BC, restriction/prolongation) at the same bar.  The mixed mode
(FP32 flux arithmetic, FP64 state) is a precision choice: its error is
measured against the oracle and bounded here (state 1e-5 relative L2 and
C_L/C_D within 1% after the run).
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def perturbed_state(mesh, seed=20261202, amp=0.05):
    """Freestream plus a smooth seeded perturbation (keeps W physical)."""
    rng = np.random.default_rng(seed)
    fv = mesh.init_state()
    x = mesh.centers
    k = rng.uniform(0.5, 2.0, size=4)
    ph = rng.uniform(0, 2 * np.pi, size=4)
    fv[:, 0] *= 1 + amp * np.sin(k[0] * x[:, 0] + ph[0]) * np.cos(k[1] * x[:, 1] + ph[1])
    fv[:, 1] += amp * np.cos(k[2] * x[:, 1] + ph[2])
    fv[:, 2] += amp * np.sin(k[3] * x[:, 0] + ph[3])
    if mesh.dim == 3:
        fv[:, 3] += amp * np.sin(k[0] * x[:, 2] * 20 + ph[0]) * np.cos(k[1] * x[:, 0])
    fv[:, 4] *= 1 + amp * np.cos(k[1] * x[:, 0] + k[2] * x[:, 1])
    for c in range(mesh.ncases):
        s, e = mesh.blk_start[c * mesh.nlevels], mesh.blk_start[(c + 1) * mesh.nlevels]
        P = mesh.phys[c]
        fv[s:e, 5] = fv[s:e, 4] / (fv[s:e, 0] * P.Rg)
        fv[s:e, 6] = P.Cv * fv[s:e, 5] + 0.5 * (fv[s:e, 1] ** 2 + fv[s:e, 2] ** 2 + fv[s:e, 3] ** 2)
    return fv


def cols(mesh):
    return [0, 1, 2, 3, 4] if mesh.dim == 3 else [0, 1, 2, 4]


@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case())


@pytest.fixture(scope="module")
def tiny3d():
    return ibgpu.Mesh(ibgpu.tiny3d_case())


@pytest.fixture(scope="module")
def pair():
    return ibgpu.Mesh([ibgpu.tiny_case(), ibgpu.tiny_case(aoa=-2.0, Ma=0.4, thickness=0.15)])


def run_oracle(mesh, fv, nsteps, nstage=3):
    orc = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(nstage=nstage, dt_var_step=1 << (mesh.nlevels - 1)))
    cpu = fv.copy()
    dt = np.array([mesh.phys[c].dt_init for c in range(mesh.ncases)])
    for ts in range(nsteps):
        code, src, linf, l2 = orc.step(cpu, dt, ts)
        assert code == 0, code
    return cpu, dt, src, linf, l2


@pytest.mark.parametrize("which", ["tiny", "tiny3d"])
def test_inviscid_faceflux_fp64(which, request):
    mesh = request.getfixturevalue(which)
    fv = perturbed_state(mesh)
    eng = ibgpu.Engine(mesh, mode="fp64")
    eng.set_state(fv)
    g = eng.faceflux()
    c = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(temporal_refine=0)).face_flux(fv, 0)
    assert rel(g[:, cols(mesh)], c[:, cols(mesh)]) < TOL


@pytest.mark.parametrize("which", ["tiny", "tiny3d"])
def test_viscous_faceflux_fp64(which, request):
    mesh = request.getfixturevalue(which)
    fv = perturbed_state(mesh)
    eng = ibgpu.Engine(mesh, mode="fp64")
    eng.set_state(fv)
    g = eng.vsflux()
    c = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(temporal_refine=0)).viscous_flux(fv, 0)
    k = [1, 2, 3, 4] if mesh.dim == 3 else [1, 2, 4]
    assert rel(g[:, k], c[:, k]) < TOL


@pytest.mark.parametrize("which,nstage", [("tiny", 3), ("tiny", 4), ("tiny3d", 3), ("pair", 3)])
def test_steps_fp64_match_oracle(which, nstage, request):
    mesh = request.getfixturevalue(which)
    fv0 = perturbed_state(mesh)
    nsteps = 8
    cpu, dt, src, linf, l2 = run_oracle(mesh, fv0, nsteps, nstage)
    eng = ibgpu.Engine(mesh, mode="fp64", nstage=nstage)
    eng.set_state(fv0)
    eng.TimeMarching(0, nsteps, use_graph=False)
    gpu = eng.get_state()
    assert rel(gpu[:, :5], cpu[:, :5]) < TOL
    assert np.allclose(eng.get_dt(), dt, rtol=1e-14, atol=0)
    glinf, gl2 = eng.ResinDomain()
    assert np.allclose(glinf, linf, rtol=1e-10, atol=1e-300)
    assert np.allclose(gl2, l2, rtol=1e-10, atol=1e-300)


@pytest.mark.parametrize("opt", [dict(limiter=1), dict(viscous=False), dict(temporal_refine=False)])
def test_steps_fp64_scheme_options(tiny, opt):
    """The van Albada limiter, the inviscid path and the march without
    temporal sub-cycling against the oracle with the same options."""
    fv0 = perturbed_state(tiny)
    nsteps, nstage = 6, 4
    tr = opt.get("temporal_refine", True)
    P = (1 << (tiny.nlevels - 1)) if tr else 1
    orc = oracle_ffi.Oracle(tiny, oracle_ffi.cfg(nstage=nstage, dt_var_step=P,
                                                 temporal_refine=1 if tr else 0,
                                                 limiter=opt.get("limiter", 0),
                                                 viscous=0 if opt.get("viscous", True) is False else 1))
    cpu = fv0.copy()
    dt = np.array([tiny.phys[0].dt_init])
    for ts in range(nsteps):
        code, _, _, _ = orc.step(cpu, dt, ts)
        assert code == 0, code
    eng = ibgpu.Engine(tiny, mode="fp64", nstage=nstage, **opt)
    eng.set_state(fv0)
    eng.TimeMarching(0, nsteps, use_graph=False)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL
    assert np.allclose(eng.get_dt(), dt, rtol=1e-14, atol=0)
    # the mixed mode with the same options stays within its bound
    mx = ibgpu.Engine(tiny, mode="mixed", nstage=nstage, **opt)
    mx.set_state(fv0)
    mx.TimeMarching(0, nsteps, use_graph=True)
    assert rel(mx.get_state()[:, :5], cpu[:, :5]) < 1e-5


def test_stagewise_source_matches_oracle(tiny):
    """Phase by phase through the C ABI; source[0] of the last stage."""
    mesh = tiny
    fv0 = perturbed_state(mesh)
    cpu, dt, src, _, _ = run_oracle(mesh, fv0, 1)
    eng = ibgpu.Engine(mesh, mode="fp64")
    eng.set_state(fv0)
    ts = 0
    eng.InterfaceExchange(ts)
    eng.IBConditions_Image()
    eng.Comptdt(ts)
    eng.BCBoxTreat(ts)
    eng.ComptViscousStress(ts)
    for k in range(3):
        eng.Stage(ts, k)
    gsrc = eng.source()
    normal = mesh.kind == 0
    c = cols(mesh)
    assert rel(gsrc[normal][:, c], src[normal][:, c]) < TOL
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL


def test_graph_replay_equals_eager(tiny):
    fv0 = perturbed_state(tiny)
    P = 1 << (tiny.nlevels - 1)
    a = ibgpu.Engine(tiny, mode="fp64")
    a.set_state(fv0)
    a.TimeMarching(0, 3 * P, use_graph=False)
    b = ibgpu.Engine(tiny, mode="fp64")
    b.set_state(fv0)
    b.TimeMarching(0, P, use_graph=True)       # capture
    b.TimeMarching(P, 2 * P, use_graph=True)   # replays
    assert np.array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("mode,turb", [("fp64", False), ("mixed", False), ("fp64", True), ("mixed", True)])
def test_fused_first_stage_equals_phase_calls(tiny, mode, turb):
    """The time march folds ComptViscousStress (the frozen viscous residual
    and the SA face term) into the first stage launch in 2D FP64; the
    phase-by-phase ABI calls keep it a launch of its own.  Same bits (the
    mixed cases check that the unfused march does too)."""
    fv0 = perturbed_state(tiny)
    nstage = 4
    P = 1 << (tiny.nlevels - 1)

    def make():
        e = ibgpu.Engine(tiny, mode=mode, nstage=nstage, turbulence=turb)
        e.set_state(fv0)
        if turb:
            e.set_nut(np.full(tiny.ncells, tiny.phys[0].nut_inf))
        return e

    a = make()
    a.TimeMarching(0, P + 3, use_graph=False)
    b = make()
    for ts in range(P + 3):
        b.InterfaceExchange(ts)
        b.IBConditions_Image()
        if ts % P == 0:
            b.Comptdt(ts)
        b.BCBoxTreat(ts)
        b.ComptViscousStress(ts)
        for k in range(nstage):
            b.Stage(ts, k)
    assert np.array_equal(a.get_state(), b.get_state())
    if turb:
        assert np.array_equal(a.get_nut(), b.get_nut())


def test_mixed_mode_error_bounded(tiny):
    fv0 = tiny.init_state()
    nsteps = 40
    cpu, dt, _, _, _ = run_oracle(tiny, fv0, nsteps)
    eng = ibgpu.Engine(tiny, mode="mixed")
    eng.set_state(fv0)
    eng.TimeMarching(0, nsteps, use_graph=True)
    gpu = eng.get_state()
    err = rel(gpu[:, :5], cpu[:, :5])
    assert err < 1e-5, err
    F = eng.ComptWallForce()
    Fo = oracle_ffi.Oracle(tiny).forces(cpu)
    assert abs(F[0] - Fo[0]) <= 0.01 * abs(Fo[0])
    assert abs(F[1] - Fo[1]) <= 0.01 * abs(Fo[1])


def test_forces_and_dt_fp64(pair):
    fv = perturbed_state(pair)
    eng = ibgpu.Engine(pair, mode="fp64")
    eng.set_state(fv)
    F = eng.ComptWallForce()
    orc = oracle_ffi.Oracle(pair)
    Fo = orc.forces(fv)
    assert rel(F, Fo) < TOL
    dt = eng.Comptdt(0)
    dto = np.array([pair.phys[c].dt_init for c in range(pair.ncases)])
    assert orc.compute_dt(fv, dto) == 0
    assert np.allclose(dt, dto, rtol=1e-14, atol=0)


@pytest.mark.parametrize("phase", ["ib", "bc", "exchange"])
def test_ghost_fills_fp64(tiny, phase):
    fv = perturbed_state(tiny)
    eng = ibgpu.Engine(tiny, mode="fp64")
    eng.set_state(fv)
    orc = oracle_ffi.Oracle(tiny)
    cpu = fv.copy()
    if phase == "ib":
        eng.IBConditions_Image()
        assert orc.ib_fill(cpu) == 0
    elif phase == "bc":
        eng.BCBoxTreat(0)
        assert orc.bc_fill(cpu, 0) == 0
    else:
        P = 1 << (tiny.nlevels - 1)
        for ts in range(P):     # every level pair is due once per period
            eng.InterfaceExchange(ts)
            orc.interface_exchange(cpu, ts)
    gpu = eng.get_state()
    assert rel(gpu[:, :5], cpu[:, :5]) < TOL
    changed = np.any(cpu[:, :5] != fv[:, :5], axis=1)
    assert changed.any()


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("fault", ["nan", "dt"])
def test_march_stops_at_first_error(tiny, fault, use_graph):
    """The march stops at its first error, as the reference aborts there: a
    NaN planted in one fluid cell is the NaN-residual abort (268) at the
    last stage of the first step; an oversized time step is the negative
    density abort (523/611/37) in its first stage.  Every later launch does
    nothing: the state (NaNs included) and dt after three requested steps
    equal those after one, with the same code and cell."""
    fv0 = perturbed_state(tiny)
    dt_var = 1 << (tiny.nlevels - 1)
    if fault == "nan":
        cand = np.flatnonzero((tiny.kind == 0) & (tiny.ibidx < 0) & (tiny.signdis > 0.5))
        fv0[cand[len(cand) // 2], 0] = np.nan
    out = []
    for n in (1, 3):
        print(f"first error: {fault}, {n} step(s), graph={use_graph}", flush=True)
        eng = ibgpu.Engine(tiny, mode="fp64", nstage=3, dt_var_step=0 if fault == "dt" else dt_var)
        eng.set_state(fv0)
        if fault == "dt":
            eng.set_dt(np.array([5.0]))
        with pytest.raises(ibgpu.IBGError) as ei:
            eng.TimeMarching(0, n, use_graph=use_graph)
        out.append((ei.value.code, ei.value.cell, eng.get_state(), eng.get_dt()))
        eng.close()
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1] >= 0
    assert out[0][0] == 268 if fault == "nan" else out[0][0] in (523, 611, 37)
    assert np.array_equal(out[0][2], out[1][2], equal_nan=True)
    assert np.array_equal(out[0][3], out[1][3])
    # a new state clears the error: the same context marches again
    eng = ibgpu.Engine(tiny, mode="fp64", nstage=3, dt_var_step=0 if fault == "dt" else dt_var)
    eng.set_state(fv0)
    if fault == "dt":
        eng.set_dt(np.array([5.0]))
    with pytest.raises(ibgpu.IBGError):
        eng.TimeMarching(0, 1, use_graph=use_graph)
    eng.set_state(perturbed_state(tiny))
    if fault == "dt":
        eng.Comptdt(0)   # back to the CFL time step
    eng.TimeMarching(0, 2, use_graph=use_graph)
    assert np.all(np.isfinite(eng.get_state()))


def test_negative_density_reports_code(tiny):
    """An oversized time step drives a density negative: the engine returns
    the reference's abort code (523/611) with the offending cell instead of
    aborting; the oracle reports the same code."""
    fv0 = perturbed_state(tiny)
    eng = ibgpu.Engine(tiny, mode="fp64", dt_var_step=0)
    eng.set_state(fv0)
    eng.set_dt(np.array([5.0]))
    with pytest.raises(ibgpu.IBGError) as ei:
        eng.TimeMarching(0, 1, use_graph=False)
    assert ei.value.code in (523, 611, 37)
    assert ei.value.cell >= 0
    orc = oracle_ffi.Oracle(tiny, oracle_ffi.cfg(dt_var_step=0))
    cpu = fv0.copy()
    code, _, _, _ = orc.step(cpu, np.array([5.0]), 0)
    assert code in (523, 611, 37)
