"""GPU parity of the Spalart-Allmaras closure (the TURBULENCE build) and of
the stencil-buffer rotation, through the C ABI against the FP64 oracle.

Bar: the plain FP64 mode within 1e-12 relative L2 of the oracle on the
state and the SA working variable after several steps, on the viscous and
SA face fluxes, the SA source of the last stage, the SA residual, the ghost
fills of the working variable (IB wall model, domain BC, AMR transfer) and
the wall-function force.  The mixed mode is bounded (state 1e-4 relative L2
for the working variable, C_L/C_D within 1 %).
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi
from test_gpu_parity import cols, perturbed_state, rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def seeded_nut(mesh, seed=11):
    rng = np.random.default_rng(seed)
    x = mesh.centers
    out = np.zeros(mesh.ncells)
    for c in range(mesh.ncases):
        s, e = mesh.blk_start[c * mesh.nlevels], mesh.blk_start[(c + 1) * mesh.nlevels]
        P = mesh.phys[c]
        out[s:e] = P.nut_inf * (1 + 0.3 * np.sin(0.9 * x[s:e, 0]) * np.cos(1.1 * x[s:e, 1]))
    return out + 0.01 * rng.random(mesh.ncells)


@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case(Re=2000.0))


@pytest.fixture(scope="module")
def tiny3d():
    return ibgpu.Mesh(ibgpu.tiny3d_case())


def run_oracle_sa(mesh, fv, nut, nsteps, nstage=3):
    orc = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(nstage=nstage, dt_var_step=1 << (mesh.nlevels - 1)))
    cpu, cnut = fv.copy(), nut.copy()
    dt = np.array([mesh.phys[c].dt_init for c in range(mesh.ncases)])
    out = None
    for ts in range(nsteps):
        code, src, msrc, norms = orc.step_sa(cpu, cnut, dt, ts)
        assert code == 0, code
        out = (src, msrc, norms)
    return cpu, cnut, dt, out


def engine(mesh, fv, nut, mode="fp64", nstage=3):
    eng = ibgpu.Engine(mesh, mode=mode, nstage=nstage, turbulence=True)
    eng.set_state(fv)
    eng.set_nut(nut)
    return eng


@pytest.mark.parametrize("which", ["tiny", "tiny3d"])
def test_sa_viscous_and_mutflux_fp64(which, request):
    mesh = request.getfixturevalue(which)
    fv, nut = perturbed_state(mesh), seeded_nut(mesh)
    eng = engine(mesh, fv, nut)
    g = eng.vsflux()
    gm = eng.mutflux()
    c, cm = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(temporal_refine=0)).viscous_flux_sa(fv, nut, 0)
    k = [1, 2, 3, 4] if mesh.dim == 3 else [1, 2, 4]
    assert rel(g[:, k], c[:, k]) < TOL
    assert rel(gm, cm) < TOL
    assert np.abs(cm).max() > 0


@pytest.mark.parametrize("which,nstage", [("tiny", 3), ("tiny", 4), ("tiny3d", 3)])
def test_sa_steps_fp64_match_oracle(which, nstage, request):
    mesh = request.getfixturevalue(which)
    fv0, nut0 = perturbed_state(mesh), seeded_nut(mesh)
    nsteps = 8
    cpu, cnut, dt, (src, msrc, (linf, l2, nlinf, nl2)) = run_oracle_sa(mesh, fv0, nut0, nsteps, nstage)
    eng = engine(mesh, fv0, nut0, nstage=nstage)
    eng.TimeMarching(0, nsteps, use_graph=False)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL
    assert rel(eng.get_nut(), cnut) < TOL
    assert np.allclose(eng.get_dt(), dt, rtol=1e-14, atol=0)
    glinf, gl2 = eng.ResinDomain()
    assert np.allclose(glinf, linf, rtol=1e-10, atol=1e-300)
    gn, gn2 = eng.ResinDomain_nut()
    assert np.allclose(gn, nlinf, rtol=1e-10, atol=1e-300)
    assert np.allclose(gn2, nl2, rtol=1e-10, atol=1e-300)


def test_sa_stagewise_source(tiny):
    fv0, nut0 = perturbed_state(tiny), seeded_nut(tiny)
    cpu, cnut, dt, (src, msrc, _) = run_oracle_sa(tiny, fv0, nut0, 1)
    eng = engine(tiny, fv0, nut0)
    ts = 0
    eng.InterfaceExchange(ts)
    eng.IBConditions_Image()
    eng.Comptdt(ts)
    eng.BCBoxTreat(ts)
    eng.ComptViscousStress(ts)
    for k in range(3):
        eng.Stage(ts, k)
    normal = tiny.kind == 0
    # the reference also evaluates mutsource on IB cells, which are never
    # updated nor counted in the residual; the engine skips them (reports 0)
    used = normal & (tiny.ibidx < 0)
    assert rel(eng.nut_source()[used], msrc[used]) < TOL
    assert np.all(eng.nut_source()[tiny.ibidx >= 0] == 0.0)
    assert rel(eng.source()[normal][:, cols(tiny)], src[normal][:, cols(tiny)]) < TOL
    assert rel(eng.get_nut(), cnut) < TOL


@pytest.mark.parametrize("phase", ["ib", "bc", "exchange"])
def test_sa_ghost_fills_fp64(tiny, phase):
    fv, nut = perturbed_state(tiny), seeded_nut(tiny)
    eng = engine(tiny, fv, nut)
    orc = oracle_ffi.Oracle(tiny)
    cpu, cnut = fv.copy(), nut.copy()
    if phase == "ib":
        eng.IBConditions_Image()
        assert orc.ib_fill_sa(cpu, cnut) == 0
    elif phase == "bc":
        eng.BCBoxTreat(0)
        assert orc.bc_fill(cpu, 0) == 0
        assert orc.sa_bc_fill(cpu, cnut, 0) == 0
    else:
        P = 1 << (tiny.nlevels - 1)
        for ts in range(P):
            eng.InterfaceExchange(ts)
            orc.interface_exchange(cpu, ts)
            orc.sa_exchange(cnut, ts)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL
    assert rel(eng.get_nut(), cnut) < TOL
    assert np.any(cnut != nut)


def test_sa_forces_fp64(tiny):
    fv = perturbed_state(tiny)
    eng = engine(tiny, fv, seeded_nut(tiny))
    F = eng.ComptWallForce()
    Fo = oracle_ffi.Oracle(tiny).forces_sa(fv)
    assert rel(F, Fo) < TOL
    assert abs(Fo[0]) > 0


def test_sa_graph_replay_equals_eager(tiny):
    fv0, nut0 = perturbed_state(tiny), seeded_nut(tiny)
    P = 1 << (tiny.nlevels - 1)
    a = engine(tiny, fv0, nut0, mode="mixed")
    a.TimeMarching(0, 3 * P, use_graph=False)
    b = engine(tiny, fv0, nut0, mode="mixed")
    b.TimeMarching(0, P, use_graph=True)
    b.TimeMarching(P, 2 * P, use_graph=True)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_nut(), b.get_nut())


def test_sa_mixed_error_bounded(tiny):
    fv0 = tiny.init_state()
    nut0 = np.full(tiny.ncells, tiny.phys[0].nut_inf)
    nsteps = 40
    cpu, cnut, _, _ = run_oracle_sa(tiny, fv0, nut0, nsteps)
    eng = engine(tiny, fv0, nut0, mode="mixed")
    eng.TimeMarching(0, nsteps, use_graph=True)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < 1e-5
    assert rel(eng.get_nut(), cnut) < 1e-4
    F = eng.ComptWallForce()
    Fo = oracle_ffi.Oracle(tiny).forces_sa(cpu)
    assert abs(F[0] - Fo[0]) <= 0.01 * abs(Fo[0])
    assert abs(F[1] - Fo[1]) <= 0.01 * abs(Fo[1])


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_stage_buffers_make_runs_reproducible(mode):
    """Stages read one stencil buffer and write another: two runs over a mesh
    of many tiles (many workgroups in flight) give identical bits, and the
    FP64 run matches the oracle."""
    mesh = ibgpu.Mesh(ibgpu.config1())
    fv0 = perturbed_state(mesh)
    runs = []
    for _ in range(2):
        eng = ibgpu.Engine(mesh, mode=mode, nstage=4)
        eng.set_state(fv0)
        eng.TimeMarching(0, 8, use_graph=True)
        runs.append(eng.get_state())
    assert np.array_equal(runs[0], runs[1])
    if mode == "fp64":
        orc = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(nstage=4, dt_var_step=1 << (mesh.nlevels - 1)))
        cpu = fv0.copy()
        dt = np.array([mesh.phys[0].dt_init])
        for ts in range(8):
            assert orc.step(cpu, dt, ts)[0] == 0
        assert rel(runs[0][:, :5], cpu[:, :5]) < TOL


def test_single_stage_copy_back(tiny):
    fv0 = perturbed_state(tiny)
    orc = oracle_ffi.Oracle(tiny, oracle_ffi.cfg(nstage=1, dt_var_step=1 << (tiny.nlevels - 1)))
    cpu = fv0.copy()
    dt = np.array([tiny.phys[0].dt_init])
    for ts in range(4):
        assert orc.step(cpu, dt, ts)[0] == 0
    eng = ibgpu.Engine(tiny, mode="fp64", nstage=1)
    eng.set_state(fv0)
    eng.TimeMarching(0, 4, use_graph=False)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL


@pytest.fixture(scope="module")
def walled():
    import dataclasses
    return ibgpu.Mesh(dataclasses.replace(ibgpu.tiny_case(), face_bc=(0, 0, -81, 0)))


def test_wall_bc_fill_fp64(walled):
    """Domain wall ghosts (BC -81, ImageViseddy_Domain) against the oracle."""
    print("domain wall fill: tiny case, wall at y = lo", flush=True)
    fv, nut = perturbed_state(walled), seeded_nut(walled)
    eng = engine(walled, fv, nut)
    orc = oracle_ffi.Oracle(walled)
    cpu, cnut = fv.copy(), nut.copy()
    eng.BCBoxTreat(0)
    assert orc.sa_bc_fill(cpu, cnut, 0) == 0
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL
    assert rel(eng.get_nut(), cnut) < TOL


def test_wall_march_fp64(walled):
    """Steps with a domain wall against the oracle with the SA closure."""
    print("domain wall march: tiny case, 6 steps", flush=True)
    fv, nut = perturbed_state(walled), seeded_nut(walled)
    orc = oracle_ffi.Oracle(walled, oracle_ffi.cfg(nstage=3, dt_var_step=1 << (walled.nlevels - 1)))
    cpu, cnut = fv.copy(), nut.copy()
    dt = np.array([walled.phys[0].dt_init])
    for ts in range(6):
        code, _, _, _ = orc.step_sa(cpu, cnut, dt, ts)
        assert code == 0, code
    eng = engine(walled, fv, nut)
    eng.TimeMarching(0, 6, use_graph=True)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL
    assert rel(eng.get_nut(), cnut) < TOL


def test_wall_needs_the_closure(walled):
    with pytest.raises(ibgpu.IBGError):
        ibgpu.Engine(walled, mode="fp64")
