"""GPU parity of the moving-body update (Body::MovingBody_Image between
regrids, Bodymove.cpp:37-767) against the oracle's incremental restatement.

Bar: bit-exact signed distances, closest patches, IB sets, image cells and
weights, attach points and wall-force patches after every move of a
sequence (integer and FP64 geometry computed in the same operation order on
both sides); the march after the bodies moved within 1e-12 relative L2 of
the oracle stepping on the moved geometry.
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi
from test_gpu_parity import TOL, perturbed_state, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny():
    spec = ibgpu.tiny_case()
    return spec, ibgpu.Mesh(spec)


def zero_z_layer(intp, coef):
    """The engine stores the zero-weight z layer of a 2D stencil as (-1, 0)."""
    intp, coef = intp.copy(), coef.copy()
    intp[:, 0::2] = -1
    coef[:, 0::2] = 0.0
    return intp, coef


def check_geometry(eng, mb, r):
    sd, cp, ib, nib = eng.geometry()
    assert np.array_equal(sd, mb.signdis)
    assert np.array_equal(cp, mb.cell_patch)
    assert np.array_equal(ib, mb.ib)
    assert nib == r["nib"]
    g = eng.ib_list(nib)
    order = np.argsort(g["cell"])
    assert np.array_equal(g["cell"][order], r["cell"])
    oi, oc = zero_z_layer(r["intp"], r["coef"])
    assert np.array_equal(g["intp"][order], oi)
    assert np.array_equal(g["coef"][order], oc)
    for k in ("hgdis", "nv", "attach", "body"):
        assert np.array_equal(g[k][order], r[k]), k
    p = eng.patches()
    assert np.array_equal(p["pc"], mb.pat_pc)
    assert np.array_equal(p["nv"], mb.pat_nv)
    assert np.array_equal(p["area"], mb.pat_area)
    pi, pc = zero_z_layer(r["pat_intp"], r["pat_coef"])
    assert np.array_equal(p["intp"], pi)
    assert np.array_equal(p["coef"], pc)
    assert np.array_equal(p["hgdis"], r["pat_hgdis"])


# pose sequences (angle step in degrees, heave step in chords) the reference's
# incremental update accepts on this mesh (larger steps near the thin
# trailing edge are its abort 314)
SEQUENCES = [(-0.3, -0.002), (0.25, 0.003)]


@pytest.mark.parametrize("da,dh", SEQUENCES)
def test_moves_equal_oracle(tiny, da, dh):
    spec, m = tiny
    print(f"moving-body geometry: tiny case, 8 moves of {da} deg, {dh} chord", flush=True)
    eng = ibgpu.Engine(m, mode="fp64")
    eng.upload_surface()
    mb = oracle_ffi.MovingBodies(m)
    for k in range(1, 9):
        pose = ibgpu.Mesh.pose(spec.aoa + da * k, dh * k)
        eng.move_bodies([pose])
        r = mb.move([pose])
        assert r["code"] == 0
        check_geometry(eng, mb, r)


def test_march_after_moves_equals_oracle(tiny):
    spec, m = tiny
    print("moving-body march: tiny case, 3 moves then 4 steps", flush=True)
    eng = ibgpu.Engine(m, mode="fp64", nstage=3)
    eng.upload_surface()
    mb = oracle_ffi.MovingBodies(m)
    for k in range(1, 4):
        pose = ibgpu.Mesh.pose(spec.aoa - 0.3 * k, -0.002 * k)
        eng.move_bodies([pose])
        r = mb.move([pose])
        assert r["code"] == 0
    fv0 = perturbed_state(m)
    mm = oracle_ffi.MovedMesh(mb, r)
    orc = oracle_ffi.Oracle(mm, oracle_ffi.cfg(nstage=3, dt_var_step=1 << (m.nlevels - 1)))
    cpu = fv0.copy()
    dt = np.array([m.phys[0].dt_init])
    nsteps = 4
    for ts in range(nsteps):
        code, _, _, _ = orc.step(cpu, dt, ts)
        assert code == 0, code
    eng.set_state(fv0)
    eng.TimeMarching(0, nsteps, use_graph=True)
    assert rel(eng.get_state()[:, :5], cpu[:, :5]) < TOL
    # the update count follows the moved masks
    upd = eng.cell_stage_updates()
    assert upd > 0


def test_too_large_a_move_is_reported(tiny):
    spec, m = tiny
    eng = ibgpu.Engine(m, mode="fp64")
    eng.upload_surface()
    with pytest.raises(ibgpu.IBGError) as e:
        eng.move_bodies([ibgpu.Mesh.pose(spec.aoa + 3.0, 0.0)])
    assert e.value.code in (263, 314)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_prescribed_motion_march(tiny, mode):
    """Pitch and heave by the device law inside the march: graph replay
    equals eager launches bit for bit; step by step, the geometry follows
    the oracle moved to the law's pose at the device's time of each step
    (the pose itself comes from device sin/cos, so within 1e-12)."""
    spec, m = tiny
    print(f"prescribed motion: tiny case, {mode}", flush=True)
    x0, y0, aoa0, amp, omega, heave = 0.0, 0.0, spec.aoa, 1.0, 10.0, 0.01
    law = [(x0, y0, aoa0, amp, omega, heave)]
    fv0 = perturbed_state(m)
    nsteps = 8
    out = []
    for use_graph in (False, True):
        eng = ibgpu.Engine(m, mode=mode, nstage=3)
        eng.upload_surface()
        eng.set_motion(law, t0=0.0)
        eng.set_state(fv0)
        eng.TimeMarching(0, nsteps, use_graph=use_graph)
        out.append((eng.get_state(), eng.geometry(), eng.get_time(), eng.cell_stage_updates()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1][0], out[1][1][0])
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]
    eng = ibgpu.Engine(m, mode=mode, nstage=3)
    eng.upload_surface()
    eng.set_motion(law, t0=0.0)
    eng.set_state(fv0)
    mb = oracle_ffi.MovingBodies(m)
    level = m.level_of()
    normal = np.asarray(m.kind)
    L = m.nlevels
    t = 0.0
    for ts in range(nsteps):
        # the march re-tags the levels the step marches (the ones its kernels
        # read); the oracle re-tags every level
        lmin = L - 1
        while lmin > 0 and ts % (1 << (L - lmin)) == 0:
            lmin -= 1
        fresh = level >= lmin
        a = np.deg2rad(aoa0 + amp * np.sin(omega * t))
        pose = [x0, y0 + heave * np.sin(omega * t), np.cos(-a), np.sin(-a)]
        r = mb.move([pose])
        assert r["code"] == 0
        eng.TimeMarching(ts, 1, use_graph=False)
        sd, cp, ib, nib = eng.geometry()
        assert np.array_equal(ib, mb.ib) and nib == r["nib"]
        assert np.allclose(sd[fresh], mb.signdis[fresh], rtol=1e-12, atol=1e-13)
        # cell-stage updates counted on the device: every marched normal cell
        # below the finest level, and the finest level's cells that are
        # neither solid nor IB cells (NS_Solver.cpp:856-983), per stage
        upd_cell = (level < L - 1) | ((sd >= 0.0) & (ib == 0))
        want = 3 * int(np.count_nonzero(fresh & (normal == 0) & upd_cell))
        assert eng.cell_stage_updates() == want, (ts, eng.cell_stage_updates(), want)
        t = eng.get_time()[0]
    assert t > 0.0
    assert np.array_equal(eng.get_state(), out[0][0])


@pytest.mark.parametrize("use_graph", [False, True])
def test_failed_move_run_ahead_equals_failed_move_in_place(tiny, use_graph):
    """In a march of several steps each step's move by the law runs ahead,
    beside the stages of the step before, with its own error word.  A law
    whose second move is far too large (the section at rest at t0, then
    tens of degrees within one step: the reference's abort) must end such a
    march with the code, state and time of single-step marches, where the
    move runs in place (the state the second step's exchange left, before
    its move; the time advanced by the first step)."""
    spec, m = tiny
    print(f"failed move ahead: tiny case, graph {use_graph}", flush=True)
    fv0 = perturbed_state(m)
    e0 = ibgpu.Engine(m, mode="fp64", nstage=3)
    e0.set_state(fv0)
    e0.TimeMarching(0, 1, use_graph=False)
    dt = e0.get_dt()[0]
    e0.close()
    # aoa(t) = aoa0 + amp sin(omega t), at rest at t0 on the mesh's pose
    amp, omega = 30.0, 1.0 / dt
    t0 = 0.5 * np.pi / omega
    law = [(0.0, 0.0, spec.aoa - amp, amp, omega, 0.0)]

    def fresh():
        e = ibgpu.Engine(m, mode="fp64", nstage=3)
        e.upload_surface()
        e.set_motion(law, t0=t0)
        e.set_state(fv0)
        return e

    a = fresh()
    with pytest.raises(ibgpu.IBGError) as ea:
        a.TimeMarching(0, 4, use_graph=use_graph)
    b = fresh()
    b.TimeMarching(0, 1, use_graph=False)
    t1 = b.get_time()
    with pytest.raises(ibgpu.IBGError) as eb:
        b.TimeMarching(1, 1, use_graph=False)
    assert ea.value.code == eb.value.code
    assert ea.value.code in (263, 314, 1183)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_time(), t1) and np.array_equal(b.get_time(), t1)
    assert t1[0] > t0


def test_regrid_during_prescribed_motion(tiny):
    """A regrid in the middle of a moving-body run: rebuild at the pose of the
    device time, state over by canonical keys (leaf totals conserved), the
    march carries on with the law's time; forces stay within 1e-3 of the run
    without the regrid."""
    import dataclasses
    from test_regrid import leaf_totals
    spec, m = tiny
    print("regrid during prescribed motion: tiny case", flush=True)
    x0, y0, aoa0, amp, omega, heave = 0.0, 0.0, spec.aoa, 1.0, 10.0, 0.01
    law = [(x0, y0, aoa0, amp, omega, heave)]
    fv0 = perturbed_state(m)
    runs = []
    for regrid in (False, True):
        eng = ibgpu.Engine(m, mode="fp64", nstage=3)
        eng.upload_surface()
        eng.set_motion(law, t0=0.0)
        eng.set_state(fv0)
        eng.TimeMarching(0, 8, use_graph=True)
        t = eng.get_time()[0]
        if regrid:
            s = np.sin(omega * t)
            m2 = ibgpu.Mesh(dataclasses.replace(spec, aoa=aoa0 + amp * s, pivot_y=y0 + heave * s))
            before = leaf_totals(eng.mesh, eng.get_state())
            eng.regrid(m2)
            after = leaf_totals(m2, eng.get_state())
            assert np.allclose(after, before, rtol=1e-12, atol=1e-12)
            assert eng.get_time()[0] == t
        eng.TimeMarching(8, 8, use_graph=True)
        st = eng.get_state()
        assert np.all(np.isfinite(st))
        runs.append(eng.ComptWallForce()[:2])
    assert np.allclose(runs[1], runs[0], rtol=1e-3), runs


@pytest.mark.parametrize("moving", [False, True])
def test_restart_from_solution_file_is_exact(tiny, tmp_path, moving):
    """Stop after 8 steps, write the Tecplot solution (state, SA variable,
    dt, motion time), restart a fresh context from it: the next 8 steps give
    the bits of the uninterrupted run."""
    spec, m = tiny
    print(f"restart from a solution file: tiny case, SA, moving={moving}", flush=True)
    law = [(0.0, 0.0, spec.aoa, 1.0, 10.0, 0.01)]
    fv0 = perturbed_state(m)
    nut0 = np.full(m.ncells, m.phys[0].nut_inf)

    def start():
        eng = ibgpu.Engine(m, mode="fp64", nstage=3, turbulence=True)
        if moving:
            eng.upload_surface()
            eng.set_motion(law, t0=0.0)
        return eng

    a = start()
    a.set_state(fv0)
    a.set_nut(nut0)
    a.TimeMarching(0, 8, use_graph=True)
    path = tmp_path / "amr_soln-8.plt"
    a.write_solution(path, 8)
    a.TimeMarching(8, 8, use_graph=True)
    b = start()
    ts = b.read_solution(path)
    assert ts == 8
    b.TimeMarching(ts, 8, use_graph=True)
    assert np.array_equal(b.get_state(), a.get_state())
    assert np.array_equal(b.get_nut(), a.get_nut())
    if moving:
        assert np.array_equal(b.get_time(), a.get_time())
