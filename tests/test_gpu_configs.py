"""The five measured configurations at their full sizes on the GPU, through
the C ABI, against the oracle (FP64) or against single-case runs (batched).

  config 1  NACA0012 M0.5 Re5000 AoA 2, 64^2 base + 3 AMR levels, the
            configuration's own scheme (van Albada, 4 stages, CFL 0.5),
            200 steps from the uniform freestream: FP64 within 1e-12 of the
            oracle (state, dt, residual norms, wall force); the mixed mode's
            error stated and bounded
  config 2  NACA0015 M0.3 Re1e4, 256^2 base + 5 AMR levels (2.6 M cells),
            SA closure, the section pitching and heaving: after every move
            the GPU geometry equals the oracle's incremental update bit for
            bit, and 4 steps on the moved geometry are within 1e-12 (FP64);
            under the device's pitch/heave law the IB set, the distances and
            the device's cell-stage count follow the oracle step by step
  config 4  the learned closure on config 2: the network inside the stage
            kernel cell by cell against the CPU mirror, inference on the
            training rows, and C_L / C_D within 1 % of the FP64 run
  config 3  NACA0012 3D quasi-span, 128^2x8 base + 4 AMR levels (21 M
            cells), SA closure: 2 steps FP64 within 1e-12 of the oracle
  config 5  32 sections, 61.8 M cells in one context: three cases of the
            ensemble equal, bit for bit, the same cases run alone (FP64 and
            mixed, one sub-cycling period, graph replay), and one case
            within 1e-12 of the oracle

Every step prints what it runs before it runs it.
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi as O
from test_gpu_closure import network_stage_check
from test_gpu_move import check_geometry
from test_gpu_parity import perturbed_state, rel
from test_gpu_sa import seeded_nut

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(msg, flush=True)


def forces_rel(a, b):
    return np.abs(a[:2] - b[:2]) / np.maximum(np.abs(b[:2]), 1e-300)


# ------------------------------------------------------------------ config 1


@pytest.mark.timeout(300)
def test_config1_own_scheme_200_steps():
    mesh = ibgpu.Mesh(ibgpu.config1())
    L = mesh.nlevels
    nsteps = 200
    say(f"config 1: {mesh.ncells} cells, van Albada, 4 stages, CFL 0.5, {nsteps} steps (oracle)")
    fv0 = mesh.init_state()
    orc = O.Oracle(mesh, O.cfg(nstage=4, dt_var_step=1 << (L - 1), limiter=1))
    cpu = fv0.copy()
    dt = np.array([mesh.phys[0].dt_init])
    for ts in range(nsteps):
        code, _, linf, l2 = orc.step(cpu, dt, ts)
        assert code == 0, (ts, code)
    F_cpu = orc.forces(cpu)
    out = {}
    for mode in ("fp64", "mixed"):
        say(f"config 1: GPU {mode}")
        eng = ibgpu.Engine(mesh, mode=mode, nstage=4, limiter=1)
        eng.set_state(fv0)
        eng.TimeMarching(0, nsteps, use_graph=True)
        out[mode] = (eng.get_state(), eng.get_dt(), eng.ResinDomain(), eng.ComptWallForce())
        eng.close()
    g, gdt, (glinf, gl2), F = out["fp64"]
    assert rel(g[:, :5], cpu[:, :5]) < TOL
    assert np.allclose(gdt, dt, rtol=1e-12, atol=0)
    assert np.allclose(glinf, linf, rtol=1e-9, atol=1e-300)
    assert np.allclose(gl2, l2, rtol=1e-9, atol=1e-300)
    assert rel(F, F_cpu) < 1e-10
    gm, _, _, Fm = out["mixed"]
    e_state = rel(gm[:, :5], cpu[:, :5])
    dF = forces_rel(Fm, F_cpu)
    say(f"config 1: mixed after {nsteps} steps: state rel L2 {e_state:.3e}, "
        f"|dC_D|/C_D {dF[0]:.3e}, |dC_L|/C_L {dF[1]:.3e}")
    assert e_state < 1e-5
    assert np.all(dF < 1e-2)


# ------------------------------------------------------------------ config 2


@pytest.fixture(scope="module")
def c2():
    spec = ibgpu.config2()
    say("config 2: building the 2.6 M-cell mesh")
    return spec, ibgpu.Mesh(spec)


@pytest.mark.timeout(600)
def test_config2_moving_section_fp64_and_mixed(c2):
    spec, m = c2
    L = m.nlevels
    nsteps = 4
    fv0, nut0 = perturbed_state(m), seeded_nut(m)
    poses = [ibgpu.Mesh.pose(spec.aoa + 0.05 * (k + 1), 0.0005 * (k + 1)) for k in range(nsteps)]
    say(f"config 2: {m.ncells} cells, SA, 4 stages, {nsteps} moves of 0.05 deg / 0.0005 chord "
        f"each followed by a step (oracle)")
    mb = O.MovingBodies(m)
    ocfg = O.cfg(nstage=4, dt_var_step=1 << (L - 1))
    cpu, cnut = fv0.copy(), nut0.copy()
    dt = np.array([m.phys[0].dt_init])
    moves = []
    for ts in range(nsteps):
        r = mb.move([poses[ts]])
        assert r["code"] == 0, r["code"]
        moves.append((mb.signdis.copy(), r["nib"]))
        orc = O.Oracle(O.MovedMesh(mb, r), ocfg)
        code, _, _, _ = orc.step_sa(cpu, cnut, dt, ts)
        assert code == 0, (ts, code)
    # the geometry of every move, bit for bit, at full size
    say("config 2: GPU geometry after each move (FP64)")
    eng = ibgpu.Engine(m, mode="fp64", nstage=4, turbulence=True)
    eng.upload_surface()
    mb2 = O.MovingBodies(m)
    eng.set_state(fv0)
    eng.set_nut(nut0)
    for ts in range(nsteps):
        eng.move_bodies([poses[ts]])
        r = mb2.move([poses[ts]])
        check_geometry(eng, mb2, r)
        eng.TimeMarching(ts, 1, use_graph=False)
    g, gn = eng.get_state(), eng.get_nut()
    eng.close()
    e, en = rel(g[:, :5], cpu[:, :5]), rel(gn, cnut)
    say(f"config 2: FP64 after {nsteps} steps: state {e:.3e}, nut {en:.3e}")
    assert e < TOL and en < TOL
    say("config 2: GPU mixed")
    eng = ibgpu.Engine(m, mode="mixed", nstage=4, turbulence=True)
    eng.upload_surface()
    eng.set_state(fv0)
    eng.set_nut(nut0)
    for ts in range(nsteps):
        eng.move_bodies([poses[ts]])
        eng.TimeMarching(ts, 1, use_graph=False)
    e, en = rel(eng.get_state()[:, :5], cpu[:, :5]), rel(eng.get_nut(), cnut)
    eng.close()
    say(f"config 2: mixed after {nsteps} steps: state {e:.3e}, nut {en:.3e}")
    assert e < 1e-5 and en < 1e-4


@pytest.mark.timeout(600)
def test_config2_device_law_follows_oracle(c2):
    """Pitch and heave by the law inside the march (the benchmark's form, at
    a higher frequency so that the IB set changes within a few steps): step
    by step the device's IB set and the distances of the marched levels
    equal the oracle moved to the law's pose at the device's time, and the
    device's cell-stage count equals a host recount from that geometry."""
    spec, m = c2
    L = m.nlevels
    law = (0.0, 0.0, spec.aoa, 2.0, 40.0, 0.01)
    nsteps = 8
    say(f"config 2: device law {law}, {nsteps} steps, geometry and update count vs the oracle")
    eng = ibgpu.Engine(m, mode="fp64", nstage=4, turbulence=True)
    eng.upload_surface()
    eng.set_motion([law], t0=0.0)
    eng.set_state(m.init_state())
    eng.set_nut(np.full(m.ncells, m.phys[0].nut_inf))
    mb = O.MovingBodies(m)
    level = m.level_of()
    normal = np.asarray(m.kind) == 0
    x0, y0, aoa0, amp, omega, heave = law
    t = 0.0
    changed = 0
    ib_prev = mb.ib.copy()
    for ts in range(nsteps):
        lmin = L - 1
        while lmin > 0 and ts % (1 << (L - lmin)) == 0:
            lmin -= 1
        fresh = level >= lmin
        a = np.deg2rad(aoa0 + amp * np.sin(omega * t))
        r = mb.move([[x0, y0 + heave * np.sin(omega * t), np.cos(-a), np.sin(-a)]])
        assert r["code"] == 0
        eng.TimeMarching(ts, 1, use_graph=False)
        sd, cp, ib, nib = eng.geometry()
        assert np.array_equal(ib, mb.ib) and nib == r["nib"], ts
        assert np.allclose(sd[fresh], mb.signdis[fresh], rtol=1e-12, atol=1e-13)
        upd_cell = (level < L - 1) | ((sd >= 0.0) & (ib == 0))
        want = 4 * int(np.count_nonzero(fresh & normal & upd_cell))
        assert eng.cell_stage_updates() == want, (ts, eng.cell_stage_updates(), want)
        changed += int(np.count_nonzero(mb.ib != ib_prev))
        ib_prev = mb.ib.copy()
        t = eng.get_time()[0]
    eng.close()
    say(f"config 2: IB membership changes over the {nsteps} steps: {changed}")
    assert changed > 0


# ------------------------------------------------------------------ config 4

MOTION = (0.0, 0.0, 0.0, 5.0, 0.2, 0.05)   # pitch 5 deg, heave 0.05 c, k = 0.1


@pytest.mark.timeout(600)
def test_config4_network_in_stage_full_size(c2):
    spec, m = c2
    say("config 4: the network inside the stage kernel on the 2.6 M-cell mesh (sampled cells)")
    fv = perturbed_state(m)
    rng = np.random.default_rng(4)
    nut = m.phys[0].nut_inf * (1.0 + 0.5 * rng.uniform(-1.0, 1.0, m.ncells))
    network_stage_check(m, fv, nut, sample=20000)


@pytest.mark.timeout(600)
def test_config4_trained_closure_forces(c2):
    """The configuration's closure: 5 M cells sampled from the plain FP64 run
    of the moving section, 20 epochs of Adam at batch 8192; the GPU
    inference on training rows against the CPU mirror; the learned form's
    C_L and C_D within 1 % of the plain FP64 form's after 256 steps, with
    both network and fallback cells in use."""
    spec, m = c2
    say("config 4: training the closure (5 M samples, 20 epochs, batch 8192)")
    prm, info = ibgpu.train_closure(m, nstage=4, motion=MOTION)
    say(f"config 4: {info}")
    assert info["samples"] >= 5_000_000 and np.all(np.isfinite(prm))
    nsteps = 256
    ref = ibgpu.plain_engine(m, 4, MOTION)
    ref.TimeMarching(0, nsteps, use_graph=True)
    F_ref, nut_ref = ref.ComptWallForce(), ref.get_nut()
    ref.close()
    eng = ibgpu.Engine(m, mode="mixed", nstage=4, turbulence=True)
    eng.set_state(m.init_state())
    eng.set_nut(np.full(m.ncells, m.phys[0].nut_inf))
    eng.upload_surface()
    eng.set_motion([MOTION], t0=0.0)
    eng.closure_init(1 << 16)
    eng.closure_set_params(prm)
    # inference on rows of this state's cells: GPU against the CPU mirror
    eng.closure_sample(0.02, 99)
    x, _, _ = eng.closure_dataset(20000)
    assert rel(eng.closure_infer(x).astype(np.float64),
               O.mlp_forward(prm, x).astype(np.float64)) < 1e-3
    eng.closure_enable(True)
    eng.closure_counts()
    eng.TimeMarching(0, nsteps, use_graph=True)
    on_net, fallback = eng.closure_counts()
    F, nut = eng.ComptWallForce(), eng.get_nut()
    eng.close()
    dF = forces_rel(F, F_ref)
    say(f"config 4: after {nsteps} steps: C_D {F_ref[0]:.6e} -> {F[0]:.6e}, C_L {F_ref[1]:.6e} -> "
        f"{F[1]:.6e}, relative {dF}, nut rel L2 {rel(nut, nut_ref):.3e}, network cells {on_net}, "
        f"fallback cells {fallback}")
    assert np.all(dF <= 1e-2), dF
    assert on_net > 0 and fallback >= 0 and fallback < 0.01 * on_net


# ------------------------------------------------------------------ config 3


@pytest.mark.timeout(900)
def test_config3_quasi_span_fp64():
    say("config 3: building the 21 M-cell 3D mesh")
    mesh = ibgpu.Mesh(ibgpu.config3())
    L = mesh.nlevels
    nsteps = 2
    fv0 = perturbed_state(mesh)
    nut0 = seeded_nut(mesh)
    say(f"config 3: {mesh.ncells} cells, SA, 4 stages, {nsteps} steps (oracle)")
    orc = O.Oracle(mesh, O.cfg(nstage=4, dt_var_step=1 << (L - 1)))
    cpu, cnut = fv0.copy(), nut0.copy()
    dt = np.array([mesh.phys[0].dt_init])
    for ts in range(nsteps):
        code, _, _, _ = orc.step_sa(cpu, cnut, dt, ts)
        assert code == 0, (ts, code)
    for mode in ("fp64", "mixed"):
        say(f"config 3: GPU {mode}")
        eng = ibgpu.Engine(mesh, mode=mode, nstage=4, turbulence=True)
        eng.set_state(fv0)
        eng.set_nut(nut0)
        eng.TimeMarching(0, nsteps, use_graph=False)
        e, en = rel(eng.get_state()[:, :5], cpu[:, :5]), rel(eng.get_nut(), cnut)
        eng.close()
        say(f"config 3: {mode}: state {e:.3e}, nut {en:.3e}")
        if mode == "fp64":
            assert e < TOL and en < TOL
        else:
            assert e < 1e-5 and en < 1e-4


# ------------------------------------------------------------------ config 5


@pytest.mark.timeout(900)
def test_config5_ensemble_equals_single_cases():
    specs = ibgpu.ensemble_cases()
    say(f"config 5: building the {len(specs)}-case ensemble")
    batch = ibgpu.Mesh(specs)
    L = batch.nlevels
    P = 1 << (L - 1)
    per_case = [int(sum(r)) for r in batch.counts()]
    say(f"config 5: {batch.ncells} cells, normal cells per case {min(per_case)}..{max(per_case)}")
    pick = [int(np.argmin(per_case)), 13, int(np.argmax(per_case))]
    fvb = batch.init_state()
    nutb = np.full(batch.ncells, batch.phys[0].nut_inf)
    got = {}
    for mode in ("fp64", "mixed"):
        say(f"config 5: ensemble {mode}, {P} steps")
        eb = ibgpu.Engine(batch, mode=mode, nstage=4, turbulence=True)
        eb.set_state(fvb)
        eb.set_nut(nutb)
        eb.TimeMarching(0, P, use_graph=True)
        st, nu = eb.get_state(), eb.get_nut()
        eb.close()
        for c in pick:
            lo, hi = batch.blk_start[c * L], batch.blk_start[(c + 1) * L]
            got[mode, c] = (st[lo:hi].copy(), nu[lo:hi].copy())
        del st, nu
    for c in pick:
        single = ibgpu.Mesh([specs[c]])
        for mode in ("fp64", "mixed"):
            say(f"config 5: case {c} alone, {mode}")
            es = ibgpu.Engine(single, mode=mode, nstage=4, turbulence=True)
            es.set_state(single.init_state())
            es.set_nut(np.full(single.ncells, single.phys[0].nut_inf))
            es.TimeMarching(0, P, use_graph=True)
            assert np.array_equal(es.get_state(), got[mode, c][0]), (mode, c)
            assert np.array_equal(es.get_nut(), got[mode, c][1]), (mode, c)
            es.close()
        if c == pick[0]:
            say(f"config 5: case {c} against the oracle, 2 steps")
            orc = O.Oracle(single, O.cfg(nstage=4, dt_var_step=1 << (L - 1)))
            cpu = single.init_state()
            cnut = np.full(single.ncells, single.phys[0].nut_inf)
            dt = np.array([single.phys[0].dt_init])
            for ts in range(2):
                assert orc.step_sa(cpu, cnut, dt, ts)[0] == 0
            es = ibgpu.Engine(single, mode="fp64", nstage=4, turbulence=True)
            es.set_state(single.init_state())
            es.set_nut(np.full(single.ncells, single.phys[0].nut_inf))
            es.TimeMarching(0, 2, use_graph=False)
            assert rel(es.get_state()[:, :5], cpu[:, :5]) < TOL
            assert rel(es.get_nut(), cnut) < TOL
            es.close()
