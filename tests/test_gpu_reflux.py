"""The optional coarse/fine flux balance on the GPU, through the C ABI.

  parity        with the balance on (no temporal refinement), the plain FP64
                mode matches the FP64 oracle within 1e-12 relative L2 on the
                state (and the SA working variable) after several steps, on
                2D (two fine faces per coarse face) and 3D meshes (two and
                four), laminar and with the SA closure; graph replay gives the
                eager bits
  conservation  on the body-free refined level set the volume integral of
                the conserved variables drifts by rounding only with the
                balance (FP64), far less than without it; in the mixed mode
                the balanced fluxes are FP32, and the drift stays well below
                the unbalanced one
"""
import dataclasses

import numpy as np
import pytest

import ibgpu
import oracle_ffi
from test_gpu_parity import perturbed_state, rel
from test_gpu_sa import seeded_nut
from test_reflux import bump_state, totals

pytestmark = pytest.mark.gpu

TOL = 1e-12


def specs():
    return {"tiny2d": ibgpu.tiny_case(Re=2000.0),
            "tiny3d_xyz": dataclasses.replace(ibgpu.tiny3d_case(), twod_from=2)}


@pytest.fixture(scope="module", params=list(specs()))
def mesh(request):
    return ibgpu.Mesh(specs()[request.param])


def oracle_march(mesh, fv, nut, nsteps, nstage, reflux=1):
    orc = oracle_ffi.Oracle(mesh, oracle_ffi.cfg(nstage=nstage, dt_var_step=1, temporal_refine=0,
                                                 reflux=reflux))
    cpu = fv.copy()
    cnut = None if nut is None else nut.copy()
    dt = np.array([mesh.phys[c].dt_init for c in range(mesh.ncases)])
    for ts in range(nsteps):
        if nut is None:
            code, *_ = orc.step(cpu, dt, ts)
        else:
            code, *_ = orc.step_sa(cpu, cnut, dt, ts)
        assert code == 0, code
    return cpu, cnut


def gpu_march(mesh, fv, nut, nsteps, nstage, mode="fp64", use_graph=False, reflux=True):
    e = ibgpu.Engine(mesh, mode=mode, nstage=nstage, temporal_refine=False,
                     turbulence=nut is not None, reflux=reflux)
    e.set_state(fv)
    if nut is not None:
        e.set_nut(nut)
    e.TimeMarching(0, nsteps, use_graph=use_graph)
    out = (e.get_state(), e.get_nut() if nut is not None else None)
    e.close()
    return out


@pytest.mark.parametrize("sa", [False, True])
def test_balanced_steps_match_oracle_fp64(mesh, sa):
    print(f"reflux parity dim {mesh.dim} sa {sa}", flush=True)
    assert mesh.nmdy > 0
    fv = perturbed_state(mesh)
    nut = seeded_nut(mesh) if sa else None
    cpu, cnut = oracle_march(mesh, fv, nut, 6, 4)
    g, gnut = gpu_march(mesh, fv, nut, 6, 4)
    assert rel(g, cpu) < TOL
    if sa:
        assert rel(gnut, cnut) < TOL
    # the balance acts: the unbalanced march differs
    off, _ = oracle_march(mesh, fv, nut, 6, 4, reflux=0)
    assert rel(off, cpu) > 1e-9


def test_balanced_graph_replay_equals_eager(mesh):
    fv = perturbed_state(mesh)
    a, _ = gpu_march(mesh, fv, None, 5, 4, use_graph=False)
    b, _ = gpu_march(mesh, fv, None, 5, 4, use_graph=True)
    assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def free_mesh():
    return ibgpu.Mesh(dataclasses.replace(ibgpu.tiny_case(base=48), refine_only=True))


@pytest.mark.parametrize("nstage,nsteps", [(1, 4), (4, 1)])
def test_balanced_march_conserves_fp64(free_mesh, nstage, nsteps):
    m = free_mesh
    fv = bump_state(m)
    t0 = totals(m, fv)
    g, _ = gpu_march(m, fv, None, nsteps, nstage)
    cpu, _ = oracle_march(m, fv, None, nsteps, nstage)
    assert rel(g, cpu) < TOL
    g_off, _ = gpu_march(m, fv, None, nsteps, nstage, reflux=False)
    on = np.abs(totals(m, g) - t0) / abs(t0[0])
    off = np.abs(totals(m, g_off) - t0) / abs(t0[0])
    print("drift on", on, "off", off, flush=True)
    assert on.max() < 2e-14
    assert off[[0, 3]].min() > 1e-9


def test_balanced_march_mixed(free_mesh):
    m = free_mesh
    fv = bump_state(m)
    t0 = totals(m, fv)
    cpu, _ = oracle_march(m, fv, None, 4, 1)
    g, _ = gpu_march(m, fv, None, 4, 1, mode="mixed")
    g_off, _ = gpu_march(m, fv, None, 4, 1, mode="mixed", reflux=False)
    # FP32 flux arithmetic: the mixed-mode state bound
    assert rel(g, cpu) < 1e-5
    on = np.abs(totals(m, g) - t0) / abs(t0[0])
    off = np.abs(totals(m, g_off) - t0) / abs(t0[0])
    print("mixed drift on", on, "off", off, flush=True)
    assert on[[0, 3]].max() < 1e-2 * off[[0, 3]].min()
