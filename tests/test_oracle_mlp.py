"""Pins for the CPU restatement of the learned closure (oracle/mlp.c).

  bf16          round-to-nearest-even on exactly representable neighbours
                (1 + 2^-8 ties to 1, 1 + 3*2^-9 rounds up, signs, powers of 2)
  FP64 network  the backward pass equals central finite differences of the
                loss for every parameter block
  BF16 mirror   computes the FP64 network's function and gradient up to the
                BF16 rounding of inputs, weights and activations
  features      the target z times the production scale is the point-wise SA
                term: the mesh-level SA source with a uniform working
                variable and no face term; a uniform flow has |Omega| = 0 and
                a vorticity share of 0
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi as O

L = O.lib()


def test_bf16_rounding():
    assert L.orc_bf16(1.0) == 1.0
    assert L.orc_bf16(1.0 + 2 ** -8) == 1.0                    # tie -> even
    assert L.orc_bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6      # tie -> even (up)
    assert L.orc_bf16(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7
    assert L.orc_bf16(-3.0) == -3.0
    assert L.orc_bf16(2.0 ** -20) == 2.0 ** -20
    x = 0.1
    b = L.orc_bf16(x)
    assert abs(b - x) <= 2 ** -9 * x


def rand_params(seed=3, scale=0.5):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, O.NPARAM)
    p[512:576] = rng.uniform(-0.1, 0.1, 64)
    return p


def test_f64_backprop_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rand_params()
    xs = rng.normal(size=(16, 8))
    t = rng.normal(size=16)
    g, loss = O.mlp_grad_f64(p, xs, t)
    assert loss == pytest.approx(O.mlp_loss_f64(p, xs, t), rel=1e-14)
    h = 1e-6
    for idx in list(rng.choice(512, 6)) + list(512 + rng.choice(64, 3)) + \
            list(576 + rng.choice(4096, 8)) + list(4672 + rng.choice(64, 3)) + \
            list(4736 + rng.choice(64, 3)) + [4800]:
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (O.mlp_loss_f64(pp, xs, t) - O.mlp_loss_f64(pm, xs, t)) / (2 * h)
        assert g[idx] == pytest.approx(fd, rel=1e-6, abs=1e-9), idx


def test_bf16_mirror_matches_f64_network():
    rng = np.random.default_rng(7)
    p = rand_params(scale=0.3)
    stats = np.concatenate([np.zeros(8), np.ones(8), -1e30 * np.ones(8), 1e30 * np.ones(8), [0.0, 1.0]])
    prm = np.concatenate([p, stats]).astype(np.float32)
    x = rng.normal(size=(256, 8)).astype(np.float32)
    t = rng.normal(size=256).astype(np.float32)
    z = O.mlp_forward(prm, x)
    gm, lm = O.mlp_grad(prm, x, t)
    g64, l64 = O.mlp_grad_f64(prm[:O.NPARAM].astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    # forward: same function up to bf16 rounding (~2^-8 relative per rounding)
    p64 = prm[:O.NPARAM].astype(np.float64)
    # y from the FP64 loss of a single row with t = 0: loss = y^2
    y64 = np.array([np.sqrt(O.mlp_loss_f64(p64, x[i:i + 1].astype(np.float64), np.zeros(1))) for i in range(8)])
    assert np.allclose(np.abs(z[:8]), y64, rtol=0.05, atol=0.02)
    assert lm == pytest.approx(l64, rel=0.05)
    assert np.linalg.norm(gm - g64) <= 0.05 * np.linalg.norm(g64)


@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case(Re=500.0))


def test_features_target_is_pointwise_sa_term(tiny):
    fv = tiny.init_state()
    x = tiny.centers
    fv[:, 1] += 0.1 * np.sin(1.3 * x[:, 1])
    fv[:, 2] += 0.05 * np.cos(0.7 * x[:, 0])
    P = tiny.phys[0]
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    nut = np.full(tiny.ncells, 3.0)
    orc = O.Oracle(tiny)
    ms = orc.mut_source(fv, nut, np.zeros(tiny.nfaces), 0)
    kappa2 = 0.41 ** 2
    checked = 0
    for g in np.nonzero((tiny.kind == 0) & (tiny.signdis > 0.05))[0][::37]:
        ok, xf, z = O.closure_features(tiny, fv, nut, g)
        assert ok
        olen = np.exp(xf[0]) - 1e-8
        d = np.exp(xf[2])
        scale = nut[g] * (olen + nut[g] / (kappa2 * d * d * P.Re))
        assert xf[1] == pytest.approx(olen / (olen + nut[g] / (kappa2 * d * d * P.Re)), rel=1e-6, abs=1e-12)
        assert z * scale == pytest.approx(ms[g], rel=1e-9, abs=1e-12)
        assert d == pytest.approx(tiny.signdis[g], rel=1e-12)
        checked += 1
    assert checked > 20


def test_features_uniform_flow(tiny):
    fv = tiny.init_state()
    nut = np.full(tiny.ncells, 5.0)
    g = int(np.nonzero((tiny.kind == 0) & (tiny.signdis > 0.5))[0][0])
    ok, xf, z = O.closure_features(tiny, fv, nut, g)
    assert ok
    assert xf[0] == pytest.approx(np.log(1e-8))
    assert xf[1] == 0.0
    assert xf[5] == 1.0 and xf[6] == 1.0
