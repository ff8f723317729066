"""State transfer of a regrid (AMR::ComptDatainNewBox AMR.H:148-215) on
canonical keys, host side (libibhost.so ibm_transfer).

Pins: a rebuild of the same level set is the identity, bit for bit; a
rebuild at another pose (cells refined and coarsened near the band edges)
conserves the mass, momentum and total energy of the leaf cells to
rounding (restriction averages conserved variables, prolongation adds
zero-mean +-1/4 gradient terms), and every cell the two level sets share
as a leaf keeps its state.
"""
import dataclasses
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from ibgpu import mesh as M  # noqa: E402
from test_move_oracle import keys_of  # noqa: E402


def smooth_state(m, seed=7):
    rng = np.random.default_rng(seed)
    fv = m.init_state()
    x = m.centers
    k = rng.uniform(0.5, 2.0, size=3)
    fv[:, 0] *= 1 + 0.05 * np.sin(k[0] * x[:, 0]) * np.cos(k[1] * x[:, 1])
    fv[:, 1] += 0.05 * np.cos(k[2] * x[:, 1])
    fv[:, 2] += 0.05 * np.sin(k[0] * x[:, 0])
    fv[:, 4] *= 1 + 0.05 * np.cos(k[1] * x[:, 0] + k[2] * x[:, 1])
    P = m.phys[0]
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2 + fv[:, 3] ** 2)
    return fv


def leaf_totals(m, fv):
    """sum over normal cells of (rho, rho u, rho v, rho E) * cell area"""
    tot = np.zeros(4)
    for b in range(m.ncases * m.nlevels):
        s, e = m.blk_start[b], m.blk_start[b] + m.blk_nnormal[b]
        h = m.blk_h[b]
        vol = h[0] * h[1]
        f = fv[s:e]
        E = f[:, 6]
        tot += vol * np.array([f[:, 0].sum(), (f[:, 0] * f[:, 1]).sum(), (f[:, 0] * f[:, 2]).sum(),
                               (f[:, 0] * E).sum()])
    return tot


@pytest.fixture(scope="module")
def spec():
    return M.tiny_case(levels=4)


def test_same_level_set_is_identity(spec):
    a, b = M.Mesh(spec), M.Mesh(spec)
    fv = smooth_state(a)
    nut = np.linspace(1.0, 5.0, a.ncells)
    out, nout = a.transfer_to(b, fv, nut)
    assert np.array_equal(out, fv) and np.array_equal(nout, nut)


@pytest.mark.parametrize("daoa,dheave", [(3.0, 0.0), (-4.0, 0.05), (0.0, 0.2)])
def test_regrid_conserves_leaf_totals(spec, daoa, dheave):
    a = M.Mesh(spec)
    b = M.Mesh(dataclasses.replace(spec, aoa=spec.aoa + daoa, pivot_y=dheave))
    ka, kb = keys_of(a), keys_of(b)
    assert ka != kb   # the level sets differ
    fv = smooth_state(a)
    nut = 1.0 + 0.1 * np.sin(a.centers[:, 0])
    out, nout = a.transfer_to(b, fv, nut)
    assert np.all(np.isfinite(out)) and np.all(out[:, 0] > 0) and np.all(out[:, 4] > 0)
    ta, tb = leaf_totals(a, fv), leaf_totals(b, out)
    assert np.allclose(tb, ta, rtol=1e-12, atol=1e-12), (ta, tb)
    # shared leaves keep their state
    ia = {k: i for i, k in enumerate(ka)}
    na, nb_ = np.asarray(a.kind), np.asarray(b.kind)
    same = [(ia[k], j) for j, k in enumerate(kb) if k in ia and na[ia[k]] == 0 and nb_[j] == 0]
    assert len(same) > 0.5 * int(np.sum(nb_ == 0))
    ii, jj = np.array(same).T
    assert np.array_equal(out[jj], fv[ii]) and np.array_equal(nout[jj], nut[ii])


def test_thermodynamics_consistent_after_transfer(spec):
    a = M.Mesh(spec)
    b = M.Mesh(dataclasses.replace(spec, aoa=spec.aoa + 3.0))
    out, _ = a.transfer_to(b, smooth_state(a))
    P = a.phys[0]
    assert np.allclose(out[:, 5], out[:, 4] / (out[:, 0] * P.Rg), rtol=1e-14)
    assert np.allclose(out[:, 6], P.Cv * out[:, 5] + 0.5 * (out[:, 1] ** 2 + out[:, 2] ** 2 + out[:, 3] ** 2),
                       rtol=1e-14)
