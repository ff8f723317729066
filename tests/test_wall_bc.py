"""The domain wall of the TURBULENCE build (BC -81, BCValues.H:89-125,
ImageViseddy_Domain Turbulence.cpp:140-195) on the host side: the builder's
wall distances and ghost neighbours, and pins of the oracle's wall fill.
"""
import dataclasses
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle_ffi as O  # noqa: E402
from ibgpu import mesh as M  # noqa: E402


@pytest.fixture(scope="module")
def walled():
    spec = dataclasses.replace(M.tiny_case(), face_bc=(0, 0, -81, 0))   # wall at y = lo
    return spec, M.Mesh(spec), M.Mesh(M.tiny_case())


def test_wall_distance_and_signdis(walled):
    spec, m, plain = walled
    v = np.ctypeslib.as_array
    dist = v(m.soa.dist_dm, (m.ncells,))
    y = m.centers[:, 1]
    normal = np.asarray(m.kind) == 0
    ib = np.asarray(m.ibidx) >= 0
    on = normal & ~ib
    assert np.array_equal(dist[on], np.abs(y[on] - spec.lo))
    assert np.all(dist[~on] == 9999.0)
    # signdis = min(signdis of the body, distance to the wall)
    assert np.array_equal(np.asarray(m.signdis)[on], np.minimum(np.asarray(plain.signdis)[on], dist[on]))
    # wall ghosts: the reference cell's inward neighbour and the inward normal
    nd = m.soa.ndmg
    ty = v(m.soa.dmg_type, (nd,))
    nbr = v(m.soa.dmg_nbr, (nd,))
    ref = v(m.soa.dmg_ref, (nd,))
    nmv = v(m.soa.dmg_nmv, (nd, 3))
    w = ty == -81
    assert w.sum() > 0 and np.all(nbr[w] >= 0) and np.all(nbr[~w] == -1)
    assert np.allclose(nmv[w], [0.0, 1.0, 0.0])
    key = np.asarray(m.key)
    assert np.array_equal(key[nbr[w], 1], key[ref[w], 1] + 1)
    assert np.array_equal(key[nbr[w], 0], key[ref[w], 0])


def wall_fill(m, fv, nut):
    orc = O.Oracle(m)
    fv, nut = fv.copy(), nut.copy()
    assert orc.sa_bc_fill(fv, nut, 0) == 0
    return fv, nut


def ghosts(m):
    v = np.ctypeslib.as_array
    nd = m.soa.ndmg
    ty = v(m.soa.dmg_type, (nd,))
    w = ty == -81
    return (v(m.soa.dmg_cell, (nd,))[w], v(m.soa.dmg_ref, (nd,))[w], v(m.soa.dmg_nbr, (nd,))[w])


def test_fluid_at_rest_gives_a_wall_at_rest(walled):
    _, m, _ = walled
    fv = m.init_state()
    fv[:, 1:4] = 0.0
    fv[:, 6] = m.phys[0].Cv * fv[:, 5]
    nut = np.full(m.ncells, 3.0)
    out, _ = wall_fill(m, fv, nut)
    g, r, a = ghosts(m)
    assert np.all(out[g, 1:4] == 0.0)
    assert np.array_equal(out[g, 4], fv[a, 4]) and np.array_equal(out[g, 5], fv[a, 5])


def test_viscous_sublayer_gives_the_linear_profile(walled):
    """A slow tangential stream (y+ of the reference neighbour in the viscous
    sublayer, u+ = y+) makes the wall function the linear profile: ghost
    u_t = u_t(neighbour) * d(reference cell) / d(neighbour), to the 0.2 %
    offset of the y+ fit, normal velocity scaled the same way; the SA
    variable is kappa y+."""
    _, m, _ = walled
    fv = m.init_state()
    fv[:, 1] = 0.002      # tangential to the wall at y = lo
    fv[:, 2] = 0.0005     # normal
    fv[:, 6] = m.phys[0].Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    nut = np.full(m.ncells, 3.0)
    out, onut = wall_fill(m, fv, nut)
    g, r, a = ghosts(m)
    dist = np.ctypeslib.as_array(m.soa.dist_dm, (m.ncells,))
    ratio = dist[r] / dist[a]
    assert np.allclose(out[g, 1], fv[a, 1] * ratio, rtol=3e-3)
    assert np.allclose(out[g, 2], fv[a, 2] * ratio, rtol=1e-14)
    assert np.all(onut[g] > 0.0) and np.all(onut[g] < 0.41 * 11.0)


def test_far_neighbour_copies_the_reference_cell(walled):
    """A reference neighbour farther than 1.0 from the wall: the ghost takes
    the reference cell's state (BCValues.H:105-107).  The distances of the
    walled mesh are raised by 1 in a copy of its flattened arrays."""
    import ctypes as C
    _, m, _ = walled
    soa = type(m.soa).from_buffer_copy(m.soa)
    dist = np.ctypeslib.as_array(m.soa.dist_dm, (m.ncells,)) + 1.0
    soa.dist_dm = dist.ctypes.data_as(O.f64p)

    class View:
        pass
    mv = View()
    mv.soa_ptr, mv.phys, mv.nlevels, mv.ncells = C.pointer(soa), m.phys, m.nlevels, m.ncells
    g, r, a = ghosts(m)
    fv = m.init_state()
    fv[:, 1] += 0.1 * np.arange(m.ncells) / m.ncells
    nut = np.linspace(1.0, 2.0, m.ncells)
    out, onut = wall_fill(mv, fv, nut)
    assert np.array_equal(out[g], fv[r]) and np.array_equal(onut[g], nut[r])
