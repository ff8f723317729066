"""Pins of the moving-body oracle (oracle/retag.c, Bodymove.cpp:37-767).

The oracle follows the reference's incremental update: a local walk from
each cell's previous closest patch (FindClosePatchtoCell Body.H:505-542),
the cross-body cycle of RenewLevelDistance, and RemoveIBCells /
InfectNewIBCells.  The host mesh builder computes the same quantities from
scratch with a different method (a scan over every segment, the initial IB
rule of Body.cpp:1016-1082, a sorted key table).  Agreement of the two on
every cell both meshes hold, after several moves, pins the oracle.
"""
import dataclasses
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle_ffi as O  # noqa: E402
from ibgpu import mesh as M  # noqa: E402


def keys_of(m):
    lv = m.level_of()
    case = np.zeros(m.ncells, np.int64)
    for b in range(m.ncases * m.nlevels):
        case[m.blk_start[b]:m.blk_start[b + 1]] = b // m.nlevels
    k = np.asarray(m.key)
    return [(int(case[i]), int(lv[i]), int(k[i, 0]), int(k[i, 1]), int(k[i, 2]))
            for i in range(m.ncells)]


@pytest.fixture(scope="module")
def tiny():
    spec = M.tiny_case(levels=3)
    return spec, M.Mesh(spec)


def test_zero_move_reproduces_the_builder(tiny):
    spec, m = tiny
    mb = O.MovingBodies(m)
    r = mb.move([M.Mesh.pose(spec.aoa, 0.0)])
    assert r["code"] == 0
    v = np.ctypeslib.as_array
    assert np.array_equal(mb.signdis, np.asarray(m.signdis))
    assert np.array_equal(r["cell"], np.asarray(m.ib_cell))
    assert np.array_equal(r["intp"], np.asarray(m.ib_intp))
    assert np.array_equal(r["coef"], np.asarray(m.ib_coef))
    assert np.array_equal(r["hgdis"], np.asarray(m.ib_hgdis))
    assert np.array_equal(mb.pat_pc, v(m.soa.pat_pc, (m.npatch, 3)))
    assert np.array_equal(mb.pat_nv, v(m.soa.pat_nv, (m.npatch, 3)))
    assert np.array_equal(r["pat_intp"], np.asarray(m.pat_intp))
    assert np.array_equal(r["pat_coef"], np.asarray(m.pat_coef))


@pytest.mark.parametrize("daoa,dheave", [(0.5, 0.0), (0.0, 0.01), (1.0, 0.02), (-2.0, -0.03)])
def test_moves_match_a_fresh_build_at_the_new_pose(tiny, daoa, dheave):
    spec, m = tiny
    mb = O.MovingBodies(m)
    steps = 4
    for k in range(1, steps + 1):
        r = mb.move([M.Mesh.pose(spec.aoa + daoa * k / steps, dheave * k / steps)])
        assert r["code"] == 0, (r["code"], r["err_cell"])
    m2 = M.Mesh(dataclasses.replace(spec, aoa=spec.aoa + daoa, pivot_y=dheave))
    k1, k2 = keys_of(m), keys_of(m2)
    i2 = {k: i for i, k in enumerate(k2)}
    v = np.ctypeslib.as_array
    # surface: bit-equal patches
    assert np.array_equal(mb.pat_pc, v(m2.soa.pat_pc, (m2.npatch, 3)))
    assert np.array_equal(mb.pat_nv, v(m2.soa.pat_nv, (m2.npatch, 3)))
    assert np.array_equal(mb.pat_area, v(m2.soa.pat_area, (m2.npatch,)))
    # signed distance and solid flag of every common cell
    sd2 = np.asarray(m2.signdis)
    common = [(i, i2[k]) for i, k in enumerate(k1) if k in i2]
    assert len(common) > 0.95 * m.ncells
    a = np.array([mb.signdis[i] for i, _ in common])
    b = np.array([sd2[j] for _, j in common])
    assert np.array_equal(a, b)
    # IB set and image points by key
    s1 = {k1[c]: j for j, c in enumerate(r["cell"].tolist())}
    s2 = {k2[c]: j for j, c in enumerate(np.asarray(m2.ib_cell).tolist())}
    assert set(s1) == set(s2)

    def as_keys(row, keys):
        return [keys[a] if a >= 0 else None for a in row]

    for k, j in s1.items():
        jj = s2[k]
        assert as_keys(r["intp"][j], k1) == as_keys(np.asarray(m2.ib_intp)[jj], k2)
        assert np.array_equal(r["coef"][j], np.asarray(m2.ib_coef)[jj])
        assert r["hgdis"][j] == np.asarray(m2.ib_hgdis)[jj]
    for p in range(m.npatch):
        assert as_keys(r["pat_intp"][p], k1) == as_keys(np.asarray(m2.pat_intp)[p], k2)
    assert np.array_equal(r["pat_coef"], np.asarray(m2.pat_coef))


def test_weights_are_a_partition_of_unity_after_moves(tiny):
    spec, m = tiny
    mb = O.MovingBodies(m)
    for k in range(1, 6):
        r = mb.move([M.Mesh.pose(spec.aoa + 0.4 * k, 0.004 * k)])
        assert r["code"] == 0
        assert np.allclose(r["coef"].sum(axis=1), 1.0, atol=1e-13)
        assert np.allclose(r["pat_coef"].sum(axis=1), 1.0, atol=1e-13)
        # image cells are neither solid nor IB cells
        used = r["intp"][r["coef"] != 0.0]
        assert np.all(mb.signdis[used] >= 0.0) and not np.any(mb.ib[used])
        # every IB cell lies on the fluid side
        assert np.all(mb.signdis[r["cell"]] >= 0.0)


def test_too_large_a_move_is_the_reference_abort(tiny):
    spec, m = tiny
    mb = O.MovingBodies(m)
    r = mb.move([M.Mesh.pose(spec.aoa + 3.0, 0.0)])
    assert r["code"] in (263, 314)
