"""Batched cases and the tile autotuner on the GPU.

  ensemble   cases of uneven AMR size advanced together in one context (one
             launch per stage over all cases, graph-captured) give, case by
             case, the same bits as each case run alone
  tiles      the tile size changes the launch shape only: every tile size
             gives the same bits, and the autotuner returns one of the
             candidates with a timing for each
"""
import numpy as np
import pytest

import ibgpu
from test_gpu_parity import perturbed_state

pytestmark = pytest.mark.gpu


def specs():
    # uneven sizes (3.5k to 5.5k cells), one level count for all cases
    return [ibgpu.tiny_case(levels=4, aoa=4.0, Ma=0.5, thickness=0.12),
            ibgpu.tiny_case(levels=4, aoa=-2.0, Ma=0.35, thickness=0.16),
            ibgpu.tiny_case(levels=4, aoa=7.0, Ma=0.6, thickness=0.10, base=24)]


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_ensemble_equals_single_runs(mode):
    sp = specs()
    L = sp[0].nlevels
    batch = ibgpu.Mesh(sp)
    fvb = perturbed_state(batch)
    eb = ibgpu.Engine(batch, mode=mode, nstage=4, turbulence=True)
    eb.set_state(fvb)
    eb.set_nut(np.full(batch.ncells, batch.phys[0].nut_inf))
    P = 1 << (L - 1)
    eb.TimeMarching(0, 2 * P, use_graph=True)
    gb, nb_ = eb.get_state(), eb.get_nut()
    for c, s in enumerate(sp):
        single = ibgpu.Mesh([s])
        lo, hi = batch.blk_start[c * L], batch.blk_start[(c + 1) * L]
        assert hi - lo == single.ncells
        es = ibgpu.Engine(single, mode=mode, nstage=4, turbulence=True)
        es.set_state(fvb[lo:hi])
        es.set_nut(np.full(single.ncells, single.phys[0].nut_inf))
        es.TimeMarching(0, 2 * P, use_graph=True)
        assert np.array_equal(es.get_state(), gb[lo:hi]), c
        assert np.array_equal(es.get_nut(), nb_[lo:hi]), c


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_tile_size_and_schedule_do_not_change_bits(mode):
    """Tile size and schedule (one workgroup per tile, or persistent
    workgroups looping over the tiles) change the launch shape only."""
    mesh = ibgpu.Mesh(ibgpu.tiny_case())
    fv = perturbed_state(mesh)
    out = []
    for tile in (64, 128, 256):
        for schedule in (0, 1):
            print(f"tile {tile}, schedule {schedule}, {mode}", flush=True)
            e = ibgpu.Engine(mesh, mode=mode, nstage=3, tile=tile, turbulence=True, schedule=schedule)
            e.set_state(fv)
            e.set_nut(np.full(mesh.ncells, 5.0))
            e.TimeMarching(0, 8, use_graph=(schedule == 1))
            out.append((e.get_state(), e.get_nut(), e.ResinDomain()))
            e.close()
    for a, b, (linf, l2) in out[1:]:
        assert np.array_equal(a, out[0][0])
        assert np.array_equal(b, out[0][1])
        assert np.array_equal(linf, out[0][2][0])
        assert np.allclose(l2, out[0][2][1], rtol=1e-13, atol=0)


def test_autotuner_picks_a_candidate():
    mesh = ibgpu.Mesh(ibgpu.config1())
    e = ibgpu.Engine(mesh, mode="mixed", nstage=4, turbulence=True)
    best, table = e.autotune_tile((64, 128, 256), reps=5)
    assert best in (64, 128, 256)
    assert set(table) == {(t, s) for t in (64, 128, 256) for s in (0, 1)}
    assert all(v > 0 for v in table.values())
    assert table[(best, e.schedule)] == min(table.values())
    e.set_state(mesh.init_state())
    e.set_nut(np.full(mesh.ncells, 5.0))
    e.TimeMarching(0, 8, use_graph=True)
    assert np.all(np.isfinite(e.get_state()))
