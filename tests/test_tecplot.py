"""Tecplot output and restart (Pltoutput::Output_Grid / Output_Soln,
Pltoutput.cpp:3-100, 417-450; InputSoln input_soln.cpp:12-180), host side.

Pins: the files carry the reference's header, variable list and zone
naming; every level is one FEBRICK zone with one element per cell and 8
corner nodes each; a solution written and read back gives the state, the
SA variable, ts, dt and t bit for bit; a file of another level set is
refused (the reference's abort 65)."""
import dataclasses
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from ibgpu import mesh as M  # noqa: E402
from test_regrid import smooth_state  # noqa: E402


@pytest.fixture(scope="module")
def tiny():
    return M.Mesh(M.tiny_case())


def test_grid_file(tiny, tmp_path):
    path = tmp_path / "amr_grid-init.plt"
    tiny.write_grid(path)
    text = path.read_text().splitlines()
    assert text[0] == 'TITLE = "Boxgrid"' and text[1] == "FILETYPE = GRID"
    zones = [l for l in text if l.startswith("zone")]
    assert len(zones) == tiny.ncases * tiny.nlevels
    for b, z in enumerate(zones):
        ne = int(z.split("ELEMENTS=")[1].split(",")[0])
        assert f'"N-0 L-{b}"' in z and "ZONETYPE=FEBRICK" in z
        assert ne == tiny.blk_start[b + 1] - tiny.blk_start[b]


def test_solution_round_trip(tiny, tmp_path):
    fv = smooth_state(tiny)
    nut = 1.0 + np.sin(tiny.centers[:, 0]) ** 2
    dt, t = np.array([1.2345678901234567e-3]), np.array([0.318])
    path = tmp_path / "amr_soln-40.plt"
    tiny.write_solution(path, fv, nut, 40, dt, t)
    head = path.read_text().splitlines()[:3]
    assert head[1] == "FILETYPE = SOLUTION" and '"viseddy"' in head[2] and '"Omegaz"' in head[2]
    fv2, nut2, ts, dt2, t2 = tiny.read_solution(path, with_nut=True)
    assert np.array_equal(fv2[:, :6], fv[:, :6]) and np.array_equal(nut2, nut)
    P = tiny.phys[0]
    assert np.array_equal(fv2[:, 6], P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2 + fv[:, 3] ** 2))
    assert ts == 40 and np.array_equal(dt2, dt) and np.array_equal(t2, t)
    # a laminar file (no viseddy) reads without the SA variable
    tiny.write_solution(path, fv)
    fv3, _, _, _, _ = tiny.read_solution(path)
    assert np.array_equal(fv3[:, :6], fv[:, :6])
    with pytest.raises(RuntimeError):
        tiny.read_solution(path, with_nut=True)


def test_other_level_set_is_refused(tiny, tmp_path):
    other = M.Mesh(dataclasses.replace(M.tiny_case(), aoa=M.tiny_case().aoa + 3.0))
    path = tmp_path / "amr_soln-0.plt"
    other.write_solution(path, other.init_state())
    with pytest.raises(RuntimeError, match="element count"):
        tiny.read_solution(path)


def test_repeated_zone_is_refused(tmp_path):
    """a file that repeats one (case, level) zone and leaves out another is
    refused, even when the element counts match"""
    two = M.Mesh([M.tiny_case(), M.tiny_case()])
    path = tmp_path / "amr_soln-0.plt"
    two.write_solution(path, two.init_state())
    text = path.read_text()
    assert text.count('"N-1 L-0"') == 1
    bad = tmp_path / "amr_soln-bad.plt"
    bad.write_text(text.replace('"N-1 L-0"', '"N-0 L-0"'))
    with pytest.raises(RuntimeError, match="repeated"):
        two.read_solution(bad)
    two.read_solution(path)
