"""The C ABI: libibgpu.so loads on a machine without a GPU and exports every
entry point include/ibgpu.h declares; the host and oracle libraries load;
calls that need a device fail with a status instead of crashing."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from ibgpu import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ibg_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared("ibgpu.h")
    assert len(names) >= 25
    lib = _ffi.gpu_lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(_ffi.LIB_DIR, "libibgpu.so")],
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ibg_[a-z_0-9]+)", out))
    assert set(names) <= exported


def test_ctypes_table_matches_header():
    assert set(_ffi.GPU_FUNCS) == set(declared("ibgpu.h"))


def test_create_without_device_returns_status():
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    lib = _ffi.gpu_lib()
    cfg = _ffi.Config()
    cfg.mode, cfg.nstage = 0, 3
    ctx = C.c_void_p()
    st = lib.ibg_create(C.byref(cfg), C.byref(ctx))
    assert st == 2 and not ctx.value


def test_bad_arguments_rejected():
    lib = _ffi.gpu_lib()
    cfg = _ffi.Config()
    cfg.mode, cfg.nstage = 0, 2          # only 1, 3 or 4 stages exist
    ctx = C.c_void_p()
    assert lib.ibg_create(C.byref(cfg), C.byref(ctx)) == 1
    # the coarse/fine balance exists only without temporal refinement
    cfg.nstage, cfg.reflux, cfg.temporal_refine = 4, 1, 1
    assert lib.ibg_create(C.byref(cfg), C.byref(ctx)) == 1
    cfg.reflux, cfg.temporal_refine = 2, 0
    assert lib.ibg_create(C.byref(cfg), C.byref(ctx)) == 1
    assert lib.ibg_destroy(None) == 1
    assert lib.ibg_check_mesh(None, None) == 1


# ---------------------------------------------------------- mesh validation
# ibg_upload_mesh runs ibg_check_mesh before any launch, so a malformed
# mesh is refused on the host; each corruption below must be caught with
# its message, and the clean mesh must pass.
@pytest.fixture(scope="module")
def tiny_mesh():
    import ibgpu
    return ibgpu.Mesh(ibgpu.tiny_case())


def check(lib, mesh):
    msg = C.c_char_p()
    st = lib.ibg_check_mesh(mesh.soa_ptr, C.byref(msg))
    return st, (msg.value.decode() if msg.value else "")


def test_clean_mesh_passes(tiny_mesh):
    lib = _ffi.gpu_lib()
    assert check(lib, tiny_mesh)[0] == 0


@pytest.mark.parametrize("what,expect", [
    ("face_cells", "face"),
    ("face_nb", "face list and neighbour table disagree"),
    # (both outer stencil cells of a face missing: every face has a normal
    # cell, whose neighbours the agreement check, run first, covers)
    ("nb_face_stencil", "face list and neighbour table disagree"),
    ("res_fi", "restriction son missing"),
    ("pro_ci", "prolongation coarse cell"),
    ("ib_intp", "IB interpolation cell"),
    ("ib_into_ib", "image point interpolates from an IB cell"),
    ("pat_intp", "patch interpolation cell missing"),
    ("mdy", "balance fine face out of range"),
    ("blk_start", "blk_start"),
])
def test_corrupted_mesh_rejected(tiny_mesh, what, expect):
    lib = _ffi.gpu_lib()
    m = tiny_mesh
    v = np.ctypeslib.as_array
    if what == "face_cells":
        arr, idx, bad = m.face_cells, (3, 1), m.ncells + 7
    elif what == "face_nb":
        # the -x neighbour of a normal cell replaced by its +x neighbour (a
        # valid cell, so only the agreement with the face list fails)
        g = 0
        arr, idx, bad = m.nb, (g, 0 * 9 + 1 * 3 + 1), int(m.nb[g, 2 * 9 + 1 * 3 + 1])
    elif what == "nb_face_stencil":
        # both the LL neighbour of an x face's low cell and the RR neighbour
        # of its high cell missing (one missing is the reference's fallback)
        f = int(np.nonzero(m.face_dir == 0)[0][0])
        lo, hi = int(m.face_cells[f, 0]), int(m.face_cells[f, 1])
        keep = m.nb[hi, 2 * 9 + 1 * 3 + 1]
        m.nb[hi, 2 * 9 + 1 * 3 + 1] = -1
        arr, idx, bad = m.nb, (lo, 0 * 9 + 1 * 3 + 1), -1
    elif what == "res_fi":
        arr, idx, bad = m.res_fi, (0, 0), -1
    elif what == "pro_ci":
        arr, idx, bad = m.pro_ci, (0,), m.ncells
    elif what == "ib_intp":
        arr, idx, bad = m.ib_intp, (0, 1), -3
    elif what == "ib_into_ib":
        arr, idx, bad = m.ib_intp, (0, 1), int(m.ib_cell[1])
    elif what == "pat_intp":
        arr, idx, bad = m.pat_intp, (0, 1), m.ncells + 1
    elif what == "mdy":
        # a two-face (x,y-refined) balance row with a third fine face
        arr, idx, bad = m.mdy, (0, 4), 5
    else:
        arr = v(m.soa.blk_start, (m.ncases * m.nlevels + 1,))
        idx, bad = (1,), -5
    old = arr[idx]
    try:
        arr[idx] = bad
        st, msg = check(lib, m)
        assert st == 1 and expect in msg, (st, msg)
    finally:
        arr[idx] = old
        if what == "nb_face_stencil":
            m.nb[hi, 2 * 9 + 1 * 3 + 1] = keep
    assert check(lib, m)[0] == 0
