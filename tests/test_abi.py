"""The C ABI: libibgpu.so loads on a machine without a GPU and exports every
entry point include/ibgpu.h declares; the host and oracle libraries load;
calls that need a device fail with a status instead of crashing."""
import ctypes as C
import os
import re
import subprocess

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
    assert lib.ibg_destroy(None) == 1
    assert lib.ibg_check_mesh(None, None) == 1
