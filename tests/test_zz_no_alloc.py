"""GPU test of ibg_closure_infer_noalloc, the closure inference entry point
that allocates no device memory (inputs and outputs staged through the
library's fixed per-device array): through the C ABI, two calls in a row,
the second with fewer rows, each equal bit for bit to ibg_closure_infer
(the same kernel on the same rows) and within the BF16 bar of the CPU
mirror; more rows than the array holds are refused before any device work.
"""
import ctypes as C

import numpy as np
import pytest

import ibgpu
from ibgpu import _ffi
import oracle_ffi as O

pytestmark = pytest.mark.gpu


def _params(seed=11, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, O.NPARAM)
    stats = np.concatenate([rng.normal(size=8), rng.uniform(0.5, 2.0, 8), -1e30 * np.ones(8),
                            1e30 * np.ones(8), [0.2, 1.7]])
    return np.concatenate([p, stats]).astype(np.float32)


def test_noalloc_inference_equals_allocating_entry():
    mesh = ibgpu.Mesh(ibgpu.tiny_case())
    e = ibgpu.Engine(mesh, mode="mixed", nstage=4, turbulence=True)
    try:
        e.closure_init(1 << 12)
        prm = _params()
        e.closure_set_params(prm)
        rng = np.random.default_rng(5)
        for n in (5000, 777):      # the second call smaller than the first
            x = rng.normal(size=(n, 8)).astype(np.float32)
            print(f"no-alloc inference: {n} rows", flush=True)
            z_ref = e.closure_infer(x)
            z = np.full(n, np.nan, dtype=np.float32)
            st = e.lib.ibg_closure_infer_noalloc(e.ctx, x.ctypes.data_as(_ffi.f32p),
                                                 z.ctypes.data_as(_ffi.f32p), C.c_int64(n))
            assert st == 0, st
            assert np.array_equal(z, z_ref), np.abs(z - z_ref).max()
            c = O.mlp_forward(prm, x).astype(np.float64)
            assert np.linalg.norm(z - c) / np.linalg.norm(c) < 1e-3
        # beyond the staging array: refused, output untouched
        big = _ffi.IBG_CLOSURE_NOALLOC_MAX_ROWS + 1
        x = np.zeros((big, 8), dtype=np.float32)
        z = np.full(big, 7.0, dtype=np.float32)
        st = e.lib.ibg_closure_infer_noalloc(e.ctx, x.ctypes.data_as(_ffi.f32p),
                                             z.ctypes.data_as(_ffi.f32p), C.c_int64(big))
        assert st == 1 and np.all(z == 7.0)
    finally:
        e.close()
