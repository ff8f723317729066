"""ctypes binding of oracle/_build/liboracle.so (the FP64 CPU restatement).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "immsered-boundary_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from ibgpu import _ffi  # noqa: E402

f64p, i32p, f32p = _ffi.f64p, _ffi.i32p, _ffi.f32p


class OrcCfg(C.Structure):
    _fields_ = [("nstage", C.c_int32), ("alpha", C.c_double * 4),
                ("dt_var_step", C.c_int32), ("temporal_refine", C.c_int32),
                ("limiter", C.c_int32), ("viscous", C.c_int32), ("reflux", C.c_int32)]


class OrcSurface(C.Structure):
    _fields_ = [("npoint", C.c_int32), ("npatch", C.c_int32), ("nbody", C.c_int32),
                ("pt_off", f64p), ("pat_pt0", i32p), ("pat_body", i32p), ("pat_case", i32p),
                ("pat_flip", C.POINTER(C.c_int8)), ("pat_dz", f64p), ("body_p0", i32p),
                ("case_origin", f64p)]


class OrcRetagOut(C.Structure):
    _fields_ = [("signdis", f64p), ("cell_patch", i32p), ("ib", C.POINTER(C.c_int8)),
                ("pat_pc", f64p), ("pat_nv", f64p), ("pat_area", f64p),
                ("ib_cap", C.c_int32), ("nib", C.c_int32),
                ("ib_cell", i32p), ("ib_intp", i32p), ("ib_body", i32p),
                ("ib_coef", f64p), ("ib_hgdis", f64p), ("ib_nv", f64p), ("ib_attach", f64p),
                ("pat_intp", i32p), ("pat_close", i32p), ("pat_coef", f64p), ("pat_hgdis", f64p),
                ("err_code", C.c_int32), ("err_cell", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(path):
        raise OSError(f"{path} missing: run `python __graft_entry__.py build`")
    L = C.CDLL(path)
    M = C.POINTER(_ffi.MeshSoA)
    P = C.POINTER(_ffi.Physics)
    G = C.POINTER(OrcCfg)
    BM = C.POINTER(_ffi.BodyMotion)
    sig = {
        "orc_marched": (C.c_int32, [C.c_int32] * 4),
        "orc_transfer_due": (C.c_int32, [C.c_int32] * 4),
        "orc_step_num": (C.c_double, [C.c_int32] * 3),
        "orc_muscl_left": (C.c_double, [C.c_double] * 3 + [C.c_int32]),
        "orc_muscl_right": (C.c_double, [C.c_double] * 3 + [C.c_int32]),
        "orc_roe_flux": (None, [f64p, f64p, f64p, f64p, C.c_int32, C.c_double, P, C.c_int32, f64p]),
        "orc_face_flux": (None, [M, P, G, f64p, C.c_int32, f64p]),
        "orc_viscous_flux": (None, [M, P, G, f64p, C.c_int32, f64p]),
        "orc_source": (None, [M, G, f64p, f64p, C.c_int32, f64p]),
        "orc_old_cons": (None, [M, G, f64p, C.c_int32, f64p]),
        "orc_update": (C.c_int32, [M, P, G, f64p, f64p, f64p, C.c_int32, C.c_int32, f64p, f64p,
                                   C.POINTER(C.c_int64)]),
        "orc_bc_fill": (C.c_int32, [M, P, G, f64p, C.c_int32]),
        "orc_interface_exchange": (None, [M, P, G, f64p, C.c_int32]),
        "orc_ib_fill": (C.c_int32, [M, P, f64p, BM]),
        "orc_compute_dt": (C.c_int32, [M, P, f64p, f64p]),
        "orc_residual": (None, [M, G, f64p, C.c_int32, f64p, f64p]),
        "orc_forces": (None, [M, P, f64p, BM, f64p]),
        "orc_step": (C.c_int32, [M, P, G, f64p, f64p, C.c_int32, f64p, f64p, f64p,
                                 C.POINTER(C.c_int64)]),
        "orc_sa_mut": (C.c_double, [C.c_double] * 3),
        "orc_sa_yplus": (C.c_double, [C.c_double]),
        "orc_sa_uplus": (C.c_double, [C.c_double]),
        "orc_sa_cell_source": (C.c_double, [C.c_double] * 7 + [f64p] * 4 + [C.c_double] * 6),
        "orc_viscous_flux_sa": (None, [M, P, G, f64p, f64p, C.c_int32, f64p, f64p]),
        "orc_mut_source": (None, [M, P, G, f64p, f64p, f64p, C.c_int32, f64p]),
        "orc_sa_update": (None, [M, P, G, f64p, f64p, f64p, C.c_int32, C.c_int32, f64p]),
        "orc_sa_bc_fill": (C.c_int32, [M, P, G, f64p, f64p, C.c_int32]),
        "orc_sa_exchange": (None, [M, G, f64p, C.c_int32]),
        "orc_ib_fill_sa": (C.c_int32, [M, P, f64p, f64p, BM]),
        "orc_forces_sa": (None, [M, P, f64p, BM, f64p]),
        "orc_residual_nut": (None, [M, G, f64p, C.c_int32, f64p, f64p]),
        "orc_step_sa": (C.c_int32, [M, P, G, f64p, f64p, f64p, C.c_int32, f64p, f64p, f64p,
                                    f64p, f64p, f64p, C.POINTER(C.c_int64)]),
        "orc_balance_flux": (None, [M, f64p]),
        "orc_bf16": (C.c_float, [C.c_float]),
        "orc_mlp_forward": (None, [f32p, f32p, C.c_int64, f32p]),
        "orc_mlp_grad": (None, [f32p, f32p, f32p, C.c_int64, f64p, f64p]),
        "orc_mlp_loss_f64": (C.c_double, [f64p, f64p, f64p, C.c_int64]),
        "orc_mlp_grad_f64": (None, [f64p, f64p, f64p, C.c_int64, f64p, f64p]),
        "orc_closure_features": (C.c_int32, [M, P, f64p, f64p, C.c_int32, f64p, f64p]),
        "orc_closure_cells": (None, [M, P, f64p, f64p, C.c_int64, i32p, f64p, f64p, f64p,
                                     C.POINTER(C.c_int8)]),
        "orc_move_bodies": (C.c_int32, [M, C.POINTER(OrcSurface), f64p, C.POINTER(OrcRetagOut)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ptr(a):
    return a.ctypes.data_as(f64p)


def fptr(a):
    return a.ctypes.data_as(f32p)


NPARAM = 4801


def mlp_forward(prm, x):
    prm = np.ascontiguousarray(prm, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    z = np.zeros(x.shape[0], dtype=np.float32)
    lib().orc_mlp_forward(fptr(prm), fptr(x), x.shape[0], fptr(z))
    return z


def mlp_grad(prm, x, z):
    prm = np.ascontiguousarray(prm, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    g = np.zeros(NPARAM)
    loss = C.c_double()
    lib().orc_mlp_grad(fptr(prm), fptr(x), fptr(z), x.shape[0], ptr(g), C.byref(loss))
    return g, loss.value


def mlp_grad_f64(p, xs, t):
    p = np.ascontiguousarray(p, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.zeros(NPARAM)
    loss = C.c_double()
    lib().orc_mlp_grad_f64(ptr(p), ptr(xs), ptr(t), xs.shape[0], ptr(g), C.byref(loss))
    return g, loss.value


def mlp_loss_f64(p, xs, t):
    p = np.ascontiguousarray(p, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    return lib().orc_mlp_loss_f64(ptr(p), ptr(xs), ptr(t), xs.shape[0])


def closure_features(mesh, fv, nut, cell):
    x = np.zeros(8)
    z = C.c_double()
    ok = lib().orc_closure_features(mesh.soa_ptr, mesh.phys, ptr(fv), ptr(nut), int(cell), ptr(x),
                                    C.byref(z))
    return bool(ok), x, z.value


def closure_cells(mesh, fv, nut, cells):
    """Features, plain-closure target and production scale of many cells
    (orc_closure_cells): x [n][8], z [n], scale [n], ok [n] (bool)."""
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    fv = np.ascontiguousarray(fv, dtype=np.float64)
    nut = np.ascontiguousarray(nut, dtype=np.float64)
    n = cells.shape[0]
    x, z, sc = np.zeros((n, 8)), np.zeros(n), np.zeros(n)
    ok = np.zeros(n, np.int8)
    lib().orc_closure_cells(mesh.soa_ptr, mesh.phys, ptr(fv), ptr(nut), n, cells.ctypes.data_as(i32p),
                            ptr(x), ptr(z), ptr(sc), ok.ctypes.data_as(C.POINTER(C.c_int8)))
    return x, z, sc, ok.astype(bool)


class MovingBodies:
    """CPU restatement of the per-step moving-body update on a Mesh: holds
    the geometry of the last move (signed distances, closest patches, IB
    set, patches) and advances it with orc_move_bodies."""

    def __init__(self, mesh, cap=None):
        m = mesh
        self.m = m
        _, a = m.surface()
        self.a = a
        self.pat_case = np.ctypeslib.as_array(m.soa.pat_case, (m.npatch,)).astype(np.int32)
        s = OrcSurface()
        s.npoint = a["pt_off"].shape[0]
        s.npatch = m.npatch
        s.nbody = len(a["body_p0"]) - 1
        s.pt_off = ptr(a["pt_off"])
        s.pat_pt0 = a["pat_pt0"].ctypes.data_as(i32p)
        s.pat_body = a["pat_body"].ctypes.data_as(i32p)
        s.pat_case = self.pat_case.ctypes.data_as(i32p)
        s.pat_flip = a["pat_flip"].ctypes.data_as(C.POINTER(C.c_int8))
        s.pat_dz = ptr(a["pat_dz"])
        s.body_p0 = a["body_p0"].ctypes.data_as(i32p)
        s.case_origin = ptr(a["case_origin"])
        self.s = s
        n = m.ncells
        self.signdis = np.array(m.signdis, dtype=np.float64)
        self.cell_patch = a["cell_patch"].copy()
        self.ib = (np.asarray(m.ibidx) >= 0).astype(np.int8)
        v = np.ctypeslib.as_array
        self.pat_pc = v(m.soa.pat_pc, (m.npatch, 3)).copy()
        self.pat_nv = v(m.soa.pat_nv, (m.npatch, 3)).copy()
        self.pat_area = v(m.soa.pat_area, (m.npatch,)).copy()
        self.cap = cap or (2 * m.nib + 4096)

    def move(self, poses):
        m = self.m
        cap, np_ = self.cap, m.npatch
        r = {"cell": np.zeros(cap, np.int32), "intp": np.zeros((cap, 8), np.int32),
             "body": np.zeros(cap, np.int32), "coef": np.zeros((cap, 8)), "hgdis": np.zeros(cap),
             "nv": np.zeros((cap, 3)), "attach": np.zeros((cap, 3)),
             "pat_intp": np.zeros((np_, 8), np.int32), "pat_close": np.zeros(np_, np.int32),
             "pat_coef": np.zeros((np_, 8)), "pat_hgdis": np.zeros(np_)}
        o = OrcRetagOut()
        o.signdis = ptr(self.signdis)
        o.cell_patch = self.cell_patch.ctypes.data_as(i32p)
        o.ib = self.ib.ctypes.data_as(C.POINTER(C.c_int8))
        o.pat_pc, o.pat_nv, o.pat_area = ptr(self.pat_pc), ptr(self.pat_nv), ptr(self.pat_area)
        o.ib_cap = cap
        o.ib_cell = r["cell"].ctypes.data_as(i32p)
        o.ib_intp = r["intp"].ctypes.data_as(i32p)
        o.ib_body = r["body"].ctypes.data_as(i32p)
        o.ib_coef, o.ib_hgdis = ptr(r["coef"]), ptr(r["hgdis"])
        o.ib_nv, o.ib_attach = ptr(r["nv"]), ptr(r["attach"])
        o.pat_intp = r["pat_intp"].ctypes.data_as(i32p)
        o.pat_close = r["pat_close"].ctypes.data_as(i32p)
        o.pat_coef, o.pat_hgdis = ptr(r["pat_coef"]), ptr(r["pat_hgdis"])
        pose = np.ascontiguousarray(np.asarray(poses, dtype=np.float64).reshape(m.ncases, 4))
        code = lib().orc_move_bodies(m.soa_ptr, C.byref(self.s), ptr(pose), C.byref(o))
        n = o.nib
        for k in ("cell", "intp", "body", "coef", "hgdis", "nv", "attach"):
            r[k] = r[k][:n]
        r["code"], r["err_cell"], r["nib"] = int(code), int(o.err_cell), n
        return r


class MovedMesh:
    """A Mesh whose immersed-boundary data are those of a move (the last
    MovingBodies.move result): for oracle calls after the bodies moved."""

    def __init__(self, mb, r):
        m = mb.m
        self._m = m
        soa = type(m.soa).from_buffer_copy(m.soa)
        n = r["nib"]
        ibidx = np.full(m.ncells, -1, np.int32)
        ibidx[r["cell"]] = np.arange(n, dtype=np.int32)
        keep = {
            "signdis": mb.signdis.copy(), "ibidx": ibidx,
            "ib_cell": np.ascontiguousarray(r["cell"]), "ib_intp": np.ascontiguousarray(r["intp"]),
            "ib_coef": np.ascontiguousarray(r["coef"]), "ib_hgdis": np.ascontiguousarray(r["hgdis"]),
            "ib_nv": np.ascontiguousarray(r["nv"]), "ib_attach": np.ascontiguousarray(r["attach"]),
            "ib_body": np.ascontiguousarray(r["body"]),
            "pat_intp": r["pat_intp"].copy(), "pat_coef": r["pat_coef"].copy(),
            "pat_hgdis": r["pat_hgdis"].copy(), "pat_nv": mb.pat_nv.copy(),
            "pat_pc": mb.pat_pc.copy(), "pat_area": mb.pat_area.copy(),
        }
        for k, a in keep.items():
            t = i32p if a.dtype == np.int32 else f64p
            setattr(soa, k, a.ctypes.data_as(t))
        soa.nib = n
        self._keep = keep
        self.soa = soa
        self.soa_ptr = C.pointer(soa)

    def __getattr__(self, k):
        return getattr(self._m, k)


class PatchedMesh:
    """A Mesh with some SoA arrays (and counts) replaced, for pins that plant
    cell classes (solid, IB, zero distance) or a synthetic wall surface.
    Arrays are passed by field name; `counts` sets integer fields."""

    def __init__(self, mesh, counts=None, **arrays):
        self._m = mesh
        soa = type(mesh.soa).from_buffer_copy(mesh.soa)
        keep = {}
        for k, a in arrays.items():
            if a.dtype == np.int32:
                a = np.ascontiguousarray(a, dtype=np.int32)
                setattr(soa, k, a.ctypes.data_as(i32p))
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                setattr(soa, k, a.ctypes.data_as(f64p))
            keep[k] = a
        for k, v in (counts or {}).items():
            setattr(soa, k, int(v))
        self._keep = keep
        self.soa = soa
        self.soa_ptr = C.pointer(soa)

    def __getattr__(self, k):
        return getattr(self._m, k)


def cfg(nstage=3, alpha=None, dt_var_step=1, temporal_refine=1, limiter=0, viscous=1, reflux=0):
    c = OrcCfg()
    c.nstage = nstage
    if alpha is None:
        alpha = {1: [1.0], 3: [0.1918, 0.4929, 1.0], 4: [0.1084, 0.2602, 0.5052, 1.0]}[nstage]
    for i, a in enumerate(alpha):
        c.alpha[i] = a
    c.dt_var_step = dt_var_step
    c.temporal_refine = temporal_refine
    c.limiter = limiter
    c.viscous = viscous
    c.reflux = reflux
    return c


class Oracle:
    """Whole-mesh oracle calls on a Mesh (ibgpu.mesh.Mesh)."""

    def __init__(self, mesh, config=None):
        self.m = mesh
        self.L = lib()
        # Comptdt period: marching_step[0] of the level schedule, as the engine
        # defaults to (mpiparams.cpp:208-210)
        self.cfg = config or cfg(dt_var_step=1 << (mesh.nlevels - 1))

    def step(self, fv, dt, ts):
        src = np.zeros((self.m.ncells, 5))
        linf = np.zeros(self.m.ncases * 5)
        l2 = np.zeros(self.m.ncases * 5)
        bad = C.c_int64(-1)
        code = self.L.orc_step(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv), ptr(dt),
                               ts, ptr(src), ptr(linf), ptr(l2), C.byref(bad))
        return code, src, linf, l2

    def step_sa(self, fv, nut, dt, ts):
        n = self.m.ncases
        src = np.zeros((self.m.ncells, 5))
        msrc = np.zeros(self.m.ncells)
        linf, l2 = np.zeros(n * 5), np.zeros(n * 5)
        nlinf, nl2 = np.zeros(n), np.zeros(n)
        bad = C.c_int64(-1)
        code = self.L.orc_step_sa(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv),
                                  ptr(nut), ptr(dt), ts, ptr(src), ptr(msrc), ptr(linf),
                                  ptr(l2), ptr(nlinf), ptr(nl2), C.byref(bad))
        return code, src, msrc, (linf, l2, nlinf, nl2)

    def viscous_flux_sa(self, fv, nut, ts):
        out = np.zeros((self.m.nfaces, 5))
        mf = np.zeros(self.m.nfaces)
        self.L.orc_viscous_flux_sa(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv),
                                   ptr(nut), ts, ptr(out), ptr(mf))
        return out, mf

    def mut_source(self, fv, nut, mutflux, ts):
        out = np.zeros(self.m.ncells)
        self.L.orc_mut_source(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv), ptr(nut),
                              ptr(mutflux), ts, ptr(out))
        return out

    def ib_fill_sa(self, fv, nut, motion=None):
        return self.L.orc_ib_fill_sa(self.m.soa_ptr, self.m.phys, ptr(fv), ptr(nut), motion)

    def sa_bc_fill(self, fv, nut, ts):
        return self.L.orc_sa_bc_fill(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv),
                                     ptr(nut), ts)

    def sa_exchange(self, nut, ts):
        self.L.orc_sa_exchange(self.m.soa_ptr, C.byref(self.cfg), ptr(nut), ts)

    def forces_sa(self, fv, motion=None):
        out = np.zeros(self.m.ncases * 6)
        self.L.orc_forces_sa(self.m.soa_ptr, self.m.phys, ptr(fv), motion, ptr(out))
        return out

    def face_flux(self, fv, ts):
        out = np.zeros((self.m.nfaces, 5))
        self.L.orc_face_flux(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv), ts, ptr(out))
        return out

    def viscous_flux(self, fv, ts):
        out = np.zeros((self.m.nfaces, 5))
        self.L.orc_viscous_flux(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv), ts, ptr(out))
        return out

    def source(self, ff, vs, ts):
        out = np.zeros((self.m.ncells, 5))
        self.L.orc_source(self.m.soa_ptr, C.byref(self.cfg), ptr(ff), ptr(vs), ts, ptr(out))
        return out

    def bc_fill(self, fv, ts):
        return self.L.orc_bc_fill(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv), ts)

    def interface_exchange(self, fv, ts):
        self.L.orc_interface_exchange(self.m.soa_ptr, self.m.phys, C.byref(self.cfg), ptr(fv), ts)

    def ib_fill(self, fv, motion=None):
        return self.L.orc_ib_fill(self.m.soa_ptr, self.m.phys, ptr(fv), motion)

    def compute_dt(self, fv, dt):
        return self.L.orc_compute_dt(self.m.soa_ptr, self.m.phys, ptr(fv), ptr(dt))

    def forces(self, fv, motion=None):
        out = np.zeros(self.m.ncases * 6)
        self.L.orc_forces(self.m.soa_ptr, self.m.phys, ptr(fv), motion, ptr(out))
        return out
