"""Case definitions and the host mesh wrapper.

A case is an airfoil immersed in a Cartesian box [-10,10]^2 (x along the
freestream).  The reference rotates its body for an angle of attack and keeps
the freestream along +x (Userfunction_cy_naca15.cpp:205-216, 356-371); so do
these cases.  Upper and lower surfaces are two bodies, as in the reference
(Userfunction_cy_naca15.cpp:205-211).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _ffi


@dataclass
class CaseSpec:
    name: str
    dim: int = 2
    nx: int = 64
    ny: int = 64
    nz: int = 1
    lo: float = -10.0
    hi: float = 10.0
    span: float = 0.1            # z extent (2D: one cell of this depth)
    nlevels: int = 4
    twod_from: int = 1
    bands: List[tuple] = field(default_factory=list)   # (lo, hi) per level < L-1
    thickness: float = 0.12
    npts: int = 100               # points per surface (cosine spaced)
    aoa: float = 2.0
    Ma: float = 0.5
    Re: float = 5000.0
    Pr: float = 0.72
    gama: float = 1.4
    cfl: float = 0.5
    backpressure: float = 0.0
    span_patches: int = 1
    pivot_y: float = 0.0          # y of the quarter chord (x is 0)
    face_bc: tuple = (0, 0, 0, 0)  # BC codes of x-, x+, y-, y+ (0: -13, -12, -50, -50)
    # SA working variable of the -13 inlet: None keeps the reference's (its
    # facevar[0].viseddy is never set, so 0); "freestream" uses the
    # freestream value.  With 0 the uniform inflow drives the first cells'
    # working variable to exactly 0, where the SA destruction term is 0/0
    # (r = nu / (S~ kappa^2 d^2) with nu = 0 and |Omega| = 0): NaN after
    # ~1400 steps of the 2.6M-cell case.
    inlet_nut: object = None
    # True: the section steers the refinement bands only (no solid or IB
    # cells, no wall patches): a body-free flow on the refined level set
    refine_only: bool = False

    def to_ctypes(self):
        c = _ffi.Case()
        c.dim = self.dim
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        dx = (self.hi - self.lo) / self.nx
        dy = (self.hi - self.lo) / self.ny
        c.dh0[0], c.dh0[1], c.dh0[2] = dx, dy, self.span / self.nz
        c.origin[0], c.origin[1], c.origin[2] = self.lo, self.lo, 0.0
        c.nlevels = self.nlevels
        c.twod_from = self.twod_from
        for i in range(_ffi.IBG_MAX_LEVELS):
            c.band_lo[i], c.band_hi[i] = (0.0, 0.0)
        for i, (a, b) in enumerate(self.bands):
            c.band_lo[i], c.band_hi[i] = a, b
        xy = np.zeros((2 * self.npts, 2), dtype=np.float64)
        host = _ffi.host_lib()
        host.ibm_naca4(self.thickness, self.npts, 0.0, self.pivot_y, self.aoa,
                       xy.ctypes.data_as(_ffi.f64p))
        self._xy = np.ascontiguousarray(xy)       # keep alive
        c.nbody = 2
        c.body_npts[0] = self.npts
        c.body_npts[1] = self.npts
        c.body_xy = self._xy.ctypes.data_as(_ffi.f64p)
        c.body_center[0], c.body_center[1], c.body_center[2] = 0.0, self.pivot_y, 0.5 * self.span
        c.span = self.span
        c.span_patches = self.span_patches
        c.gama, c.Ma, c.Re, c.Pr = self.gama, self.Ma, self.Re, self.Pr
        c.cfl = self.cfl
        c.backpressure = self.backpressure
        c.dt_init = 1.0e30
        c.refine_only = 1 if self.refine_only else 0
        for j in range(4):
            c.face_bc[j] = int(self.face_bc[j])
        return c


def config1(levels=4, base=64):
    """Plain method, small: NACA0012, 2D, Mach 0.5, AoA 2, Re 5000."""
    # refinement bands (signed distance, chord units) per level
    bands = [(-0.6, 2.0), (-0.3, 1.0), (-0.15, 0.5), (-0.08, 0.25), (-0.04, 0.12)]
    return CaseSpec(name="naca0012_m05_re5000", dim=2, nx=base, ny=base, nz=1,
                    nlevels=levels, bands=bands[:levels - 1], thickness=0.12,
                    npts=100, aoa=2.0, Ma=0.5, Re=5000.0)


def config2(levels=6, base=256):
    """Mid 2D case: NACA0015, Mach 0.3, Re 10000, 256^2 base + 5 AMR levels."""
    bands = [(-0.3, 6.0), (-0.15, 4.0), (-0.08, 2.5), (-0.04, 1.6), (-0.02, 1.0)]
    return CaseSpec(name="naca0015_m03_re10000", dim=2, nx=base, ny=base, nz=1,
                    nlevels=levels, bands=bands[:levels - 1], thickness=0.15,
                    npts=200, aoa=0.0, Ma=0.3, Re=10000.0, inlet_nut="freestream")


def tiny_case(levels=3, base=32, aoa=4.0, Ma=0.5, Re=2000.0, thickness=0.12):
    """A few-thousand-cell case for parity tests that the oracle runs in seconds."""
    bands = [(-1.0, 1.5), (-0.5, 0.6), (-0.25, 0.3), (-0.12, 0.15)]
    return CaseSpec(name="tiny", dim=2, nx=base, ny=base, nz=1, lo=-4.0, hi=4.0,
                    nlevels=levels, bands=bands[:levels - 1], thickness=thickness,
                    npts=60, aoa=aoa, Ma=Ma, Re=Re)


def tiny3d_case(levels=3, base=24, nz=4, aoa=4.0, Ma=0.5, Re=2000.0):
    """Quasi-span 3D case (periodic span, levels refine x,y only) for parity tests."""
    bands = [(-1.0, 1.5), (-0.5, 0.6), (-0.25, 0.3)]
    return CaseSpec(name="tiny3d", dim=3, nx=base, ny=base, nz=nz, lo=-4.0, hi=4.0, span=0.4,
                    nlevels=levels, twod_from=1, bands=bands[:levels - 1], thickness=0.12,
                    npts=40, aoa=aoa, Ma=Ma, Re=Re, span_patches=nz)


def config3(levels=5, base=128, nz=8):
    """3D quasi-span: NACA0012 extruded over a periodic span of 0.1 chord,
    128x128x8 base, 4 AMR levels, Mach 0.6, AoA 4, Re 5e4 (~20M leaf cells).
    The first two refinements also halve the span spacing (0.0125 -> 0.003);
    the finer ones refine x,y only, where the cells have become near
    isotropic."""
    bands = [(-0.3, 6.0), (-0.15, 4.0), (-0.08, 3.2), (-0.04, 2.8)]
    return CaseSpec(name="naca0012_3d_m06_re5e4", dim=3, nx=base, ny=base, nz=nz, span=0.1,
                    nlevels=levels, twod_from=3, bands=bands[:levels - 1], thickness=0.12,
                    npts=100, aoa=4.0, Ma=0.6, Re=50000.0, span_patches=nz,
                    inlet_nut="freestream")


def ensemble_cases(n=32, seed=20261202, base=256, levels=6):
    """Batched ensemble: NACA 4-digit sections, thickness U[0.09,0.18], AoA
    U[-4,8] deg, Mach U[0.3,0.7], Re 1e4 (seeded); the refinement extent of
    each case is drawn too (uneven AMR, ~0.6M to ~3M leaf cells per case on
    the 256^2 base with 5 AMR levels)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = float(rng.uniform(0.09, 0.18))
        a = float(rng.uniform(-4.0, 8.0))
        m = float(rng.uniform(0.3, 0.7))
        w = float(rng.uniform(0.3, 1.15))   # uneven refinement extent
        bands = [(-0.3, 6.0 * w), (-0.15, 4.0 * w), (-0.08, 2.5 * w), (-0.04, 1.6 * w), (-0.02, 1.0 * w)]
        out.append(CaseSpec(name=f"ens{i:02d}", dim=2, nx=base, ny=base, nz=1, nlevels=levels,
                            bands=bands[:levels - 1], thickness=t, npts=100, aoa=a, Ma=m,
                            Re=10000.0, inlet_nut="freestream"))
    return out


class Mesh:
    """Owns an ibm_handle; exposes the flattened mesh as numpy views."""

    def __init__(self, specs):
        if isinstance(specs, CaseSpec):
            specs = [specs]
        self.specs = list(specs)
        host = _ffi.host_lib()
        arr = (_ffi.Case * len(self.specs))()
        for i, s in enumerate(self.specs):
            arr[i] = s.to_ctypes()
        self._cases = arr
        h = host.ibm_build(arr, len(self.specs))
        if not h:
            raise RuntimeError("mesh build failed: " + host.ibm_last_error().decode())
        self._h = C.c_void_p(h)
        self._host = host
        self.soa_ptr = host.ibm_mesh(self._h)
        self.soa = self.soa_ptr.contents
        self.phys = (_ffi.Physics * len(self.specs))()
        for i in range(len(self.specs)):
            host.ibm_physics(C.byref(arr[i]), C.byref(self.phys[i]))
            nin = self.specs[i].inlet_nut
            if nin is not None:
                self.phys[i].nut_inlet = self.phys[i].nut_inf if nin == "freestream" else float(nin)
        m = self.soa
        self.dim, self.ncases, self.nlevels = m.dim, m.ncases, m.nlevels
        self.ncells, self.nfaces = m.ncells, m.nfaces
        nb = self.ncases * self.nlevels
        v = np.ctypeslib.as_array
        self.blk_start = v(m.blk_start, (nb + 1,)).copy()
        self.blk_nnormal = v(m.blk_nnormal, (nb,)).copy()
        self.blk_h = v(m.blk_h, (nb, 3)).copy()
        self.blk_face_start = v(m.blk_face_start, (nb + 1,)).copy()
        n = self.ncells
        self.key = v(m.key, (n, 3))
        self.kind = v(m.kind, (n,))
        self.nb = v(m.nb, (n, 27))
        self.cell_faces = v(m.cell_faces, (n, 6))
        self.signdis = v(m.signdis, (n,))
        self.ibidx = v(m.ibidx, (n,))
        self.face_cells = v(m.face_cells, (self.nfaces, 2)) if self.nfaces else np.zeros((0, 2), np.int32)
        self.face_dir = v(m.face_dir, (self.nfaces,)) if self.nfaces else np.zeros(0, np.int8)
        self.dmg = (v(m.dmg_cell, (m.ndmg,)), v(m.dmg_ref, (m.ndmg,)), v(m.dmg_type, (m.ndmg,))) if m.ndmg else None
        self.nres, self.npro, self.nib, self.npatch = m.nres, m.npro, m.nib, m.npatch
        self.res_ci = v(m.res_ci, (m.nres,)) if m.nres else np.zeros(0, np.int32)
        self.res_fi = v(m.res_fi, (m.nres, 8)) if m.nres else np.zeros((0, 8), np.int32)
        self.pro_ci = v(m.pro_ci, (m.npro,)) if m.npro else np.zeros(0, np.int32)
        self.pro_fi = v(m.pro_fi, (m.npro, 8)) if m.npro else np.zeros((0, 8), np.int32)
        self.ib_cell = v(m.ib_cell, (m.nib,)) if m.nib else np.zeros(0, np.int32)
        self.ib_intp = v(m.ib_intp, (m.nib, 8)) if m.nib else np.zeros((0, 8), np.int32)
        self.ib_coef = v(m.ib_coef, (m.nib, 8)) if m.nib else np.zeros((0, 8))
        self.ib_hgdis = v(m.ib_hgdis, (m.nib,)) if m.nib else np.zeros(0)
        self.pat_intp = v(m.pat_intp, (m.npatch, 8)) if m.npatch else np.zeros((0, 8), np.int32)
        self.pat_coef = v(m.pat_coef, (m.npatch, 8)) if m.npatch else np.zeros((0, 8))
        # coarse/fine flux-balance faces: form, coarse face, fine faces
        self.nmdy = m.nmdy
        self.mdy = v(m.mdy, (m.nmdy, 6)) if m.nmdy else np.zeros((0, 6), np.int32)
        self.centers = np.zeros((n, 3))
        host.ibm_cell_centers(self._h, self.centers.ctypes.data_as(_ffi.f64p))

    def block(self, case, level):
        return case * self.nlevels + level

    def level_of(self):
        """level index per cell"""
        lv = np.zeros(self.ncells, np.int32)
        for b in range(self.ncases * self.nlevels):
            lv[self.blk_start[b]:self.blk_start[b + 1]] = b % self.nlevels
        return lv

    def surface(self, ib_capacity=0):
        """ibg_surface of the mesh's 2D sections for moving bodies (ibm_surface):
        the zero-pose point offsets from each case's quarter chord
        (ibm_naca4_offsets) with the builder's patch segments, patch ranges,
        closest patch of every cell and image cell of every patch.  Returns
        (Surface, arrays) with the arrays as numpy views."""
        if self.dim != 2:
            raise ValueError("moving bodies: 2D cases only")
        host = self._host
        offs = []
        for sp in self.specs:
            o = np.zeros((2 * sp.npts, 2))
            host.ibm_naca4_offsets(sp.thickness, sp.npts, o.ctypes.data_as(_ffi.f64p))
            offs.append(o)
        pt_off = np.ascontiguousarray(np.concatenate(offs))
        s = _ffi.Surface()
        if host.ibm_surface(self._h, pt_off.ctypes.data_as(_ffi.f64p), pt_off.shape[0],
                            C.byref(s)) != 0:
            raise ValueError(host.ibm_last_error().decode())
        s.ib_capacity = int(ib_capacity)
        v = np.ctypeslib.as_array
        nb = len([0 for sp in self.specs for _ in range(2)])
        a = {
            "pt_off": pt_off,
            "pat_pt0": v(s.pat_pt0, (self.npatch,)), "pat_body": v(s.pat_body, (self.npatch,)),
            "pat_flip": v(s.pat_flip, (self.npatch,)), "pat_dz": v(s.pat_dz, (self.npatch,)),
            "body_p0": v(s.body_p0, (nb + 1,)),
            "case_origin": v(s.case_origin, (self.ncases, 3)),
            "cell_patch": v(s.cell_patch, (self.ncells,)),
            "pat_close": v(s.pat_close, (max(self.npatch, 1),)),
        }
        s._keep = a
        return s, a

    @staticmethod
    def pose(aoa_deg, pivot_y=0.0):
        """{pivot x, pivot y, cos a, sin a} of a section at angle of attack
        aoa_deg with its quarter chord at (0, pivot_y) (ibm_pose)."""
        p = np.zeros(4)
        _ffi.host_lib().ibm_pose(0.0, float(pivot_y), float(aoa_deg), p.ctypes.data_as(_ffi.f64p))
        return p

    def transfer_to(self, new, fv, nut=None):
        """State on `new` (a rebuild of the same cases after a regrid) from the
        state `fv` [ncells][7] (and SA variable `nut`) on this level set:
        AMR::ComptDatainNewBox (AMR.H:148-215) on canonical keys."""
        fv = np.ascontiguousarray(fv, dtype=np.float64)
        out = np.zeros((new.ncells, _ffi.IBG_NFV))
        nut_in = None if nut is None else np.ascontiguousarray(nut, dtype=np.float64)
        nut_out = None if nut is None else np.zeros(new.ncells)
        rc = self._host.ibm_transfer(self._h, self.phys, fv.ctypes.data_as(_ffi.f64p),
                                     None if nut_in is None else nut_in.ctypes.data_as(_ffi.f64p),
                                     new._h, out.ctypes.data_as(_ffi.f64p),
                                     None if nut_out is None else nut_out.ctypes.data_as(_ffi.f64p))
        if rc != 0:
            raise RuntimeError("state transfer failed: " + self._host.ibm_last_error().decode())
        return out, nut_out

    def write_grid(self, path):
        """Tecplot grid file of every level (Pltoutput::Output_Grid)."""
        if self._host.ibm_write_grid(self._h, str(path).encode()) != 0:
            raise RuntimeError(self._host.ibm_last_error().decode())

    def write_solution(self, path, fv, nut=None, ts=0, dt=None, t=None):
        """Tecplot solution file (Pltoutput::Output_Soln), also a restart."""
        fv = np.ascontiguousarray(fv, dtype=np.float64)
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        nut, dt, t = arr(nut), arr(dt), arr(t)
        p = lambda a: None if a is None else a.ctypes.data_as(_ffi.f64p)
        if self._host.ibm_write_solution(self._h, self.phys, p(fv), p(nut), int(ts), p(dt), p(t),
                                         str(path).encode()) != 0:
            raise RuntimeError(self._host.ibm_last_error().decode())

    def read_solution(self, path, with_nut=False):
        """State (and SA variable, ts, dt, t) from a solution file of this
        level set (InputSoln)."""
        fv = self.init_state()
        nut = np.zeros(self.ncells) if with_nut else None
        ts = C.c_int32(0)
        dt = np.full(self.ncases, np.nan)
        t = np.full(self.ncases, np.nan)
        rc = self._host.ibm_read_solution(self._h, self.phys, str(path).encode(),
                                          fv.ctypes.data_as(_ffi.f64p),
                                          None if nut is None else nut.ctypes.data_as(_ffi.f64p),
                                          C.byref(ts), dt.ctypes.data_as(_ffi.f64p),
                                          t.ctypes.data_as(_ffi.f64p))
        if rc != 0:
            raise RuntimeError("solution file rejected: " + self._host.ibm_last_error().decode())
        return fv, nut, ts.value, dt, t

    def init_state(self):
        fv = np.zeros((self.ncells, _ffi.IBG_NFV))
        self._host.ibm_init_state(self._h, self.phys, fv.ctypes.data_as(_ffi.f64p))
        return fv

    def counts(self):
        out = []
        for c in range(self.ncases):
            row = []
            for l in range(self.nlevels):
                b = self.block(c, l)
                row.append(int(self.blk_nnormal[b]))
            out.append(row)
        return out

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._host.ibm_free(self._h)
                self._h = None
        except Exception:
            pass
