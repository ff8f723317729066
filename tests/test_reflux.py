"""The optional coarse/fine flux balance (refluxing) on the CPU.

The reference builds the face pairs (AMR::FindModifyFace Fluxmodification.cpp:
5-289) and the balance (BalanceFaceFlux :291-358) but leaves both call sites
commented (NS_Solver.H:672, 686), so no reference output exists for it.  The
checks here are what the balance has to satisfy whatever its weights:

  pairs        every face between a normal cell and a first-layer restriction
               ghost is a coarse balance face, every face between a normal
               fine cell and a prolongation ghost is one of its fine faces,
               each exactly once; the fine faces cover the coarse face (area)
               and lie on the coarse face's plane, between the coarse cell's
               sons and the restriction ghost's sons
  balance      after orc_balance_flux the coarse flux equals the sum of its
               fine fluxes; a flux field that is already consistent
               (fine = coarse / count) is left as it is; other faces keep
               their bits
  conservation on a body-free refined level set (refine_only) the volume
               integral of the conserved variables changes only by rounding
               over steps with the balance (1 and 4 stages), and by far more
               without it; only the cells of balance faces see the change
"""
import dataclasses

import numpy as np
import pytest

import ibgpu
import oracle_ffi

NORMAL, RES1, PRO = 0, 2, 4


def meshes():
    return {
        "tiny2d": ibgpu.tiny_case(),
        "tiny3d": ibgpu.tiny3d_case(),
        # level 1 refined in z as well: the four-face form
        "tiny3d_xyz": dataclasses.replace(ibgpu.tiny3d_case(), twod_from=2),
    }


@pytest.fixture(scope="module", params=list(meshes()))
def mesh(request):
    return ibgpu.Mesh(meshes()[request.param])


def face_area(m, f):
    lo = m.face_cells[f, 0]
    b = int(np.searchsorted(m.blk_start, lo, side="right") - 1)
    h = m.blk_h[b]
    d = int(m.face_dir[f])
    return float(np.prod([h[k] for k in range(3) if k != d]))


def test_pairs_cover_every_interface(mesh):
    m = mesh
    assert m.nmdy > 0
    lv = m.level_of()
    kind = m.kind
    fc = m.face_cells
    coarse = set()
    fine = set()
    parent_of_pro = {}
    for i in range(m.npro):
        for s in m.pro_fi[i]:
            if s >= 0:
                parent_of_pro[int(s)] = int(m.pro_ci[i])
    res_sons = {}
    for i in range(m.nres):
        res_sons[int(m.res_ci[i])] = {int(s) for s in m.res_fi[i] if s >= 0}
    for form, cf, *ffs in m.mdy:
        nf = 2 if form == 1 else 4
        assert all(f >= 0 for f in ffs[:nf]) and all(f == -1 for f in ffs[nf:])
        a, b = int(fc[cf, 0]), int(fc[cf, 1])
        kinds = {int(kind[a]), int(kind[b])}
        assert kinds == {NORMAL, RES1}, kinds
        C = a if kind[a] == NORMAL else b
        P = b if C == a else a
        assert lv[C] == lv[P]
        assert cf not in coarse
        coarse.add(int(cf))
        area = 0.0
        d = int(m.face_dir[cf])
        for f in ffs[:nf]:
            x, y = int(fc[f, 0]), int(fc[f, 1])
            assert {int(kind[x]), int(kind[y])} == {NORMAL, PRO}
            g = x if kind[x] == PRO else y
            n = y if g == x else x
            assert int(m.face_dir[f]) == d
            assert lv[g] == lv[C] + 1
            assert parent_of_pro[g] == C           # a son of the coarse cell
            assert n in res_sons[P]                # a son of the restriction ghost
            # on the coarse face's plane
            xc = 0.5 * (m.centers[a, d] + m.centers[b, d])
            xf = 0.5 * (m.centers[x, d] + m.centers[y, d])
            assert abs(xc - xf) < 1e-12 * max(1.0, abs(xc))
            assert f not in fine
            fine.add(int(f))
            area += face_area(m, f)
        assert abs(area - face_area(m, cf)) < 1e-12 * face_area(m, cf)
        # the form: two faces when the fine level refines x,y only (2D)
        if m.dim == 2:
            assert form == 1
    # nothing missed: every normal/RES1 face and every normal/PRO face
    k0, k1 = kind[fc[:, 0]], kind[fc[:, 1]]
    want_c = set(np.nonzero(((k0 == NORMAL) & (k1 == RES1)) | ((k0 == RES1) & (k1 == NORMAL)))[0].tolist())
    want_f = set(np.nonzero(((k0 == NORMAL) & (k1 == PRO)) | ((k0 == PRO) & (k1 == NORMAL)))[0].tolist())
    assert coarse == want_c
    assert fine == want_f


def test_balance_conserves_and_fixes_consistent_fluxes(mesh):
    m = mesh
    L = oracle_ffi.lib()
    rng = np.random.default_rng(5)
    flux = rng.normal(size=(m.nfaces, 5))
    out = flux.copy()
    L.orc_balance_flux(m.soa_ptr, oracle_ffi.ptr(out))
    touched = np.zeros(m.nfaces, bool)
    for form, cf, *ffs in m.mdy:
        nf = 2 if form == 1 else 4
        fi = list(ffs[:nf])
        touched[[cf] + fi] = True
        # the coarse flux now equals the sum of its fine fluxes
        assert np.allclose(out[cf], out[fi].sum(axis=0), rtol=1e-14, atol=1e-14)
        # it moved towards the fine side: neither side kept its value
        assert not np.array_equal(out[cf], flux[cf])
    assert np.array_equal(out[~touched], flux[~touched])
    # a consistent field (every fine flux = coarse / count) is a fixed point
    cons = flux.copy()
    for form, cf, *ffs in m.mdy:
        nf = 2 if form == 1 else 4
        for f in ffs[:nf]:
            cons[f] = cons[cf] / nf
    out = cons.copy()
    L.orc_balance_flux(m.soa_ptr, oracle_ffi.ptr(out))
    assert np.allclose(out, cons, rtol=1e-15, atol=1e-15)


def bump_state(m, amp=0.08, radius=1.1, center=(0.25, 0.0)):
    """Freestream plus a compact cos^2 bump in density and pressure: exactly
    freestream beyond `radius` of `center`."""
    fv = m.init_state()
    r = np.hypot(m.centers[:, 0] - center[0], m.centers[:, 1] - center[1])
    w = np.where(r < radius, np.cos(0.5 * np.pi * r / radius) ** 2, 0.0)
    fv[:, 0] *= 1.0 + amp * w
    fv[:, 4] *= 1.0 + 1.5 * amp * w
    fv[:, 1] += 0.3 * amp * w
    P = m.phys[0]
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2 + fv[:, 3] ** 2)
    return fv


def totals(m, fv):
    """Volume integral of rho, rho u, rho v, rho E over the normal cells."""
    out = np.zeros(4)
    for b in range(m.ncases * m.nlevels):
        s, n = m.blk_start[b], m.blk_nnormal[b]
        vol = float(np.prod(m.blk_h[b]))
        f = fv[s:s + n]
        rho = f[:, 0]
        out += vol * np.array([rho.sum(), (rho * f[:, 1]).sum(), (rho * f[:, 2]).sum(),
                               (rho * f[:, 6]).sum()])
    return out


@pytest.fixture(scope="module")
def free_mesh():
    # the section places the refinement bands; no body in the flow
    return ibgpu.Mesh(dataclasses.replace(ibgpu.tiny_case(base=48), refine_only=True))


def march(m, nstage, nsteps, reflux):
    orc = oracle_ffi.Oracle(m, oracle_ffi.cfg(nstage=nstage, dt_var_step=1, temporal_refine=0,
                                              reflux=reflux))
    fv = bump_state(m)
    dt = np.array([m.phys[0].dt_init])
    t0 = totals(m, fv)
    for ts in range(nsteps):
        code, *_ = orc.step(fv, dt, ts)
        assert code == 0
    return fv, totals(m, fv) - t0, t0


@pytest.mark.parametrize("nstage,nsteps", [(1, 4), (4, 1)])
def test_balance_conserves_on_a_body_free_level_set(free_mesh, nstage, nsteps):
    m = free_mesh
    assert m.nib == 0 and m.nmdy > 0
    fv0 = bump_state(m)
    # the bump stays clear of the domain boundary (exact freestream there)
    far = np.hypot(m.centers[:, 0] - 0.25, m.centers[:, 1]) > 3.0
    fv1, drift_on, t0 = march(m, nstage, nsteps, 1)
    assert np.array_equal(fv1[far & (m.kind == NORMAL)], fv0[far & (m.kind == NORMAL)])
    _, drift_off, _ = march(m, nstage, nsteps, 0)
    scale = np.abs(t0)
    scale[1:3] = abs(t0[0])   # the momentum totals are near zero
    on, off = np.abs(drift_on) / scale, np.abs(drift_off) / scale
    assert on.max() < 2e-14, on
    assert off[[0, 3]].min() > 1e-9, off
    assert off.max() > 1e4 * on.max()


def test_balance_changes_only_the_cells_of_balance_faces(free_mesh):
    m = free_mesh
    f_on, _, _ = march(m, 1, 1, 1)
    f_off, _, _ = march(m, 1, 1, 0)
    diff = np.any(f_on != f_off, axis=1)
    cells = set()
    for form, cf, *ffs in m.mdy:
        nf = 2 if form == 1 else 4
        for f in [cf] + list(ffs[:nf]):
            for c in m.face_cells[f]:
                if m.kind[c] == NORMAL:
                    cells.add(int(c))
    allowed = np.zeros(m.ncells, bool)
    allowed[list(cells)] = True
    assert diff.any()
    assert not np.any(diff & ~allowed)
