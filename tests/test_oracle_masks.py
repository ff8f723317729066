"""Pins for the oracle's per-cell masks and reductions, CPU only.

Each test plants values in chosen cell classes and checks what the oracle
does with them against the rules as the reference writes them, so that a
wrong mask, a wrong level time step, a missing clamp or reset, a dropped
copy or a mixed-up case shows up:

  orc_residual      ResinDomain NS_Solver.H:264-378: max|source[0]| over the
                    marched non-finest levels (every normal cell, solid and
                    IB ones included), the finest level always, there only
                    cells with infectbox == -1 and signdis > 0 (strict);
                    ghosts never; per case; 2D: the w component is zero
  orc_residual_nut  the mutsource part of the same loops (NS_Solver.H:297,
                    321), same cells
  orc_sa_update     UpdateConsVars NS_Solver.cpp:857-876 (non-finest,
                    marched: every normal cell, level time step
                    alpha*dt*2^(L-1-l), max(., min_viseddy = 0)) and
                    :934-957 (finest: only !solid && infectbox == -1), then
                    ComptNewPrimVar :28-35, :159-165 (m_data = initvar
                    resets viseddy where signdis <= 0; finest IB cells are
                    left alone)
  orc_sa_bc_fill    SetBCType_13/12/50 BCValues.H:191-220 for the SA
                    variable: inlet ghost = facevar viseddy, outlet and
                    outflow ghosts copy the reference cell (bcvar = refvar);
                    the outlet keeps the reference cell's E (no Get_E)
"""
import ctypes as C

import numpy as np
import pytest

import ibgpu
import oracle_ffi as O

L = O.lib()
ptr = O.ptr


@pytest.fixture(scope="module")
def mesh():
    return ibgpu.Mesh(ibgpu.tiny_case())


@pytest.fixture(scope="module")
def mesh2():
    """two cases of different sizes in one index space"""
    return ibgpu.Mesh([ibgpu.tiny_case(aoa=4.0), ibgpu.tiny_case(aoa=-2.0, base=24)])


def normal_range(m, c, l):
    b = m.block(c, l)
    s = int(m.blk_start[b])
    return s, s + int(m.blk_nnormal[b])


def marched(L_, l, ts):
    """level l is marched when ts is a multiple of 2^(L-1-l) (mpiparams.cpp:186-212)"""
    return ts % (1 << (L_ - 1 - l)) == 0


def pick(m, c, l, cond):
    s, e = normal_range(m, c, l)
    idx = np.nonzero(cond(np.arange(s, e)))[0]
    assert idx.size, "no cell of this class"
    return s + int(idx[idx.size // 2])


def residual(m, src, ts, cfg=None):
    cfg = cfg or O.cfg()
    linf = np.zeros(m.ncases * 5)
    l2 = np.zeros(m.ncases * 5)
    src = np.ascontiguousarray(src, dtype=np.float64)
    L.orc_residual(m.soa_ptr, C.byref(cfg), ptr(src), ts, ptr(linf), ptr(l2))
    return linf.reshape(m.ncases, 5), l2.reshape(m.ncases, 5)


def residual_nut(m, ms, ts, cfg=None):
    cfg = cfg or O.cfg()
    linf = np.zeros(m.ncases)
    l2 = np.zeros(m.ncases)
    ms = np.ascontiguousarray(ms, dtype=np.float64)
    L.orc_residual_nut(m.soa_ptr, C.byref(cfg), ptr(ms), ts, ptr(linf), ptr(l2))
    return linf, l2


def planted_mesh(m):
    """tiny mesh with, on the middle level, one normal cell made solid, one
    given an IB index, and on the finest level one fluid cell with signdis
    exactly 0; returns the patched mesh and the planted cells"""
    sd = np.array(m.signdis, dtype=np.float64)
    ib = np.array(m.ibidx, dtype=np.int32)
    Lm = m.nlevels
    s1, e1 = normal_range(m, 0, 1)
    c_solid1, c_ib1 = s1 + 10, s1 + 20
    sd[c_solid1] = -0.05
    ib[c_ib1] = 3
    s2, e2 = normal_range(m, 0, Lm - 1)
    f = s2 + int(np.nonzero((sd[s2:e2] > 0) & (ib[s2:e2] < 0))[0][0])
    sd[f] = 0.0
    # an IB cell on the solid side (the mirror branch of IBConditions_Image)
    fi = s2 + int(np.nonzero(ib[s2:e2] >= 0)[0][-1])
    sd[fi] = -0.01
    pm = O.PatchedMesh(m, signdis=sd, ibidx=ib)
    return pm, {"solid1": c_solid1, "ib1": c_ib1, "zero_fine": f, "ib_neg": fi}


def cell_classes(m, pm, pm_cells):
    Lm = m.nlevels
    sd, ib = pm._keep["signdis"], pm._keep["ibidx"]
    fin = lambda cond: pick(m, 0, Lm - 1, cond)
    s0, e0 = normal_range(m, 0, 0)
    ghost = int(m.blk_start[m.block(0, 0)] + m.blk_nnormal[m.block(0, 0)])
    assert ghost < m.blk_start[m.block(0, 1)], "level 0 holds ghost cells"
    return {
        # (cell, level, included when its level is marched)
        "coarse": (s0 + 7, 0, True),
        "mid": (normal_range(m, 0, 1)[0] + 30, 1, True),
        "solid1": (pm_cells["solid1"], 1, True),     # non-finest: no solid mask
        "ib1": (pm_cells["ib1"], 1, True),           # non-finest: no IB mask
        "fine_fluid": (fin(lambda g: (sd[g] > 0) & (ib[g] < 0)), Lm - 1, True),
        "fine_solid": (fin(lambda g: sd[g] < 0), Lm - 1, False),
        "fine_ib": (fin(lambda g: ib[g] >= 0), Lm - 1, False),
        "zero_fine": (pm_cells["zero_fine"], Lm - 1, False),  # signdis > 0 is strict
        "ib_neg": (pm_cells["ib_neg"], Lm - 1, False),
        "ghost": (ghost, 0, False),
    }


@pytest.mark.parametrize("ts", [0, 1, 2, 3])
def test_residual_planted_cells(mesh, ts):
    pm, planted = planted_mesh(mesh)
    Lm = mesh.nlevels
    for name, (g, lvl, inc) in cell_classes(mesh, pm, planted).items():
        for v in (0, 2, 4):
            src = np.zeros((mesh.ncells, 5))
            src[g, v] = -3.5 - v
            linf, l2 = residual(pm, src, ts)
            want = inc and (lvl == Lm - 1 or marched(Lm, lvl, ts))
            exp = np.zeros(5)
            if want:
                exp[v] = 3.5 + v
            assert np.array_equal(linf[0], exp), (name, ts, v, linf[0])
            assert np.array_equal(l2[0], exp), (name, ts, v, l2[0])


@pytest.mark.parametrize("ts", [0, 1, 2])
def test_residual_nut_planted_cells(mesh, ts):
    pm, planted = planted_mesh(mesh)
    Lm = mesh.nlevels
    for name, (g, lvl, inc) in cell_classes(mesh, pm, planted).items():
        ms = np.zeros(mesh.ncells)
        ms[g] = -0.25
        linf, l2 = residual_nut(pm, ms, ts)
        want = inc and (lvl == Lm - 1 or marched(Lm, lvl, ts))
        assert linf[0] == (0.25 if want else 0.0), (name, ts)
        assert l2[0] == (0.25 if want else 0.0), (name, ts)


def included_mask(m, ts):
    """the cells ResinDomain reads, from per-cell level and distance data"""
    Lm = m.nlevels
    lvl = m.level_of()
    inc = np.zeros(m.ncells, bool)
    for c in range(m.ncases):
        for l in range(Lm):
            s, e = normal_range(m, c, l)
            if l == Lm - 1:
                g = np.arange(s, e)
                inc[s:e] = (np.asarray(m.ibidx)[g] == -1) & (np.asarray(m.signdis)[g] > 0.0)
            elif marched(Lm, l, ts):
                inc[s:e] = True
    assert np.all(lvl[inc] >= 0)
    return inc


@pytest.mark.parametrize("ts", [0, 1, 2, 3])
def test_residual_random_recount(mesh2, ts):
    m = mesh2
    rng = np.random.default_rng(7 + ts)
    src = rng.normal(size=(m.ncells, 5)) * rng.uniform(0.1, 10.0, size=(m.ncells, 1))
    ms = rng.normal(size=m.ncells)
    linf, l2 = residual(m, src, ts)
    nlinf, nl2 = residual_nut(m, ms, ts)
    inc = included_mask(m, ts)
    for c in range(m.ncases):
        s, e = int(m.blk_start[m.block(c, 0)]), int(m.blk_start[m.block(c, m.nlevels - 1) + 1])
        sel = np.zeros(m.ncells, bool)
        sel[s:e] = inc[s:e]
        assert sel.sum() > 100
        assert np.array_equal(linf[c], np.abs(src[sel]).max(axis=0))
        np.testing.assert_allclose(l2[c], np.sqrt((src[sel] ** 2).sum(axis=0)), rtol=1e-13)
        assert nlinf[c] == np.abs(ms[sel]).max()
        np.testing.assert_allclose(nl2[c], np.sqrt((ms[sel] ** 2).sum()), rtol=1e-13)
    # the other case's spike never leaks into this case
    src2 = np.zeros((m.ncells, 5))
    g1 = normal_range(m, 1, 0)[0] + 3
    src2[g1, 1] = 9.0
    linf2, _ = residual(m, src2, 0)
    assert linf2[0, 1] == 0.0 and linf2[1, 1] == 9.0


def test_residual_2d_w_component_is_zero(mesh):
    """ResinDomain sets res_sum[3] = 0 in 2D (NS_Solver.H:345-347): the 2D
    source has no w component, so its norms are exactly zero after steps"""
    fv = mesh.init_state()
    orc = O.Oracle(mesh)
    dt = np.array([mesh.phys[0].dt_init])
    for ts in range(3):
        code, src, linf, l2 = orc.step(fv, dt, ts)
        assert code == 0
        assert linf[3] == 0.0 and l2[3] == 0.0
        assert linf[0] > 0.0 and linf[4] > 0.0
        assert not np.any(src[:, 3])


# ----------------------------------------------------------------- SA update
def sa_update(m, nut_old, ms, dt, ts, subt, nut_in, cfg):
    nut = np.array(nut_in, dtype=np.float64)
    L.orc_sa_update(m.soa_ptr, m.phys, C.byref(cfg), ptr(np.ascontiguousarray(nut_old)),
                    ptr(np.ascontiguousarray(ms)), ptr(np.ascontiguousarray(dt)), ts, subt,
                    ptr(nut))
    return nut


@pytest.mark.parametrize("ts,subt", [(0, 0), (0, 2), (1, 1), (2, 3)])
def test_sa_update_clamp_reset_and_masks(mesh, ts, subt):
    pm, planted = planted_mesh(mesh)
    m = mesh
    Lm = m.nlevels
    cfg = O.cfg(nstage=4)
    alpha = [0.1084, 0.2602, 0.5052, 1.0][subt]
    dt = np.array([3.0e-3])
    nut_inf = m.phys[0].nut_inf
    sentinel = -7.0
    rng = np.random.default_rng(ts * 10 + subt)
    nut_old = 2.0 + rng.uniform(0, 1, m.ncells)
    # a source that drives half the cells below zero on every level's step
    ms = np.where(rng.uniform(size=m.ncells) < 0.5, -1.0e4, rng.uniform(-50.0, 50.0, m.ncells))
    out = sa_update(pm, nut_old, ms, dt, ts, subt, np.full(m.ncells, sentinel), cfg)
    sd = np.asarray(pm._keep["signdis"])
    ib = np.asarray(pm._keep["ibidx"])
    n_clamped = 0
    for l in range(Lm):
        s, e = normal_range(m, 0, l)
        g = np.arange(s, e)
        step = alpha * dt[0] * (1 << (Lm - 1 - l))     # dt0 * step_num[l]
        upd = nut_old[g] + ms[g] * step
        clamped = np.where(upd < 0.0, 0.0, upd)
        if l < Lm - 1:
            if not marched(Lm, l, ts):
                assert np.all(out[g] == sentinel), l
                continue
            exp = np.where(sd[g] > 0.0, clamped, nut_inf)       # initvar where signdis <= 0
        else:
            solid, isib = sd[g] < 0.0, ib[g] >= 0
            exp = np.where(~solid & ~isib, clamped, sentinel)   # IB and solid: not updated
            exp = np.where(~isib & ~(sd[g] > 0.0), nut_inf, exp)  # reset unless IB
        n_clamped += int(np.sum((upd < 0) & (exp == 0.0)))
        np.testing.assert_allclose(out[g], exp, rtol=1e-15, atol=0)
        assert np.array_equal(out[g] == 0.0, exp == 0.0)
    assert n_clamped > 100
    # the planted cells: a middle-level solid cell is updated, then reset
    if marched(Lm, 1, ts):
        assert out[planted["solid1"]] == nut_inf
        # a middle-level cell with an IB index is updated like any other
        gi = planted["ib1"]
        v = nut_old[gi] + ms[gi] * alpha * dt[0] * (1 << (Lm - 2))
        assert out[gi] == pytest.approx(max(v, 0.0), rel=1e-15)
    assert out[planted["zero_fine"]] == nut_inf
    assert out[planted["ib_neg"]] == sentinel       # IB cells are never reset
    # ghosts are never touched
    for b in range(m.ncases * Lm):
        s = int(m.blk_start[b] + m.blk_nnormal[b])
        assert np.all(out[s:int(m.blk_start[b + 1])] == sentinel)


def test_sa_update_level_time_steps(mesh):
    """a uniform source on a uniform working variable: each level advances by
    its own alpha*dt*2^(L-1-l), so the increments of levels l and l+1 differ by
    exactly a factor 2 (temporal sub-cycling)"""
    m = mesh
    Lm = m.nlevels
    cfg = O.cfg(nstage=3)
    dt = np.array([1.0e-3])
    out = sa_update(m, np.ones(m.ncells), np.full(m.ncells, 2.0), dt, 0, 1, np.zeros(m.ncells), cfg)
    inc = []
    for l in range(Lm):
        s, e = normal_range(m, 0, l)
        sel = np.arange(s, e)[np.asarray(m.signdis)[s:e] > 0.0]
        if l == Lm - 1:
            sel = sel[np.asarray(m.ibidx)[sel] < 0]
        d = out[sel] - 1.0
        assert np.ptp(d) == 0.0
        inc.append(d[0])
    for l in range(Lm - 1):
        assert inc[l] == pytest.approx(2.0 * inc[l + 1], rel=1e-14)
    assert inc[-1] == pytest.approx(2.0 * 0.4929 * 1.0e-3, rel=1e-14)


# -------------------------------------------------------------- SA BC fill
def test_sa_bc_fill_branches(mesh):
    m = mesh
    P = m.phys[0]
    rng = np.random.default_rng(3)
    fv = m.init_state()
    fv[:, 1] *= 1.0 + 0.1 * rng.uniform(-1, 1, m.ncells)
    fv[:, 2] += 0.05 * rng.uniform(-1, 1, m.ncells)
    fv[:, 5] *= 1.0 + 0.02 * rng.uniform(-1, 1, m.ncells)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    nut = rng.uniform(1.0, 9.0, m.ncells)
    before_fv, before_nut = fv.copy(), nut.copy()
    orc = O.Oracle(m)
    assert orc.sa_bc_fill(fv, nut, 0) == 0
    cell, ref, ty = (np.asarray(a) for a in m.dmg)
    seen = set()
    for g, r, t in zip(cell, ref, ty):
        seen.add(int(t))
        if t == -13:
            assert nut[g] == P.nut_inlet
        elif t == -12:
            assert nut[g] == nut[r]
            assert fv[g, 6] == fv[r, 6]            # E of the reference cell kept
            assert fv[g, 4] == P.outlet_p
            assert fv[g, 1] == fv[r, 1] and fv[g, 2] == fv[r, 2] and fv[g, 5] == fv[r, 5]
            assert fv[g, 0] == pytest.approx(P.outlet_p / (P.Rg * fv[r, 5]), rel=1e-15)
        elif t == -50:
            assert nut[g] == nut[r]
            assert np.array_equal(fv[g], fv[r])
    assert {-13, -12, -50} <= seen
    # normal cells are never written
    normal = np.concatenate([np.arange(*normal_range(m, 0, l)) for l in range(m.nlevels)])
    assert np.array_equal(nut[normal], before_nut[normal])
    assert np.array_equal(fv[normal], before_fv[normal])


def test_sa_bc_fill_skips_unmarched_levels(mesh):
    m = mesh
    rng = np.random.default_rng(4)
    fv = m.init_state()
    nut = rng.uniform(1.0, 9.0, m.ncells)
    nut0 = nut.copy()
    orc = O.Oracle(m)
    assert orc.sa_bc_fill(fv, nut, 1) == 0     # ts = 1: only the finest level is marched
    cell = np.asarray(m.dmg[0])
    lvl = m.level_of()[cell]
    Lm = m.nlevels
    assert np.array_equal(nut[cell[lvl < Lm - 1]], nut0[cell[lvl < Lm - 1]])
    if np.any(lvl == Lm - 1):
        assert not np.array_equal(nut[cell[lvl == Lm - 1]], nut0[cell[lvl == Lm - 1]])
