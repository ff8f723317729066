"""Pins for the FP64 oracle (oracle/oracle.c), CPU only.

The reference ships no tests, fixtures or golden vectors, and its build
needs MPI (treated as unbuildable), so every oracle function is pinned here
by mathematics that does not restate its formula:

  limiter       van Leer values at known r, linear data -> exact midpoint,
                extremum -> first order, left/right mirror symmetry
  Roe flux      consistency F(W,W) = F_phys(W) n A; first-order flux equals
                1/2(F_L+F_R)A - A sum|lambda_k| alpha_k r_k with the wave
                strengths from a linear solve on the Roe eigenvectors (f=1
                and the all-speed eigenvalues for f<1); x-mirror symmetry;
                x/y rotation symmetry
  face->cell    uniform state -> source exactly 0 (free-stream
                preservation); per-level telescoping of sum(source)
  viscous       linear u, v, T fields give the exact stress / heat flux
                with the Sutherland law in closed form (mu(1) = 1)
  stage update  the multistage scheme on dU/dt = lambda U reproduces
                1 + a_S z (1 + a_{S-1} z (...)), z = lambda dt
  CFL           uniform state closed form
  BC            isentropic total-state inlet reproduces the freestream;
                outflow copies its reference cell
  AMR transfer  restriction conserves mass and momentum; prolongation
                conserves the coarse conserved state and is exact for
                linear fields
  IB fill       bilinear weights reproduce linear fields at the image point;
                LinearDist scaling d/(d+h)
  force         uniform pressure on a closed body gives zero net force
  schedule      the level/transfer tables of the temporal sub-cycling
"""
import ctypes as C

import numpy as np
import pytest

import ibgpu
import oracle_ffi as O

L = O.lib()
ptr = O.ptr


def phys_for(Ma=0.5, Re=2000.0):
    m = ibgpu.Mesh(ibgpu.tiny_case(Ma=Ma, Re=Re))
    return m.phys[0]


# ------------------------------------------------------------------ limiter
@pytest.mark.parametrize("r,phi", [(0.25, 0.4), (0.5, 2.0 / 3.0), (1.0, 1.0), (2.0, 4.0 / 3.0),
                                   (4.0, 1.6)])
def test_vanleer_known_values(r, phi):
    # q_i = 0, qi = 1, qi1 = 1 + r  ->  dqi = 1, dqii = r, q_L = 1 + phi/2
    got = L.orc_muscl_left(0.0, 1.0, 1.0 + r, 0)
    assert got == pytest.approx(1.0 + 0.5 * phi, rel=1e-15)


def test_limiter_linear_data_gives_midpoint():
    for a, s in [(0.3, 0.1), (-2.0, 0.7), (5.0, -0.25)]:
        q = [a - s, a, a + s, a + 2 * s]
        assert L.orc_muscl_left(q[0], q[1], q[2], 0) == pytest.approx(a + 0.5 * s, rel=1e-15)
        assert L.orc_muscl_right(q[1], q[2], q[3], 0) == pytest.approx(a + 0.5 * s, rel=1e-15)


def test_limiter_extremum_is_first_order():
    assert L.orc_muscl_left(1.0, 2.0, 1.5, 0) == 2.0
    assert L.orc_muscl_right(1.5, 2.0, 1.0, 0) == 2.0
    assert L.orc_muscl_left(3.0, 3.0, 5.0, 0) == 3.0       # dqi = 0
    assert L.orc_muscl_left(3.0, 3.0, 3.0, 0) == 3.0       # 0/0


def test_limiter_mirror_symmetry():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b, c = rng.normal(size=3)
        assert L.orc_muscl_right(a, b, c, 0) == L.orc_muscl_left(c, b, a, 0)


def test_vanalbada_known_value():
    # phi(r) = (r^2 + r)/(r^2 + 1): phi(1) = 1, phi(2) = 6/5
    assert L.orc_muscl_left(0.0, 1.0, 3.0, 1) == pytest.approx(1.0 + 0.5 * 1.2, rel=1e-15)


@pytest.mark.parametrize("r,phi", [(0.5, 0.6), (1.0, 1.0), (3.0, 1.2), (0.25, 5.0 / 17.0)])
def test_vanalbada_values_and_midpoint(r, phi):
    # (r^2 + r)/(r^2 + 1) at r = 1/2, 1, 3, 1/4; the right state mirrors the left one
    assert L.orc_muscl_left(0.0, 1.0, 1.0 + r, 1) == pytest.approx(1.0 + 0.5 * phi, rel=1e-15)
    assert L.orc_muscl_right(1.0 + r, 1.0, 0.0, 1) == pytest.approx(1.0 + 0.5 * phi, rel=1e-15)
    for a, sl in [(0.3, 0.1), (-2.0, 0.7)]:
        q = [a - sl, a, a + sl]
        assert L.orc_muscl_left(q[0], q[1], q[2], 1) == pytest.approx(a + 0.5 * sl, rel=1e-15)


# ----------------------------------------------------------------- Roe flux
def fv_of(P, r, u, v, w, p):
    T = p / (r * P.Rg)
    return np.array([r, u, v, w, p, T, P.Cv * T + 0.5 * (u * u + v * v + w * w)])


def roe(P, LL, Lc, Rc, RR, d=0, area=0.7, lim=0):
    out = np.zeros(5)
    L.orc_roe_flux(ptr(LL), ptr(Lc), ptr(Rc), ptr(RR), d, area, C.byref(P), lim, ptr(out))
    return out


def phys_flux(P, q, d):
    r, u, v, w, p = q[:5]
    n = np.eye(3)[d]
    un = np.dot([u, v, w], n)
    E = q[6]
    return np.array([r * un, r * u * un + p * n[0], r * v * un + p * n[1], r * w * un + p * n[2],
                     (r * E + p) * un])


def test_roe_consistency():
    P = phys_for()
    for d in range(3):
        q = fv_of(P, 1.2, 0.4, -0.3, 0.2, P.initvar[4] * 1.1)
        got = roe(P, q, q, q, q, d, 0.7)
        assert np.allclose(got, phys_flux(P, q, d) * 0.7, rtol=1e-14, atol=1e-15)


def roe_dissipation_eig(P, qL, qR, f_all_speed=None):
    """sum |lambda_k| alpha_k r_k from a linear solve on the Roe eigenvectors"""
    g = P.gama
    rl, rr = qL[0], qR[0]
    HL = qL[6] + qL[4] / rl
    HR = qR[6] + qR[4] / rr
    s = np.sqrt(rr / rl)
    u = (qL[1:4] + qR[1:4] * s) / (1 + s)
    H = (HL + HR * s) / (1 + s)
    c = np.sqrt((g - 1) * (H - 0.5 * np.dot(u, u)))
    un = u[0]
    R = np.zeros((5, 5))
    R[:, 0] = [1, u[0] - c, u[1], u[2], H - un * c]
    R[:, 1] = [1, u[0], u[1], u[2], 0.5 * np.dot(u, u)]
    R[:, 2] = [0, 0, 1, 0, u[1]]
    R[:, 3] = [0, 0, 0, 1, u[2]]
    R[:, 4] = [1, u[0] + c, u[1], u[2], H + un * c]

    def cons(q):
        return np.array([q[0], q[0] * q[1], q[0] * q[2], q[0] * q[3], q[0] * q[6]])

    alpha = np.linalg.solve(R, cons(qR) - cons(qL))
    if f_all_speed is None:
        lam = np.abs([un - c, un, un, un, un + c])
    else:
        f = f_all_speed
        l0 = np.sqrt(4 * c * c * f + (1 - f) ** 2 * un * un)
        lam = np.abs([0.5 * ((1 + f) * un - l0), un, un, un, 0.5 * ((1 + f) * un + l0)])
    return R @ (lam * alpha)


@pytest.mark.parametrize("Ma_inf", [1.0, 0.3])
def test_roe_dissipation_matches_eigendecomposition(Ma_inf):
    P = phys_for()
    P.Ma = Ma_inf
    rng = np.random.default_rng(7)
    for _ in range(20):
        qL = fv_of(P, rng.uniform(0.8, 1.2), rng.uniform(0.2, 0.9), rng.uniform(-0.2, 0.2),
                   rng.uniform(-0.1, 0.1), P.initvar[4] * rng.uniform(0.8, 1.2))
        qR = fv_of(P, rng.uniform(0.8, 1.2), rng.uniform(0.2, 0.9), rng.uniform(-0.2, 0.2),
                   rng.uniform(-0.1, 0.1), P.initvar[4] * rng.uniform(0.8, 1.2))
        A = 0.37
        # LL = L and RR = R: first-order reconstruction
        got = roe(P, qL, qL, qR, qR, 0, A)
        # face Mach of the reference's all-speed switch
        TL, TR = qL[5], qR[5]
        Ma_face = abs(0.5 * (qL[1] + qR[1]) / np.sqrt(P.gama * P.Rg * 0.5 * (TL + TR)))
        f = min(max(Ma_face, Ma_inf), 1.0)
        D = roe_dissipation_eig(P, qL, qR, None if f == 1.0 else f)
        want = 0.5 * (phys_flux(P, qL, 0) + phys_flux(P, qR, 0)) * A - A * D
        assert np.allclose(got, want, rtol=1e-11, atol=1e-14)


def test_roe_mirror_and_rotation_symmetry():
    P = phys_for()
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = [fv_of(P, rng.uniform(0.8, 1.2), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 0.0,
                   P.initvar[4] * rng.uniform(0.8, 1.2)) for _ in range(4)]

        def mx(a):
            b = a.copy()
            b[1] = -b[1]
            return b

        f = roe(P, *q, d=0)
        g = roe(P, mx(q[3]), mx(q[2]), mx(q[1]), mx(q[0]), d=0)
        # mass and energy flip sign, x-momentum keeps it, y-momentum flips
        assert np.allclose([f[0], f[1], f[2], f[4]], [-g[0], g[1], -g[2], -g[4]], rtol=1e-12, atol=1e-15)

        def sw(a):
            b = a.copy()
            b[1], b[2] = a[2], a[1]
            return b

        fy = roe(P, *[sw(x) for x in q], d=1)
        assert np.allclose([fy[0], fy[2], fy[1], fy[4]], [f[0], f[1], f[2], f[4]], rtol=1e-12, atol=1e-15)


# ------------------------------------------------------ face -> cell source
@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case())


def test_freestream_preservation(tiny):
    orc = O.Oracle(tiny, O.cfg(temporal_refine=0))
    fv = tiny.init_state()
    ff = orc.face_flux(fv, 0)
    vs = orc.viscous_flux(fv, 0)
    src = orc.source(ff, vs, 0)
    assert np.count_nonzero(src) == 0
    assert np.all(vs == 0.0)


def test_source_telescopes_per_level(tiny):
    rng = np.random.default_rng(5)
    orc = O.Oracle(tiny, O.cfg(temporal_refine=0))
    fv = tiny.init_state()
    fv[:, 0] *= 1 + 0.05 * rng.standard_normal(tiny.ncells)
    fv[:, 4] *= 1 + 0.05 * rng.standard_normal(tiny.ncells)
    P = tiny.phys[0]
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    ff = orc.face_flux(fv, 0)
    vs = np.zeros_like(ff)
    src = orc.source(ff, vs, 0)
    normal = tiny.kind == 0
    lo, hi = tiny.face_cells[:, 0], tiny.face_cells[:, 1]
    for b in range(tiny.nlevels):
        s, e = tiny.blk_start[b], tiny.blk_start[b + 1]
        fs, fe = tiny.blk_face_start[b], tiny.blk_face_start[b + 1]
        cells = np.arange(s, e)
        total = src[cells][normal[cells]].sum(axis=0)
        fl, fh = lo[fs:fe], hi[fs:fe]
        inflow = ff[fs:fe][(normal[fh]) & (~normal[fl])].sum(axis=0)
        outflow = ff[fs:fe][(normal[fl]) & (~normal[fh])].sum(axis=0)
        want = inflow - outflow
        scale = np.abs(ff[fs:fe]).sum(axis=0) + 1e-300
        assert np.all(np.abs(total - want) <= 1e-12 * scale)


# ------------------------------------------------------------------ viscous
def test_viscous_exact_for_linear_fields(tiny):
    P = tiny.phys[0]
    orc = O.Oracle(tiny, O.cfg(temporal_refine=0))
    x = tiny.centers
    gu, gv, gT = 0.3, -0.2, 0.05
    fv = tiny.init_state()
    fv[:, 1] = 0.5 + gu * x[:, 1]          # u = u0 + gu y
    fv[:, 2] = 0.1 + gv * x[:, 0]          # v = v0 + gv x
    fv[:, 5] = 1.0 + gT * x[:, 0]          # T linear in x
    fv[:, 0] = 1.0
    fv[:, 4] = fv[:, 0] * P.Rg * fv[:, 5]
    vs = orc.viscous_flux(fv, 0)
    S = P.S_over_T_ref

    def mu(T):
        return T ** 1.5 * (1 + S) / (T + S)

    lo, hi = tiny.face_cells[:, 0], tiny.face_cells[:, 1]
    d = tiny.face_dir
    for b in range(tiny.nlevels):
        fs, fe = tiny.blk_face_start[b], tiny.blk_face_start[b + 1]
        h = tiny.blk_h[b]
        for f in range(fs, fe, 7):
            Tf = 0.5 * (fv[lo[f], 5] + fv[hi[f], 5])
            m = mu(Tf)
            tauxy = m / P.Re * (gu + gv)
            fu = 0.5 * (fv[lo[f], 1] + fv[hi[f], 1])
            fvv = 0.5 * (fv[lo[f], 2] + fv[hi[f], 2])
            if d[f] == 0:
                A = h[1] * h[2]
                want = [0.0, 0.0, tauxy * A, fvv * tauxy * A + P.qk * m * gT * A]
            else:
                A = h[2] * h[0]
                want = [tauxy * A, 0.0, fu * tauxy * A]
                want = [want[0], want[1], want[2]]
                got = vs[f, [1, 2, 4]]
                assert np.allclose(got, want, rtol=1e-9, atol=1e-15)
                continue
            got = vs[f, [1, 2, 4]]
            assert np.allclose(got, want[1:], rtol=1e-9, atol=1e-15)


def test_sutherland_unity_at_reference_temperature(tiny):
    # at T = 1 the Sutherland factor is exactly 1: mu(1) = (1+S)/(1+S)
    P = tiny.phys[0]
    orc = O.Oracle(tiny, O.cfg(temporal_refine=0))
    x = tiny.centers
    fv = tiny.init_state()
    g = 0.25
    fv[:, 1] = 0.4 + g * x[:, 1]
    vs = orc.viscous_flux(fv, 0)
    f = np.nonzero(tiny.face_dir == 1)[0][5]
    b = np.searchsorted(tiny.blk_face_start, f, side="right") - 1
    h = tiny.blk_h[b]
    assert vs[f, 1] == pytest.approx(1.0 / P.Re * g * h[2] * h[0], rel=1e-12)


# ------------------------------------------------------------- stage update
@pytest.mark.parametrize("nstage", [1, 3, 4])
def test_multistage_linear_ode(tiny, nstage):
    cfg = O.cfg(nstage=nstage, temporal_refine=0)
    fv = tiny.init_state()
    N = tiny.ncells
    cv_old = np.zeros((N, 5))
    L.orc_old_cons(tiny.soa_ptr, C.byref(cfg), ptr(fv), 0, ptr(cv_old))
    dt = np.array([1e-3])
    lam = -40.0
    vol = np.zeros(N)
    for b in range(tiny.nlevels):
        h = tiny.blk_h[b]
        vol[tiny.blk_start[b]:tiny.blk_start[b + 1]] = h[0] * h[1] * h[2]
    cv_new = np.zeros((N, 5))
    cur = cv_old.copy()
    for k in range(nstage):
        src = lam * cur * vol[:, None]
        src[:, 3] = 0.0
        code = L.orc_update(tiny.soa_ptr, tiny.phys, C.byref(cfg), ptr(cv_old), ptr(src), ptr(dt), 0, k,
                            ptr(cv_new), ptr(fv), None)
        assert code == 0
        cur = np.zeros((N, 5))
        L.orc_old_cons(tiny.soa_ptr, C.byref(cfg), ptr(fv), 0, ptr(cur))
    z = lam * dt[0]
    a = {1: [1.0], 3: [0.1918, 0.4929, 1.0], 4: [0.1084, 0.2602, 0.5052, 1.0]}[nstage]
    amp = 1.0
    for ak in a:
        amp = 1.0 + ak * z * amp
    # cells recovered from U (signdis > 0, not IB, normal) follow the polynomial
    upd = (tiny.kind == 0) & (tiny.signdis > 0) & (tiny.ibidx < 0)
    ratio = cur[upd, 0] / cv_old[upd, 0]
    assert np.allclose(ratio, amp, rtol=1e-13)


# ---------------------------------------------------------------------- CFL
def test_cfl_uniform_closed_form(tiny):
    P = tiny.phys[0]
    orc = O.Oracle(tiny)
    fv = tiny.init_state()
    dt = np.array([1e30])
    assert orc.compute_dt(fv, dt) == 0
    c = np.sqrt(P.gama * P.Rg * P.initvar[5])
    u = P.initvar[1:4]
    best = min(P.cfl / sum((abs(u[d]) + c) / tiny.blk_h[b][d] for d in range(3))
               for b in range(tiny.nlevels))
    assert dt[0] == pytest.approx(best, rel=1e-14)


# ----------------------------------------------------------------------- BC
def test_inlet_is_isentropic_freestream(tiny):
    orc = O.Oracle(tiny)
    fv = tiny.init_state()
    ref = fv.copy()
    assert orc.bc_fill(fv, 0) == 0
    cells, refc, types = tiny.dmg
    assert set(np.unique(types)) <= {-13, -12, -50}
    assert np.allclose(fv[cells, :5], ref[cells, :5], rtol=1e-13, atol=1e-15)


def test_outflow_copies_reference_cell(tiny):
    orc = O.Oracle(tiny)
    rng = np.random.default_rng(2)
    fv = tiny.init_state()
    fv[:, 1] += 0.1 * rng.standard_normal(tiny.ncells)
    assert orc.bc_fill(fv, 0) == 0
    cells, refc, types = tiny.dmg
    sel = types == -50
    edge = tiny.kind[refc] != 1
    assert np.array_equal(fv[cells[sel & edge]], fv[refc[sel & edge]])


# ------------------------------------------------------------- AMR transfer
def smooth_state(mesh, lin=False):
    P = mesh.phys[0]
    x = mesh.centers
    fv = mesh.init_state()
    if lin:
        fv[:, 0] = 1.0
        fv[:, 1] = 0.5 + 0.1 * x[:, 0] - 0.05 * x[:, 1]
        fv[:, 2] = 0.02 * x[:, 0] + 0.03 * x[:, 1]
    else:
        fv[:, 0] = 1.0 + 0.1 * np.sin(x[:, 0]) * np.cos(0.7 * x[:, 1])
        fv[:, 1] = 0.6 + 0.1 * np.cos(x[:, 1])
        fv[:, 2] = 0.05 * np.sin(x[:, 0])
    fv[:, 4] = P.initvar[4] * (1 + 0.05 * np.cos(0.5 * x[:, 0]))
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    return fv


def cons(P, q):
    return np.stack([q[:, 0], q[:, 0] * q[:, 1], q[:, 0] * q[:, 2], q[:, 0] * q[:, 6]], axis=1)


def test_restriction_conserves_and_prolongation_conserves(tiny):
    P = tiny.phys[0]
    orc = O.Oracle(tiny)
    fv = smooth_state(tiny)
    before = fv.copy()
    for ts in range(1 << (tiny.nlevels - 1)):
        orc.interface_exchange(fv, ts)
    # restriction: coarse = mean of the 4 sons in (roe, roe u, roe v, roe e)
    sons = tiny.res_fi[:, [0, 2, 4, 6]]
    for i in range(tiny.nres):
        want = cons(P, fv[sons[i]]).mean(axis=0)
        got = cons(P, fv[tiny.res_ci[i]][None, :])[0]
        assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    # prolongation: mean over the sons of the conserved state = coarse state
    psons = tiny.pro_fi[:, [0, 2, 4, 6]]
    for i in range(tiny.npro):
        got = cons(P, fv[psons[i]]).mean(axis=0)
        want = cons(P, fv[tiny.pro_ci[i]][None, :])[0]
        assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    assert not np.array_equal(before, fv)


def test_prolongation_exact_for_linear_conserved_fields():
    mesh = ibgpu.Mesh(ibgpu.tiny_case())
    P = mesh.phys[0]
    x = mesh.centers
    # conserved variables linear in x, y on every cell
    r = 1.0 + 0.05 * x[:, 0] + 0.02 * x[:, 1]
    mu_ = 0.5 + 0.03 * x[:, 0] - 0.01 * x[:, 1]
    mv = 0.01 * x[:, 0] + 0.02 * x[:, 1]
    me = 2.0 * P.initvar[6] + 0.02 * x[:, 0] + 0.01 * x[:, 1]
    fv = mesh.init_state()
    fv[:, 0] = r
    fv[:, 1] = mu_ / r
    fv[:, 2] = mv / r
    fv[:, 6] = me / r
    fv[:, 4] = (fv[:, 6] - 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)) * (P.gama - 1) * r
    fv[:, 5] = fv[:, 4] / (r * P.Rg)
    orc = O.Oracle(mesh)
    pro = np.unique(mesh.pro_fi[:, [0, 2, 4, 6]].ravel())
    for ts in range(1 << (mesh.nlevels - 1)):
        orc.interface_exchange(fv, ts)
    got = cons(P, fv[pro])
    want = np.stack([r[pro], mu_[pro], mv[pro], me[pro]], axis=1)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------------ IB fill
def test_ib_bilinear_reproduces_linear_fields_and_lineardist(tiny):
    P = tiny.phys[0]
    m = tiny.soa
    x = tiny.centers
    fv = tiny.init_state()
    fv[:, 1] = 0.7 + 0.2 * x[:, 0] - 0.1 * x[:, 1]
    fv[:, 2] = 0.1 + 0.05 * x[:, 0] + 0.3 * x[:, 1]
    fv[:, 5] = 1.0 + 0.02 * x[:, 0] + 0.01 * x[:, 1]
    fv[:, 4] = P.initvar[4] * (1.0 + 0.03 * x[:, 0] - 0.02 * x[:, 1])
    fv[:, 0] = fv[:, 4] / (P.Rg * fv[:, 5])
    orc = O.Oracle(tiny)
    new = fv.copy()
    assert orc.ib_fill(new) == 0
    nv = np.ctypeslib.as_array(m.ib_nv, (m.nib, 3))
    for i in range(tiny.nib):
        c = tiny.ib_cell[i]
        img = x[c] + nv[i] * tiny.ib_hgdis[i]
        u_img = 0.7 + 0.2 * img[0] - 0.1 * img[1]
        p_img = P.initvar[4] * (1.0 + 0.03 * img[0] - 0.02 * img[1])
        T_img = 1.0 + 0.02 * img[0] + 0.01 * img[1]
        sd = tiny.signdis[c]
        r0 = sd / (sd + tiny.ib_hgdis[i])
        assert new[c, 1] == pytest.approx(u_img * r0, rel=1e-12, abs=1e-14)
        assert new[c, 4] == pytest.approx(p_img, rel=1e-12)
        assert new[c, 5] == pytest.approx(T_img, rel=1e-12)


# -------------------------------------------------------------------- force
def test_uniform_pressure_gives_zero_net_force(tiny):
    orc = O.Oracle(tiny)
    fv = tiny.init_state()
    fv[:, 1:4] = 0.0
    fv[:, 6] = tiny.phys[0].Cv * fv[:, 5]
    F = orc.forces(fv)
    scale = tiny.phys[0].initvar[4] * 2.0 * 0.1
    assert np.all(np.abs(F[:2]) < 1e-13 * scale * 10)


# ----------------------------------------------------------------- schedule
def test_temporal_subcycling_tables():
    Lv = 3
    marched = [[L.orc_marched(Lv, l, ts, 1) for l in range(Lv)] for ts in range(8)]
    assert marched == [[1, 1, 1], [0, 0, 1], [0, 1, 1], [0, 0, 1],
                       [1, 1, 1], [0, 0, 1], [0, 1, 1], [0, 0, 1]]
    due = [[L.orc_transfer_due(Lv, l, ts, 1) for l in range(Lv - 1)] for ts in range(8)]
    # pair (1,2) at odd steps, pair (0,1) at ts % 4 == 2
    assert due == [[0, 0], [0, 1], [1, 0], [0, 1], [0, 0], [0, 1], [1, 0], [0, 1]]
    assert [L.orc_step_num(Lv, l, 1) for l in range(Lv)] == [4.0, 2.0, 1.0]
