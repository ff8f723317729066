"""Pins for the Spalart-Allmaras part of the FP64 oracle, CPU only.

The reference's TURBULENCE build (Turbulence.cpp, NS_Solver.cpp:632-840)
ships no tests or fixtures; each oracle function is pinned here by facts
that do not restate its formula:

  Viseddy2mut     X = cv1 gives fv1 = 1/2 exactly; X < 0 is clamped; X -> inf
                  gives mu_t -> roe nu
  wall fits       y+(Re_y) and u+(y+) are inverse fits of one law of the
                  wall: y+(y+ u+(y+)) = y+ within 1 % over 0.5..2000;
                  u+ = y+ in the sublayer; u+ ~ ln(y+)/0.41 + B in the log
                  layer with B in [4, 6.5]
  SA source       log-layer equilibrium nu = kappa u_tau d, S = u_tau/(kappa d),
                  div(nu grad nu)/sigma = (kappa u_tau)^2/sigma: production =
                  destruction + diffusion (the definition of cw1); fw(r=1) = 1;
                  the advection term is -u.grad(nu) (linear in u); the face
                  term passes through; the cb2 term is even in grad(nu); the
                  r >= 10 cap gives fw = (1 + cw3^6)^(1/6)
  viscous + SA    nu = 0 reproduces the laminar viscous flux exactly and a
                  zero face flux of nu; nu linear in x gives the exact face flux
                  (nu_f + mu/roe) dnu/dx / sigma * A
  mesh source     one cell of the mesh loop equals the pure cell function on
                  the gathered neighbours
  IB wall model   with the image point in the viscous sublayer (u+ = y+) the
                  wall function reproduces the laminar LinearDist velocity
                  (to the 0.2 % offset of the y+ fit); df = 0 keeps T = T_hg
  AMR transfer    restriction of nu is the mean of the sons; prolongation is
                  exact for nu linear in x,y
"""
import ctypes as C
import math

import numpy as np
import pytest

import ibgpu
import oracle_ffi as O

L = O.lib()
ptr = O.ptr
KAPPA, SIGMA, CB1, CB2, CV1, CW2, CW3 = 0.41, 2.0 / 3.0, 0.1355, 0.622, 7.1, 0.3, 2.0


def vec(*a):
    return np.array(a, dtype=np.float64)


def cell_src(u=0.0, v=0.0, w=0.0, T=1.0, roe=1.0, nu=1.0, sd=1.0, g=(0, 0, 0), gu=(0, 0, 0),
             gv=(0, 0, 0), gw=(0, 0, 0), k=(1.0, 1.0, 1.0), SoT=0.5, Re=1.0, mutf=0.0):
    arrs = [vec(*g), vec(*gu), vec(*gv), vec(*gw)]
    return L.orc_sa_cell_source(u, v, w, T, roe, nu, sd, ptr(arrs[0]), ptr(arrs[1]),
                                ptr(arrs[2]), ptr(arrs[3]), k[0], k[1], k[2], SoT, Re, mutf)


# ------------------------------------------------------------ Viseddy2mut
def test_mut_half_at_cv1():
    mu, roe = 0.7, 1.3
    nu = CV1 * mu / roe                      # X = cv1 -> fv1 = 1/2
    assert L.orc_sa_mut(nu, roe, mu) == pytest.approx(0.5 * roe * nu, rel=1e-14)


def test_mut_limits():
    assert L.orc_sa_mut(-2.0, 1.0, 1.0) == 0.0
    assert L.orc_sa_mut(1e6, 1.0, 1.0) == pytest.approx(1e6, rel=1e-9)


# ------------------------------------------------------------ wall-function fits
@pytest.mark.parametrize("yp", [0.5, 1, 2, 3, 5, 8, 10, 20, 40, 100, 200, 300, 500, 1000, 2000])
def test_wall_fits_are_inverse(yp):
    up = L.orc_sa_uplus(yp)
    assert L.orc_sa_yplus(yp * up) == pytest.approx(yp, rel=1e-2)


def test_sublayer_and_log_law():
    for yp in (0.1, 1.0, 2.9):
        assert L.orc_sa_uplus(yp) == yp
    for yp in (100.0, 300.0, 1000.0):
        B = L.orc_sa_uplus(yp) - math.log(yp) / KAPPA
        assert 4.0 <= B <= 6.5, (yp, B)


# ------------------------------------------------------------ SA source
def log_layer(ut=1.0, d=2.0, h=0.01, mu_scale=1e-3):
    """Cell in the log layer of a wall at distance d (Re = 1)."""
    nu = KAPPA * ut * d
    S = ut / (KAPPA * d)
    # the reference's vorticity magnitude is sqrt(2)|curl u / 2| = |du/dy|/sqrt(2)
    dudy = math.sqrt(2.0) * S
    # small laminar viscosity: T such that mu(T) << nu (X large: ft2 -> 0)
    T = mu_scale
    return dict(nu=nu, sd=d, gu=(0.0, dudy * h, 0.0), g=(0.0, KAPPA * ut * h, 0.0),
                k=(1.0 / h, 1.0 / h, 1.0 / h), T=T, SoT=0.5, Re=1.0,
                mutf=(KAPPA * ut) ** 2 / SIGMA), CB1 * S * nu


def test_log_layer_equilibrium():
    kw, prod = log_layer()
    s = cell_src(**kw)
    assert abs(s) < 1e-2 * prod, (s, prod)
    # dropping the face term (diffusion) breaks the balance by (kappa u)^2/sigma
    kw2 = dict(kw, mutf=0.0)
    assert cell_src(**kw2) == pytest.approx(s - kw["mutf"], rel=1e-12)


def test_fw_one_at_r_one():
    # in the log layer r = nu/(S kappa^2 d^2) = 1: the destruction term equals
    # cw1 (nu/d)^2 with cw1 = cb1/kappa^2 + (1+cb2)/sigma
    kw, prod = log_layer()
    cw1 = CB1 / KAPPA ** 2 + (1 + CB2) / SIGMA
    dest = cw1 * (kw["nu"] / kw["sd"]) ** 2
    # production + cb2 + face terms - destruction = s
    cb2t = CB2 * (KAPPA * 1.0) ** 2 / SIGMA
    s = cell_src(**kw)
    assert s == pytest.approx(prod + cb2t + kw["mutf"] - dest, abs=1e-2 * prod)


def test_advection_linear_in_velocity():
    # nu = 0 removes production and destruction; some vorticity keeps r = 0
    # finite (with nu = 0 and no vorticity the reference's r is 0/0)
    base = dict(nu=0.0, g=(0.3, -0.2, 0.1), gu=(0.0, 0.5, 0.0), k=(2.0, 3.0, 4.0), T=1.0, sd=0.5)
    s0 = cell_src(**base)
    s1 = cell_src(u=1.5, v=-0.5, w=2.0, **base)
    grad = np.array([0.3 * 2.0, -0.2 * 3.0, 0.1 * 4.0])
    assert s1 - s0 == pytest.approx(-(1.5 * grad[0] - 0.5 * grad[1] + 2.0 * grad[2]), rel=1e-12)


def test_cb2_term_even_and_face_term_passthrough():
    base = dict(nu=0.0, gu=(0.0, 0.5, 0.0), k=(2.0, 3.0, 4.0), T=1.0, sd=0.5, Re=7.0)
    sp = cell_src(g=(0.3, -0.2, 0.1), **base)
    sm = cell_src(g=(-0.3, 0.2, -0.1), **base)
    assert sp == pytest.approx(sm, rel=1e-14)
    grad2 = (0.3 * 2) ** 2 + (0.2 * 3) ** 2 + (0.1 * 4) ** 2
    assert sp == pytest.approx(CB2 * grad2 / SIGMA / 7.0, rel=1e-12)
    assert cell_src(g=(0.3, -0.2, 0.1), mutf=0.25, **base) - sp == pytest.approx(0.25, rel=1e-12)


def test_destruction_cap():
    # no vorticity, tiny distance: r capped at 10 -> fw = (1+cw3^6)^(1/6)
    nu, d = 1.0, 1e-3
    s = cell_src(nu=nu, sd=d, T=1e-6, SoT=0.5)
    fw = (1 + CW3 ** 6) ** (1.0 / 6.0)
    g = 10 + CW2 * (10 ** 6 - 10)
    fw_exact = g * ((1 + CW3 ** 6) / (g ** 6 + CW3 ** 6)) ** (1.0 / 6.0)
    assert fw_exact == pytest.approx(fw, rel=1e-6)
    cw1 = CB1 / KAPPA ** 2 + (1 + CB2) / SIGMA
    # production with S = nu fv2/(kappa d)^2 (fv2 ~ 1/(1+X) tiny for X large) is
    # negligible next to the destruction cw1 fw (nu/d)^2
    assert s == pytest.approx(-cw1 * fw * (nu / d) ** 2, rel=1e-3)


# ------------------------------------------------------------ mesh level
@pytest.fixture(scope="module")
def tiny():
    return ibgpu.Mesh(ibgpu.tiny_case(Re=500.0))


def seeded(mesh, seed=7):
    rng = np.random.default_rng(seed)
    fv = mesh.init_state()
    x = mesh.centers
    fv[:, 1] += 0.05 * np.sin(1.3 * x[:, 1] + 0.2)
    fv[:, 2] += 0.05 * np.cos(0.7 * x[:, 0])
    P = mesh.phys[0]
    fv[:, 5] = fv[:, 4] / (fv[:, 0] * P.Rg)
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2 + fv[:, 3] ** 2)
    nut = P.nut_inf * (1 + 0.3 * np.sin(0.9 * x[:, 0]) * np.cos(1.1 * x[:, 1])) + 0.01 * rng.random(mesh.ncells)
    return fv, nut


def test_viscous_flux_sa_zero_nu_is_laminar(tiny):
    fv, _ = seeded(tiny)
    orc = O.Oracle(tiny)
    lam = orc.viscous_flux(fv, 0)
    sa, mf = orc.viscous_flux_sa(fv, np.zeros(tiny.ncells), 0)
    assert np.array_equal(lam, sa)
    assert np.all(mf == 0.0)


def test_mutflux_exact_for_linear_nu(tiny):
    fv, _ = seeded(tiny)
    x = tiny.centers
    a = 0.37
    nut = 2.0 + a * x[:, 0]
    orc = O.Oracle(tiny)
    _, mf = orc.viscous_flux_sa(fv, nut, 0)
    P = tiny.phys[0]
    fc, fd = tiny.face_cells, tiny.face_dir
    chk = 0
    for f in range(tiny.nfaces):
        if fd[f] != 0:
            continue
        lo, hi = fc[f]
        b = int(np.searchsorted(tiny.blk_face_start, f, side="right") - 1)
        h = tiny.blk_h[b]
        T = 0.5 * (fv[lo, 5] + fv[hi, 5])
        mu = T ** 1.5 * (1 + P.S_over_T_ref) / (T + P.S_over_T_ref)
        roe = 0.5 * (fv[lo, 0] + fv[hi, 0])
        nuf = 0.5 * (nut[lo] + nut[hi])
        want = (nuf + mu / roe) * ((nut[hi] - nut[lo]) / h[0]) / SIGMA * (h[1] * h[2])
        assert mf[f] == pytest.approx(want, rel=1e-10)
        assert (nut[hi] - nut[lo]) / h[0] == pytest.approx(a, rel=1e-9)
        chk += 1
    assert chk > 100


def test_mesh_source_matches_cell_function(tiny):
    fv, nut = seeded(tiny)
    orc = O.Oracle(tiny)
    _, mf = orc.viscous_flux_sa(fv, nut, 0)
    ms = orc.mut_source(fv, nut, mf, 0)
    P = tiny.phys[0]
    nb = tiny.nb.reshape(-1, 3, 3, 3)
    rng = np.random.default_rng(3)
    L_ = tiny.nlevels
    checked = 0
    for b in range(tiny.ncases * L_):
        s, n = tiny.blk_start[b], tiny.blk_nnormal[b]
        h = tiny.blk_h[b]
        for g in rng.choice(np.arange(s, s + n), size=min(n, 20), replace=False):
            if tiny.signdis[g] < 0:
                assert ms[g] == 0.0
                continue
            xl, xr = nb[g, 0, 1, 1], nb[g, 2, 1, 1]
            yl, yr = nb[g, 1, 0, 1], nb[g, 1, 2, 1]
            zl, zr = nb[g, 1, 1, 0], nb[g, 1, 1, 2]
            d = lambda arr: ((arr[xr] - arr[xl]) / 2, (arr[yr] - arr[yl]) / 2, (arr[zr] - arr[zl]) / 2)
            tot = 0.0
            for di in range(2):
                for f0 in range(2):
                    tot += (-1.0 if f0 == 0 else 1.0) * mf[tiny.cell_faces[g, di * 2 + f0]]
            vol = h[0] * h[1] * h[2]
            want = cell_src(u=fv[g, 1], v=fv[g, 2], w=fv[g, 3], T=fv[g, 5], roe=fv[g, 0], nu=nut[g],
                            sd=tiny.signdis[g], g=d(nut), gu=d(fv[:, 1]), gv=d(fv[:, 2]),
                            gw=d(fv[:, 3]), k=tuple(1.0 / h), SoT=P.S_over_T_ref, Re=P.Re,
                            mutf=tot / vol / P.Re)
            assert ms[g] == want
            checked += 1
    assert checked > 30


def test_ib_wall_model_sublayer_matches_lineardist():
    # very low Re: every image point sits in the viscous sublayer (Re_y <= 9)
    mesh = ibgpu.Mesh(ibgpu.tiny_case(Re=5.0))
    fv, nut = seeded(mesh)
    orc = O.Oracle(mesh)
    a = fv.copy()
    orc.ib_fill(a)
    b = fv.copy()
    nb_ = nut.copy()
    orc.ib_fill_sa(b, nb_)
    ib = mesh.ib_cell
    pos = mesh.signdis[ib] >= 0
    assert pos.sum() > 10
    ua, ub = a[ib[pos], 1:3], b[ib[pos], 1:3]
    assert np.abs(ub - ua).max() <= 5e-3 * np.abs(ua).max()
    # nu = kappa y+ >= 0 only where the fit is positive; finite everywhere
    assert np.all(np.isfinite(nb_[ib]))
    neg = ~pos
    if neg.any():
        assert np.all(nb_[ib[neg]] == mesh.phys[0].nut_inf)


def test_ib_wall_model_no_slip_temperature():
    mesh = ibgpu.Mesh(ibgpu.tiny_case(Re=500.0))
    fv = mesh.init_state()
    fv[:, 1:4] = 0.0     # fluid at rest: df = 0
    P = mesh.phys[0]
    fv[:, 6] = P.Cv * fv[:, 5]
    nut = np.full(mesh.ncells, P.nut_inf)
    orc = O.Oracle(mesh)
    orc.ib_fill_sa(fv, nut)
    ib = mesh.ib_cell[mesh.signdis[mesh.ib_cell] >= 0]
    assert np.allclose(fv[ib, 5], P.initvar[5], rtol=1e-14)
    assert np.allclose(fv[ib, 1:4], 0.0, atol=1e-15)


def test_sa_exchange_linear_and_mean():
    mesh = ibgpu.Mesh(ibgpu.tiny_case())
    x = mesh.centers
    nut = 1.0 + 0.2 * x[:, 0] - 0.1 * x[:, 1]
    before = nut.copy()
    orc = O.Oracle(mesh, O.cfg(temporal_refine=0))
    orc.sa_exchange(nut, 0)
    # restriction: mean of the sons
    for i in range(mesh.nres):
        fi = mesh.res_fi[i]
        sons = [fi[q] for q in range(8) if not (q & 1)]
        assert nut[mesh.res_ci[i]] == pytest.approx(np.mean(before[sons]), rel=1e-14)
    # prolongation of a field linear in x, y reproduces it at the sons
    for i in range(mesh.npro):
        for q in range(8):
            if q & 1:
                continue
            f = mesh.pro_fi[i][q]
            assert nut[f] == pytest.approx(before[f], rel=1e-12, abs=1e-13)


# ------------------------------------------------------------ the 0/0 of the SA source
def test_source_is_undefined_at_zero_nu_in_uniform_flow():
    """The reference's SA source (NS_Solver.cpp:700-767) with nu = 0 in a
    uniform stream (|Omega| = 0): S~ = max(|Omega| + ..., 0.3 |Omega|) = 0,
    r = nu / (S~ kappa^2 d^2 Re) = 0/0, and the destruction term is NaN
    (multiplying (nu/d)^2 = 0 does not save it).  The reference's -13 inlet
    carries nu = 0 (its facevar[0].viseddy is never set), so the uniform
    inflow reaches this state; the measured configurations set the inlet
    value to the freestream one (Mesh inlet_nut)."""
    assert math.isnan(cell_src(u=0.3, nu=0.0, sd=5.0, Re=1e4))
    # any nonzero nu, or any vorticity, keeps it finite
    assert math.isfinite(cell_src(u=0.3, nu=1e-12, sd=5.0, Re=1e4))
    assert math.isfinite(cell_src(u=0.3, nu=0.0, sd=5.0, gu=(0, 1e-3, 0), Re=1e4))


def test_inlet_nut_option():
    spec = ibgpu.tiny_case()
    import dataclasses
    m0 = ibgpu.Mesh(spec)
    m1 = ibgpu.Mesh(dataclasses.replace(spec, inlet_nut="freestream"))
    m2 = ibgpu.Mesh(dataclasses.replace(spec, inlet_nut=0.25))
    assert m0.phys[0].nut_inlet == 0.0                  # the reference's value
    assert m1.phys[0].nut_inlet == m1.phys[0].nut_inf > 0.0
    assert m2.phys[0].nut_inlet == 0.25
    assert ibgpu.config2().inlet_nut == "freestream"
