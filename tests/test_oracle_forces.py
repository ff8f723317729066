"""Pins for the oracle's wall forces, CPU only.

orc_forces (Body::ComptWallForce, Body.cpp:439-558, laminar build) and
orc_forces_sa (the TURBULENCE build: Patchut Turbulence.cpp:197-226 and the
Cf branch Body.cpp:514-528) are checked on a synthetic flat wall: straight
patches at y = y_w with normal +y, each with its image point at height h
above the wall, interpolated bilinearly from four far-field cells of the
tiny mesh (weights computed here from the cell centres).

  Couette flow      u = gamma (y - y_w), v = beta (y - y_w), uniform p, T,
                    wall at rest: the shear on the wall is mu gamma per unit
                    area along +x whatever h and beta are (the normal
                    velocity gradient gives no shear); pressure -p A along
                    -y; the moment about the body centre is the sum of
                    r x F of the patch forces; each part carries its
                    reference scale (p_ref area_ref, shear_ref area_ref)
  moving wall       a wall sliding with a uniform stream shears nothing; the
                    same wall at rest is sheared by mu U / h per unit area
  sublayer limit    with the image points in the viscous sublayer
                    (y+ < 3, u+ = y+), the wall-function force of the SA
                    build equals the laminar shear: Cf = rho_w u_tau^2 with
                    u_tau^2 = nu u_t / h, and shear_ref = p_ref / Re
                    (FlowConstants.cpp: shear_ref = mu_ref u_ref / L_ref,
                    p_ref = rho_ref u_ref^2, Re = rho_ref u_ref L_ref /
                    mu_ref), to the accuracy of the y+ fit
"""
import numpy as np
import pytest

import ibgpu
import oracle_ffi as O
from ibgpu import _ffi


@pytest.fixture(scope="module")
def mesh():
    return ibgpu.Mesh(ibgpu.tiny_case())


def level0_lookup(m):
    s = int(m.blk_start[m.block(0, 0)])
    e = s + int(m.blk_nnormal[m.block(0, 0)])
    cen = m.centers[s:e]
    h = m.blk_h[m.block(0, 0)]
    return s, cen, h


def bilinear(m, pt):
    """four level-0 normal cells around pt and their bilinear weights"""
    s, cen, h = level0_lookup(m)
    i0 = np.floor((pt[0] - cen[:, 0].min()) / h[0])
    j0 = np.floor((pt[1] - cen[:, 1].min()) / h[1])
    x0 = cen[:, 0].min() + i0 * h[0]
    y0 = cen[:, 1].min() + j0 * h[1]
    cells, w = [], []
    fx, fy = (pt[0] - x0) / h[0], (pt[1] - y0) / h[1]
    for dj in (0, 1):
        for di in (0, 1):
            c = np.nonzero((np.abs(cen[:, 0] - (x0 + di * h[0])) < 1e-9 * h[0])
                           & (np.abs(cen[:, 1] - (y0 + dj * h[1])) < 1e-9 * h[1]))[0]
            assert c.size == 1, "image point outside the level-0 far field"
            cells.append(s + int(c[0]))
            w.append((fx if di else 1 - fx) * (fy if dj else 1 - fy))
    return cells, w


def flat_wall(m, y_w=2.0, n=9, seed=1, hrange=(0.05, 0.3)):
    rng = np.random.default_rng(seed)
    xs = np.linspace(-3.3, -1.1, n)
    hs = rng.uniform(hrange[0], hrange[1], n)
    areas = rng.uniform(0.01, 0.03, n)
    intp = np.full((n, 8), -1, np.int32)
    coef = np.zeros((n, 8))
    for i in range(n):
        c, w = bilinear(m, (xs[i], y_w + hs[i]))
        intp[i, :4] = c
        coef[i, :4] = w
    pc = np.stack([xs, np.full(n, y_w), np.full(n, 0.05)], axis=1)
    nv = np.tile([0.0, 1.0, 0.0], (n, 1))
    pm = O.PatchedMesh(m, counts={"npatch": n},
                       pat_intp=intp, pat_coef=coef, pat_hgdis=hs, pat_nv=nv, pat_pc=pc,
                       pat_area=areas, pat_case=np.zeros(n, np.int32))
    return pm, pc, hs, areas


def linear_field(m, gamma, beta, y_w, p=None, T=1.0):
    P = m.phys[0]
    fv = m.init_state()
    p = P.initvar[4] if p is None else p
    y = m.centers[:, 1]
    fv[:, 1] = gamma * (y - y_w)
    fv[:, 2] = beta * (y - y_w)
    fv[:, 3] = 0.0
    fv[:, 4] = p
    fv[:, 5] = T
    fv[:, 0] = p / (P.Rg * T)
    fv[:, 6] = P.Cv * T + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    return fv, p


def body_centre(m):
    soa = m.soa
    bc = np.ctypeslib.as_array(soa.body_center, (m.soa.nbody, 3))
    bcase = np.ctypeslib.as_array(soa.body_case, (m.soa.nbody,))
    return bc[int(np.nonzero(bcase == 0)[0][0])]


def expected(m, pc, areas, tau_x, p, scale_tau):
    P = m.phys[0]
    bc = body_centre(m)
    F = np.zeros(3)
    M = np.zeros(3)
    for i in range(len(areas)):
        ft = np.array([tau_x * areas[i], 0.0, 0.0]) * scale_tau
        fp = np.array([0.0, -p * areas[i], 0.0]) * P.p_ref_area
        r = pc[i] - bc
        F += ft + fp
        M += np.cross(r, ft) + np.cross(r, fp)
    return F, M


@pytest.mark.parametrize("gamma,beta", [(0.8, 0.0), (0.8, 0.5), (-1.7, -0.3)])
def test_couette_shear_closed_form(mesh, gamma, beta):
    m = mesh
    P = m.phys[0]
    pm, pc, hs, areas = flat_wall(m)
    fv, p = linear_field(m, gamma, beta, 2.0)
    out = O.Oracle(pm).forces(fv)
    F, M = expected(m, pc, areas, P.mu * gamma, p, P.shear_ref_area)
    np.testing.assert_allclose(out[:3], F, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(out[3:6], M, rtol=1e-12, atol=1e-15)
    # the shear part alone: mu gamma per area along +x, the sign of gamma
    shear = out[0] / P.shear_ref_area
    assert shear == pytest.approx(P.mu * gamma * areas.sum(), rel=1e-12)


def test_couette_sliding_wall_no_shear(mesh):
    """u = U everywhere and the wall translating at U: no shear, only
    pressure (Surface_Point_Velocity Body.H:709-721 enters du)"""
    m = mesh
    P = m.phys[0]
    pm, pc, hs, areas = flat_wall(m)
    fv, p = linear_field(m, 0.0, 0.0, 2.0)
    fv[:, 1] += 0.37
    fv[:, 6] = P.Cv * fv[:, 5] + 0.5 * (fv[:, 1] ** 2 + fv[:, 2] ** 2)
    nb = m.soa.nbody
    mo = (_ffi.BodyMotion * nb)()
    for b in range(nb):
        mo[b].tra_vel[0] = 0.37
    out = O.Oracle(pm).forces(fv, mo)
    F, M = expected(m, pc, areas, 0.0, p, P.shear_ref_area)
    np.testing.assert_allclose(out[:3], F, rtol=1e-12, atol=1e-13)
    # the wall at rest under the same uniform stream: shear mu U / h per area
    out0 = O.Oracle(pm).forces(fv)
    shear0 = out0[0] / P.shear_ref_area
    assert shear0 == pytest.approx(P.mu * 0.37 * np.sum(areas / hs), rel=1e-12)


@pytest.mark.parametrize("gamma", [0.05, 0.1, -0.08])
def test_wall_function_force_sublayer_limit(mesh, gamma):
    m = mesh
    P = m.phys[0]
    # image points at a local Reynolds number rho u_t h Re / mu in [0.5, 9]:
    # u+ = y+ holds and the reference's fit y+ = 1.002 sqrt(Re) - 0.001223
    # (Turbulence.cpp:34-35) is within 0.4 % of sqrt(Re) there
    pm, pc, hs, areas = flat_wall(m, hrange=(0.08, 0.2))
    fv, p = linear_field(m, gamma, 0.1, 2.0)
    orc = O.Oracle(pm)
    lam = orc.forces(fv)
    sa = orc.forces_sa(fv)
    # the pressure parts are the same; the shear parts agree in the sublayer
    press = -p * areas.sum() * P.p_ref_area
    assert lam[1] == pytest.approx(press, rel=1e-12)
    assert sa[1] == pytest.approx(press, rel=1e-12)
    tau_lam = lam[0]
    tau_sa = sa[0]
    assert np.sign(tau_sa) == np.sign(gamma)
    assert tau_sa == pytest.approx(tau_lam, rel=8e-3)
    # the image points lie where the fit holds to 0.4 % in y+ (0.8 % in Cf)
    re_loc = np.abs(gamma) * hs * hs * P.Re / P.mu * (p / (P.Rg * 1.0))
    assert re_loc.min() > 0.5 and re_loc.max() < 9.0
    assert np.sqrt(re_loc).max() < 3.0          # u+ = y+ below y+ = 3
    assert sa[2] == 0.0 and lam[2] == 0.0


def test_wall_function_force_exceeds_laminar_in_log_layer(mesh):
    """away from the sublayer the wall function's u_tau grows faster than the
    laminar estimate mu u / h (u+ < y+ on the log law), so its shear is larger"""
    m = mesh
    pm, pc, hs, areas = flat_wall(m)
    fv, p = linear_field(m, 3.0, 0.0, 2.0)
    orc = O.Oracle(pm)
    lam, sa = orc.forces(fv), orc.forces_sa(fv)
    assert sa[0] > 1.5 * lam[0] > 0.0
