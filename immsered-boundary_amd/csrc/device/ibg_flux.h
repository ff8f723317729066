// ibg_flux.h — per-face arithmetic of the residual engine (device side).
//
// Templated on the arithmetic type R (double: plain mode; float: mixed mode)
// and on the dimension D (2: the w component is identically zero and all of
// its terms are dropped, which is exact because every one of them is a
// product with w = 0 or a zero metric component).
#pragma once
#include <hip/hip_runtime.h>

namespace ibg {

__device__ __forceinline__ double r_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float r_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double r_abs(double x) { return fabs(x); }
__device__ __forceinline__ float r_abs(float x) { return fabsf(x); }
__device__ __forceinline__ bool r_isnan(double x) { return x != x; }
__device__ __forceinline__ bool r_isnan(float x) { return x != x; }

// Physical constants of one case in the arithmetic type of the kernel.
template <typename R>
struct PhysR {
  R gama, Rg, Cv, Ma;
};

template <typename R, int D>
struct Prim {
  R r, u, v, w, p;   // w unused when D == 2
};

// van Leer phi = 2r/(1+r) (Slopelimiter.H:17-20); van Albada (lim == 1)
// phi = (r^2+r)/(r^2+1) is an option outside the reference.
template <typename R>
__device__ __forceinline__ R limiter_phi(R qr, int lim) {
  if (lim == 1) return (qr * qr + qr) / (qr * qr + R(1.0));
  return R(2.0) * qr / (R(1) + qr);
}

// SlopeCorrection_Left + comptqleft (Slopelimiter.H:22-37, NS_Solver.H:25-27)
template <typename R>
__device__ __forceinline__ R muscl_left(R q_i, R qi, R qi1, int lim) {
  R dqi = qi - q_i;
  R dqii = qi1 - qi;
  R qr = dqii / dqi;
  if (r_abs(qr) > R(1.0e10) || r_isnan(qr)) qr = (dqii > R(0.0)) ? R(1000000.0) : R(-1.0);
  if (qr <= R(0.0)) qr = R(0.0);
  return qi + R(0.5) * dqi * limiter_phi(qr, lim);
}

// SlopeCorrection_Right + comptqright (Slopelimiter.H:39-58, NS_Solver.H:21-23)
template <typename R>
__device__ __forceinline__ R muscl_right(R q_i, R qi, R qi1, int lim) {
  R dqi = qi - q_i;
  R dqii = qi1 - qi;
  R qr = dqi / dqii;
  if (r_abs(qr) > R(1.0e10) || r_isnan(qr)) qr = (dqi > R(0.0)) ? R(1000000.0) : R(-1.0);
  if (qr <= R(0.0)) qr = R(0.0);
  return qi - R(0.5) * dqii * limiter_phi(qr, lim);
}

// All-speed Roe flux of one axis-aligned face from the 4-cell stencil
// (t0,t1 | t2,t3), the body of ComputeFaceFlux NS_Solver.H:766-794 with
// FaceLeftRight (:31-45), Get_inv_flux (FlowVariables.H:59-64), Roeaverage
// (:51-62), roe0c_coef ROE_ALLSPEED (:66-77) and ComptFaceFlux (:87-113),
// the upwind term without the factor 1/2 as in the reference.
// Output: D == 2: flux[0..3] = (mass, x-mom, y-mom, energy);
//         D == 3: flux[0..4] = (mass, x, y, z momentum, energy).
template <typename R, int D>
__device__ __forceinline__ void roe_flux(const Prim<R, D> &t0, const Prim<R, D> &t1,
                                         const Prim<R, D> &t2, const Prim<R, D> &t3,
                                         int dir, R area, const PhysR<R> &ph, int lim,
                                         R *flux) {
  const R gama = ph.gama, Rg = ph.Rg, Cv = ph.Cv;
  R lr = muscl_left(t0.r, t1.r, t2.r, lim), rr = muscl_right(t1.r, t2.r, t3.r, lim);
  R lu = muscl_left(t0.u, t1.u, t2.u, lim), ru = muscl_right(t1.u, t2.u, t3.u, lim);
  R lvv = muscl_left(t0.v, t1.v, t2.v, lim), rvv = muscl_right(t1.v, t2.v, t3.v, lim);
  R lw = R(0), rw = R(0);
  if (D == 3) { lw = muscl_left(t0.w, t1.w, t2.w, lim); rw = muscl_right(t1.w, t2.w, t3.w, lim); }
  R lp = muscl_left(t0.p, t1.p, t2.p, lim), rp = muscl_right(t1.p, t2.p, t3.p, lim);
  // Get_T_ideal_gas, Get_E (FlowVariables.H:23,33)
  R lT = lp / (lr * Rg);
  R rT = rp / (rr * Rg);
  R lke = (D == 3) ? R(0.5) * (lu * lu + lvv * lvv + lw * lw) : R(0.5) * (lu * lu + lvv * lvv);
  R rke = (D == 3) ? R(0.5) * (ru * ru + rvv * rvv + rw * rw) : R(0.5) * (ru * ru + rvv * rvv);
  R le = Cv * lT + lke;
  R re = Cv * rT + rke;
  // face normal is the unit axis: the normal velocity is one component
  R lv = dir == 0 ? lu : (dir == 1 ? lvv : lw);
  R rv = dir == 0 ? ru : (dir == 1 ? rvv : rw);
  const R kx = dir == 0 ? R(1) : R(0);
  const R ky = dir == 1 ? R(1) : R(0);
  const R kz = dir == 2 ? R(1) : R(0);
  // Get_inv_flux
  R pf0[5], pf1[5];
  pf0[0] = lr * lv;
  pf0[1] = lr * lu * lv + lp * kx;
  pf0[2] = lr * lvv * lv + lp * ky;
  pf0[3] = lr * lw * lv + lp * kz;
  pf0[4] = (lp + le * lr) * lv;
  pf1[0] = rr * rv;
  pf1[1] = rr * ru * rv + rp * kx;
  pf1[2] = rr * rvv * rv + rp * ky;
  pf1[3] = rr * rw * rv + rp * kz;
  pf1[4] = (rp + re * rr) * rv;
  const R droe = rr - lr;
  const R dp = rp - lp;
  const R du0 = ru - lu, du1 = rvv - lvv, du2 = rw - lw;
  const R dubar = rv - lv;
  // Roeaverage
  R r2lroe = r_sqrt(rr / lr);
  R r2l0 = R(1.0) / (R(1.0) + r2lroe);
  R h_left = le + lp / lr;
  R h_right = re + rp / rr;
  R a_roe = r_sqrt(lr * rr);
  R a_u = (lu + ru * r2lroe) * r2l0;
  R a_v = (lvv + rvv * r2lroe) * r2l0;
  R a_w = (D == 3) ? (lw + rw * r2lroe) * r2l0 : R(0);
  R a_h = (h_left + h_right * r2lroe) * r2l0;
  R q2 = (D == 3) ? (a_u * a_u + a_v * a_v + a_w * a_w) : (a_u * a_u + a_v * a_v);
  R a_c = r_sqrt((gama - R(1.0)) * (a_h - q2 / R(2.0)));
  R a_ubar = dir == 0 ? a_u : (dir == 1 ? a_v : a_w);
  // ComptFaceFlux
  const R csquare = a_c * a_c;
  const R namda3 = r_abs(a_ubar);
  const R coef12 = R(1.0) / (R(2.0) * csquare) * area;
  const R roe0cxdubar = a_roe * a_c * dubar;
  const R faceareaxabsubar = area * namda3;
  const R halffacearea = R(0.5) * area;
  R al0 = faceareaxabsubar * (droe - dp / csquare);
  // roe0c_coef (all-speed): f = min(max(Ma_face, Ma_inf), 1)
  R T_a = (lT + rT) * R(0.5);
  R V_a = (lv + rv) * R(0.5);
  R Ma_a = r_abs(V_a / r_sqrt(gama * Rg * T_a));
  R f_Ma = Ma_a < ph.Ma ? ph.Ma : Ma_a;
  f_Ma = f_Ma > R(1.0) ? R(1.0) : f_Ma;
  R s0 = (R(1.0) - f_Ma) * (R(1.0) - f_Ma);
  R namda0 = r_sqrt(R(4.0) * a_c * a_c * f_Ma + s0 * a_ubar * a_ubar);
  R namda1 = r_abs(R(0.5) * ((R(1.0) + f_Ma) * a_ubar + namda0));
  R namda2 = r_abs(R(0.5) * ((R(1.0) + f_Ma) * a_ubar - namda0));
  R al1 = coef12 * namda1 * (dp + roe0cxdubar);
  R al2 = coef12 * namda2 * (dp - roe0cxdubar);
  R al3 = al0 + al1 + al2;
  R al4 = a_c * (al1 - al2);
  R al5 = faceareaxabsubar * (a_roe * (du0 - kx * dubar));
  R al6 = faceareaxabsubar * (a_roe * (du1 - ky * dubar));
  R al7 = (D == 3) ? faceareaxabsubar * (a_roe * (du2 - kz * dubar)) : R(0);
  R f0 = -al3;
  R f1 = -(a_u * al3 + (kx)*al4 + al5);
  R f2 = -(a_v * al3 + (ky)*al4 + al6);
  R f4;
  if (D == 3)
    f4 = -(a_h * al3 + a_ubar * al4 + a_u * al5 + a_v * al6 + a_w * al7 -
           (csquare / (gama - R(1.0))) * al0);
  else
    f4 = -(a_h * al3 + a_ubar * al4 + a_u * al5 + a_v * al6 - (csquare / (gama - R(1.0))) * al0);
  f0 += (pf0[0] + pf1[0]) * halffacearea;
  f1 += (pf0[1] + pf1[1]) * halffacearea;
  f2 += (pf0[2] + pf1[2]) * halffacearea;
  f4 += (pf0[4] + pf1[4]) * halffacearea;
  flux[0] = f0;
  flux[1] = f1;
  flux[2] = f2;
  if (D == 3) {
    R f3 = -(a_w * al3 + (kz)*al4 + al7);
    f3 += (pf0[3] + pf1[3]) * halffacearea;
    flux[3] = f3;
    flux[4] = f4;
  } else {
    flux[3] = f4;
  }
}

// Viscous face flux of ComptViscousStress (NS_Solver.cpp:426-581, laminar).
// Stencil of the face between L (low side, the reference's abox) and Rr:
// t[0..3] = L+t1, Rr+t1, L-t1, Rr-t1 and, in 3D, t[4..7] the same for t2,
// where t1 < t2 are the two tangential axes.  Gradients: normal difference
// (pfpkesai0) and the four-point tangential difference (pfpkesai1), times
// the cell metric 1/dh (VarReflect with keisa = diag(1/dh)); Sutherland mu
// (:531-533); tau and heat flux (:544-557); output A*[0, tau.n, (u.tau+q).n].
template <typename R, int D>
struct ViscPhys {
  R Rg, Re, qk, SoT;
  R invh[3];
  R area[3];
};

template <typename R, int D>
__device__ __forceinline__ void visc_flux(const Prim<R, D> &L, const Prim<R, D> &Rr,
                                          const Prim<R, D> *t, int d, const ViscPhys<R, D> &vp,
                                          R *out) {
  auto T_of = [&](const Prim<R, D> &q) -> R { return q.p / (q.r * vp.Rg); };
  const R TL = T_of(L), TR = T_of(Rr);
  // g[q][k]: d(u,v,w,T)/d(kesai_k)
  R g[4][3];
  R Tt[8];
  const int nt = (D == 3) ? 8 : 4;
#pragma unroll
  for (int k = 0; k < 8; ++k) Tt[k] = (k < nt) ? T_of(t[k]) : R(0);
  auto val = [&](const Prim<R, D> &q, int var, R T) -> R {
    return var == 0 ? q.u : (var == 1 ? q.v : (var == 2 ? q.w : T));
  };
  const int t1 = (d == 0) ? 1 : 0;
  const int t2 = (d == 2) ? 1 : 2;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (D == 2 && q == 2) { g[q][0] = g[q][1] = g[q][2] = R(0); continue; }
    g[q][d] = val(Rr, q, TR) - val(L, q, TL);
    g[q][t1] = R(0.25) * (val(t[0], q, Tt[0]) + val(t[1], q, Tt[1]) - val(t[2], q, Tt[2]) - val(t[3], q, Tt[3]));
    if (D == 3)
      g[q][t2] = R(0.25) * (val(t[4], q, Tt[4]) + val(t[5], q, Tt[5]) - val(t[6], q, Tt[6]) - val(t[7], q, Tt[7]));
    else
      g[q][2] = R(0);
  }
  const R dux = g[0][0] * vp.invh[0], duy = g[0][1] * vp.invh[1], duz = g[0][2] * vp.invh[2];
  const R dvx = g[1][0] * vp.invh[0], dvy = g[1][1] * vp.invh[1], dvz = g[1][2] * vp.invh[2];
  const R dwx = g[2][0] * vp.invh[0], dwy = g[2][1] * vp.invh[1], dwz = g[2][2] * vp.invh[2];
  const R dTx = g[3][0] * vp.invh[0], dTy = g[3][1] * vp.invh[1], dTz = g[3][2] * vp.invh[2];
  const R ftemp = (TL + TR) * R(0.5);
  const R mu0 = r_sqrt(ftemp * ftemp * ftemp) * ((R(1.0) + vp.SoT) / (ftemp + vp.SoT));
  const R qk0 = vp.qk * mu0;
  const R Re = vp.Re;
  const R divu = (D == 3) ? (dux + dvy + dwz) : (dux + dvy);
  const R tauxx = R(2.0) * mu0 * (dux - divu / R(3.0)) / Re;
  const R tauyy = R(2.0) * mu0 * (dvy - divu / R(3.0)) / Re;
  const R tauxy = mu0 / Re * (duy + dvx);
  const R fu = (L.u + Rr.u) * R(0.5), fv = (L.v + Rr.v) * R(0.5);
  const R area = vp.area[d];
  if (D == 2) {
    if (d == 0) {
      const R bx = fu * tauxx + fv * tauxy + qk0 * dTx;
      out[0] = R(0); out[1] = tauxx * area; out[2] = tauxy * area; out[3] = bx * area;
    } else {
      const R by = fu * tauxy + fv * tauyy + qk0 * dTy;
      out[0] = R(0); out[1] = tauxy * area; out[2] = tauyy * area; out[3] = by * area;
    }
  } else {
    const R tauzz = R(2.0) * mu0 * (dwz - divu / R(3.0)) / Re;
    const R tauxz = mu0 / Re * (duz + dwx);
    const R tauyz = mu0 / Re * (dvz + dwy);
    const R fw = (L.w + Rr.w) * R(0.5);
    if (d == 0) {
      const R bx = fu * tauxx + fv * tauxy + fw * tauxz + qk0 * dTx;
      out[0] = R(0); out[1] = tauxx * area; out[2] = tauxy * area; out[3] = tauxz * area; out[4] = bx * area;
    } else if (d == 1) {
      const R by = fu * tauxy + fv * tauyy + fw * tauyz + qk0 * dTy;
      out[0] = R(0); out[1] = tauxy * area; out[2] = tauyy * area; out[3] = tauyz * area; out[4] = by * area;
    } else {
      const R bz = fu * tauxz + fv * tauyz + fw * tauzz + qk0 * dTz;
      out[0] = R(0); out[1] = tauxz * area; out[2] = tauyz * area; out[3] = tauzz * area; out[4] = bz * area;
    }
  }
}

}  // namespace ibg
