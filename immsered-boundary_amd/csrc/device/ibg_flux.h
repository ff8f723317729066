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

// ---------------------------------------------------------------------------
// Mixed-mode (FP32) form of the same face flux.  The arithmetic is the same
// expression set, evaluated with the hardware reciprocal and square root
// (v_rcp_f32, v_sqrt_f32, ~1 ulp) and with the van Leer limiter in its
// one-reciprocal form: for r = dqii/dqi > 0,
//   qi + 0.5 dqi phi(r) = qi + dqi dqii / (dqi + dqii),
// and qi otherwise (the reference's 1e10/1e6 guards only matter when
// |dqi| < 1e-10 |dqii|, where both forms give qi to within 1e-10 dqii).
// The error of this mode against the FP64 oracle is measured, not assumed.
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// van Albada in FP32 through the ratio r (as the reference forms it, with
// its guards): r > 0 here, phi = (r^2 + r)/(r^2 + 1).  The one-reciprocal
// form pr (dqi + dqii) / (dqi^2 + dqii^2) overflows in FP32 once both
// differences fall below ~1e-19 (the square sum below 2^-126).
__device__ __forceinline__ float fast_albada(float num, float den) {
  float r = num * f_rcp(den);
  if (!(r <= 1.0e10f)) r = (num > 0.0f) ? 1000000.0f : 0.0f;   // also NaN / inf
  return (r * r + r) * f_rcp(r * r + 1.0f);
}

__device__ __forceinline__ float fast_left(float q_i, float qi, float qi1, int lim) {
  const float dqi = qi - q_i, dqii = qi1 - qi;
  const float pr = dqi * dqii;
  if (!(pr > 0.0f)) return qi;
  if (lim == 1) return qi + 0.5f * dqi * fast_albada(dqii, dqi);
  return qi + pr * f_rcp(dqi + dqii);
}

__device__ __forceinline__ float fast_right(float q_i, float qi, float qi1, int lim) {
  const float dqi = qi - q_i, dqii = qi1 - qi;
  const float pr = dqi * dqii;
  if (!(pr > 0.0f)) return qi;
  if (lim == 1) return qi - 0.5f * dqii * fast_albada(dqi, dqii);
  return qi - pr * f_rcp(dqi + dqii);
}

template <int D>
__device__ __forceinline__ void roe_flux_fast(const Prim<float, D> &t0, const Prim<float, D> &t1,
                                              const Prim<float, D> &t2, const Prim<float, D> &t3,
                                              int dir, float area, const PhysR<float> &ph, int lim,
                                              float *flux) {
  const float gama = ph.gama, Rg = ph.Rg, Cv = ph.Cv;
  const float lr = fast_left(t0.r, t1.r, t2.r, lim), rr = fast_right(t1.r, t2.r, t3.r, lim);
  const float lu = fast_left(t0.u, t1.u, t2.u, lim), ru = fast_right(t1.u, t2.u, t3.u, lim);
  const float lvv = fast_left(t0.v, t1.v, t2.v, lim), rvv = fast_right(t1.v, t2.v, t3.v, lim);
  float lw = 0.0f, rw = 0.0f;
  if (D == 3) { lw = fast_left(t0.w, t1.w, t2.w, lim); rw = fast_right(t1.w, t2.w, t3.w, lim); }
  const float lp = fast_left(t0.p, t1.p, t2.p, lim), rp = fast_right(t1.p, t2.p, t3.p, lim);
  const float ilr = f_rcp(lr), irr = f_rcp(rr);
  const float iRg = f_rcp(Rg);
  const float lT = lp * ilr * iRg, rT = rp * irr * iRg;
  const float lke = (D == 3) ? 0.5f * (lu * lu + lvv * lvv + lw * lw) : 0.5f * (lu * lu + lvv * lvv);
  const float rke = (D == 3) ? 0.5f * (ru * ru + rvv * rvv + rw * rw) : 0.5f * (ru * ru + rvv * rvv);
  const float le = Cv * lT + lke, re = Cv * rT + rke;
  const float lv = dir == 0 ? lu : (dir == 1 ? lvv : lw);
  const float rv = dir == 0 ? ru : (dir == 1 ? rvv : rw);
  const float kx = dir == 0 ? 1.0f : 0.0f, ky = dir == 1 ? 1.0f : 0.0f, kz = dir == 2 ? 1.0f : 0.0f;
  // central part 1/2 (F_L + F_R) A
  const float ha = 0.5f * area;
  const float p0 = (lr * lv + rr * rv) * ha;
  const float p1 = (lr * lu * lv + lp * kx + rr * ru * rv + rp * kx) * ha;
  const float p2 = (lr * lvv * lv + lp * ky + rr * rvv * rv + rp * ky) * ha;
  const float p3 = (lr * lw * lv + lp * kz + rr * rw * rv + rp * kz) * ha;
  const float p4 = ((lp + le * lr) * lv + (rp + re * rr) * rv) * ha;
  const float droe = rr - lr, dp = rp - lp;
  const float du0 = ru - lu, du1 = rvv - lvv, du2 = rw - lw;
  const float dubar = rv - lv;
  // Roe average
  const float r2lroe = f_sqrt(rr * ilr);
  const float r2l0 = f_rcp(1.0f + r2lroe);
  const float h_left = le + lp * ilr, h_right = re + rp * irr;
  const float a_roe = f_sqrt(lr * rr);
  const float a_u = (lu + ru * r2lroe) * r2l0;
  const float a_v = (lvv + rvv * r2lroe) * r2l0;
  const float a_w = (D == 3) ? (lw + rw * r2lroe) * r2l0 : 0.0f;
  const float a_h = (h_left + h_right * r2lroe) * r2l0;
  const float q2 = (D == 3) ? (a_u * a_u + a_v * a_v + a_w * a_w) : (a_u * a_u + a_v * a_v);
  const float csquare = (gama - 1.0f) * (a_h - 0.5f * q2);
  const float a_c = f_sqrt(csquare);
  const float a_ubar = dir == 0 ? a_u : (dir == 1 ? a_v : a_w);
  const float icsq = f_rcp(csquare);
  const float namda3 = fabsf(a_ubar);
  const float coef12 = 0.5f * icsq * area;
  const float roe0cxdubar = a_roe * a_c * dubar;
  const float faa = area * namda3;
  const float al0 = faa * (droe - dp * icsq);
  // all-speed eigenvalues (NS_Solver.H:66-77)
  const float T_a = (lT + rT) * 0.5f;
  const float V_a = (lv + rv) * 0.5f;
  float f_Ma = fabsf(V_a) * f_rcp(f_sqrt(gama * Rg * T_a));
  f_Ma = f_Ma < ph.Ma ? ph.Ma : f_Ma;
  f_Ma = f_Ma > 1.0f ? 1.0f : f_Ma;
  const float s0 = (1.0f - f_Ma) * (1.0f - f_Ma);
  const float namda0 = f_sqrt(4.0f * csquare * f_Ma + s0 * a_ubar * a_ubar);
  const float namda1 = fabsf(0.5f * ((1.0f + f_Ma) * a_ubar + namda0));
  const float namda2 = fabsf(0.5f * ((1.0f + f_Ma) * a_ubar - namda0));
  const float al1 = coef12 * namda1 * (dp + roe0cxdubar);
  const float al2 = coef12 * namda2 * (dp - roe0cxdubar);
  const float al3 = al0 + al1 + al2;
  const float al4 = a_c * (al1 - al2);
  const float al5 = faa * (a_roe * (du0 - kx * dubar));
  const float al6 = faa * (a_roe * (du1 - ky * dubar));
  const float al7 = (D == 3) ? faa * (a_roe * (du2 - kz * dubar)) : 0.0f;
  const float ig1 = f_rcp(gama - 1.0f);
  flux[0] = p0 - al3;
  flux[1] = p1 - (a_u * al3 + kx * al4 + al5);
  flux[2] = p2 - (a_v * al3 + ky * al4 + al6);
  if (D == 3) {
    flux[3] = p3 - (a_w * al3 + kz * al4 + al7);
    flux[4] = p4 - (a_h * al3 + a_ubar * al4 + a_u * al5 + a_v * al6 + a_w * al7 - csquare * ig1 * al0);
  } else {
    flux[3] = p4 - (a_h * al3 + a_ubar * al4 + a_u * al5 + a_v * al6 - csquare * ig1 * al0);
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
  R ih0, ih1, ih2;   // 1/dh
  R ar0, ar1, ar2;   // face areas
  R qkt;             // Re*Pr_tur*(gama-1)*Ma^2 (SA closure)
};

// Spalart-Allmaras constants, Turbulence.cpp:3-19
struct SA {
  static constexpr double cb1 = 0.1355;
  static constexpr double segma = 2.0 / 3.0;
  static constexpr double cb2 = 0.622;
  static constexpr double kappa = 0.41;
  static constexpr double cw2 = 0.3;
  static constexpr double cw3 = 2.0;
  static constexpr double cv1 = 7.1;
  static constexpr double ct3 = 1.2;
  static constexpr double ct4 = 0.5;
  static constexpr double cw1 = cb1 / (kappa * kappa) + (1.0 + cb2) / segma;
  static constexpr double cv1three = cv1 * cv1 * cv1;
};

// Viseddy2mut (Turbulence.cpp:23-30): mu_t = roe fv1 nu, X = nu/(mu/roe)
// clamped at 0, fv1 = X^3/(X^3 + cv1^3).
template <typename R>
__device__ __forceinline__ R sa_mut(R nu, R roe, R mu);

template <typename R> __device__ __forceinline__ R v_div(R a, R b) { return a / b; }
template <> __device__ __forceinline__ float v_div<float>(float a, float b) { return a * f_rcp(b); }
template <typename R> __device__ __forceinline__ R v_sqrt(R a) { return r_sqrt(a); }
template <> __device__ __forceinline__ float v_sqrt<float>(float a) { return f_sqrt(a); }

template <typename R>
__device__ __forceinline__ R sa_mut(R nu, R roe, R mu) {
  R bigx = v_div(nu, v_div(mu, roe));
  if (bigx < R(0.0)) bigx = R(0.0);
  const R b3 = bigx * bigx * bigx;
  const R fv1 = v_div(b3, b3 + R(SA::cv1three));
  return roe * fv1 * nu;
}

// four-point tangential difference pfpkesai1 (NS_Solver.H:153-154)
template <typename R>
__device__ __forceinline__ R pf1(R up1, R up2, R dn1, R dn2) {
  return R(0.25) * (up1 + up2 - dn1 - dn2);
}

// With turb set, the SA terms of NS_Solver.cpp:534-542: the face eddy
// viscosity mu_t = Viseddy2mut(nu_f, roe_f, mu) joins mu and the heat
// conduction, and mflux receives the face flux of the working variable
// (nu_f + mu/roe_f) (d nu / d n) / sigma * A (the face keisa is the unit
// axis, so the dot product picks the normal derivative).
template <typename R, int D>
__device__ __forceinline__ void visc_flux(const Prim<R, D> &L, const Prim<R, D> &Rr,
                                          const Prim<R, D> *t, int d, const ViscPhys<R, D> &vp,
                                          R *out, bool turb = false, R nuL = R(0), R nuR = R(0),
                                          R *mflux = nullptr) {
  // T = p/(roe Rg) of every stencil cell (Get_T_ideal_gas)
  const R TL = v_div(L.p, L.r * vp.Rg), TR = v_div(Rr.p, Rr.r * vp.Rg);
  const R T0 = v_div(t[0].p, t[0].r * vp.Rg), T1 = v_div(t[1].p, t[1].r * vp.Rg);
  const R T2 = v_div(t[2].p, t[2].r * vp.Rg), T3 = v_div(t[3].p, t[3].r * vp.Rg);
  // normal differences (pfpkesai0) and the first tangential axis
  const R nu = Rr.u - L.u, nv = Rr.v - L.v, nw = Rr.w - L.w, nT = TR - TL;
  const R au = pf1(t[0].u, t[1].u, t[2].u, t[3].u);
  const R av = pf1(t[0].v, t[1].v, t[2].v, t[3].v);
  const R aw = (D == 3) ? pf1(t[0].w, t[1].w, t[2].w, t[3].w) : R(0);
  const R aT = pf1(T0, T1, T2, T3);
  R bu = R(0), bv = R(0), bw = R(0), bT = R(0);
  if (D == 3) {
    const R T4 = v_div(t[4 % (D == 3 ? 8 : 4)].p, t[4 % (D == 3 ? 8 : 4)].r * vp.Rg);
    const R T5 = v_div(t[5 % (D == 3 ? 8 : 4)].p, t[5 % (D == 3 ? 8 : 4)].r * vp.Rg);
    const R T6 = v_div(t[6 % (D == 3 ? 8 : 4)].p, t[6 % (D == 3 ? 8 : 4)].r * vp.Rg);
    const R T7 = v_div(t[7 % (D == 3 ? 8 : 4)].p, t[7 % (D == 3 ? 8 : 4)].r * vp.Rg);
    const Prim<R, D> &s4 = t[4 % (D == 3 ? 8 : 4)], &s5 = t[5 % (D == 3 ? 8 : 4)];
    const Prim<R, D> &s6 = t[6 % (D == 3 ? 8 : 4)], &s7 = t[7 % (D == 3 ? 8 : 4)];
    bu = pf1(s4.u, s5.u, s6.u, s7.u);
    bv = pf1(s4.v, s5.v, s6.v, s7.v);
    bw = pf1(s4.w, s5.w, s6.w, s7.w);
    bT = pf1(T4, T5, T6, T7);
  }
  // gradients in kesai: normal along d, first tangential t1, second t2
  R gu[3], gv[3], gw[3], gT[3];
  const int t1 = (d == 0) ? 1 : 0;
  const int t2 = (d == 2) ? 1 : 2;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool isn = (k == d), is1 = (k == t1), is2 = (D == 3) && (k == t2);
    gu[k] = isn ? nu : (is1 ? au : (is2 ? bu : R(0)));
    gv[k] = isn ? nv : (is1 ? av : (is2 ? bv : R(0)));
    gw[k] = isn ? nw : (is1 ? aw : (is2 ? bw : R(0)));
    gT[k] = isn ? nT : (is1 ? aT : (is2 ? bT : R(0)));
  }
  if (D == 2) { gw[0] = gw[1] = gw[2] = R(0); }
  // VarReflect with keisa = diag(1/dh)
  const R dux = gu[0] * vp.ih0, duy = gu[1] * vp.ih1, duz = gu[2] * vp.ih2;
  const R dvx = gv[0] * vp.ih0, dvy = gv[1] * vp.ih1, dvz = gv[2] * vp.ih2;
  const R dwx = gw[0] * vp.ih0, dwy = gw[1] * vp.ih1, dwz = gw[2] * vp.ih2;
  const R dTx = gT[0] * vp.ih0, dTy = gT[1] * vp.ih1, dTz = gT[2] * vp.ih2;
  const R ftemp = (TL + TR) * R(0.5);
  R mu0 = v_sqrt(ftemp * ftemp * ftemp) * v_div(R(1.0) + vp.SoT, ftemp + vp.SoT);
  R qk0 = vp.qk * mu0;
  // (x*1 + 0 + 0 is exact; written as a sum so that no per-lane table is built)
  const R area = (d == 0 ? R(1) : R(0)) * vp.ar0 + (d == 1 ? R(1) : R(0)) * vp.ar1 +
                 (d == 2 ? R(1) : R(0)) * vp.ar2;
  if (turb) {
    const R ih = (d == 0 ? R(1) : R(0)) * vp.ih0 + (d == 1 ? R(1) : R(0)) * vp.ih1 +
                 (d == 2 ? R(1) : R(0)) * vp.ih2;
    const R fvis = (nuL + nuR) * R(0.5);
    const R froe = (L.r + Rr.r) * R(0.5);
    const R faceroexvis = fvis + v_div(mu0, froe);
    const R mut0 = sa_mut(fvis, froe, mu0);
    mu0 += mut0;
    qk0 += v_div(mut0, vp.qkt);
    *mflux = v_div(faceroexvis * ((nuR - nuL) * ih), R(SA::segma)) * area;
  }
  const R Re = vp.Re;
  const R divu = (D == 3) ? (dux + dvy + dwz) : (dux + dvy);
  const R tauxx = v_div(R(2.0) * mu0 * (dux - v_div(divu, R(3.0))), Re);
  const R tauyy = v_div(R(2.0) * mu0 * (dvy - v_div(divu, R(3.0))), Re);
  const R tauxy = v_div(mu0, Re) * (duy + dvx);
  const R fu = (L.u + Rr.u) * R(0.5), fv = (L.v + Rr.v) * R(0.5);
  if (D == 2) {
    if (d == 0) {
      const R bx = fu * tauxx + fv * tauxy + qk0 * dTx;
      out[0] = R(0); out[1] = tauxx * area; out[2] = tauxy * area; out[3] = bx * area;
    } else {
      const R by = fu * tauxy + fv * tauyy + qk0 * dTy;
      out[0] = R(0); out[1] = tauxy * area; out[2] = tauyy * area; out[3] = by * area;
    }
  } else {
    const R tauzz = v_div(R(2.0) * mu0 * (dwz - v_div(divu, R(3.0))), Re);
    const R tauxz = v_div(mu0, Re) * (duz + dwx);
    const R tauyz = v_div(mu0, Re) * (dvz + dwy);
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

// exp, x^6 and x^(1/6) of the SA functions: the libm forms in FP64 (the
// reference's pow/exp); in the mixed mode x^6 = (x^2)^3 and the hardware
// exp2/log2 (v_exp_f32, v_log_f32, ~1 ulp).
template <typename R> __device__ __forceinline__ R r_exp(R x);
template <> __device__ __forceinline__ double r_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float r_exp<float>(float x) {
  return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
}
template <typename R> __device__ __forceinline__ R r_pow6(R x);
template <> __device__ __forceinline__ double r_pow6<double>(double x) { return pow(x, 6.0); }
template <> __device__ __forceinline__ float r_pow6<float>(float x) {
  const float x2 = x * x;
  return x2 * x2 * x2;
}
template <typename R> __device__ __forceinline__ R r_root6(R x);
template <> __device__ __forceinline__ double r_root6<double>(double x) { return pow(x, 1.0 / 6.0); }
template <> __device__ __forceinline__ float r_root6<float>(float x) {
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 6.0f));
}

// ComptMutSource (NS_Solver.cpp:648-808) for one non-solid cell: central
// differences over the face neighbours (nl/nr: working variable, ul..wr:
// velocity, low and high neighbour per axis), vorticity, Sutherland mu,
// fv1/fv2, the modified vorticity S~ (floored at 0.3|Omega|sqrt2), ft2, r, g,
// fw, then advection + production - destruction + cb2 diffusion.  mutf is
// the frozen face-flux term sum(dirsign * mutflux)/V/Re of the step.  The
// frame velocity of the reference is zero here.
template <typename R, int D>
__device__ __forceinline__ R sa_source(R u, R v, R w, R T, R roe, R nu, R sd, const R *nl,
                                       const R *nr, const R *ul, const R *ur, const R *vl,
                                       const R *vr, const R *wl, const R *wr, R ih0, R ih1, R ih2,
                                       R SoT, R Re, R mutf) {
  const R h = R(0.5);
  const R dvisx = ((nr[0] - nl[0]) * h) * ih0;
  const R dvisy = ((nr[1] - nl[1]) * h) * ih1;
  const R dvisz = (D == 3) ? ((nr[2] - nl[2]) * h) * ih2 : R(0);
  const R duy = ((ur[1] - ul[1]) * h) * ih1;
  const R dvx = ((vr[0] - vl[0]) * h) * ih0;
  R duz = R(0), dvz = R(0), dwx = R(0), dwy = R(0);
  if (D == 3) {
    duz = ((ur[2] - ul[2]) * h) * ih2;
    dvz = ((vr[2] - vl[2]) * h) * ih2;
    dwx = ((wr[0] - wl[0]) * h) * ih0;
    dwy = ((wr[1] - wl[1]) * h) * ih1;
  }
  const R ox = R(0.5) * (dwy - dvz), oy = R(0.5) * (duz - dwx), oz = R(0.5) * (dvx - duy);
  const R mu0 = v_sqrt(T * T * T) * v_div(R(1.0) + SoT, T + SoT);
  const R vis0 = v_div(mu0, roe);
  const R bigx = v_div(nu, vis0);
  const R b3 = bigx * bigx * bigx;
  const R fv1 = v_div(b3, b3 + R(SA::cv1three));
  const R fv2 = R(1.0) - v_div(bigx, R(1.0) + bigx * fv1);
  R segdis, vis2d;
  if (r_abs(sd) < R(0.00000001)) {
    segdis = R(1.0);
    vis2d = R(1.0);
  } else {
    segdis = R(SA::kappa) * sd;
    vis2d = v_div(nu, sd);
  }
  segdis *= segdis;
  const R olen = v_sqrt(R(2.0)) * v_sqrt(ox * ox + oy * oy + oz * oz);
  R bigs = olen + v_div(v_div(nu * fv2, segdis), Re);
  bigs = (R(0.3) * olen < bigs) ? bigs : R(0.3) * olen;
  const R ft2 = R(SA::ct3) * r_exp(-R(SA::ct4) * bigx * bigx);
  R rr = v_div(v_div(nu, bigs * segdis), Re);
  rr = (R(10.0) < rr) ? R(10.0) : rr;
  const R rsix = r_pow6(rr);
  const R g = rr + R(SA::cw2) * (rsix - rr);
  const R gsix = r_pow6(g);
  const R cw3six = R(64.0);
  const R fw = g * r_root6(v_div(R(1.0) + cw3six, gsix + cw3six));
  R src = -(u * dvisx + v * dvisy + w * dvisz);
  src += R(SA::cb1) * (R(1.0) - ft2) * bigs * nu;
  src -= v_div((R(SA::cw1) * fw - v_div(R(SA::cb1) * ft2, R(SA::kappa * SA::kappa))) * (vis2d * vis2d), Re);
  src += v_div(v_div(R(SA::cb2) * (dvisx * dvisx + dvisy * dvisy + dvisz * dvisz), R(SA::segma)), Re);
  src += mutf;
  return src;
}

// Wall-function fits of Turbulence.cpp:32-55: y+ as a function of the local
// Reynolds number and u+ as a function of y+ (piecewise fits in sqrt and the
// fourth root; FP64 in both modes, as all wall interpolation).
__device__ __forceinline__ double sa_yplus(double re) {
  const double sq = sqrt(re);
  if (re <= 9.0) return 1.002 * sq - 0.001223;
  if (re <= 37.0) return 1.272 * sq - 0.9881 * sqrt(sq) + 0.9054;
  if (re <= 98.0) return 0.2715 * sq + 1.144 * sqrt(sq) + 0.04718 * re;
  if (re <= 250.0) return 0.04921 * sq + 1.576 * sqrt(sq) + 0.05582 * re;
  if (re <= 790.0) return 0.5382 * sq + 0.2478 * sqrt(sq) + 0.04593 * re;
  if (re <= 2800.0) return 1.105 * sq - 1.782 * sqrt(sq) + 0.03933 * re;
  return 2.643 * sq - 10.52 * sqrt(sq) + 0.03207 * re;
}

__device__ __forceinline__ double sa_uplus(double y) {
  if (y < 3.0) return y;
  const double lg = log(y), sq = sqrt(y), qu = sqrt(sq);
  if (y < 8.5) return -6.645 * lg + 25.59 * sq - 21.6 * qu - 1.864 * y;
  if (y < 50.0) return 12.04 * lg - 4.079 * sq - 4.504 * qu + 0.1715 * y;
  if (y < 300.0) return 3.031 * lg - 0.6252 * sq + 2.639 * qu + 0.007553 * y;
  if (y < 1000.0) return -3.535 * lg - 1.104 * sq + 13.89 * qu + 0.004011 * y;
  return 16.66 * lg + 1.408 * sq - 23.82 * qu - 0.002909 * y;
}

}  // namespace ibg
