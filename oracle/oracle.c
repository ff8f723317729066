/*
 * oracle.c — FP64 CPU restatement of the reference's per-step residual path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Not part of the product: the HIP
 * engine is checked against this code, never routed through it.
 *
 * Pinning: every function here is checked in tests/test_oracle_*.py against
 * closed forms and invariants (limiter values and symmetry, Roe flux
 * consistency and the |A|dU eigen-decomposition, free-stream preservation,
 * exact viscous fluxes of linear fields, Sutherland, multilinear weights,
 * restriction/prolongation identities, conservation); see the header of each
 * test file for the list.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NF IBG_NFV
enum { RO = 0, UU = 1, VV = 2, WW = 3, PP = 4, TT = 5, EE = 6 };

static inline int blk_of(const ibg_mesh_soa *m, int c, int l) { return c * m->nlevels + l; }

/* Spalart-Allmaras constants, Turbulence.cpp:3-19 */
#define ORC_SA_CB1 0.1355
#define ORC_SA_SEGMA (2.0 / 3.0)
#define ORC_SA_CB2 0.622
#define ORC_SA_KAPPA 0.41
#define ORC_SA_CW2 0.3
#define ORC_SA_CW3 2.0
#define ORC_SA_CV1 7.1
#define ORC_SA_CT3 1.2
#define ORC_SA_CT4 0.5
static const double orc_sa_cw1 = ORC_SA_CB1 / (ORC_SA_KAPPA * ORC_SA_KAPPA) + (1.0 + ORC_SA_CB2) / ORC_SA_SEGMA;
static const double orc_sa_cv1three = ORC_SA_CV1 * ORC_SA_CV1 * ORC_SA_CV1;

/* Viseddy2mut Turbulence.cpp:23-30 */
double orc_sa_mut(double nu, double roe, double mu) {
  double bigx = nu / (mu / roe);
  if (bigx < 0.0) bigx = 0.0;
  double b3 = bigx * bigx * bigx;
  double fv1 = b3 / (b3 + orc_sa_cv1three);
  return roe * fv1 * nu;
}

/* yplus_as_function_of_re / uplus_as_function_of_yplus Turbulence.cpp:32-55 */
double orc_sa_yplus(double re) {
  double sq = sqrt(re);
  if (re <= 9.0) return 1.002 * sq - 0.001223;
  if (re <= 37.0) return 1.272 * sq - 0.9881 * sqrt(sq) + 0.9054;
  if (re <= 98.0) return 0.2715 * sq + 1.144 * sqrt(sq) + 0.04718 * re;
  if (re <= 250.0) return 0.04921 * sq + 1.576 * sqrt(sq) + 0.05582 * re;
  if (re <= 790.0) return 0.5382 * sq + 0.2478 * sqrt(sq) + 0.04593 * re;
  if (re <= 2800.0) return 1.105 * sq - 1.782 * sqrt(sq) + 0.03933 * re;
  return 2.643 * sq - 10.52 * sqrt(sq) + 0.03207 * re;
}

double orc_sa_uplus(double y) {
  if (y < 3.0) return y;
  double lg = log(y), sq = sqrt(y), qu = sqrt(sq);
  if (y < 8.5) return -6.645 * lg + 25.59 * sq - 21.6 * qu - 1.864 * y;
  if (y < 50.0) return 12.04 * lg - 4.079 * sq - 4.504 * qu + 0.1715 * y;
  if (y < 300.0) return 3.031 * lg - 0.6252 * sq + 2.639 * qu + 0.007553 * y;
  if (y < 1000.0) return -3.535 * lg - 1.104 * sq + 13.89 * qu + 0.004011 * y;
  return 16.66 * lg + 1.408 * sq - 23.82 * qu - 0.002909 * y;
}

/* ComptMutSource NS_Solver.cpp:648-808 for one non-solid cell.
   g[k] = d(nu)/d(kesai_k), gu/gv/gw the velocity differences (central, /2),
   keisa = diag(k0,k1,k2); mutf = the face-flux term sum(dirsign mutflux)/V/Re. */
double orc_sa_cell_source(double u, double v, double w, double T, double roe, double nu,
                          double sd, const double *g, const double *gu, const double *gv,
                          const double *gw, double k0, double k1, double k2, double SoT,
                          double Re, double mutf) {
  double dvisx = g[0] * k0, dvisy = g[1] * k1, dvisz = g[2] * k2;
  double duy = gu[1] * k1, duz = gu[2] * k2;
  double dvx = gv[0] * k0, dvz = gv[2] * k2;
  double dwx = gw[0] * k0, dwy = gw[1] * k1;
  double ox = 0.5 * (dwy - dvz), oy = 0.5 * (duz - dwx), oz = 0.5 * (dvx - duy);
  double mu0 = sqrt(pow(T, 3)) * ((1.0 + SoT) / (T + SoT));
  double vis0 = mu0 / roe;
  double bigx = nu / vis0;
  double b3 = bigx * bigx * bigx;
  double fv1 = b3 / (b3 + orc_sa_cv1three);
  double fv2 = 1.0 - bigx / (1.0 + bigx * fv1);
  double segdis, vis2d;
  if (fabs(sd) < 0.00000001) {
    segdis = 1.0;
    vis2d = 1.0;
  } else {
    segdis = ORC_SA_KAPPA * sd;
    vis2d = nu / sd;
  }
  segdis *= segdis;
  double olen = sqrt(2.0) * sqrt(ox * ox + oy * oy + oz * oz);
  double bigs = olen + nu * fv2 / segdis / Re;
  bigs = (0.3 * olen < bigs) ? bigs : 0.3 * olen;            /* std::max */
  double ft2 = ORC_SA_CT3 * exp(-ORC_SA_CT4 * bigx * bigx);
  double r = nu / (bigs * segdis) / Re;
  r = (10.0 < r) ? 10.0 : r;                                  /* std::min */
  double rsix = pow(r, 6);
  double gg = r + ORC_SA_CW2 * (rsix - r);
  double gsix = pow(gg, 6);
  double cw3six = pow(ORC_SA_CW3, 6);
  double fw = gg * pow((1.0 + cw3six) / (gsix + cw3six), 1.0 / 6.0);
  double src = -(u * dvisx + v * dvisy + w * dvisz);
  src += ORC_SA_CB1 * (1.0 - ft2) * bigs * nu;
  src -= (orc_sa_cw1 * fw - ORC_SA_CB1 * ft2 / (ORC_SA_KAPPA * ORC_SA_KAPPA)) * (vis2d * vis2d) / Re;
  src += ORC_SA_CB2 * (dvisx * dvisx + dvisy * dvisy + dvisz * dvisz) / ORC_SA_SEGMA / Re;
  src += mutf;
  return src;
}

/* ---------------------------------------------------------------- schedule
 * mpiparams.cpp:186-212: marching_step[l] = 2^(L-1-l), marching_left = 0;
 * average/prolongation_step[l] = 2^(L-l), left = 2^(L-l-1); step_num =
 * marching_step under TEMPORAL_REFINE, 1 otherwise.                      */
int32_t orc_marched(int32_t L, int32_t level, int32_t ts, int32_t tr) {
  if (!tr) return 1;
  int ms = 1 << (L - 1 - level);
  return (ts % ms) == 0;
}

int32_t orc_transfer_due(int32_t L, int32_t cl, int32_t ts, int32_t tr) {
  if (!tr) return 1;
  int i = cl + 1;   /* the schedule is indexed by the fine level */
  int st = 1 << (L - i);
  int left = 1 << (L - i - 1);
  return (ts % st) == left;
}

double orc_step_num(int32_t L, int32_t level, int32_t tr) {
  return tr ? (double)(1 << (L - 1 - level)) : 1.0;
}

/* ----------------------------------------------------------------- limiter
 * SlopeCorrection_Left / _Right (Slopelimiter.H:22-58) and comptqleft /
 * comptqright (NS_Solver.H:21-27).  VanLeer phi = 2r/(1+r) (Slopelimiter.H:
 * 17-20); van Albada phi = (r^2+r)/(r^2+1) is an option outside the
 * reference.                                                              */
static inline double phi_of(double qr, int lim) {
  if (lim == 1) return (qr * qr + qr) / (qr * qr + 1.0);
  return 2.0 * qr / (1 + qr);
}

double orc_muscl_left(double q_i, double qi, double qi1, int32_t lim) {
  double dqi = qi - q_i;
  double dqii = qi1 - qi;
  double qr = dqii / dqi;
  if (fabs(qr) > 1.0e10 || isnan(qr)) {
    if (dqii > 0.0) qr = 1000000.0;
    else qr = -1.0;
  }
  if (qr <= 0.0) qr = 0.0;
  double limitvalue = phi_of(qr, lim);
  return qi + 0.5 * dqi * limitvalue;
}

double orc_muscl_right(double q_i, double qi, double qi1, int32_t lim) {
  double dqi = qi - q_i;
  double dqii = qi1 - qi;
  double qr = dqi / dqii;
  if (fabs(qr) > 1.0e10 || isnan(qr)) {
    if (dqi > 0.0) qr = 1000000.0;
    else qr = -1.0;
  }
  if (qr <= 0.0) qr = 0.0;
  double limitvalue = phi_of(qr, lim);
  return qi - 0.5 * dqii * limitvalue;
}

/* ------------------------------------------------------------- Roe flux
 * ComputeFaceFlux body NS_Solver.H:766-794 with FaceLeftRight (:31-45),
 * Get_inv_flux (FlowVariables.H:59-64), Roeaverage (:51-62), the all-speed
 * roe0c_coef (:66-77) and ComptFaceFlux (:87-113).  Frame velocity is zero
 * (frame_rotate_vel = frame_trans_vel = 0, Userfunction_cy_naca15.cpp:42-44),
 * so its terms (+0) are dropped.                                          */
void orc_roe_flux(const double *t0, const double *t1, const double *t2,
                  const double *t3, int32_t dir, double area,
                  const ibg_physics *ph, int32_t lim, double *a_flux) {
  const double gama = ph->gama, Rg = ph->Rg, Cv = ph->Cv, Ma = ph->Ma;
  double left[NF], right[NF];
  static const int var[5] = {RO, UU, VV, WW, PP};
  for (int k = 0; k < 5; ++k) {
    int v = var[k];
    left[v] = orc_muscl_left(t0[v], t1[v], t2[v], lim);
    right[v] = orc_muscl_right(t1[v], t2[v], t3[v], lim);
  }
  left[TT] = left[PP] / (left[RO] * Rg);
  left[EE] = Cv * left[TT] + (0.5 * (left[UU] * left[UU] + left[VV] * left[VV] + left[WW] * left[WW]));
  right[TT] = right[PP] / (right[RO] * Rg);
  right[EE] = Cv * right[TT] + (0.5 * (right[UU] * right[UU] + right[VV] * right[VV] + right[WW] * right[WW]));
  const double kx = dir == 0 ? 1.0 : 0.0;
  const double ky = dir == 1 ? 1.0 : 0.0;
  const double kz = dir == 2 ? 1.0 : 0.0;
  const double lv = left[UU] * kx + left[VV] * ky + left[WW] * kz;
  const double rv = right[UU] * kx + right[VV] * ky + right[WW] * kz;
  double primflux[2][5];
  primflux[0][0] = left[RO] * lv;
  primflux[0][1] = left[RO] * left[UU] * lv + left[PP] * kx;
  primflux[0][2] = left[RO] * left[VV] * lv + left[PP] * ky;
  primflux[0][3] = left[RO] * left[WW] * lv + left[PP] * kz;
  primflux[0][4] = (left[PP] + left[EE] * left[RO]) * lv;
  primflux[1][0] = right[RO] * rv;
  primflux[1][1] = right[RO] * right[UU] * rv + right[PP] * kx;
  primflux[1][2] = right[RO] * right[VV] * rv + right[PP] * ky;
  primflux[1][3] = right[RO] * right[WW] * rv + right[PP] * kz;
  primflux[1][4] = (right[PP] + right[EE] * right[RO]) * rv;
  const double droe = right[RO] - left[RO];
  const double dp = right[PP] - left[PP];
  const double du[3] = {right[UU] - left[UU], right[VV] - left[VV], right[WW] - left[WW]};
  const double dubar = rv - lv;
  /* Roeaverage */
  double r2lroe = sqrt(right[RO] / left[RO]);
  double r2l0 = 1.0 / (1.0 + r2lroe);
  double h_left = left[EE] + left[PP] / left[RO];
  double h_right = right[EE] + right[PP] / right[RO];
  double r_roe = sqrt(left[RO] * right[RO]);
  double r_u = (left[UU] + right[UU] * r2lroe) * r2l0;
  double r_v = (left[VV] + right[VV] * r2lroe) * r2l0;
  double r_w = (left[WW] + right[WW] * r2lroe) * r2l0;
  double r_h = (h_left + h_right * r2lroe) * r2l0;
  double r_c = sqrt((gama - 1.0) * (r_h - (r_u * r_u + r_v * r_v + r_w * r_w) / 2.0));
  double r_ubar = r_u * kx + r_v * ky + r_w * kz;
  /* ComptFaceFlux */
  const double csquare = r_c * r_c;
  const double namda3 = fabs(r_ubar);
  const double coef12 = 1.0 / (2.0 * csquare) * area;
  const double roe0cxdubar = r_roe * r_c * dubar;
  const double faceareaxabsubar = area * namda3;
  const double halffacearea = 0.5 * area;
  double alpha[8];
  alpha[0] = faceareaxabsubar * (droe - dp / csquare);
  /* roe0c_coef, ROE_ALLSPEED */
  double T_a = (left[TT] + right[TT]) * 0.5;
  double V_a = (lv + rv) * 0.5;
  double Ma_a = fabs(V_a / sqrt(gama * Rg * T_a));
  double f_Ma = fmin(fmax(Ma_a, Ma), 1.0);
  double s0 = (1.0 - f_Ma) * (1.0 - f_Ma);
  double namda0 = sqrt(4.0 * r_c * r_c * f_Ma + s0 * r_ubar * r_ubar);
  double namda1 = fabs(0.5 * ((1.0 + f_Ma) * r_ubar + namda0));
  double namda2 = fabs(0.5 * ((1.0 + f_Ma) * r_ubar - namda0));
  alpha[1] = coef12 * namda1 * (dp + roe0cxdubar);
  alpha[2] = coef12 * namda2 * (dp - roe0cxdubar);
  alpha[3] = alpha[0] + alpha[1] + alpha[2];
  alpha[4] = r_c * (alpha[1] - alpha[2]);
  alpha[5] = faceareaxabsubar * (r_roe * (du[0] - kx * dubar));
  alpha[6] = faceareaxabsubar * (r_roe * (du[1] - ky * dubar));
  alpha[7] = faceareaxabsubar * (r_roe * (du[2] - kz * dubar));
  a_flux[0] = -alpha[3];
  a_flux[1] = -(r_u * alpha[3] + (kx)*alpha[4] + alpha[5]);
  a_flux[2] = -(r_v * alpha[3] + (ky)*alpha[4] + alpha[6]);
  a_flux[3] = -(r_w * alpha[3] + (kz)*alpha[4] + alpha[7]);
  a_flux[4] = -(r_h * alpha[3] + r_ubar * alpha[4] + r_u * alpha[5] + r_v * alpha[6] +
                r_w * alpha[7] - (csquare / (gama - 1.0)) * alpha[0]);
  a_flux[0] += (primflux[0][0] + primflux[1][0]) * halffacearea;
  a_flux[1] += (primflux[0][1] + primflux[1][1]) * halffacearea;
  a_flux[2] += (primflux[0][2] + primflux[1][2]) * halffacearea;
  a_flux[3] += (primflux[0][3] + primflux[1][3]) * halffacearea;
  a_flux[4] += (primflux[0][4] + primflux[1][4]) * halffacearea;
}

static inline double face_area(const double *h, int d) {
  /* ComptFaceArea MeshOp.H:17-36 */
  if (d == 0) return h[1] * h[2];
  if (d == 1) return h[2] * h[0];
  return h[0] * h[1];
}

static inline int nbr(const ibg_mesh_soa *m, int cell, int di, int dj, int dk) {
  return m->nb[(size_t)cell * 27 + (di + 1) * 9 + (dj + 1) * 3 + (dk + 1)];
}

/* ComputeFaceFlux NS_Solver.H:704-832: stencil LL,L,R,RR through Box::neib
 * with the fallback of :743-750, then the flux of every face of every marched
 * level (2D skips z faces; none are built).                               */
void orc_face_flux(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                   const double *fv, int32_t ts, double *faceflux) {
  const int L = m->nlevels;
  for (int c = 0; c < m->ncases; ++c)
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      const double *h = &m->blk_h[3 * b];
#pragma omp parallel for schedule(static)
      for (int f = m->blk_face_start[b]; f < m->blk_face_start[b + 1]; ++f) {
        int d = m->face_dir[f];
        int dn[3] = {0, 0, 0};
        dn[d] = 1;
        int lo = m->face_cells[2 * f], hi = m->face_cells[2 * f + 1];
        int LL = nbr(m, lo, -dn[0], -dn[1], -dn[2]);
        if (LL < 0) LL = nbr(m, hi, dn[0], dn[1], dn[2]);
        int RR = nbr(m, hi, dn[0], dn[1], dn[2]);
        if (RR < 0) RR = nbr(m, lo, -dn[0], -dn[1], -dn[2]);
        orc_roe_flux(&fv[(size_t)LL * NF], &fv[(size_t)lo * NF], &fv[(size_t)hi * NF],
                     &fv[(size_t)RR * NF], d, face_area(h, d), &ph[c], cfg->limiter,
                     &faceflux[(size_t)f * 5]);
      }
    }
}

/* BalanceFaceFlux Fluxmodification.cpp:291-358.  Per Facepair (mdy row:
 * form, coarse face, fine faces), FlowVec arithmetic in the reference's
 * order:
 *   form 0 (fine level refined in x,y,z, :299-318): c_old = F_c*0.25;
 *     F_c *= 0.5; per fine face F_f *= 0.5, F_c += F_f, F_f += c_old*0.5
 *   form 1 (x,y-refined fine level, x/y face, :324-334): c_old = F_c;
 *     F_c *= 0.5; per fine face (two) F_f *= 0.5, F_c += F_f,
 *     F_f += c_old*0.25
 *   form 2 (x,y-refined fine level, z face, :336-348): as form 1 over four
 *     fine faces with F_f += c_old*0.125.                                  */
void orc_balance_flux(const ibg_mesh_soa *m, double *flux5) {
  for (int i = 0; i < m->nmdy; ++i) {
    const int32_t *r = m->mdy + (size_t)i * 6;
    const int form = r[0];
    const int nf = form == 1 ? 2 : 4;
    const double w = form == 0 ? 0.5 : (form == 1 ? 0.25 : 0.125);
    double *fc = &flux5[(size_t)r[1] * 5];
    double cold[5];
    for (int v = 0; v < 5; ++v) cold[v] = form == 0 ? fc[v] * 0.25 : fc[v];
    for (int v = 0; v < 5; ++v) fc[v] *= 0.5;
    for (int k = 0; k < nf; ++k) {
      double *ff = &flux5[(size_t)r[2 + k] * 5];
      for (int v = 0; v < 5; ++v) {
        ff[v] *= 0.5;
        fc[v] += ff[v];
        ff[v] += cold[v] * w;
      }
    }
  }
}

/* ComptViscousStress NS_Solver.cpp:365-629 (laminar) with VarReflect /
 * pfpkesai0 / pfpkesai1 of NS_Solver.H:115-154.  The cell metric keisa is
 * diag(1/dh) (ComptCellParams MeshOp.H:10-12); its zero off-diagonal terms
 * contribute exact zeros and are dropped.                                 */
static void visc_core(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                      const double *fv, const double *nut, int32_t ts, double *vsflux,
                      double *mutflux) {
  const int L = m->nlevels;
  for (int c = 0; c < m->ncases; ++c) {
    const ibg_physics *P = &ph[c];
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      const double *h = &m->blk_h[3 * b];
      const double k0 = 1.0 / h[0], k1 = 1.0 / h[1], k2 = 1.0 / h[2];
#pragma omp parallel for schedule(static)
      for (int f = m->blk_face_start[b]; f < m->blk_face_start[b + 1]; ++f) {
        int d = m->face_dir[f];
        int abox = m->face_cells[2 * f];
        /* vsflowptr[i][j][k] = neighbour (i-1,j-1,k-1) of abox */
#define VP(i, j, k, v) fv[(size_t)m->nb[(size_t)abox * 27 + (i) * 9 + (j) * 3 + (k)] * NF + (v)]
#define PF1(up1, up2, dn1, dn2) (0.25 * ((up1) + (up2) - (dn1) - (dn2)))
        double g[4][3];   /* d(u,v,w,T)/d(kesai0,1,2) */
        static const int vv[4] = {UU, VV, WW, TT};
        for (int q = 0; q < 4; ++q) {
          int v = vv[q];
          if (d == 0) {
            g[q][0] = VP(2, 1, 1, v) - VP(1, 1, 1, v);
            g[q][1] = PF1(VP(1, 2, 1, v), VP(2, 2, 1, v), VP(1, 0, 1, v), VP(2, 0, 1, v));
            g[q][2] = PF1(VP(1, 1, 2, v), VP(2, 1, 2, v), VP(1, 1, 0, v), VP(2, 1, 0, v));
          } else if (d == 1) {
            g[q][1] = VP(1, 2, 1, v) - VP(1, 1, 1, v);
            g[q][0] = PF1(VP(2, 1, 1, v), VP(2, 2, 1, v), VP(0, 1, 1, v), VP(0, 2, 1, v));
            g[q][2] = PF1(VP(1, 1, 2, v), VP(1, 2, 2, v), VP(1, 1, 0, v), VP(1, 2, 0, v));
          } else {
            g[q][2] = VP(1, 1, 2, v) - VP(1, 1, 1, v);
            g[q][0] = PF1(VP(2, 1, 1, v), VP(2, 1, 2, v), VP(0, 1, 1, v), VP(0, 1, 2, v));
            g[q][1] = PF1(VP(1, 2, 1, v), VP(1, 2, 2, v), VP(1, 0, 1, v), VP(1, 0, 2, v));
          }
        }
        double dux = g[0][0] * k0, duy = g[0][1] * k1, duz = g[0][2] * k2;
        double dvx = g[1][0] * k0, dvy = g[1][1] * k1, dvz = g[1][2] * k2;
        double dwx = g[2][0] * k0, dwy = g[2][1] * k1, dwz = g[2][2] * k2;
        double dTx = g[3][0] * k0, dTy = g[3][1] * k1, dTz = g[3][2] * k2;
        int rb = m->face_cells[2 * f + 1];   /* vsflowptr[neibx][neiby][neibz] */
        const double *FL = &fv[(size_t)abox * NF], *FR = &fv[(size_t)rb * NF];
        double ftemp = (FL[TT] + FR[TT]) * 0.5;
        double mu0 = sqrt(pow(ftemp, 3)) * ((1.0 + P->S_over_T_ref) / (ftemp + P->S_over_T_ref));
        double qk0 = P->qk * mu0;
        const double Re = P->Re;
        if (nut) {
          /* NS_Solver.cpp:534-542: mu_t of the face joins mu and the heat
             conduction; mutflux = (nu_f + mu/roe_f) (d nu/d n)/sigma * A with
             the face keisa the unit axis (dvis along d = (nu_R - nu_L)/dh) */
          double fvis = (nut[abox] + nut[rb]) * 0.5;
          double froe = (FL[RO] + FR[RO]) * 0.5;
          double faceroexvis = fvis + mu0 / froe;
          double mut0 = orc_sa_mut(fvis, froe, mu0);
          double kd = d == 0 ? k0 : (d == 1 ? k1 : k2);
          double dvis = (nut[rb] - nut[abox]) * kd;
          mu0 += mut0;
          qk0 += mut0 / (Re * P->Pr_tur * (P->gama - 1) * pow(P->Ma, 2));
          mutflux[f] = faceroexvis * dvis / ORC_SA_SEGMA;
          mutflux[f] *= face_area(h, d);
        }
        double tauxx = 2.0 * mu0 * (dux - (dux + dvy + dwz) / 3.0) / Re;
        double tauyy = 2.0 * mu0 * (dvy - (dux + dvy + dwz) / 3.0) / Re;
        double tauzz = 2.0 * mu0 * (dwz - (dux + dvy + dwz) / 3.0) / Re;
        double tauxy = mu0 / Re * (duy + dvx);
        double tauxz = mu0 / Re * (duz + dwx);
        double tauyz = mu0 / Re * (dvz + dwy);
        double qx = qk0 * dTx, qy = qk0 * dTy, qz = qk0 * dTz;
        double fu = (FL[UU] + FR[UU]) * 0.5, fvv = (FL[VV] + FR[VV]) * 0.5, fw = (FL[WW] + FR[WW]) * 0.5;
        double bbbx = fu * tauxx + fvv * tauxy + fw * tauxz + qx;
        double bbby = fu * tauxy + fvv * tauyy + fw * tauyz + qy;
        double bbbz = fu * tauxz + fvv * tauyz + fw * tauzz + qz;
        double area = face_area(h, d);
        double *o = &vsflux[(size_t)f * 5];
        o[0] = 0.0;
        /* keisa of the face is the unit axis: the product picks one term */
        if (d == 0) { o[1] = tauxx * area; o[2] = tauxy * area; o[3] = tauxz * area; o[4] = bbbx * area; }
        else if (d == 1) { o[1] = tauxy * area; o[2] = tauyy * area; o[3] = tauyz * area; o[4] = bbby * area; }
        else { o[1] = tauxz * area; o[2] = tauyz * area; o[3] = tauzz * area; o[4] = bbbz * area; }
        if (m->dim == 2) o[3] = 0.0;   /* unused in 2D (NS_Solver.cpp:320-324) */
#undef VP
#undef PF1
      }
    }
  }
}

void orc_viscous_flux(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                      const double *fv, int32_t ts, double *vsflux) {
  visc_core(m, ph, cfg, fv, NULL, ts, vsflux, NULL);
}

void orc_viscous_flux_sa(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                         const double *fv, const double *nut, int32_t ts, double *vsflux,
                         double *mutflux) {
  visc_core(m, ph, cfg, fv, nut, ts, vsflux, mutflux);
}

/* ComptSource NS_Solver.cpp:282-363: (F_lo - F_hi) + (Fv_hi - Fv_lo) per
 * direction, summed over directions; 2D leaves component 3 at zero.       */
void orc_source(const ibg_mesh_soa *m, const orc_cfg *cfg, const double *ff,
                const double *vs, int32_t ts, double *source) {
  const int L = m->nlevels;
  const int D = m->dim;
  for (int c = 0; c < m->ncases; ++c)
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
#pragma omp parallel for schedule(static)
      for (int g = s; g < e; ++g) {
        double sinc[3][5];
        memset(sinc, 0, sizeof(sinc));
        for (int fi = 0; fi < D; ++fi) {
          int bf1 = m->cell_faces[(size_t)g * 6 + fi * 2 + 0];
          int bf2 = m->cell_faces[(size_t)g * 6 + fi * 2 + 1];
          for (int v = 0; v < 5; ++v) {
            if (D == 2 && v == 3) continue;
            sinc[fi][v] = ff[(size_t)bf1 * 5 + v] - ff[(size_t)bf2 * 5 + v];
            if (cfg->viscous) sinc[fi][v] += vs[(size_t)bf2 * 5 + v] - vs[(size_t)bf1 * 5 + v];
          }
        }
        for (int v = 0; v < 5; ++v) source[(size_t)g * 5 + v] = sinc[0][v] + sinc[1][v] + sinc[2][v];
      }
    }
}

static inline void get_cons(const double *fv, double *cv) {
  /* Get_Cons_Vars FlowVariables.H:43-48 */
  cv[0] = fv[RO];
  cv[1] = fv[RO] * fv[UU];
  cv[2] = fv[RO] * fv[VV];
  cv[3] = fv[RO] * fv[WW];
  cv[4] = fv[RO] * fv[EE];
}

static inline void get_prim(double *fv, const double *cv, const ibg_physics *P) {
  /* Get_Prim_Vars FlowVariables.H:50-57 with Get_p_energy and
     Get_T_ideal_gas (:21-35) */
  fv[RO] = cv[0];
  fv[UU] = cv[1] / cv[0];
  fv[VV] = cv[2] / cv[0];
  fv[WW] = cv[3] / cv[0];
  fv[EE] = cv[4] / cv[0];
  fv[PP] = (fv[EE] - (0.5 * (fv[UU] * fv[UU] + fv[VV] * fv[VV] + fv[WW] * fv[WW]))) * (P->gama - 1.0) * fv[RO];
  fv[TT] = fv[PP] / (fv[RO] * P->Rg);
}

void orc_old_cons(const ibg_mesh_soa *m, const orc_cfg *cfg, const double *fv, int32_t ts,
                  double *cv_old) {
  const int L = m->nlevels;
  for (int c = 0; c < m->ncases; ++c)
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
#pragma omp parallel for schedule(static)
      for (int g = s; g < e; ++g) get_cons(&fv[(size_t)g * NF], &cv_old[(size_t)g * 5]);
    }
}

/* UpdateConsVars NS_Solver.cpp:844-985 and ComptNewPrimVar :3-194.  The
 * finest level (bodies present) updates only cells that are neither solid
 * nor IB cells; every other level updates all normal cells.  Primitive
 * recovery uses signdis > 0, otherwise the freestream initvar.           */
int32_t orc_update(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                   const double *cv_old, const double *source, const double *dt, int32_t ts,
                   int32_t subt, double *cv_new, double *fv, int64_t *bad_cell) {
  const int L = m->nlevels;
  int32_t code = 0;
  int64_t badc = -1;
  for (int c = 0; c < m->ncases; ++c) {
    const ibg_physics *P = &ph[c];
    const double dt0 = cfg->alpha[subt] * dt[c];
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      const double *h = &m->blk_h[3 * b];
      const double vol = h[0] * h[1] * h[2];
      const double level_ts = dt0 * orc_step_num(L, l, cfg->temporal_refine);
      const int finest = (l == L - 1) && (m->nbody > 0);
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
#pragma omp parallel for schedule(static)
      for (int g = s; g < e; ++g) {
        const int solid = m->signdis[g] < 0.0;
        const int ib = m->ibidx[g] >= 0;
        if (!finest || (!solid && !ib)) {
          for (int v = 0; v < 5; ++v)
            cv_new[(size_t)g * 5 + v] = cv_old[(size_t)g * 5 + v] + source[(size_t)g * 5 + v] * level_ts / vol;
          if (cv_new[(size_t)g * 5] < 0.0) {
#pragma omp critical(orc_err)
            if (!code || g < badc) {
              code = finest ? 611 : 523;
              badc = g;
            }
          }
        }
      }
#pragma omp parallel for schedule(static)
      for (int g = s; g < e; ++g) {
        const int ib = m->ibidx[g] >= 0;
        double *w = &fv[(size_t)g * NF];
        if (!finest) {
          if (m->signdis[g] > 0.0) get_prim(w, &cv_new[(size_t)g * 5], P);
          else memcpy(w, P->initvar, sizeof(double) * NF);
          if (w[TT] < 0.0) {
#pragma omp critical(orc_err)
            if (!code || g < badc) { code = 37; badc = g; }
          }
        } else {
          if (m->signdis[g] > 0.0 && !ib) {
            get_prim(w, &cv_new[(size_t)g * 5], P);
            if (w[TT] < 0.0) {
#pragma omp critical(orc_err)
              if (!code || g < badc) { code = 37; badc = g; }
            }
          } else if (!ib) {
            memcpy(w, P->initvar, sizeof(double) * NF);
          }
        }
      }
    }
  }
  if (bad_cell && code) *bad_cell = badc;
  return code;
}

/* BCBoxTreat BCValues.H:51-188 for the types the cases use: -13 pressure
 * inlet (:191-206), -12 pressure outlet (:208-214), -50 outflow (:216-220),
 * -99 periodic (no-op).  Applied on marched levels only.                   */
static int32_t wall_fill(const ibg_mesh_soa *m, const ibg_physics *P, int i, double *fv,
                         double *nut);

/* nut == NULL: the laminar build, which has no -81 branch (abort 52).  With
   nut, the SA working variable of every ghost too, in the same pass (a
   corner ghost reads an edge ghost that may be a wall ghost).            */
static int32_t bc_fill(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                       double *fv, double *nut, int32_t ts) {
  const int L = m->nlevels;
  for (int i = 0; i < m->ndmg; ++i) {
    int g = m->dmg_cell[i], r = m->dmg_ref[i], ty = m->dmg_type[i];
    int c = 0, l = 0;
    for (int b = 0; b < m->ncases * L; ++b)
      if (g >= m->blk_start[b] && g < m->blk_start[b + 1]) { c = b / L; l = b % L; break; }
    if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
    const ibg_physics *P = &ph[c];
    double *bc = &fv[(size_t)g * NF];
    const double *rf = &fv[(size_t)r * NF];
    if (ty == -99) {
    } else if (ty == -50) {
      memcpy(bc, rf, sizeof(double) * NF);
      if (nut) nut[g] = nut[r];
    } else if (ty == -12) {
      memcpy(bc, rf, sizeof(double) * NF);
      bc[PP] = P->outlet_p;
      bc[RO] = bc[PP] / (P->Rg * bc[TT]);
      if (nut) nut[g] = nut[r];
    } else if (ty == -13) {
      if (nut) nut[g] = P->nut_inlet;
      double speed0 = sqrt(rf[UU] * rf[UU] + rf[VV] * rf[VV] + rf[WW] * rf[WW]);
      bc[UU] = speed0 * cos(P->inlet_angle);
      bc[VV] = speed0 * sin(P->inlet_angle);
      bc[WW] = 0.0;
      bc[TT] = P->inlet_T0 - (0.5 * (bc[UU] * bc[UU] + bc[VV] * bc[VV] + bc[WW] * bc[WW])) / P->Cp;
      bc[PP] = P->inlet_p0 * pow((bc[TT] / P->inlet_T0), 3.5);
      bc[RO] = bc[PP] / (P->Rg * bc[TT]);
      bc[EE] = P->Cv * bc[TT] + (0.5 * (bc[UU] * bc[UU] + bc[VV] * bc[VV] + bc[WW] * bc[WW]));
    } else if (ty == -81 && nut) {
      const int32_t code = wall_fill(m, P, i, fv, nut);
      if (code) return code;
    } else {
      return 52;   /* unidentified boundary type, BCValues.H:170-176 */
    }
  }
  return 0;
}

int32_t orc_bc_fill(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                    double *fv, int32_t ts) {
  return bc_fill(m, ph, cfg, fv, NULL, ts);
}

/* AMR::InterfaceExchange AMRInterface.cpp:3-35: restriction of every due
 * level pair (PairRestriction[_2d] TwoLevelOp.H:217-231, AverageCell[_2d]
 * :45-87), then prolongation (PairProlongation[_2d] :183-215 with
 * GradientInaCell / ComptFineCellData :16-43, TwoLevelOp ctor :348-359).  */
void orc_interface_exchange(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                            double *fv, int32_t ts) {
  const int L = m->nlevels;
  for (int l = 0; l + 1 < L; ++l) {
    if (!orc_transfer_due(L, l, ts, cfg->temporal_refine)) continue;
    for (int i = 0; i < m->nres; ++i) {
      int b = m->res_blk[i];
      if (b % L != l) continue;
      const ibg_physics *P = &ph[b / L];
      const int twod = m->blk_twod[b + 1];
      const int *fi = &m->res_fi[(size_t)i * 8];
      double *cv = &fv[(size_t)m->res_ci[i] * NF];
      double r = 0, u = 0, v = 0, w = 0, e = 0;
      for (int q = 0; q < 8; ++q) {
        if (twod && (q & 1)) continue;
        const double *s = &fv[(size_t)fi[q] * NF];
        r += s[RO];
        u += s[UU] * s[RO];
        v += s[VV] * s[RO];
        w += s[WW] * s[RO];
        e += s[EE] * s[RO];
      }
      cv[UU] = u / r; cv[VV] = v / r; cv[WW] = w / r; cv[EE] = e / r;
      cv[RO] = r * (twod ? 0.25 : 0.125);
      cv[PP] = (cv[EE] - (0.5 * (cv[UU] * cv[UU] + cv[VV] * cv[VV] + cv[WW] * cv[WW]))) * (P->gama - 1.0) * cv[RO];
      cv[TT] = cv[PP] / (cv[RO] * P->Rg);
    }
  }
  /* DataExchange_alllevels(ts+1): no copies on one device (2D z-neighbours
     alias the cell, periodic z neighbours index the source cell directly) */
  const double fincrem[2] = {-0.25, 0.25};
  for (int l = 0; l + 1 < L; ++l) {
    if (!orc_transfer_due(L, l, ts, cfg->temporal_refine)) continue;
    for (int i = 0; i < m->npro; ++i) {
      int b = m->pro_blk[i];
      if (b % L != l) continue;
      const ibg_physics *P = &ph[b / L];
      const int twod = m->blk_twod[b + 1];
      int ci = m->pro_ci[i];
      double U[3][3][3][5];
      static const int cn[7][3] = {{1, 1, 0}, {1, 1, 2}, {1, 0, 1}, {1, 2, 1}, {0, 1, 1}, {2, 1, 1}, {1, 1, 1}};
      for (int k = 0; k < 7; ++k) {
        int a = m->nb[(size_t)ci * 27 + cn[k][0] * 9 + cn[k][1] * 3 + cn[k][2]];
        get_cons(&fv[(size_t)a * NF], U[cn[k][0]][cn[k][1]][cn[k][2]]);
      }
      double gx[5], gy[5], gz[5];
      for (int t = 0; t < 5; ++t) {
        gx[t] = (U[2][1][1][t] - U[0][1][1][t]) * 0.5;
        gy[t] = (U[1][2][1][t] - U[1][0][1][t]) * 0.5;
        gz[t] = (U[1][1][2][t] - U[1][1][0][t]) * 0.5;
      }
      for (int si = 0; si < 2; ++si)
        for (int sj = 0; sj < 2; ++sj)
          for (int sk = 0; sk < (twod ? 1 : 2); ++sk) {
            int f = m->pro_fi[(size_t)i * 8 + si * 4 + sj * 2 + sk];
            double fc[5];
            for (int t = 0; t < 5; ++t) {
              if (twod)
                fc[t] = U[1][1][1][t] + fincrem[si] * gx[t] + fincrem[sj] * gy[t];
              else
                fc[t] = U[1][1][1][t] + fincrem[si] * gx[t] + fincrem[sj] * gy[t] + fincrem[sk] * gz[t];
            }
            get_prim(&fv[(size_t)f * NF], fc, P);
          }
    }
  }
}

static void body_velocity(const ibg_mesh_soa *m, const ibg_body_motion *mo, int body,
                          const double *pt, double *vel) {
  /* Surface_Point_Velocity Body.H:709-721 with zero frame velocity */
  vel[0] = vel[1] = vel[2] = 0.0;
  if (!mo || body < 0) return;
  const ibg_body_motion *M = &mo[body];
  const double *bc = &m->body_center[3 * body];
  double r1[3] = {pt[0] - bc[0], pt[1] - bc[1], pt[2] - bc[2]};
  vel[0] = M->tra_vel[0] + (M->rot_vel[1] * r1[2] - M->rot_vel[2] * r1[1]);
  vel[1] = M->tra_vel[1] + (M->rot_vel[2] * r1[0] - M->rot_vel[0] * r1[2]);
  vel[2] = M->tra_vel[2] + (M->rot_vel[0] * r1[1] - M->rot_vel[1] * r1[0]);
}

static void intp_hg(const double *fv, const int32_t *intp, const double *coef,
                    const ibg_physics *P, double *hg) {
  /* Mesh::IntpHGgcellVar Mesh.cpp:1532-1547 */
  memset(hg, 0, sizeof(double) * NF);
  for (int q = 0; q < 8; ++q) {
    if (intp[q] < 0) continue;   /* 2D: the z layer with zero weight */
    const double *a = &fv[(size_t)intp[q] * NF];
    hg[TT] += coef[q] * a[TT];
    hg[PP] += coef[q] * a[PP];
    hg[UU] += coef[q] * a[UU];
    hg[VV] += coef[q] * a[VV];
    hg[WW] += coef[q] * a[WW];
  }
  hg[RO] = hg[PP] / (P->Rg * hg[TT]);
}

/* Body::IBConditions_Image Body.cpp:1347-1504 (laminar): image state, then
 * LinearDist (Body.cpp:183-204) for signdis >= 0 or the mirror of
 * :1385-1393 otherwise; all IB cells are computed before any is written.  */
int32_t orc_ib_fill(const ibg_mesh_soa *m, const ibg_physics *ph, double *fv,
                    const ibg_body_motion *motion) {
  int32_t code = 0;
  double *tmp = (double *)malloc(sizeof(double) * NF * (size_t)(m->nib > 0 ? m->nib : 1));
  const int L = m->nlevels;
  for (int i = 0; i < m->nib; ++i) {
    int ci = m->ib_cell[i];
    int c = 0;
    for (int b = 0; b < m->ncases * L; ++b)
      if (ci >= m->blk_start[b] && ci < m->blk_start[b + 1]) { c = b / L; break; }
    const ibg_physics *P = &ph[c];
    double hg[NF];
    intp_hg(fv, &m->ib_intp[(size_t)i * 8], &m->ib_coef[(size_t)i * 8], P, hg);
    double patv[3];
    body_velocity(m, motion, m->ib_body[i], &m->ib_attach[3 * (size_t)i], patv);
    double *o = &tmp[(size_t)i * NF];
    memcpy(o, &fv[(size_t)ci * NF], sizeof(double) * NF);
    double sd = m->signdis[ci];
    if (sd < 0.0) {
      o[UU] = 2.0 * patv[0] - hg[UU];
      o[VV] = 2.0 * patv[1] - hg[VV];
      o[WW] = 2.0 * patv[2] - hg[WW];
      o[PP] = hg[PP];
      o[TT] = hg[TT];
    } else {
      double rt = sd + m->ib_hgdis[i];
      double r0 = sd / rt;
      double du[3] = {hg[UU] - patv[0], hg[VV] - patv[1], hg[WW] - patv[2]};
      o[UU] = patv[0] + du[0] * r0;
      o[VV] = patv[1] + du[1] * r0;
      o[WW] = patv[2] + du[2] * r0;
      o[PP] = hg[PP];
      if (o[PP] < 0.0) o[PP] = hg[PP];
      o[TT] = hg[TT];
    }
    o[RO] = o[PP] / (P->Rg * o[TT]);
    o[EE] = P->Cv * o[TT] + (0.5 * (o[UU] * o[UU] + o[VV] * o[VV] + o[WW] * o[WW]));
    if (hg[TT] < 0.0 && !code) code = 1050;
  }
  for (int i = 0; i < m->nib; ++i)
    memcpy(&fv[(size_t)m->ib_cell[i] * NF], &tmp[(size_t)i * NF], sizeof(double) * NF);
  free(tmp);
  return code;
}

/* Comptdt NS_Solver.cpp:196-280: min over every normal cell of every level
 * of CFL*V/sum_d (|u.n_d| + c) A_d (all three directions, also in 2D), then
 * dt = min(CFL_dt, dt*dt_ratio).                                          */
int32_t orc_compute_dt(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                       double *dt) {
  const int L = m->nlevels;
  int32_t code = 0;
  for (int c = 0; c < m->ncases; ++c) {
    const ibg_physics *P = &ph[c];
    double CFL_dt = 10000.0;
    for (int l = 0; l < L; ++l) {
      int b = blk_of(m, c, l);
      const double *h = &m->blk_h[3 * b];
      const double vol = h[0] * h[1] * h[2];
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
      for (int g = s; g < e; ++g) {
        const double *w = &fv[(size_t)g * NF];
        double cs = sqrt(P->gama * P->Rg * w[TT]);
        double convc[3];
        for (int di = 0; di < 3; ++di) {
          double un = di == 0 ? w[UU] : (di == 1 ? w[VV] : w[WW]);
          convc[di] = (fabs(un) + cs) * face_area(h, di);
        }
        double d0 = P->cfl * vol / (convc[0] + convc[1] + convc[2]);
        if (d0 < 0.0 && !code) code = 424;
        CFL_dt = fmin(CFL_dt, d0);
      }
    }
    dt[c] = fmin(CFL_dt, dt[c] * P->dt_ratio);
  }
  return code;
}

/* ResinDomain NS_Solver.H:264-378 (before normalisation): max |source[0]|
 * over the marched non-finest levels, and on the finest level over cells
 * with infectbox == -1 and signdis > 0; the L2 norm over the same cells.  */
void orc_residual(const ibg_mesh_soa *m, const orc_cfg *cfg, const double *source,
                  int32_t ts, double *linf, double *l2) {
  const int L = m->nlevels;
  for (int c = 0; c < m->ncases; ++c) {
    double mx[5] = {0, 0, 0, 0, 0}, sq[5] = {0, 0, 0, 0, 0};
    for (int l = 0; l < L; ++l) {
      const int finest = (l == L - 1) && m->nbody > 0;
      if (!finest && !orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
      for (int g = s; g < e; ++g) {
        if (finest && !(m->ibidx[g] == -1 && m->signdis[g] > 0.0)) continue;
        for (int v = 0; v < 5; ++v) {
          double a = source[(size_t)g * 5 + v];
          mx[v] = fmax(fabs(a), mx[v]);
          sq[v] += a * a;
        }
      }
    }
    for (int v = 0; v < 5; ++v) {
      if (linf) linf[c * 5 + v] = mx[v];
      if (l2) l2[c * 5 + v] = sqrt(sq[v]);
    }
  }
}

/* Body::ComptWallForce Body.cpp:439-558 (laminar) summed over the bodies of
 * a case (Bodymove.cpp:49-55): pressure -p n A, shear mu (du_t/hgdis) A.    */
void orc_forces(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                const ibg_body_motion *motion, double *out) {
  for (int c = 0; c < m->ncases; ++c) {
    double pf[3] = {0, 0, 0}, tf[3] = {0, 0, 0}, pm[3] = {0, 0, 0}, tm[3] = {0, 0, 0};
    const ibg_physics *P = &ph[c];
    int body0 = -1;
    for (int bi = 0; bi < m->nbody; ++bi) if (m->body_case[bi] == c) { body0 = bi; break; }
    for (int i = 0; i < m->npatch; ++i) {
      if (m->pat_case[i] != c) continue;
      double hg[NF];
      intp_hg(fv, &m->pat_intp[(size_t)i * 8], &m->pat_coef[(size_t)i * 8], P, hg);
      const double *pc = &m->pat_pc[3 * (size_t)i];
      const double *nv = &m->pat_nv[3 * (size_t)i];
      double area = m->pat_area[i];
      double patv[3];
      body_velocity(m, motion, body0, pc, patv);
      double du[3] = {hg[UU] - patv[0], hg[VV] - patv[1], hg[WW] - patv[2]};
      const double *bc = body0 >= 0 ? &m->body_center[3 * body0] : pc;
      double dx[3] = {pc[0] - bc[0], pc[1] - bc[1], pc[2] - bc[2]};
      double dn = du[0] * nv[0] + du[1] * nv[1] + du[2] * nv[2];
      double dt_[3];
      for (int k = 0; k < 3; ++k) dt_[k] = (du[k] - nv[k] * dn) / m->pat_hgdis[i];
      double df0[3], df1[3];
      for (int k = 0; k < 3; ++k) {
        df0[k] = dt_[k] * P->mu * area;
        df1[k] = nv[k] * (-hg[PP] * area);
      }
      double dm0[3] = {dx[1] * df0[2] - dx[2] * df0[1], dx[2] * df0[0] - dx[0] * df0[2], dx[0] * df0[1] - dx[1] * df0[0]};
      double dm1[3] = {dx[1] * df1[2] - dx[2] * df1[1], dx[2] * df1[0] - dx[0] * df1[2], dx[0] * df1[1] - dx[1] * df1[0]};
      for (int k = 0; k < 3; ++k) { pf[k] += df1[k]; tf[k] += df0[k]; pm[k] += dm1[k]; tm[k] += dm0[k]; }
    }
    for (int k = 0; k < 3; ++k) {
      out[c * 6 + k] = pf[k] * P->p_ref_area + tf[k] * P->shear_ref_area;
      out[c * 6 + 3 + k] = pm[k] * P->p_ref_area + tm[k] * P->shear_ref_area;
    }
  }
}

int32_t orc_step(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                 double *fv, double *dt, int32_t ts, double *source_out, double *linf,
                 double *l2, int64_t *bad_cell) {
  const size_t N = (size_t)m->ncells, F = (size_t)m->nfaces;
  double *ff = (double *)calloc(F * 5 + 1, sizeof(double));
  double *vs = (double *)calloc(F * 5 + 1, sizeof(double));
  double *src = (double *)calloc(N * 5, sizeof(double));
  double *cvo = (double *)calloc(N * 5, sizeof(double));
  double *cvn = (double *)calloc(N * 5, sizeof(double));
  int32_t code = 0;
  orc_interface_exchange(m, ph, cfg, fv, ts);
  code = orc_ib_fill(m, ph, fv, NULL);
  if (!code && cfg->dt_var_step > 0 && ts % cfg->dt_var_step == 0) code = orc_compute_dt(m, ph, fv, dt);
  if (!code) {
    orc_old_cons(m, cfg, fv, ts, cvo);
    code = orc_bc_fill(m, ph, cfg, fv, ts);
  }
  if (!code && cfg->viscous) {
    orc_viscous_flux(m, ph, cfg, fv, ts, vs);
    if (cfg->reflux) orc_balance_flux(m, vs);   /* NS_Solver.H:672 */
  }
  for (int st = 0; st < cfg->nstage && !code; ++st) {
    orc_face_flux(m, ph, cfg, fv, ts, ff);
    if (cfg->reflux && st == cfg->nstage - 1) orc_balance_flux(m, ff);   /* NS_Solver.H:686 */
    orc_source(m, cfg, ff, vs, ts, src);
    code = orc_update(m, ph, cfg, cvo, src, dt, ts, st, cvn, fv, bad_cell);
  }
  if (!code) orc_residual(m, cfg, src, ts, linf, l2);
  if (source_out) memcpy(source_out, src, sizeof(double) * N * 5);
  free(ff); free(vs); free(src); free(cvo); free(cvn);
  return code;
}

/* ===================================================== SA closure (TURBULENCE)
 * The -D TURBULENCE build of the reference: FlowVariables carries the SA
 * working variable viseddy (here the separate array nut[ncells]).          */

/* ComptMutSource NS_Solver.cpp:632-840 over the non-solid normal cells of
 * the marched levels; mutsource of every other cell is left at 0 (the
 * reference never writes it).                                             */
void orc_mut_source(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                    const double *fv, const double *nut, const double *mutflux, int32_t ts,
                    double *mutsource) {
  const int L = m->nlevels;
  const int D = m->dim;
  for (int c = 0; c < m->ncases; ++c) {
    const ibg_physics *P = &ph[c];
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      const double *h = &m->blk_h[3 * b];
      const double k0 = 1.0 / h[0], k1 = 1.0 / h[1], k2 = 1.0 / h[2];
      const double vol = h[0] * h[1] * h[2];
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
#pragma omp parallel for schedule(static)
      for (int g = s; g < e; ++g) {
        mutsource[g] = 0.0;
        if (m->signdis[g] < 0.0) continue;   /* solid */
        int xl = nbr(m, g, -1, 0, 0), xr = nbr(m, g, 1, 0, 0);
        int yl = nbr(m, g, 0, -1, 0), yr = nbr(m, g, 0, 1, 0);
        int zl = nbr(m, g, 0, 0, -1), zr = nbr(m, g, 0, 0, 1);
        double gn[3] = {(nut[xr] - nut[xl]) / 2.0, (nut[yr] - nut[yl]) / 2.0, (nut[zr] - nut[zl]) / 2.0};
        double gu[3], gv[3], gw[3];
        const double *FX0 = &fv[(size_t)xl * NF], *FX1 = &fv[(size_t)xr * NF];
        const double *FY0 = &fv[(size_t)yl * NF], *FY1 = &fv[(size_t)yr * NF];
        const double *FZ0 = &fv[(size_t)zl * NF], *FZ1 = &fv[(size_t)zr * NF];
        gu[0] = (FX1[UU] - FX0[UU]) / 2.0; gu[1] = (FY1[UU] - FY0[UU]) / 2.0; gu[2] = (FZ1[UU] - FZ0[UU]) / 2.0;
        gv[0] = (FX1[VV] - FX0[VV]) / 2.0; gv[1] = (FY1[VV] - FY0[VV]) / 2.0; gv[2] = (FZ1[VV] - FZ0[VV]) / 2.0;
        gw[0] = (FX1[WW] - FX0[WW]) / 2.0; gw[1] = (FY1[WW] - FY0[WW]) / 2.0; gw[2] = (FZ1[WW] - FZ0[WW]) / 2.0;
        /* total_face_flux over the cell's faces, dirsign = {-1, +1} */
        double tot = 0.0;
        for (int di = 0; di < D; ++di)
          for (int f0 = 0; f0 < 2; ++f0) {
            int fi = m->cell_faces[(size_t)g * 6 + di * 2 + f0];
            tot += (f0 == 0 ? -1.0 : 1.0) * mutflux[fi];
          }
        const double *W = &fv[(size_t)g * NF];
        mutsource[g] = orc_sa_cell_source(W[UU], W[VV], W[WW], W[TT], W[RO], nut[g], m->signdis[g],
                                          gn, gu, gv, gw, k0, k1, k2, P->S_over_T_ref, P->Re,
                                          tot / vol / P->Re);
      }
    }
  }
}

/* The viseddy part of UpdateConsVars (NS_Solver.cpp:873-876, 954-957:
 * max(c_vis_old + mutsource*level_ts, 0)) and of ComptNewPrimVar (m_data =
 * initvar resets viseddy to initvar.viseddy, :32-54, :170-176), on the same
 * cells as orc_update.                                                    */
void orc_sa_update(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                   const double *nut_old, const double *mutsource, const double *dt,
                   int32_t ts, int32_t subt, double *nut) {
  const int L = m->nlevels;
  for (int c = 0; c < m->ncases; ++c) {
    const ibg_physics *P = &ph[c];
    const double dt0 = cfg->alpha[subt] * dt[c];
    for (int l = 0; l < L; ++l) {
      if (!orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      const double level_ts = dt0 * orc_step_num(L, l, cfg->temporal_refine);
      const int finest = (l == L - 1) && (m->nbody > 0);
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
#pragma omp parallel for schedule(static)
      for (int g = s; g < e; ++g) {
        const int solid = m->signdis[g] < 0.0;
        const int ib = m->ibidx[g] >= 0;
        if (!finest || (!solid && !ib)) {
          double v = nut_old[g] + mutsource[g] * level_ts;
          nut[g] = (v < 0.0) ? 0.0 : v;   /* max(v, min_viseddy[_wall] = 0) */
        }
        if (!finest) {
          if (!(m->signdis[g] > 0.0)) nut[g] = P->nut_inf;
        } else if (!ib && !(m->signdis[g] > 0.0)) {
          nut[g] = P->nut_inf;
        }
      }
    }
  }
}

/* A -81 wall ghost (BCValues.H:89-125): a copy of the reference cell when
 * its inward neighbour lies more than 1.0 from the wall, else
 * ImageViseddy_Domain (Turbulence.cpp:140-195) from that neighbour at the
 * reference cell's wall distance, wall at rest.                          */
static int32_t wall_fill(const ibg_mesh_soa *m, const ibg_physics *P, int i, double *fv,
                         double *nut) {
  if (!m->dmg_nbr || !m->dist_dm) return 52;
  const int g = m->dmg_cell[i], r = m->dmg_ref[i], a = m->dmg_nbr[i];
  double *bc = &fv[(size_t)g * NF];
  if (m->dist_dm[a] > 1.0) {
    memcpy(bc, &fv[(size_t)r * NF], sizeof(double) * NF);
    nut[g] = nut[r];
    return 0;
  }
  const double *rv = &fv[(size_t)a * NF];
  const double *nm = &m->dmg_nmv[(size_t)i * 3];
  const double refdis = m->dist_dm[a], targetdis = m->dist_dm[r];
  const double ib_rdis = targetdis / refdis;
  const double hvn = rv[UU] * nm[0] + rv[VV] * nm[1] + rv[WW] * nm[2];
  const double hn[3] = {hvn * nm[0], hvn * nm[1], hvn * nm[2]};
  const double ht[3] = {rv[UU] - hn[0], rv[VV] - hn[1], rv[WW] - hn[2]};
  const double df_velt = sqrt(ht[0] * ht[0] + ht[1] * ht[1] + ht[2] * ht[2]);
  const double T = rv[PP] / (rv[RO] * P->Rg);
  const double hg_mu = sqrt(T * T * T) * ((1.0 + P->S_over_T_ref) / (T + P->S_over_T_ref));
  const double hg_vis = hg_mu / rv[RO];
  const double hg_re = df_velt * refdis / hg_vis * P->Re;
  const double hg_yplus = orc_sa_yplus(hg_re);
  const double ut = hg_yplus * hg_vis / refdis / P->Re;
  const double yplus = ib_rdis * hg_yplus;
  const double ib_velt = orc_sa_uplus(yplus) * ut;
  const double ib_r0 = (fabs(df_velt) < 0.00000001) ? 0.0 : ib_velt / df_velt;
  bc[UU] = hn[0] * ib_rdis + ht[0] * ib_r0;
  bc[VV] = hn[1] * ib_rdis + ht[1] * ib_r0;
  bc[WW] = m->dim == 3 ? hn[2] * ib_rdis + ht[2] * ib_r0 : 0.0;
  bc[PP] = rv[PP];
  bc[TT] = T;
  bc[RO] = bc[PP] / (P->Rg * bc[TT]);
  bc[EE] = P->Cv * bc[TT] + 0.5 * (bc[UU] * bc[UU] + bc[VV] * bc[VV] + bc[WW] * bc[WW]);
  nut[g] = ORC_SA_KAPPA * yplus;
  return 0;
}

/* BCBoxTreat of the TURBULENCE build: the state (as orc_bc_fill) and the
 * viseddy of every domain ghost (BCValues.H:203-205 inlet = facevar, -12
 * and -50 copy the reference cell), and the -81 walls, in one pass.      */
int32_t orc_sa_bc_fill(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                       double *fv, double *nut, int32_t ts) {
  return bc_fill(m, ph, cfg, fv, nut, ts);
}

/* viseddy of the AMR transfer: AverageCell[_2d] TwoLevelOp.H:129-179 (plain
 * mean of the sons) and GradientInaCell / ComptFineCellData[_2d] :91-127
 * (central gradient over the cneib neighbours, sons at +-1/4), on the same
 * schedule as orc_interface_exchange.                                     */
void orc_sa_exchange(const ibg_mesh_soa *m, const orc_cfg *cfg, double *nut, int32_t ts) {
  const int L = m->nlevels;
  for (int l = 0; l + 1 < L; ++l) {
    if (!orc_transfer_due(L, l, ts, cfg->temporal_refine)) continue;
    for (int i = 0; i < m->nres; ++i) {
      int b = m->res_blk[i];
      if (b % L != l) continue;
      const int twod = m->blk_twod[b + 1];
      const int *fi = &m->res_fi[(size_t)i * 8];
      double v = 0.0;
      for (int q = 0; q < 8; ++q) {
        if (twod && (q & 1)) continue;
        v += nut[fi[q]];
      }
      nut[m->res_ci[i]] = v * (twod ? 0.25 : 0.125);
    }
  }
  const double fincrem[2] = {-0.25, 0.25};
  for (int l = 0; l + 1 < L; ++l) {
    if (!orc_transfer_due(L, l, ts, cfg->temporal_refine)) continue;
    for (int i = 0; i < m->npro; ++i) {
      int b = m->pro_blk[i];
      if (b % L != l) continue;
      const int twod = m->blk_twod[b + 1];
      int ci = m->pro_ci[i];
      double g0 = (nut[nbr(m, ci, 1, 0, 0)] - nut[nbr(m, ci, -1, 0, 0)]) * 0.5;
      double g1 = (nut[nbr(m, ci, 0, 1, 0)] - nut[nbr(m, ci, 0, -1, 0)]) * 0.5;
      double g2 = (nut[nbr(m, ci, 0, 0, 1)] - nut[nbr(m, ci, 0, 0, -1)]) * 0.5;
      for (int si = 0; si < 2; ++si)
        for (int sj = 0; sj < 2; ++sj)
          for (int sk = 0; sk < (twod ? 1 : 2); ++sk) {
            int f = m->pro_fi[(size_t)i * 8 + si * 4 + sj * 2 + sk];
            if (twod)
              nut[f] = nut[ci] + fincrem[si] * g0 + fincrem[sj] * g1;
            else
              nut[f] = nut[ci] + fincrem[si] * g0 + fincrem[sj] * g1 + fincrem[sk] * g2;
          }
    }
  }
}

/* Body::IBConditions_Image with TURBULENCE (Body.cpp:1393-1426):
 * signdis < 0 -> initvar with the wall velocity; otherwise ImageViseddy
 * (Turbulence.cpp:57-138).  All IB cells are computed before any is written. */
int32_t orc_ib_fill_sa(const ibg_mesh_soa *m, const ibg_physics *ph, double *fv, double *nut,
                       const ibg_body_motion *motion) {
  int32_t code = 0;
  const size_t n = (size_t)(m->nib > 0 ? m->nib : 1);
  double *tmp = (double *)malloc(sizeof(double) * (NF + 1) * n);
  const int L = m->nlevels;
  for (int i = 0; i < m->nib; ++i) {
    int ci = m->ib_cell[i];
    int c = 0;
    for (int b = 0; b < m->ncases * L; ++b)
      if (ci >= m->blk_start[b] && ci < m->blk_start[b + 1]) { c = b / L; break; }
    const ibg_physics *P = &ph[c];
    double hg[NF];
    intp_hg(fv, &m->ib_intp[(size_t)i * 8], &m->ib_coef[(size_t)i * 8], P, hg);
    double patv[3];
    body_velocity(m, motion, m->ib_body[i], &m->ib_attach[3 * (size_t)i], patv);
    const double *nmv = &m->ib_nv[3 * (size_t)i];
    double *o = &tmp[(size_t)i * (NF + 1)];
    double sd = m->signdis[ci];
    if (sd < 0.0) {
      memcpy(o, P->initvar, sizeof(double) * NF);
      o[UU] = patv[0]; o[VV] = patv[1]; o[WW] = patv[2];
      o[NF] = P->nut_inf;
    } else {
      double rt0 = m->ib_hgdis[i] + sd;
      double ib_rdis = sd / rt0;
      double hvn = hg[UU] * nmv[0] + hg[VV] * nmv[1] + hg[WW] * nmv[2];
      double hn[3] = {hvn * nmv[0], hvn * nmv[1], hvn * nmv[2]};
      double ht[3] = {hg[UU] - hn[0], hg[VV] - hn[1], hg[WW] - hn[2]};
      double wvn = patv[0] * nmv[0] + patv[1] * nmv[1] + patv[2] * nmv[2];
      double wn[3] = {nmv[0] * wvn, nmv[1] * wvn, nmv[2] * wvn};
      double wt[3] = {patv[0] - wn[0], patv[1] - wn[1], patv[2] - wn[2]};
      double df[3] = {ht[0] - wt[0], ht[1] - wt[1], ht[2] - wt[2]};
      double df_velt = sqrt(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
      double hg_mu = sqrt(pow(hg[TT], 3)) * ((1.0 + P->S_over_T_ref) / (hg[TT] + P->S_over_T_ref));
      double hg_vis = hg_mu / hg[RO];
      double hg_re = df_velt * rt0 / hg_vis * P->Re;
      double hg_yplus = orc_sa_yplus(hg_re);
      double ut = hg_yplus * hg_vis / rt0 / P->Re;
      double yplus = ib_rdis * hg_yplus;
      double uplus = orc_sa_uplus(yplus);
      double ib_velt = uplus * ut;
      double ib_r0 = (fabs(df_velt) < 0.00000001) ? 0.0 : ib_velt / df_velt;
      for (int k = 0; k < 3; ++k) {
        double ibn = wn[k] + (hn[k] - wn[k]) * ib_rdis;
        double ibt = wt[k] + df[k] * ib_r0;
        o[UU + k] = ibn + ibt;
      }
      o[PP] = hg[PP];
      o[TT] = hg[TT] + 0.5 * pow(P->Pr, 1.0 / 3.0) / P->Cp * (pow(df_velt, 2) - pow(ib_velt, 2));
      o[RO] = o[PP] / (P->Rg * o[TT]);
      o[EE] = P->Cv * o[TT] + (0.5 * (o[UU] * o[UU] + o[VV] * o[VV] + o[WW] * o[WW]));
      o[NF] = ORC_SA_KAPPA * yplus;
    }
    if (hg[TT] < 0.0 && !code) code = 1050;
  }
  for (int i = 0; i < m->nib; ++i) {
    memcpy(&fv[(size_t)m->ib_cell[i] * NF], &tmp[(size_t)i * (NF + 1)], sizeof(double) * NF);
    nut[m->ib_cell[i]] = tmp[(size_t)i * (NF + 1) + NF];
  }
  free(tmp);
  return code;
}

/* Body::ComptWallForce with TURBULENCE (Body.cpp:505-553): Patchut
 * (Turbulence.cpp:197-226) wall shear roe_w ut^2 along the slip direction,
 * pressure -p n A (FiltSinglePatchPressure is the identity: Filt.cpp:5-6),
 * both scaled by p_ref*area_ref.                                           */
void orc_forces_sa(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                   const ibg_body_motion *motion, double *out) {
  for (int c = 0; c < m->ncases; ++c) {
    double pf[3] = {0, 0, 0}, tf[3] = {0, 0, 0}, pm[3] = {0, 0, 0}, tm[3] = {0, 0, 0};
    const ibg_physics *P = &ph[c];
    int body0 = -1;
    for (int bi = 0; bi < m->nbody; ++bi) if (m->body_case[bi] == c) { body0 = bi; break; }
    for (int i = 0; i < m->npatch; ++i) {
      if (m->pat_case[i] != c) continue;
      double hg[NF];
      intp_hg(fv, &m->pat_intp[(size_t)i * 8], &m->pat_coef[(size_t)i * 8], P, hg);
      const double *pc = &m->pat_pc[3 * (size_t)i];
      const double *nv = &m->pat_nv[3 * (size_t)i];
      double area = m->pat_area[i], hgdis = m->pat_hgdis[i];
      double patv[3];
      body_velocity(m, motion, body0, pc, patv);
      const double *bc = body0 >= 0 ? &m->body_center[3 * body0] : pc;
      double dx[3] = {pc[0] - bc[0], pc[1] - bc[1], pc[2] - bc[2]};
      double hv = hg[UU] * nv[0] + hg[VV] * nv[1] + hg[WW] * nv[2];
      double pn = patv[0] * nv[0] + patv[1] * nv[1] + patv[2] * nv[2];
      double df[3];
      df[0] = (hg[UU] - nv[0] * hv) - (patv[0] - nv[0] * pn);
      df[1] = (hg[VV] - nv[1] * hv) - (patv[1] - nv[1] * pn);
      df[2] = (hg[WW] - nv[2] * hv) - (patv[2] - nv[2] * pn);
      double dv = sqrt(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
      double vt[3] = {0.0, 0.0, 0.0};
      if (!(dv < 1e-8)) { vt[0] = df[0] / dv; vt[1] = df[1] / dv; vt[2] = df[2] / dv; }
      double hg_mu = sqrt(pow(hg[TT], 3)) * ((1.0 + P->S_over_T_ref) / (hg[TT] + P->S_over_T_ref));
      double hg_re = hg[RO] * dv * hgdis / hg_mu * P->Re;
      double yplus = orc_sa_yplus(hg_re);
      double ut = yplus * (hg_mu / hg[RO]) / hgdis / P->Re;
      double wall_T = hg[TT] + 0.5 * pow(P->Pr, 1.0 / 3.0) / P->Cp * pow(dv, 2);
      double wall_roe = hg[PP] / (P->Rg * wall_T);
      double Cf = wall_roe * pow(ut, 2);
      double df0[3], df1[3];
      for (int k = 0; k < 3; ++k) {
        df0[k] = vt[k] * Cf * area;
        df1[k] = nv[k] * (-hg[PP] * area);
      }
      double dm0[3] = {dx[1] * df0[2] - dx[2] * df0[1], dx[2] * df0[0] - dx[0] * df0[2], dx[0] * df0[1] - dx[1] * df0[0]};
      double dm1[3] = {dx[1] * df1[2] - dx[2] * df1[1], dx[2] * df1[0] - dx[0] * df1[2], dx[0] * df1[1] - dx[1] * df1[0]};
      for (int k = 0; k < 3; ++k) { pf[k] += df1[k]; tf[k] += df0[k]; pm[k] += dm1[k]; tm[k] += dm0[k]; }
    }
    for (int k = 0; k < 3; ++k) {
      out[c * 6 + k] = pf[k] * P->p_ref_area + tf[k] * P->p_ref_area;
      out[c * 6 + 3 + k] = pm[k] * P->p_ref_area + tm[k] * P->p_ref_area;
    }
  }
}

/* ResinDomain's res_vis_sum (NS_Solver.H:297, 321) and the L2 norm, over the
 * cells of orc_residual.                                                  */
void orc_residual_nut(const ibg_mesh_soa *m, const orc_cfg *cfg, const double *mutsource,
                      int32_t ts, double *linf, double *l2) {
  const int L = m->nlevels;
  for (int c = 0; c < m->ncases; ++c) {
    double mx = 0.0, sq = 0.0;
    for (int l = 0; l < L; ++l) {
      const int finest = (l == L - 1) && m->nbody > 0;
      if (!finest && !orc_marched(L, l, ts, cfg->temporal_refine)) continue;
      int b = blk_of(m, c, l);
      int s = m->blk_start[b], e = s + m->blk_nnormal[b];
      for (int g = s; g < e; ++g) {
        if (finest && !(m->ibidx[g] == -1 && m->signdis[g] > 0.0)) continue;
        double a = mutsource[g];
        mx = fmax(fabs(a), mx);
        sq += a * a;
      }
    }
    if (linf) linf[c] = mx;
    if (l2) l2[c] = sqrt(sq);
  }
}

/* One step of the TURBULENCE build (main.cpp:84-99, TimeMarching
 * NS_Solver.H:657-702 with ComptMutSource after ComptSource).              */
int32_t orc_step_sa(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                    double *fv, double *nut, double *dt, int32_t ts, double *source_out,
                    double *mutsource_out, double *linf, double *l2, double *nlinf,
                    double *nl2, int64_t *bad_cell) {
  const size_t N = (size_t)m->ncells, F = (size_t)m->nfaces;
  double *ff = (double *)calloc(F * 5 + 1, sizeof(double));
  double *vs = (double *)calloc(F * 5 + 1, sizeof(double));
  double *mf = (double *)calloc(F + 1, sizeof(double));
  double *src = (double *)calloc(N * 5, sizeof(double));
  double *msrc = (double *)calloc(N + 1, sizeof(double));
  double *cvo = (double *)calloc(N * 5, sizeof(double));
  double *cvn = (double *)calloc(N * 5, sizeof(double));
  double *nuo = (double *)calloc(N + 1, sizeof(double));
  int32_t code = 0;
  orc_interface_exchange(m, ph, cfg, fv, ts);
  orc_sa_exchange(m, cfg, nut, ts);
  code = orc_ib_fill_sa(m, ph, fv, nut, NULL);
  if (!code && cfg->dt_var_step > 0 && ts % cfg->dt_var_step == 0) code = orc_compute_dt(m, ph, fv, dt);
  if (!code) {
    orc_old_cons(m, cfg, fv, ts, cvo);
    memcpy(nuo, nut, sizeof(double) * N);   /* c_vis_old */
    code = bc_fill(m, ph, cfg, fv, nut, ts);
  }
  if (!code) {
    orc_viscous_flux_sa(m, ph, cfg, fv, nut, ts, vs, mf);
    if (cfg->reflux) orc_balance_flux(m, vs);   /* NS_Solver.H:672 */
  }
  for (int st = 0; st < cfg->nstage && !code; ++st) {
    orc_face_flux(m, ph, cfg, fv, ts, ff);
    if (cfg->reflux && st == cfg->nstage - 1) orc_balance_flux(m, ff);   /* NS_Solver.H:686 */
    orc_source(m, cfg, ff, vs, ts, src);
    orc_mut_source(m, ph, cfg, fv, nut, mf, ts, msrc);
    code = orc_update(m, ph, cfg, cvo, src, dt, ts, st, cvn, fv, bad_cell);
    orc_sa_update(m, ph, cfg, nuo, msrc, dt, ts, st, nut);
  }
  if (!code) {
    orc_residual(m, cfg, src, ts, linf, l2);
    orc_residual_nut(m, cfg, msrc, ts, nlinf, nl2);
  }
  if (source_out) memcpy(source_out, src, sizeof(double) * N * 5);
  if (mutsource_out) memcpy(mutsource_out, msrc, sizeof(double) * N);
  free(ff); free(vs); free(mf); free(src); free(msrc); free(cvo); free(cvn); free(nuo);
  return code;
}
