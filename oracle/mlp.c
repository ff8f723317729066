/*
 * mlp.c — CPU restatement of the learned SA closure (features, network
 * forward and backward) used to check the GPU kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): tests/ load it as the checker.
 *
 * The reference has no learned component; the network replaces the
 * point-wise SA term of ComptMutSource (NS_Solver.cpp:700-767), which this
 * file takes from oracle.c.  Two restatements of the network:
 *   orc_mlp_forward / orc_mlp_grad     the GPU arithmetic: BF16 rounding of
 *                                      inputs, weights and activations at the
 *                                      same points, sums in double;
 *   orc_mlp_grad_f64 / orc_mlp_loss_f64 plain FP64 (no rounding), pinned by
 *                                      finite differences in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define NP 4801
#define OW1 0
#define OB1 512
#define OW2 576
#define OB2 4672
#define OW3 4736
#define OB3 4800

/* round-to-nearest-even FP32 -> BF16 -> FP32 */
float orc_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;   /* inf / nan */
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  float y;
  memcpy(&y, &u, 4);
  return y;
}

/* The hidden units are ReLU units; the derivative is taken as 0 at z = 0
 * (as the GPU's z > 0 test). */
static float relu_f(float z) { return z > 0.0f ? z : 0.0f; }
static float drelu_f(float z) { return z > 0.0f ? 1.0f : 0.0f; }

/* prm: parameters then fmean[8] fstd[8] flo[8] fhi[8] zmean zstd */
static void fwd_row(const float *prm, const float *xraw, float *xb, float *h1, float *s1,
                    float *h2, float *s2, float *y) {
  const float *fm = prm + NP, *fs = prm + NP + 8;
  /* standardised with the FP32 reciprocal of the deviation, as the GPU */
  for (int k = 0; k < 8; ++k) {
    const float inv = 1.0f / fs[k];
    xb[k] = orc_bf16((xraw[k] - fm[k]) * inv);
  }
  for (int j = 0; j < 64; ++j) {
    double a = 0.0;
    for (int k = 0; k < 8; ++k) a += (double)xb[k] * (double)orc_bf16(prm[OW1 + k * 64 + j]);
    float z = (float)(a + prm[OB1 + j]);
    h1[j] = orc_bf16(relu_f(z));
    s1[j] = orc_bf16(drelu_f(z));
  }
  double yy = 0.0;
  for (int j = 0; j < 64; ++j) {
    double a = 0.0;
    for (int i = 0; i < 64; ++i) a += (double)h1[i] * (double)orc_bf16(prm[OW2 + i * 64 + j]);
    float z = (float)(a + prm[OB2 + j]);
    h2[j] = orc_bf16(relu_f(z));
    s2[j] = orc_bf16(drelu_f(z));
    yy += (double)h2[j] * (double)orc_bf16(prm[OW3 + j]);
  }
  *y = (float)(yy + prm[OB3]);
}

void orc_mlp_forward(const float *prm, const float *x, int64_t n, float *z) {
  const float zm = prm[NP + 32], zs = prm[NP + 33];
  for (int64_t r = 0; r < n; ++r) {
    float xb[8], h1[64], s1[64], h2[64], s2[64], y;
    fwd_row(prm, &x[r * 8], xb, h1, s1, h2, s2, &y);
    z[r] = zm + zs * y;
  }
}

void orc_mlp_grad(const float *prm, const float *x, const float *z, int64_t n, double *grad,
                  double *loss) {
  const float zm = prm[NP + 32], zs = prm[NP + 33];
  memset(grad, 0, sizeof(double) * NP);
  double L = 0.0;
  const float invB = 1.0f / (float)n;
  for (int64_t r = 0; r < n; ++r) {
    float xb[8], h1[64], s1[64], h2[64], s2[64], y;
    fwd_row(prm, &x[r * 8], xb, h1, s1, h2, s2, &y);
    float t = (z[r] - zm) / zs;
    float e = y - t;
    L += (double)e * e;
    float dy = 2.0f * e * invB;
    float dz2[64], dz1[64];
    for (int j = 0; j < 64; ++j) {
      dz2[j] = orc_bf16(dy * orc_bf16(prm[OW3 + j]) * s2[j]);
      grad[OW3 + j] += (double)h2[j] * dy;
      grad[OB2 + j] += dz2[j];
    }
    grad[OB3] += dy;
    for (int i = 0; i < 64; ++i) {
      double a = 0.0;
      for (int j = 0; j < 64; ++j) {
        grad[OW2 + i * 64 + j] += (double)h1[i] * dz2[j];
        a += (double)dz2[j] * orc_bf16(prm[OW2 + i * 64 + j]);
      }
      dz1[i] = orc_bf16((float)a * s1[i]);
      grad[OB1 + i] += dz1[i];
    }
    for (int k = 0; k < 8; ++k)
      for (int i = 0; i < 64; ++i) grad[OW1 + k * 64 + i] += (double)xb[k] * dz1[i];
  }
  if (loss) *loss = L / (double)n;
}

/* plain FP64 network on standardised inputs xs[n][8] and targets t[n] */
static double relu_d(double z) { return z > 0.0 ? z : 0.0; }
static double drelu_d(double z) { return z > 0.0 ? 1.0 : 0.0; }

double orc_mlp_loss_f64(const double *p, const double *xs, const double *t, int64_t n) {
  double L = 0.0;
  for (int64_t r = 0; r < n; ++r) {
    double h1[64], y = p[OB3];
    for (int j = 0; j < 64; ++j) {
      double a = p[OB1 + j];
      for (int k = 0; k < 8; ++k) a += xs[r * 8 + k] * p[OW1 + k * 64 + j];
      h1[j] = relu_d(a);
    }
    for (int j = 0; j < 64; ++j) {
      double a = p[OB2 + j];
      for (int i = 0; i < 64; ++i) a += h1[i] * p[OW2 + i * 64 + j];
      y += relu_d(a) * p[OW3 + j];
    }
    L += (y - t[r]) * (y - t[r]);
  }
  return L / (double)n;
}

void orc_mlp_grad_f64(const double *p, const double *xs, const double *t, int64_t n, double *grad,
                      double *loss) {
  memset(grad, 0, sizeof(double) * NP);
  double L = 0.0;
  for (int64_t r = 0; r < n; ++r) {
    double z1[64], h1[64], z2[64], h2[64], y = p[OB3];
    for (int j = 0; j < 64; ++j) {
      double a = p[OB1 + j];
      for (int k = 0; k < 8; ++k) a += xs[r * 8 + k] * p[OW1 + k * 64 + j];
      z1[j] = a;
      h1[j] = relu_d(a);
    }
    for (int j = 0; j < 64; ++j) {
      double a = p[OB2 + j];
      for (int i = 0; i < 64; ++i) a += h1[i] * p[OW2 + i * 64 + j];
      z2[j] = a;
      h2[j] = relu_d(a);
      y += h2[j] * p[OW3 + j];
    }
    double e = y - t[r];
    L += e * e;
    double dy = 2.0 * e / (double)n;
    double dz2[64];
    for (int j = 0; j < 64; ++j) {
      grad[OW3 + j] += h2[j] * dy;
      dz2[j] = dy * p[OW3 + j] * drelu_d(z2[j]);
      grad[OB2 + j] += dz2[j];
    }
    grad[OB3] += dy;
    for (int i = 0; i < 64; ++i) {
      double a = 0.0;
      for (int j = 0; j < 64; ++j) {
        grad[OW2 + i * 64 + j] += h1[i] * dz2[j];
        a += dz2[j] * p[OW2 + i * 64 + j];
      }
      double dz1 = a * drelu_d(z1[i]);
      grad[OB1 + i] += dz1;
      for (int k = 0; k < 8; ++k) grad[OW1 + k * 64 + i] += xs[r * 8 + k] * dz1;
    }
  }
  if (loss) *loss = L / (double)n;
}

/* Features and the plain-closure target of one cell (the sampler's
 * arithmetic in FP64 from the AoS state): central differences over the
 * Box::neib neighbours, |Omega| as ComptMutSource forms it, the point-wise
 * SA term of orc_sa_cell_source with the derivative terms removed.
 * Returns 0 when the cell is outside the network's domain.                */
int32_t orc_closure_features(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                             const double *nut, int32_t cell, double *x, double *z) {
  double scale;
  return orc_closure_cell(m, ph, fv, nut, cell, x, z, &scale);
}

/* The same, also returning the production scale nu * S~ that turns the
 * network's output into the point-wise SA term (term = z * scale).        */
int32_t orc_closure_cell(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                         const double *nut, int32_t cell, double *x, double *z, double *scale_out) {
  const int L = m->nlevels;
  int b = 0;
  for (int q = 0; q < m->ncases * L; ++q)
    if (cell >= m->blk_start[q] && cell < m->blk_start[q + 1]) { b = q; break; }
  const ibg_physics *P = &ph[b / L];
  const double *h = &m->blk_h[3 * b];
  const int32_t *nb = &m->nb[(size_t)cell * 27];
  const double *W = &fv[(size_t)cell * IBG_NFV];
  const double *XL = &fv[(size_t)nb[0 * 9 + 1 * 3 + 1] * IBG_NFV], *XR = &fv[(size_t)nb[2 * 9 + 1 * 3 + 1] * IBG_NFV];
  const double *YL = &fv[(size_t)nb[1 * 9 + 0 * 3 + 1] * IBG_NFV], *YR = &fv[(size_t)nb[1 * 9 + 2 * 3 + 1] * IBG_NFV];
  const double *ZL = &fv[(size_t)nb[1 * 9 + 1 * 3 + 0] * IBG_NFV], *ZR = &fv[(size_t)nb[1 * 9 + 1 * 3 + 2] * IBG_NFV];
  double dux = (XR[1] - XL[1]) * 0.5 / h[0], dvx = (XR[2] - XL[2]) * 0.5 / h[0], dwx = (XR[3] - XL[3]) * 0.5 / h[0];
  double duy = (YR[1] - YL[1]) * 0.5 / h[1], dvy = (YR[2] - YL[2]) * 0.5 / h[1], dwy = (YR[3] - YL[3]) * 0.5 / h[1];
  double duz = 0, dvz = 0, dwz = 0;
  if (m->dim == 3) {
    duz = (ZR[1] - ZL[1]) * 0.5 / h[2]; dvz = (ZR[2] - ZL[2]) * 0.5 / h[2]; dwz = (ZR[3] - ZL[3]) * 0.5 / h[2];
  }
  double ox = 0.5 * (dwy - dvz), oy = 0.5 * (duz - dwx), oz = 0.5 * (dvx - duy);
  double olen = sqrt(2.0) * sqrt(ox * ox + oy * oy + oz * oz);
  double roe = W[0], p = W[4];
  double T = p / (roe * P->Rg);
  double mu0 = sqrt(T * T * T) * ((1.0 + P->S_over_T_ref) / (T + P->S_over_T_ref));
  double nul = mu0 / roe, nu = nut[cell], d = m->signdis[cell];
  double eps = 1e-8;
  double speed = sqrt(W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
  const double kap2 = 0.41 * 0.41;
  double shat = olen + nu / (kap2 * d * d * P->Re);
  x[0] = log(olen + eps);
  x[1] = shat > 0.0 ? olen / shat : 0.0;
  x[2] = log(d > eps ? d : eps);
  x[3] = log(nu / nul + eps);
  x[4] = log(speed * d * P->Re / nul + eps);
  x[5] = roe;
  x[6] = p / P->initvar[4];
  x[7] = log(nul);
  double scale = nu * shat;
  *scale_out = scale;
  if (!(d > eps && nu > eps && scale > 0.0)) return 0;
  /* point-wise SA term: the full cell source minus its derivative terms,
     evaluated with zero gradients of nu and zero face term               */
  double g0[3] = {0, 0, 0};
  double gu[3] = {dux * h[0], duy * h[1], duz * h[2]};
  double gv[3] = {dvx * h[0], dvy * h[1], dvz * h[2]};
  double gw[3] = {dwx * h[0], dwy * h[1], dwz * h[2]};
  double q = orc_sa_cell_source(W[1], W[2], W[3], T, roe, nu, d, g0, gu, gv, gw, 1.0 / h[0],
                                1.0 / h[1], 1.0 / h[2], P->S_over_T_ref, P->Re, 0.0);
  *z = q / scale;
  return 1;
}

/* orc_closure_cell over a list of cells (ok[k] = its return value). */
void orc_closure_cells(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                       const double *nut, int64_t n, const int32_t *cells, double *x, double *z,
                       double *scale, int8_t *ok) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k)
    ok[k] = (int8_t)orc_closure_cell(m, ph, fv, nut, cells[k], x + 8 * k, z + k, scale + k);
}
