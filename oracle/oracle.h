/*
 * oracle.h — FP64 CPU restatement of the reference's per-step residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (libibgpu.so and its host code) never links or calls it.
 *
 * State is the reference's AoS FlowVariables (roe,u,v,w,p,T,e; 7 doubles per
 * cell, FlowVariables.H:320-326) over the flattened mesh of ibgpu.h.  Every
 * function follows the reference's arithmetic in the same operation order;
 * each cites the lines it restates.
 */
#ifndef ORACLE_H_
#define ORACLE_H_

#include <stdint.h>
#include "ibgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg {
  int32_t nstage;            /* sub_time_step  NS_Solver.H:200            */
  double alpha[4];           /* tp[]           NS_Solver.H:202-231        */
  int32_t dt_var_step;       /* NS_Solver.H:661                            */
  int32_t temporal_refine;   /* TEMPORAL_REFINE                            */
  int32_t limiter;           /* 0 van Leer, 1 van Albada                   */
  int32_t viscous;           /* VISCOUSITY                                 */
  int32_t reflux;            /* optional coarse/fine balance (off in the
                                reference: its call sites are commented);
                                meaningful only without TEMPORAL_REFINE    */
} orc_cfg;

/* schedule (mpiparams.cpp:186-212) */
int32_t orc_marched(int32_t L, int32_t level, int32_t ts, int32_t temporal_refine);
int32_t orc_transfer_due(int32_t L, int32_t coarse_level, int32_t ts, int32_t temporal_refine);
double orc_step_num(int32_t L, int32_t level, int32_t temporal_refine);

/* limiter on one variable (Slopelimiter.H:22-58 + NS_Solver.H:21-27)     */
double orc_muscl_left(double q_i, double qi, double qi1, int32_t limiter);
double orc_muscl_right(double q_i, double qi, double qi1, int32_t limiter);
/* inviscid face flux from a 4-cell stencil of FlowVariables (7 doubles)
   NS_Solver.H:51-113, 766-794                                            */
void orc_roe_flux(const double *fLL, const double *fL, const double *fR,
                  const double *fRR, int32_t dir, double area,
                  const ibg_physics *ph, int32_t limiter, double *flux5);

/* BalanceFaceFlux Fluxmodification.cpp:291-358 on a face-flux array
   [nfaces][5] over the mesh's flux-balance faces (mdy)                    */
void orc_balance_flux(const ibg_mesh_soa *m, double *flux5);

/* per-phase functions over the whole mesh; only marched blocks are touched */
void orc_face_flux(const ibg_mesh_soa *m, const ibg_physics *ph,
                   const orc_cfg *cfg, const double *fv, int32_t ts,
                   double *faceflux);
void orc_viscous_flux(const ibg_mesh_soa *m, const ibg_physics *ph,
                      const orc_cfg *cfg, const double *fv, int32_t ts,
                      double *vsflux);
void orc_source(const ibg_mesh_soa *m, const orc_cfg *cfg,
                const double *faceflux, const double *vsflux, int32_t ts,
                double *source);
/* U_old <- cons(W) (NS_Solver.H:392-424) */
void orc_old_cons(const ibg_mesh_soa *m, const orc_cfg *cfg, const double *fv,
                  int32_t ts, double *cv_old);
/* UpdateConsVars + ComptNewPrimVar; returns 0 or the abort code; *bad_cell */
int32_t orc_update(const ibg_mesh_soa *m, const ibg_physics *ph,
                   const orc_cfg *cfg, const double *cv_old,
                   const double *source, const double *dt, int32_t ts,
                   int32_t subt, double *cv_new, double *fv, int64_t *bad_cell);
int32_t orc_bc_fill(const ibg_mesh_soa *m, const ibg_physics *ph,
                    const orc_cfg *cfg, double *fv, int32_t ts);
void orc_interface_exchange(const ibg_mesh_soa *m, const ibg_physics *ph,
                            const orc_cfg *cfg, double *fv, int32_t ts);
int32_t orc_ib_fill(const ibg_mesh_soa *m, const ibg_physics *ph, double *fv,
                    const ibg_body_motion *motion);
int32_t orc_compute_dt(const ibg_mesh_soa *m, const ibg_physics *ph,
                       const double *fv, double *dt /* [ncases] in/out */);
void orc_residual(const ibg_mesh_soa *m, const orc_cfg *cfg,
                  const double *source, int32_t ts, double *linf, double *l2);
void orc_forces(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                const ibg_body_motion *motion, double *out /* [ncases*6] */);
/* One full step of main.cpp:84-99: InterfaceExchange, IBConditions_Image,
   TimeMarching.  fv, dt updated in place; source (may be NULL) receives the
   last stage's source[0]; linf/l2 (may be NULL) the residual.             */
int32_t orc_step(const ibg_mesh_soa *m, const ibg_physics *ph,
                 const orc_cfg *cfg, double *fv, double *dt, int32_t ts,
                 double *source, double *linf, double *l2, int64_t *bad_cell);

/* ---------------------------------------------- SA closure (TURBULENCE)
 * nut[ncells] is FlowVariables::viseddy of the -D TURBULENCE build.       */
double orc_sa_mut(double nu, double roe, double mu);          /* Turbulence.cpp:23-30 */
double orc_sa_yplus(double re);                                /* Turbulence.cpp:32-42 */
double orc_sa_uplus(double yplus);                             /* Turbulence.cpp:44-55 */
/* one cell of ComptMutSource (NS_Solver.cpp:648-808); g, gu, gv, gw are the
   central differences /2 along kesai of nu, u, v, w; mutf the face term   */
double orc_sa_cell_source(double u, double v, double w, double T, double roe, double nu,
                          double sd, const double *g, const double *gu, const double *gv,
                          const double *gw, double k0, double k1, double k2, double SoT,
                          double Re, double mutf);
void orc_viscous_flux_sa(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                         const double *fv, const double *nut, int32_t ts, double *vsflux,
                         double *mutflux /* [nfaces] */);
void orc_mut_source(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                    const double *fv, const double *nut, const double *mutflux, int32_t ts,
                    double *mutsource /* [ncells] */);
void orc_sa_update(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                   const double *nut_old, const double *mutsource, const double *dt,
                   int32_t ts, int32_t subt, double *nut);
int32_t orc_sa_bc_fill(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                       double *fv, double *nut, int32_t ts);
void orc_sa_exchange(const ibg_mesh_soa *m, const orc_cfg *cfg, double *nut, int32_t ts);
int32_t orc_ib_fill_sa(const ibg_mesh_soa *m, const ibg_physics *ph, double *fv, double *nut,
                       const ibg_body_motion *motion);
void orc_forces_sa(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                   const ibg_body_motion *motion, double *out);
void orc_residual_nut(const ibg_mesh_soa *m, const orc_cfg *cfg, const double *mutsource,
                      int32_t ts, double *linf, double *l2);
int32_t orc_step_sa(const ibg_mesh_soa *m, const ibg_physics *ph, const orc_cfg *cfg,
                    double *fv, double *nut, double *dt, int32_t ts, double *source_out,
                    double *mutsource_out, double *linf, double *l2, double *nlinf,
                    double *nl2, int64_t *bad_cell);

/* ------------------------------------------- learned closure (mlp.c)
 * prm = the 4801 network parameters then fmean[8] fstd[8] flo[8] fhi[8]
 * zmean zstd, as the engine's parameter vector.                           */
float orc_bf16(float x);
void orc_mlp_forward(const float *prm, const float *x, int64_t n, float *z);
void orc_mlp_grad(const float *prm, const float *x, const float *z, int64_t n, double *grad,
                  double *loss);
double orc_mlp_loss_f64(const double *p, const double *xs, const double *t, int64_t n);
void orc_mlp_grad_f64(const double *p, const double *xs, const double *t, int64_t n,
                      double *grad, double *loss);
int32_t orc_closure_features(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                             const double *nut, int32_t cell, double *x, double *z);
int32_t orc_closure_cell(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                         const double *nut, int32_t cell, double *x, double *z, double *scale);
void orc_closure_cells(const ibg_mesh_soa *m, const ibg_physics *ph, const double *fv,
                       const double *nut, int64_t n, const int32_t *cells, double *x, double *z,
                       double *scale, int8_t *ok);

/* ------------------------------------------- moving bodies (retag.c)
 * Surface of 2D sections: point offsets from the pivot at zero pose, the
 * segment of each patch (mesh pat_* order) and the patch range per body.  */
typedef struct orc_surface {
  int32_t npoint, npatch, nbody;
  const double *pt_off;        /* [npoint][2]                               */
  const int32_t *pat_pt0;      /* [npatch] first point of the segment       */
  const int32_t *pat_body;     /* [npatch] global body                      */
  const int32_t *pat_case;     /* [npatch]                                  */
  const int8_t *pat_flip;      /* [npatch] odd body: normal reversed        */
  const double *pat_dz;        /* [npatch] z extent of the patch            */
  const int32_t *body_p0;      /* [nbody+1] patch range of each body        */
  const double *case_origin;   /* [ncases][3]                               */
} orc_surface;

/* In/out state of one move (caller-allocated).  signdis, cell_patch, ib
 * (0/1 per cell) and pat_pc/pat_nv/pat_area hold the previous step on
 * entry and the new one on return; the IB list comes back in cell order.  */
typedef struct orc_retag_out {
  double *signdis;
  int32_t *cell_patch;
  int8_t *ib;
  double *pat_pc, *pat_nv, *pat_area;
  int32_t ib_cap, nib;
  int32_t *ib_cell, *ib_intp, *ib_body;
  double *ib_coef, *ib_hgdis, *ib_nv, *ib_attach;
  int32_t *pat_intp, *pat_close;
  double *pat_coef, *pat_hgdis;
  int32_t err_code;
  int64_t err_cell;
} orc_retag_out;

/* Body::MovingBody_Image Bodymove.cpp:37-125 between regrids; pose
 * [ncases][4] = {pivot x, pivot y, cos a, sin a}.  Returns 0 or the
 * reference's abort code (263, 289, 314), 1082 (IB cell on a ghost), 1090
 * (list capacity), 1183 / 1575 (no image point).                           */
int32_t orc_move_bodies(const ibg_mesh_soa *m, const orc_surface *s, const double *pose,
                        orc_retag_out *o);

#ifdef __cplusplus
}
#endif
#endif
