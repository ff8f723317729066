/*
 * ibgpu.h — C ABI of the MI355X residual engine (libibgpu.so).
 *
 * The engine replaces, for one GPU holding the whole mesh, the per-stage
 * residual path of the Cartesian-AMR immersed-boundary solver: MUSCL
 * reconstruction with the van Leer limiter, the all-speed Roe inviscid flux,
 * the viscous face flux, the image-point immersed-boundary ghost fill, the
 * domain-boundary ghosts, the coarse/fine transfer and the low-storage
 * multistage update over the AMR level set.
 *
 * Every entry point below names the reference interface it replaces
 * (file:line in the reference tree).  All entry points are stream ordered on
 * the context's HIP stream, are not re-entrant per context, and return an
 * ibg_status.  No torch types cross this boundary: plain pointers and sizes.
 *
 * Data contract.  The host (a C++ NS_Solver-shaped class, or a Python/ctypes
 * driver) hands over one flattened mesh, ibg_mesh_soa, whose arrays stay owned
 * by the caller; the library copies what it needs into device memory.  Cell
 * state crosses the boundary as the reference's AoS FlowVariables layout
 * (roe,u,v,w,p,T,e as 7 doubles per cell, FlowVariables.H:320-326).
 */
#ifndef IBGPU_H_
#define IBGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBG_MAX_LEVELS 12
#define IBG_MAX_CASES 64
#define IBG_NFV 7          /* doubles per FlowVariables (laminar) */

/* ------------------------------------------------------------------ status */
typedef enum ibg_status {
  IBG_OK = 0,
  IBG_ERR_ARG = 1,          /* bad argument / size mismatch                 */
  IBG_ERR_HIP = 2,          /* HIP runtime error                            */
  IBG_ERR_STATE = 3,        /* call out of order (e.g. no mesh uploaded)    */
  /* numerical aborts of the reference, returned instead of MPI_Abort       */
  IBG_ERR_NEG_DENSITY = 523, /* NS_Solver.cpp:888-894 (611 finest level)    */
  IBG_ERR_NEG_T = 37,        /* NS_Solver.cpp:55-66, 139-155                 */
  IBG_ERR_NAN_RESIDUAL = 268,/* NS_Solver.H:355-375                          */
  IBG_ERR_LIMITER = 57,      /* Slopelimiter.H:54-57                         */
  IBG_ERR_NEG_DT = 424,      /* NS_Solver.cpp:253-258                        */
  IBG_ERR_NEG_HG_T = 1050    /* Body.cpp:1429-1437                           */
} ibg_status;

/* cell kinds on a level (Box.H:63-66 types plus the AMR pair role) */
enum ibg_cell_kind {
  IBG_NORMAL = 0,   /* marched leaf cell (m_box [0,num_without_ghost))       */
  IBG_DMGHOST = 1,  /* domain-boundary ghost (Dmghost)                       */
  IBG_RES1 = 2,     /* coarse cell under fine cells, f_res_ghost[0]          */
  IBG_RES2 = 3,     /* coarse cell under fine cells, f_res_ghost[1]          */
  IBG_PRO = 4       /* fine cell over coarse cells, son in f_pro_ghost       */
};

/* precision modes */
enum ibg_mode {
  IBG_MODE_FP64 = 0,   /* plain: every operation in FP64                     */
  IBG_MODE_MIXED = 1   /* FP32 limiter/Roe/viscous arithmetic, FP64 state,
                          accumulation, norms and wall interpolation          */
};

/* ------------------------------------------------------------- mesh (SoA)
 * One or more cases (an ensemble) flattened into one cell index space.
 * Cells of case c, level l occupy [blk_start[b], blk_start[b+1]) with
 * b = c*nlevels + l; the first blk_nnormal[b] of them are normal cells
 * (Morton order), the rest are ghosts.  Neighbour indices are global.
 * Canonical keys: (case, level, ix, iy, iz).                              */
typedef struct ibg_mesh_soa {
  int32_t dim;                 /* 2 or 3                                      */
  int32_t ncases;
  int32_t nlevels;             /* max_mesh_level, same for every case         */
  int32_t ncells;
  const int32_t *blk_start;    /* [ncases*nlevels+1]                          */
  const int32_t *blk_nnormal;  /* [ncases*nlevels]                            */
  const double  *blk_h;        /* [ncases*nlevels][3] cell size dh[l]         */
  const int32_t *blk_twod;     /* [ncases*nlevels] level refined in x,y only  */
  /* per cell */
  const int32_t *key;          /* [ncells][3]  (ix,iy,iz) on its level        */
  const int8_t  *kind;         /* [ncells]     ibg_cell_kind                  */
  const int32_t *nb;           /* [ncells][27] Box::neib[i][j][k] at i*9+j*3+k*/
  const int32_t *cell_faces;   /* [ncells][6]  Box::faces[d][s] at d*2+s      */
  const double  *signdis;      /* [ncells]     BoxtoWall::signdis             */
  const int32_t *ibidx;        /* [ncells]     Mesh::infectbox (-1 = none)    */
  /* faces, grouped per block (case,level); Face{bi[2],fnv} Face.H:9-15     */
  int32_t nfaces;
  const int32_t *blk_face_start; /* [ncases*nlevels+1]                        */
  const int32_t *face_cells;   /* [nfaces][2] bi[0]=low side, bi[1]=high side */
  const int8_t  *face_dir;     /* [nfaces]    fnv                             */
  /* domain ghosts, Domainghost{cell,refcell,nbface} GhostPair.H:20-38       */
  int32_t ndmg;
  const int32_t *dmg_cell, *dmg_ref, *dmg_type;  /* type = BC code (-13,...)  */
  /* coarse/fine pairs CtoFPair{ci, fi.son[2][2][2]} GhostPair.H:6-18        */
  int32_t nres;                /* f_res_ghost[0] then [1], per coarse level   */
  const int32_t *res_ci, *res_fi;   /* res_fi [nres][8], son[i][j][k] i*4+j*2+k*/
  const int32_t *res_blk;      /* [nres] coarse block b                       */
  int32_t npro;                /* f_pro_ghost                                 */
  const int32_t *pro_ci, *pro_fi, *pro_blk;
  /* immersed-boundary cells, IBCell/HGCell Box.H:331-395 (finest level)    */
  int32_t nib;
  const int32_t *ib_cell;      /* [nib]     IBCell::ci                        */
  const int32_t *ib_intp;      /* [nib][8]  HGCell::intpcell                  */
  const double  *ib_coef;      /* [nib][8]  HGCell::intpcoef                  */
  const double  *ib_hgdis;     /* [nib]     HGCell::hgdis                     */
  const double  *ib_nv;        /* [nib][3]  closest patch normal              */
  const double  *ib_attach;    /* [nib][3]  HGCell::attach_pt                 */
  const int32_t *ib_body;      /* [nib]     BoxtoWall::body                   */
  /* surface patches used for the wall force, Surfpatch Box.H:497-544       */
  int32_t npatch;
  const int32_t *pat_intp;     /* [npatch][8] */
  const double  *pat_coef;     /* [npatch][8] */
  const double  *pat_hgdis;    /* [npatch]    */
  const double  *pat_nv;       /* [npatch][3] */
  const double  *pat_pc;       /* [npatch][3] */
  const double  *pat_area;     /* [npatch]    */
  const int32_t *pat_case;     /* [npatch]    */
  /* bodies: rigid motion for the wall velocity, Body.H:709-721              */
  int32_t nbody;
  const int32_t *body_case;    /* [nbody]     */
  const double  *body_center;  /* [nbody][3]  bodycenter                     */
  /* domain walls (BC -81 of the TURBULENCE build, BCValues.H:89-125); all
     three NULL without walls                                               */
  const int32_t *dmg_nbr;      /* [ndmg] the reference cell's neighbour one
                                  step further inside (-1: not a wall ghost)  */
  const double  *dmg_nmv;      /* [ndmg][3] Domainghost::nmv, wall normal     */
  const double  *dist_dm;      /* [ncells] BoxtoWall::distance_to_dm
                                  (Mesh.cpp:1638-1757), 9999 away from walls  */
  /* coarse/fine flux-balance faces: Facepair{cface, fface[2][2]}
     GhostPair.H:67-84 as AMR::FindModifyFace Fluxmodification.cpp:5-289
     builds them, one per face between a normal cell and a coarse ghost
     whose sons are normal cells.  mdy [nmdy][6]: the balance form (0: fine
     level refined in x, y and z, four fine faces; 1: fine level refined in
     x and y only, x or y face, two fine faces; 2: the same, z face, four
     fine faces), the coarse face, then the fine faces (faces of the coarse
     cell's sons on that side, the first transverse axis outer; -1 past the
     count).  Used only by the optional balance (ibg_config.reflux).        */
  int32_t nmdy;
  const int32_t *mdy;
} ibg_mesh_soa;

/* Nondimensional constants of one case, FlowConstants.cpp:5-44 and the case
 * globals of Userfunction_cy_naca15.cpp:18-65, 356-425.                    */
typedef struct ibg_physics {
  double gama, Ma, Re, Pr, Rg, Cv, Cp, mu, qk, S_over_T_ref;
  double initvar[IBG_NFV];     /* freestream FlowVariables                    */
  double inlet_T0, inlet_p0;   /* facevar[0].T / .p  (total state at inlet)   */
  double inlet_angle;          /* inlet_flow_angle                            */
  double outlet_p;             /* facevar[1].p + backpressure                 */
  double cfl, dt_ratio;        /* CFL, dt_ratio                               */
  double dt_init;              /* dt before the first Comptdt                 */
  double force_ref;            /* F_ref for C_L/C_D                           */
  double moment_ref;           /* M_ref                                       */
  double p_ref_area, shear_ref_area;  /* p_ref*area_ref, shear_ref*area_ref   */
  /* SA closure (used when ibg_config.turbulence = 1)                        */
  double nut_inf;              /* initvar.viseddy = far_vis_ratio*viseddy
                                  (Userfunction_cy_naca15.cpp:368-370)         */
  double nut_inlet;            /* facevar[0].viseddy of the -13 inlet
                                  (BCValues.H:203-205); the reference never
                                  sets it (BCValues.H:32), 0 is the value a
                                  fresh allocation holds                      */
  double Pr_tur;               /* FlowConstants.cpp:11                         */
} ibg_physics;

/* Rigid body motion of one body at the current step (Body.H:709-721).     */
typedef struct ibg_body_motion {
  double tra_vel[3];           /* free_tra_vel + force_tra_vel + frame_trans  */
  double rot_vel[3];           /* force_rot_vel + free_rot_vel (about center) */
} ibg_body_motion;

typedef struct ibg_config {
  int32_t mode;                /* ibg_mode                                    */
  int32_t nstage;              /* sub_time_step: 1, 3 or 4 (NS_Solver.H:200)  */
  double  alpha[4];            /* tp[] stage coefficients (NS_Solver.H:202-231)*/
  int32_t dt_var_step;         /* Comptdt period (NS_Solver.H:661)            */
  int32_t temporal_refine;     /* TEMPORAL_REFINE schedule (mpiparams.cpp:186)*/
  int32_t limiter;             /* 0 van Leer (Slopelimiter.H:17-20), 1 van Albada */
  int32_t viscous;             /* VISCOUSITY (NS_Solver.H:15)                 */
  int32_t tile;                /* cells per stage-kernel tile (0 = default)   */
  int32_t turbulence;          /* 1: Spalart-Allmaras closure of the
                                  -D TURBULENCE build (Turbulence.cpp,
                                  NS_Solver.cpp:632-840); 0: laminar          */
  int32_t schedule;            /* stage-kernel schedule: 0 one workgroup per
                                  tile; 1 persistent workgroups (as many as
                                  the device holds at once) looping over the
                                  tiles, the learned closure's weights staged
                                  once per workgroup                          */
  int32_t reflux;              /* 1: coarse/fine flux balance, BalanceFaceFlux
                                  Fluxmodification.cpp:291-358, on the viscous
                                  fluxes once per step and on the inviscid
                                  fluxes of the last stage (the call sites the
                                  reference leaves commented, NS_Solver.H:672,
                                  686).  The reference compiles it only
                                  without TEMPORAL_REFINE: requires
                                  temporal_refine = 0.  Default 0 (off).      */
} ibg_config;

typedef struct ibg_ctx ibg_ctx;

/* Library / context lifetime.  Replaces the NS_Solver constructor
 * NS_Solver(Mesh&, Domain&, int) NS_Solver.H:195-237 (stage table, solver
 * containers) and AMR::AllocateDataMem AMR.H:116-124.                      */
ibg_status ibg_create(const ibg_config *cfg, ibg_ctx **out);
ibg_status ibg_destroy(ibg_ctx *ctx);
const char *ibg_version(void);

/* Check every index the kernels will dereference (neighbour, stencil, pair,
 * image-point and patch lists) without touching the GPU.  The reference
 * asserts the same facts under -D DEBUG (NS_Solver.H:741, NS_Solver.cpp:421,
 * TwoLevelOp.H:188, Mesh.cpp:1539).  ibg_upload_mesh runs it first.        */
ibg_status ibg_check_mesh(const ibg_mesh_soa *mesh, const char **msg);

/* Upload the flattened mesh (after construction and after every regrid).
 * Replaces CheckDataContainer NS_Solver.H:549-569 (container sizing) and the
 * pair lists consumed by AMR::InterfaceExchange AMRInterface.cpp:3-110.     */
ibg_status ibg_upload_mesh(ibg_ctx *ctx, const ibg_mesh_soa *mesh,
                           const ibg_physics *phys /* [ncases] */);

/* State in/out as AoS FlowVariables [ncells][7].  Replaces direct access to
 * Meshlevel::m_data Meshlevel.H:233 (and InitFlow NS_Solver.H:432-446).    */
ibg_status ibg_set_state(ibg_ctx *ctx, const double *fv_aos, int64_t ncells);
ibg_status ibg_get_state(ibg_ctx *ctx, double *fv_aos, int64_t ncells);
/* SA working variable per cell, FlowVariables::viseddy (FlowVariables.H:
 * 327-329, present with -D TURBULENCE); [ncells].  IBG_ERR_STATE when the
 * context runs without the closure.                                        */
ibg_status ibg_set_nut(ibg_ctx *ctx, const double *nut, int64_t ncells);
ibg_status ibg_get_nut(ibg_ctx *ctx, double *nut, int64_t ncells);

/* Per-step ghost fills.
 * BCValues::BCBoxTreat(Mesh&) BCValues.H:51-188                            */
ibg_status ibg_bc_fill(ibg_ctx *ctx, int32_t ts);
/* AMR::InterfaceExchange() AMRInterface.cpp:3-35 (restriction, periodic
 * copy, prolongation on their sub-cycling schedules)                       */
ibg_status ibg_interface_exchange(ibg_ctx *ctx, int32_t ts);
/* Body::IBConditions_Image(Mesh&, vector<Body>&) Body.cpp:1347-1504.
 * motion: [nbody] or NULL for stationary bodies.                           */
ibg_status ibg_ib_fill(ibg_ctx *ctx, const ibg_body_motion *motion);

/* Solver phases (NS_Solver public methods).
 * NS_Solver::Comptdt() NS_Solver.cpp:196-280; dt_out [ncases] (may be NULL) */
ibg_status ibg_compute_dt(ibg_ctx *ctx, int32_t ts, double *dt_out);
/* NS_Solver::ComptOldConsVar() NS_Solver.H:392-424                         */
ibg_status ibg_old_cons(ibg_ctx *ctx, int32_t ts);
/* NS_Solver::ComptViscousStress() NS_Solver.cpp:365-629 (frozen per step) */
ibg_status ibg_viscous_flux(ibg_ctx *ctx, int32_t ts);
/* One stage: ComputeFaceFlux NS_Solver.H:704-832 + ComptSource
 * NS_Solver.cpp:282-363 + UpdateConsVars NS_Solver.cpp:844-985 +
 * ComptNewPrimVar NS_Solver.cpp:3-194 (+ periodic copy NS_Solver.H:698).   */
ibg_status ibg_stage(ibg_ctx *ctx, int32_t ts, int32_t subt);
/* NS_Solver::TimeMarching(AMR&) NS_Solver.H:657-702 (dt, U_old, BC, viscous,
 * all stages, residual) for nsteps steps starting at ts0, each preceded by
 * InterfaceExchange and IBConditions_Image as in main.cpp:84-99.  When
 * use_graph is nonzero the steps are replayed from captured HIP graphs, one
 * per run of steps up to a boundary of the level schedule's period
 * (2^(L-1) steps; any ts0 and nsteps), captured on first use; with a
 * dt_var_step that is neither 1 nor a multiple of the period the steps run
 * as direct launches.                                                      */
ibg_status ibg_time_march(ibg_ctx *ctx, int32_t ts0, int32_t nsteps,
                          int32_t use_graph);
/* Capture (without running anything) the graphs ibg_time_march(ts0, nsteps,
 * 1) will replay, so that the call itself is pure replay.  No counterpart
 * in the reference (its march has no launch sequence to prepare).          */
ibg_status ibg_prepare_march(ibg_ctx *ctx, int32_t ts0, int32_t nsteps);
/* Residual of the last stage.  NS_Solver::ResinDomain() NS_Solver.H:264-378:
 * linf[case*5+v] = max |source[0][v]| over the marched cells (unnormalised),
 * l2[case*5+v]   = sqrt(sum source[0][v]^2) over the same cells.            */
ibg_status ibg_residual(ibg_ctx *ctx, double *linf, double *l2);
/* SA residual of the last stage: res_vis_sum of ResinDomain NS_Solver.H:
 * 297, 321 (max |mutsource| over the same cells, unnormalised) and its L2
 * norm; [ncases] each.                                                     */
ibg_status ibg_residual_nut(ibg_ctx *ctx, double *linf, double *l2);
/* Body::ComptWallForce(Mesh&) Body.cpp:439-558, summed over the bodies of
 * each case (Bodymove.cpp:49-55): out[case*6 + {Fx,Fy,Fz,Mx,My,Mz}] in the
 * reference's dimensional units (force, moment about body 0 center).       */
ibg_status ibg_forces(ibg_ctx *ctx, const ibg_body_motion *motion,
                      double *out);

/* Parity probes: copy device-side intermediate arrays to the host.
 * faceflux / vsflux [nfaces][5]: NS_Solver::faceflux / vsflux NS_Solver.H:850-851
 * source [ncells][5] (zero for non-marched / ghost cells): source[0]       */
ibg_status ibg_get_faceflux(ibg_ctx *ctx, double *out, int64_t nfaces);
ibg_status ibg_get_vsflux(ibg_ctx *ctx, double *out, int64_t nfaces);
ibg_status ibg_get_source(ibg_ctx *ctx, double *out, int64_t ncells);
/* SA closure probes: mutflux [nfaces] of the current state
 * (NS_Solver::mutflux NS_Solver.H:867, NS_Solver.cpp:541-542) and the
 * mutsource [ncells] of the last stage (NS_Solver.H:868) on the cells the
 * update consumes it (0 on IB cells, which the reference evaluates but
 * never uses).                                                             */
ibg_status ibg_get_mutflux(ibg_ctx *ctx, double *out, int64_t nfaces);
ibg_status ibg_get_nut_source(ibg_ctx *ctx, double *out, int64_t ncells);
ibg_status ibg_get_dt(ibg_ctx *ctx, double *dt /* [ncases] */);
ibg_status ibg_set_dt(ibg_ctx *ctx, const double *dt /* [ncases] */);

/* First offending cell of the last numerical abort (global cell index), or
 * -1.  Replaces the printf + MPI_Abort diagnostics of NS_Solver.cpp:55-66,
 * 888-894.                                                                  */
ibg_status ibg_last_error(ibg_ctx *ctx, int64_t *cell, int32_t *code);

/* ------------------------------------------------------------ moving bodies
 * Rigid motion of the bodies of 2D cases on a fixed level set: the per-step
 * part of Body::MovingBody_Image (Bodymove.cpp:37-125) that runs between
 * regrids.  One move recomputes the surface patches (ComptNewLocation /
 * ComptPatchParams :3-35), the closest patch and signed distance of every
 * cell (RenewLevelDistance :127-258), the solid flags, the IB set
 * (RemoveIBCells / InfectNewIBCells :345-767), the image points and weights
 * of the IB cells (FindImageCell Body.cpp:999-1190) and of the wall-force
 * patches (FindPatchHangingNode Body.cpp:1506-1575), and the stage masks.
 * The reference point of every body of a case becomes the case's pivot.   */
typedef struct ibg_surface {
  int32_t npoint;
  const double  *pt_off;       /* [npoint][2] surface point minus the pivot
                                  at zero pose (Body::surfptoff)              */
  const int32_t *pat_pt0;      /* [npatch] first point of the patch segment
                                  (the mesh's pat_* order)                    */
  const int32_t *pat_body;     /* [npatch] body of the patch                  */
  const int8_t  *pat_flip;     /* [npatch] 1: outward normal is the right-
                                  hand normal of the segment (lower surface)  */
  const double  *pat_dz;       /* [npatch] z extent: area = length * dz       */
  const int32_t *body_p0;      /* [nbody+1] patches of body b are
                                  [body_p0[b], body_p0[b+1]), one polyline    */
  const double  *case_origin;  /* [ncases][3] low corner of key (0,0,0)       */
  const int32_t *cell_patch;   /* [ncells] BoxtoWall::patch, -1: no body      */
  const int32_t *pat_close;    /* [npatch] HGCell::closecell of the patch     */
  int32_t ib_capacity;         /* IB list capacity (0: 2*nib + 4096)          */
} ibg_surface;

/* Pose of the bodies of one case: point = pivot + R(a) * offset,
 * R(a) = [[cosa, -sina], [sina, cosa]].                                    */
typedef struct ibg_body_pose {
  double pivot[2];
  double cosa, sina;
} ibg_body_pose;

/* Prescribed pitch and heave of one case (the FSI configuration's motion):
 * aoa(t) = aoa0 + amp sin(omega t) degrees nose up about the pivot,
 * pivot y(t) = y0 + heave sin(omega t); the rotation is a = -aoa.        */
typedef struct ibg_pitch_heave {
  double x0, y0, aoa0_deg, amp_deg, omega, heave;
} ibg_pitch_heave;

/* Enable moving bodies for the uploaded mesh (2D only).  mesh is the mesh
 * given to ibg_upload_mesh; the surface describes its bodies.             */
ibg_status ibg_upload_surface(ibg_ctx *ctx, const ibg_mesh_soa *mesh,
                              const ibg_surface *surf);
/* Move the bodies to pose[ncases] now (stream-ordered; returns the
 * reference's abort code 263 / 289 / 314 when the move is too large for
 * the incremental update, 1082 for an IB cell on a ghost, 1090 when the
 * IB list outgrows its capacity, 1183 / 1575 when no image point exists). */
ibg_status ibg_move_bodies(ibg_ctx *ctx, const ibg_body_pose *pose);
/* Move the bodies by law[ncases] at the start of every step of
 * ibg_time_march (the pose at each case's time t, t0 at this call, t += dt
 * after each step) with the law's wall velocity in the IB fill and, for
 * motion == NULL, in ibg_forces.  Inside a march each step's move runs
 * ahead, beside the stages of the step before, on a second stream of the
 * context; results and abort codes are those of the move in place.
 * law == NULL stops it.                                                    */
ibg_status ibg_set_motion(ibg_ctx *ctx, const ibg_pitch_heave *law, double t0);
ibg_status ibg_get_time(ibg_ctx *ctx, double *t /* [ncases] */);
/* The law's time of every case (carried over a regrid).                    */
ibg_status ibg_set_time(ibg_ctx *ctx, const double *t /* [ncases] */);
/* Probes of the moving-body geometry (host cell and patch indices).
 * signdis / cell_patch / ib: [ncells] each (any may be NULL).             */
ibg_status ibg_get_geometry(ibg_ctx *ctx, double *signdis, int32_t *cell_patch, int8_t *ib,
                            int32_t *nib);
/* The IB list (cap entries at most, cell order): IBCell::ci, intpcell[8],
 * intpcoef[8], hgdis, patch normal, attach point, body.                    */
ibg_status ibg_get_ib_list(ibg_ctx *ctx, int32_t cap, int32_t *cell, int32_t *intp,
                           double *coef, double *hgdis, double *nv, double *attach,
                           int32_t *body);
/* Wall-force patches: centre, normal, area, image cells and weights.       */
ibg_status ibg_get_patches(ibg_ctx *ctx, double *pc, double *nv, double *area, int32_t *intp,
                           double *coef, double *hgdis);

/* Wait for all work of the context.                                        */
ibg_status ibg_synchronize(ibg_ctx *ctx);

/* Time (ms, HIP events on the context stream) of the last ibg_time_march:
 * whole region and the per-stage kernel's average duration.                */
ibg_status ibg_last_timing(ibg_ctx *ctx, double *total_ms,
                           double *stage_kernel_avg_ms, int64_t *stage_launches);

/* Measurement probe with no counterpart in the reference: launch the stage
 * kernel of step ts, stage subt, reps times back to back on the context
 * stream between two HIP events.  avg_ms = elapsed / reps; cells and faces
 * are the normal cells and tile faces one launch processes.  The state is
 * advanced by the repeated stages (use it after the run being measured).   */
ibg_status ibg_bench_stage(ibg_ctx *ctx, int32_t ts, int32_t subt, int32_t reps,
                           double *avg_ms, int64_t *cells, int64_t *faces);

/* Measurement probe with no counterpart in the reference: a plain device
 * copy of nbytes (read nbytes, write nbytes) repeated reps times on the
 * context stream; avg_ms per copy.  Calibrates HBM counters and bandwidth. */
ibg_status ibg_bench_copy(ibg_ctx *ctx, int64_t nbytes, int32_t reps, double *avg_ms);

/* Tile and schedule autotuner with no counterpart in the reference: for
 * each tile candidate (a multiple of 64 up to 256 cells per tile) rebuild
 * the tiles and time the stage kernel of a full-march step on the
 * freestream under both schedules (ibg_config.schedule), and keep the
 * fastest pair (ms[2k+s]: candidate k under schedule s; best = its tile,
 * the schedule is left in the context).  The mesh is re-uploaded and the
 * state reset: call it before ibg_set_state.                              */
ibg_status ibg_autotune_tile(ibg_ctx *ctx, const ibg_mesh_soa *mesh, const ibg_physics *phys,
                             const int32_t *candidates, int32_t ncand, int32_t reps,
                             int32_t *best, double *ms);

/* Counters for the metric: normal cells updated per step, summed over the
 * steps of the last ibg_time_march, times the stage count (the cell-stage
 * update count of NS_Solver.cpp:868-872, 948-953).                        */
ibg_status ibg_cell_stage_updates(ibg_ctx *ctx, int64_t *count);

/* ------------------------------------------------------- learned closure
 * A small network stands in for the point-wise part of the SA source that
 * ComptMutSource evaluates per cell (NS_Solver.cpp:700-767); the reference
 * has no learned component, so these calls replace nothing there and are
 * the training / inference surface of the closure (the reference's
 * per-cell evaluation stays the fallback and the measure of its error).
 * Network 8 -> 64 -> 64 -> 1 with ReLU units, BF16
 * weights and activations, FP32 accumulation, trained on cells sampled
 * from this context's own state.
 * Parameter vector (floats): W1[8][64] b1[64] W2[64][64] b2[64] w3[64] b3,
 * then the statistics fmean[8] fstd[8] flo[8] fhi[8] zmean zstd
 * (IBG_CLOSURE_NPARAM + 34 values).                                        */
#define IBG_CLOSURE_NPARAM 4801
#define IBG_CLOSURE_NFEAT 8
/* allocate the sample store (capacity rows) and initialise the weights */
ibg_status ibg_closure_init(ibg_ctx *ctx, int64_t capacity, uint64_t seed);
/* append the features and plain-closure targets of the current state's
 * fluid cells, each kept with probability frac (seeded); total = rows held */
ibg_status ibg_closure_sample(ibg_ctx *ctx, double frac, uint64_t seed, int64_t *total);
/* standardisation and the input range (min/max widened by margin * span) */
ibg_status ibg_closure_fit_stats(ibg_ctx *ctx, double margin);
/* Adam steps on random batches of the sample store, the learning rate
 * decayed from lr to lr/100 (cosine) over the call; loss of the last batch */
ibg_status ibg_closure_train(ibg_ctx *ctx, int32_t steps, int32_t batch, double lr,
                             uint64_t seed, double *loss);
/* use the network in the stage (mixed mode with the SA closure only)       */
ibg_status ibg_closure_enable(ibg_ctx *ctx, int32_t on);
ibg_status ibg_closure_params(ibg_ctx *ctx, float *out, int64_t n);
ibg_status ibg_closure_set_params(ibg_ctx *ctx, const float *in, int64_t n);
/* network output z (destandardised) for n raw feature rows x[n][8]          */
ibg_status ibg_closure_infer(ibg_ctx *ctx, const float *x, float *z, int64_t n);
/* ibg_closure_infer without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * x and z are host buffers as in ibg_closure_infer; they are staged through
 * a fixed per-device array of the library (2.25 MiB, held until the process
 * ends) instead of per-call allocations.  n is at most
 * IBG_CLOSURE_NOALLOC_MAX_ROWS (IBG_ERR_ARG otherwise).  Every context on a
 * device shares that array: calls that overlap on one device race.  Same
 * kernel and results as ibg_closure_infer.                                  */
#define IBG_CLOSURE_NOALLOC_MAX_ROWS 65536
ibg_status ibg_closure_infer_noalloc(ibg_ctx *ctx, const float *x, float *z, int64_t n);
/* gradient of the batch loss (1/n) sum (y - (z - zmean)/zstd)^2 w.r.t. the
 * parameter vector for n explicit rows; loss = that mean                    */
ibg_status ibg_closure_grad(ibg_ctx *ctx, const float *x, const float *z, int64_t n,
                            float *grad /* [IBG_CLOSURE_NPARAM] */, double *loss);
/* first n rows of the sample store: features, targets, global cell index   */
ibg_status ibg_closure_dataset(ibg_ctx *ctx, float *x, float *z, int64_t *cell, int64_t n);
/* cells that used the network / the plain closure in the last stages since
 * the previous call (then reset)                                           */
ibg_status ibg_closure_counts(ibg_ctx *ctx, int64_t *on_net, int64_t *fallback);

#ifdef __cplusplus
}
#endif
#endif /* IBGPU_H_ */
