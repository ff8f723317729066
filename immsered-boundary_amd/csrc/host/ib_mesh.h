/*
 * ib_mesh.h — host-side AMR + immersed-boundary mesh builder (libibhost.so).
 *
 * Builds, for one or more airfoil cases, the flattened level set that the
 * residual engine consumes (ibg_mesh_soa in ibgpu.h): per-level normal and
 * ghost cells, the 27-neighbour tables, faces, domain-boundary ghosts, the
 * coarse/fine restriction and prolongation pairs, signed distances, the
 * immersed-boundary cell list with image-point interpolation weights, and the
 * surface patches with their own image points for the wall force.
 *
 * The construction rules follow the reference's initial refinement:
 *   band tagging            Userfunction_cy_naca15.cpp:533-605 (rs < signdis < re)
 *   restriction / pro rings AMRGhost_refine.cpp:30-98
 *   domain ghosts           Domain.H:372-446
 *   signed distance         Body.cpp:791-821
 *   IB tagging              Body.cpp:1016-1096
 *   image points + weights  Body.cpp:999-1190, Mesh.cpp:1330-1442
 *   patch image points      Body.cpp:1506-1600
 */
#ifndef IB_MESH_H_
#define IB_MESH_H_

#include <stdint.h>
#include "ibgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IBM_MAX_BODIES 8

typedef struct ibm_case {
  int32_t dim;                  /* 2 or 3                                     */
  int32_t nx, ny, nz;           /* base grid (2D: nz = 1)                     */
  double dh0[3];                /* base cell size                             */
  double origin[3];             /* physical position of the low corner        */
  int32_t nlevels;              /* max_mesh_level                             */
  int32_t twod_from;            /* levels >= twod_from refine x,y only        */
  double band_lo[IBG_MAX_LEVELS];/* level l cells with band_lo<signdis<band_hi */
  double band_hi[IBG_MAX_LEVELS];/*   are refined to l+1 (l = 0..nlevels-2)   */
  /* bodies: closed airfoils split in two open polylines (upper, lower) as
     the reference splits its bodies (Userfunction_cy_naca15.cpp:205-211)    */
  int32_t nbody;
  int32_t body_npts[IBM_MAX_BODIES];
  const double *body_xy;        /* concatenated (x,y) points                  */
  double body_center[3];        /* moment reference point                     */
  double span;                  /* z extent of the surface patches            */
  int32_t span_patches;         /* patches per segment along z (>= 1)         */
  /* flow */
  double gama, Ma, Re, Pr, cfl, backpressure, dt_init;
  /* BC code of the domain faces x-, x+, y-, y+; 0 keeps the case file's
     -13 inlet, -12 outlet, -50, -50 (Userfunction_cy_naca15.cpp:283-346);
     -81 is the wall of the TURBULENCE build (BCValues.H:89-125)          */
  int32_t face_bc[4];
  /* 1: the bodies steer the refinement bands only -- no solid or immersed-
     boundary cells, no wall patches, signdis 9999 (a body-free flow on the
     refined level set, e.g. for conservation checks); 0: immersed bodies   */
  int32_t refine_only;
} ibm_case;

typedef struct ibm_handle ibm_handle;

ibm_handle *ibm_build(const ibm_case *cases, int32_t ncases);
const ibg_mesh_soa *ibm_mesh(const ibm_handle *h);
/* physics constants of case c, FlowConstants.cpp:5-44 semantics            */
int32_t ibm_physics(const ibm_case *c, ibg_physics *out);
/* initial state: every cell set to the case's freestream (InitFlow
   NS_Solver.H:432-446), AoS [ncells][7]                                    */
int32_t ibm_init_state(const ibm_handle *h, const ibg_physics *phys,
                       double *fv_aos);
/* closest surface patch of every cell (BoxtoWall::patch, Box.H:68-106) as
   an index into the mesh's pat_* arrays, -1 for cases without bodies        */
int32_t ibm_cell_patch(const ibm_handle *h, int32_t *out /* [ncells] */);
/* The ibg_surface of the mesh's 2D sections for ibg_upload_surface: pt_off
   (caller-owned, [npoint][2]) are the zero-pose offsets of every body's
   points from its pivot in the order the bodies were given (for the NACA
   sections: ibm_naca4_offsets); the other arrays live in the handle.     */
int32_t ibm_surface(const ibm_handle *h, const double *pt_off, int32_t npoint,
                    ibg_surface *out);
/* per surface patch: the cell holding its image point (HGCell::closecell) */
int32_t ibm_patch_close(const ibm_handle *h, int32_t *out /* [npatch] */);
/* State of a new level set of the same cases from the state of the old one
   after a regrid (AMR::ComptDatainNewBox AMR.H:148-215): AoS FlowVariables
   [ncells][7] in and out, the SA variable [ncells] in and out or both
   NULL.  Returns 0, or 1 with ibm_last_error.                             */
int32_t ibm_transfer(const ibm_handle *old_mesh, const ibg_physics *phys /* [ncases] */,
                     const double *fv_old, const double *nut_old,
                     const ibm_handle *new_mesh, double *fv_new, double *nut_new);
/* Tecplot output (Pltoutput::Output_Grid / Output_Soln, Pltoutput.cpp:3-100,
   417-450): one FEBRICK zone "N-<case> L-<level>" per level, every cell of
   the level an element (ghosts included), the solution cell-centred with
   the reference's variables (roe u v w p T [viseddy] Ma dis dmdis i j k
   Omegaz).  17 significant digits, ts / dt / t as DATASETAUXDATA: the
   solution file is also an exact restart (ibm_read_solution, InputSoln
   input_soln.cpp:12-180).  nut, dt, t may be NULL.                        */
int32_t ibm_write_grid(const ibm_handle *h, const char *path);
int32_t ibm_write_solution(const ibm_handle *h, const ibg_physics *phys, const double *fv,
                           const double *nut, int32_t ts, const double *dt, const double *t,
                           const char *path);
int32_t ibm_read_solution(const ibm_handle *h, const ibg_physics *phys, const char *path,
                          double *fv, double *nut, int32_t *ts, double *dt, double *t);
/* cell centers [ncells][3]                                                 */
int32_t ibm_cell_centers(const ibm_handle *h, double *xyz);
/* error text of the last failed build                                      */
const char *ibm_last_error(void);
void ibm_free(ibm_handle *h);

/* NACA 4-digit symmetric section (closed trailing edge), npts cosine-spaced
   points per surface, chord 1, quarter chord at (xc,yc), rotated nose-up
   by aoa_deg.  Writes upper then lower surface, each LE -> TE, into xy
   [2*npts][2].                                                             */
void ibm_naca4(double t, int32_t npts, double xc, double yc, double aoa_deg,
               double *xy);
/* The same points before the pose: offsets from the quarter chord at zero
   angle, [2*npts][2]; ibm_naca4 = pivot + R(aoa) * offset with the pivot
   and rotation that ibm_pose writes as {xc, yc, cos a, sin a}.             */
void ibm_naca4_offsets(double t, int32_t npts, double *off);
void ibm_pose(double xc, double yc, double aoa_deg, double *pose /* [4] */);

#ifdef __cplusplus
}
#endif
#endif
