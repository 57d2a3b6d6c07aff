/*
 * oracle.h -- CPU restatement of the reference pressure-Poisson path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker and as the timed CPU baseline.
 *
 * Every routine restates one routine of the reference Fortran (cited as
 * file:line under IBM/) in plain C, in the reference's own index space:
 *
 *   - 3-D fields are (0:nx+1, 0:ny+1, 0:nz+1) with i fastest, where
 *     nx = nxc+1 is the reference's node count (MAIN.F90:519-521);
 *     linear index = i + (nx+2)*(j + (ny+2)*k).
 *   - Flags (iblank, ghostCellMark, ium..kup, fresh_cell) are int8_t
 *     (INTEGER(1), AMODULES.F90:388-399).
 *   - All arithmetic is double (CGREAL=8, AMODULES.F90:6).
 *
 * Parity pins for each part are listed in the header of the .c file that
 * implements it.  Functions that no golden vector, closed form, invariant or
 * brute-force check pins are marked "parity unpinned" there.
 */
#ifndef FSI_ORACLE_H
#define FSI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

/* ---- grid (CORE_GRID.F90:29-151) ------------------------------------ */
/* xn/yn/zn: node coordinates x(1..nx) (nx = nxc+1 values each). */
or_ctx *or_create(int nxc, int nyc, int nzc, const double *xn,
                  const double *yn, const double *zn);
void or_destroy(or_ctx *c);
/* dir 0/1/2; out holds 6 arrays of (n+2): x, xc, dx, dxinv, dxc, dxcinv */
void or_get_metrics(const or_ctx *c, int dir, double *out);
long or_field_size(const or_ctx *c); /* (nx+2)(ny+2)(nz+2) */

/* ---- flags and ghost marking ---------------------------------------- */
/* iblank in reference layout.  Sets outer-face flags (BOUNDARY_SET.F90:54-92),
 * ghostCellMark and the i-fastest ghost list (GCM_SET_BOUNDARY.F90:49-118). */
int or_set_iblank(or_ctx *c, const int8_t *iblank);
int or_nghost(const or_ctx *c);
void or_get_flags(const or_ctx *c, int8_t *ghost, int8_t *ium, int8_t *iup,
                  int8_t *jum, int8_t *jup, int8_t *kum, int8_t *kup);
void or_get_ghost_list(const or_ctx *c, int32_t *ig, int32_t *jg, int32_t *kg);
/* IP cell indices (reference indexing), coeffGCMN(8,n) column-major and
 * probeLength(n) -- the output of the geometry pipeline. */
void or_set_ghost_stencils(or_ctx *c, const int32_t *ic, const int32_t *jc,
                           const int32_t *kc, const double *coeff,
                           const double *probe);
void or_set_fresh(or_ctx *c, const int8_t *fresh);

/* ---- single-level operators ----------------------------------------- */
/* calc_residual (CORE_POISSON.F90:1111-1222) with nlw0 == 0.  Writes the
 * masked residual field into res (may be NULL), returns max|res| (resm) and
 * its argmax (loc, reference indices) and the L2 norm. */
double or_calc_residual(const or_ctx *c, const double *var, const double *r,
                        double *res, int32_t *loc, double *l2);
/* subtracting_average_pressure (CORE_POISSON.F90:419-490) */
double or_mean_remove(const or_ctx *c, double *p);
/* MG_itsolv_Point_Jacobi (CORE_MG.F90:3716-3800) on the finest level: one
 * weighted point-Jacobi sweep of var (inout) against r, ghost rows held */
void or_point_jacobi(const or_ctx *c, double *var, const double *r, double omega);
/* set_outer_ghost_pres for all-Neumann outer faces
 * (CORE_SET_OUTER_GHOST.F90:66-213) */
void or_outer_ghost(const or_ctx *c, double *p);
/* GCM_p_set_bc_internal (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:250-384).
 * g (uNormal per ghost) may be NULL (viscous flow: 0).  Returns sweeps. */
int or_ghost_rows(const or_ctx *c, double *p, const double *g, int itermax,
                  double restol);
/* rhs_poisson non-cut-cell branch (CORE_POISSON.F90:281-288) */
double or_divergence(const or_ctx *c, const double *fu, const double *fv,
                     const double *fw, double dtinv, double *nlu0);
/* tdma (CORE_TDMA.F90:10-34); arrays indexed 0..n2 as in the reference */
void or_tdma(const double *a, const double *b, const double *c,
             const double *r, double *u, int n1, int n2);

/* ---- reference multigrid (CORE_MG.F90) ------------------------------ */
typedef struct or_mg_opts {
  int iter_max;        /* iterMax_Poisson */
  double restol;       /* restol_Poisson: absolute L-inf stop (reference) */
  double rel_tol;      /* >0: also stop when ||r||2/||b||2 <= rel_tol */
  int iter_finest;     /* iterFinest */
  int iter_coarsest;   /* iterCoarsest */
  double omega;        /* SOR weight */
  int itermax_ad;      /* iterMax_ad for ghost rows */
  double restol_ad;    /* restol_ad for ghost rows */
  int mean_remove;     /* call subtracting_average_pressure at the start */
  int zebra;           /* 0: the reference's line order (iRedBlack = 0, one
                          thread); 1: two-colour line order, lines and cell
                          loops on all OpenMP threads (same fixed point) */
} or_mg_opts;

/* OpenMP threads of the zebra variant */
int or_threads(void);

typedef struct or_mg_stats {
  int iterations;
  double linf;         /* final abs L-inf residual */
  double l2;           /* final L2 residual */
  double b_l2;         /* ||rhs||2 over fluid cells */
  int32_t loc[3];
  int levels[3];       /* mgLevels_X/Y/Z */
} or_mg_stats;

/* solve_poisson MG branch (CORE_POISSON.F90:493-676); p in/out, rhs = nlu0 */
int or_mg_solve(or_ctx *c, double *p, const double *rhs,
                const or_mg_opts *o, or_mg_stats *s);
/* One MG_Solver call (CORE_MG.F90:281-334) followed by GCM_correct_pressure */
void or_mg_cycle(or_ctx *c, double *p, const double *rhs, const or_mg_opts *o);

/* ---- FP64 operator applies for checks ------------------------------- */
/* y = A p over the fluid rows (the calc_residual operator without rhs),
 * ghost rows y_g = p_g - sum_r c_r p_r; other cells 0. */
void or_apply_A(const or_ctx *c, const double *p, double *y);
/* y = A^T x for the same operator (fluid + ghost unknowns). */
void or_apply_AT(const or_ctx *c, const double *x, double *y);

/* ---- geometry (oracle/or_geom.c) ------------------------------------ */
/* Canonical ellipsoid surface triangulation (BOUNDARY_MARKER_INIT.F90:115-205)
 * and triangle normals/centroids (BOUNDARY_SET_ARCLENGTH_NORM.F90:69-167),
 * normals oriented inward.  Returns the marker / element counts; arrays are
 * owned by the oracle and read back with or_geom_get_surface. */
int or_geom_ellipsoid(or_ctx *c, int n_theta, double xc0, double yc0,
                      double zc0, double ra, double rb, double rc,
                      double alpha_deg);  /* markers, or -1 bad body, -2 body around the origin */
void or_geom_clear_bodies(or_ctx *c);
int or_geom_nbodies(const or_ctx *c);
void or_geom_surface_sizes(const or_ctx *c, int body, int *nmarker,
                           int *nelem);
void or_geom_get_surface(const or_ctx *c, int body, double *xyz, int32_t *tri,
                         double *cent, double *norm);
/* set_internal_boundary for canonical solid bodies (BOUNDARY_SET.F90:264-410):
 * iblank by set_iblank_canonical_body_fast over all bodies, fresh cells
 * against iblank_prev (may be NULL: none), find_isolatedIblank(1),
 * fill_cell, ghost marking, body intercepts (GCM_calc_bodyIntercept_Unstruc),
 * image points and Neumann coefficients.  Returns 0 or an error code:
 * 1 image-point cell not found, 2 intercept too long, 3 singular
 * Vandermonde matrix, 4 no unblocked marker among the 7 closest, 5 no
 * element selected, 7 ghost cell without a body. */
int or_geom_set_boundary(or_ctx *c, const int8_t *iblank_prev,
                         double probe_length_normalized);
void or_geom_get_iblank(const or_ctx *c, int8_t *iblank, int8_t *fresh);
/* per ghost: image-point cell, coeffGCMN(8,n), probeLength, body
 * intercept (3), image point (3), closest element (1-based) and body
 * (1-based); any pointer may be NULL */
void or_geom_get_ghost_geometry(const or_ctx *c, int32_t *ic, int32_t *jc,
                                int32_t *kc, double *coeff, double *probe,
                                double *bi, double *ip, int32_t *elem,
                                int32_t *body);

/* GCM_set_freshcell (GCM_SET_FRESHCELL.F90:1-262) for the fresh cells of
 * the last or_geom_set_boundary (or or_set_fresh): the i-fastest list,
 * image-point stencils and coeffGCMFreshD; returns 0 or 10 (image point
 * outside the 2x2x2 stencil), 11 (intercept at the cell centre), 3
 * (singular), 4/5 (intercept), 7 (no body).  or_geom_get_fresh reads the
 * tables (n = or_nfresh_tables; any pointer may be NULL; body and element
 * 1-based).  or_set_fresh_tables installs caller tables instead (BI, body
 * and element left 0).  or_set_fresh and or_geom_set_boundary clear them. */
int or_geom_set_freshcell(or_ctx *c, double probe_length_normalized);
int or_nfresh_tables(const or_ctx *c);
void or_geom_get_fresh(const or_ctx *c, int32_t *ifr, int32_t *jfr,
                       int32_t *kfr, int32_t *ic, int32_t *jc, int32_t *kc,
                       double *coeffD, double *bi, int32_t *body,
                       int32_t *elem);
void or_set_fresh_tables(or_ctx *c, int n, const int32_t *ifr,
                         const int32_t *jfr, const int32_t *kfr,
                         const int32_t *ic, const int32_t *jc,
                         const int32_t *kc, const double *coeffD);

/* coeffGCMD(8,n), column-major, of the last or_geom_set_boundary (the
 * velocity set of GCM_Calc_VanMatrixDN, GCM_SET_BOUNDARY.F90:307-494) */
void or_geom_get_dirichlet(const or_ctx *c, double *coeffD);

/* ---- fractional step either side of the solve (oracle/or_frac.c) ----- */
/* coeffGCMD(8,n) for the current ghost list and the GCM weights
 * bodyInterceptWeight, imagePointWeight (MAIN.F90:913-914) */
void or_set_vel_ghosts(or_ctx *c, const double *coeffD, double bi_weight, double ip_weight);
/* GCM_vel_set_bc_internal (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:126-247);
 * vbi: (u,v,w) at each body intercept, NULL = 0.  Returns sweeps. */
int or_vel_ghosts(const or_ctx *c, double *u, double *v, double *w, const double *vbi, int itermax,
                  double restol);
/* correct_vel (CORE_FRAC_STEP_ROUTINES.F90:244-452), non-periodic, no cut
 * cells: cell velocities, ghost velocities, face velocities, solid faces */
void or_correct_vel(const or_ctx *c, double dt, const double *p, double *u, double *v, double *w, double *fu,
                    double *fv, double *fw, const double *bcxu, const double *bcyv, const double *bczw,
                    const double *vbi, int itermax_ad, double restol_ad);
/* update_pressure (CORE_FRAC_STEP_ROUTINES.F90:949-967): p += p' (van_kan)
 * or p = p', then GCM_p_set_bc_internal on p */
/* face_vel (CORE_FRAC_STEP_ROUTINES.F90:6-240), GCM, non-periodic: face
 * velocities from u, v, w (uTilde = u + dt grad p for VAN_KAN), outer faces
 * from bcxu/bcyv/bczw, VAN_KAN face pressure correction and ghost-cell
 * faces from u itself, solid faces zero; fu, fv, fw are written. */
void or_face_vel(const or_ctx *c, int van_kan, double dt, const double *p, const double *u, const double *v,
                 const double *w, const double *bcxu, const double *bcyv, const double *bczw, double *fu, double *fv,
                 double *fw);
void or_update_pressure(const or_ctx *c, int van_kan, const double *pprime, double *p, const double *g, int itermax,
                        double restol);

/* ---- implicit momentum solve (oracle/or_ad.c) ------------------------ */
/* set_solve_ad (CORE_SET_SOLVERS.F90:20-223), fresh rows from the fresh
 * tables (none without them): coef =
 * amx_ad, apx_ad, amy_ad, apy_ad, amz_ad, apz_ad (6 fields).  visc is
 * viscTot; bxv, byv, bzv the boundary-face viscosities bcx/y/zvisc. */
void or_set_solve_ad(const or_ctx *c, double dt, const double *visc, const double *bxv, const double *byv,
                     const double *bzv, double *coef);
/* calc_residual_ad (CORE_SOLVE_AD.F90:766-907): returns max |res| over
 * the cells, loc = its 1-based (i,j,k), res (may be NULL) the field.
 * half_dt = dt/2 for CRANK_NICOLSON1, 0 for ADAMS_BASHFORTH2. */
double or_calc_residual_ad(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                           const double *fw, const double *var, const double *r, double *res, int32_t *loc);
/* itsolv_ad (CORE_SOLVE_AD.F90:272-762): x, y and z line sweeps in the
 * reference's loop order, Thomas algorithm per line, in place */
void or_itsolv_ad(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                  const double *fw, double *var, const double *r);
/* solve_ad, GCM branch (CORE_SOLVE_AD.F90:200-267): returns iterations,
 * *maxres the last maximum residual of (u, v, w) */
int or_solve_ad(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                const double *fw, double *u, double *v, double *w, const double *nlu, const double *nlv,
                const double *nlw, const double *vbi, int itermax, double restol, double *maxres);
/* solve_ad with CRANK_NICOLSON2 (CORE_SOLVE_AD.F90:218-247): face_vel after
 * every ghost update (fu, fv, fw in/out); half_dt = dt/2. */
int or_solve_ad_cn2(const or_ctx *c, double dt, const double *coef, double *fu, double *fv, double *fw, double *u,
                    double *v, double *w, const double *nlu, const double *nlv, const double *nlw, const double *vbi,
                    int van_kan, const double *p, const double *bcxu, const double *bcyv, const double *bczw,
                    int itermax, double restol, double *maxres_out);

#ifdef __cplusplus
}
#endif
#endif
