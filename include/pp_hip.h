/*
 * pp_hip.h -- C-ABI of the MI355X pressure-Poisson solve.
 *
 * Drop-in replacement for the pressure-Poisson path of the reference
 * ghost-cell immersed-boundary solver (VICAR3D):
 *
 *   SUBROUTINE solve_poisson()            IBM/CORE_POISSON.F90:493-896
 *     MG_Solver(pPrime,nlu0,MG_itsolv,MG_Residual,compTime)
 *                                          IBM/CORE_MG.F90:281-334
 *     GCM_correct_pressure                 IBM/CORE_POISSON.F90:1794-1823
 *     calc_residual                        IBM/CORE_POISSON.F90:1111-1222
 *
 * Plain pointers and sizes only.  3-D fields are passed in the reference
 * layout: (0:nx+1, 0:ny+1, 0:nz+1) with i fastest and nx = nxc+1 node
 * counts, i.e. (nxc+3)*(nyc+3)*(nzc+3) entries, linear index
 * i + (nxc+3)*(j + (nyc+3)*k) with interior cells i = 1..nxc.
 * Flag arrays are int8_t (INTEGER(1)); values are double (REAL(8)).
 *
 * Every entry point returns an int32_t status (PP_OK = 0) and never aborts;
 * pp_last_error() returns the text of the calling thread's last error.
 * One host thread per context; calls are synchronous on return.
 */
#ifndef PP_HIP_H
#define PP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PP_OK 0
#define PP_NOT_CONVERGED 1   /* iteration limit reached (reference prints and continues) */
#define PP_BAD_ARG 2
#define PP_BAD_GEOMETRY 3    /* e.g. an image point outside the grid */
#define PP_HIP_ERROR 4
#define PP_NONFINITE 5       /* NaN/Inf in the RHS or the residual */
#define PP_NO_GEOMETRY 6     /* solve called before pp_set_geometry */

/* outer pressure boundary condition per face (pbcx1..pbcz2,
 * AMODULES.F90:187-333); only PP_BC_NEUMANN is implemented */
#define PP_BC_NEUMANN 2

/* precision of the multigrid cycles */
#define PP_PREC_FP64 0       /* FP64 V-cycles (the GPU FP64 multigrid) */
#define PP_PREC_MIXED 1      /* FP32 V-cycles inside FP64 iterative refinement */

typedef struct pp_ctx pp_ctx;

/* Grid: make_grid (CORE_GRID.F90:29-151) inputs.  x/y/z hold the node
 * coordinates x(1..nxc+1) (uniform or read from xgrid.dat in the reference);
 * the library derives x(0), x(nx+1), centres, widths and inverses exactly as
 * make_grid does. */
typedef struct pp_grid_desc {
  int32_t nxc, nyc, nzc;
  const double *x, *y, *z;
  int32_t bc[6];             /* pbcx1, pbcx2, pbcy1, pbcy2, pbcz1, pbcz2 */
  int32_t device;            /* HIP device ordinal */
} pp_grid_desc;

/* Geometry: the outputs of set_boundary (BOUNDARY_SET.F90:8-51) that the
 * pressure path reads (AMODULES.F90:378-436, 507-579).  Host pointers in the
 * reference layout / 1-based reference indices; not retained. */
typedef struct pp_geometry {
  const int8_t *iblank;          /* required */
  const int8_t *fresh_cell;      /* may be NULL */
  int32_t n_ghost;               /* nGhost */
  const int32_t *i_ghost, *j_ghost, *k_ghost;   /* iGhost/jGhost/kGhost */
  const int32_t *i_cell, *j_cell, *k_cell;      /* iCellIndex/... (IP cell) */
  const double *coeff;           /* coeffGCMN(8, nGhost), column-major */
  const double *probe_length;    /* probeLength(nGhost) */
} pp_geometry;

/* Solve options (input.dat knobs, MAIN.F90:476-489, plus the build's). */
typedef struct pp_solve_opts {
  int32_t precision;         /* PP_PREC_FP64 or PP_PREC_MIXED */
  double rel_tol;            /* stop when ||b-Ap||_2/||b||_2 <= rel_tol (if > 0) */
  double abs_linf_tol;       /* restol_Poisson: stop when max|b-Ap| <= tol (if > 0) */
  int32_t max_outer;         /* iterMax_Poisson: outer iterations */
  int32_t cycles_per_outer;  /* V-cycles per outer iteration */
  int32_t nu_pre, nu_post;   /* smoothing steps before/after coarse correction */
  double omega;              /* relaxation weight (MAIN.F90:478); default 1.3 for the
                                zebra line smoother, use < 1 with point Jacobi */
  int32_t mean_remove;       /* subtract the fluid volume mean on return (0/1) */
  int32_t project_rhs;       /* compatibility projection of the RHS (0/1) */
  int32_t use_graph;         /* replay the outer iteration as a hipGraph (0/1) */
  int32_t profile;           /* time every kernel launch with HIP events (0/1) */
  int32_t krylov;            /* 0: FP64 iterative refinement around the V-cycles;
                                m in 1..16: flexible GMRES(m) with the V-cycles
                                as right preconditioner (pp_solve, pp_adjoint) */
  int32_t nu_coarse;         /* smoothing steps before and after the coarse
                                correction on levels below the finest
                                (<= 0: nu_pre / nu_post everywhere) */
  int32_t smoother;          /* PP_SMOOTH_LINE or PP_SMOOTH_POINT_JACOBI */
} pp_solve_opts;

/* smoothers (pp_solve_opts.smoother, pp_smooth) */
#define PP_SMOOTH_LINE 0          /* zebra line relaxation x, y, z (MG_itsolv,
                                     CORE_MG.F90:1883-2270, pp_solver_type 3) */
#define PP_SMOOTH_POINT_JACOBI 1  /* weighted point Jacobi (MG_itsolv_Point_Jacobi,
                                     CORE_MG.F90:3716-3800, pp_solver_type 8) */
#define PP_SMOOTH_LINE_NATURAL 2  /* PP_SMOOTH_LINE with the finest level's FP32 arrays
                                     kept in the natural layout (the default colour-split
                                     copy gives the same results; pp_solve_opts only) */

/* kernel classes of pp_solve_stats.kern_* (profile = 1) */
#define PP_K_OUTER_RESIDUAL 0  /* FP64 b - A p on the reference layout */
#define PP_K_UPDATE 1          /* FP64 p += e */
#define PP_K_XLINE 2           /* x-line zebra pass (wavefront PCR) */
#define PP_K_YLINE 3           /* y-line zebra pass (partitioned Thomas) */
#define PP_K_ZLINE 4           /* z-line zebra pass (partitioned Thomas) */
#define PP_K_RESIDUAL 5        /* level residual */
#define PP_K_RESTRICT 6
#define PP_K_PROLONG 7
#define PP_K_GHOST 8           /* ghost-row gathers/scatters */
#define PP_K_COARSE 9          /* coarsest-level dense solve */
#define PP_K_OTHER 10
#define PP_K_KRYLOV 11         /* flexible-GMRES apply / orthogonalise / update */
#define PP_K_JACOBI 12         /* point-Jacobi sweep */
#define PP_K_CLASSES 13

typedef struct pp_solve_stats {
  int32_t status;
  int32_t outer_iterations;
  int32_t vcycles;
  double rel_l2;             /* final ||b-Ap||_2/||b||_2 over fluid cells */
  double abs_linf;           /* final max|b-Ap| (calc_residual's maxres) */
  int32_t loc[3];            /* argmax, reference indices */
  double b_l2;               /* ||b||_2 after projection */
  double rhs_shift;          /* constant removed from the fluid RHS */
  double mean_removed;       /* fluid volume mean removed from p */
  double ms;                 /* device time of the solve (HIP events) */
  double bytes;              /* algorithmic bytes of the launched kernels */
  double history[64];        /* rel_l2 after each outer iteration */
  /* profile = 1: per kernel class, summed device time of the launches on
   * the finest level (level 0) and over all levels, algorithmic bytes and
   * launch counts */
  double kern_ms[PP_K_CLASSES];
  double kern_bytes[PP_K_CLASSES];
  int32_t kern_launches[PP_K_CLASSES];
  double kern_ms_l0[PP_K_CLASSES];
  double kern_bytes_l0[PP_K_CLASSES];
  int32_t kern_launches_l0[PP_K_CLASSES];
} pp_solve_stats;

/* Library defaults for pp_solve_opts. */
void pp_default_opts(pp_solve_opts *o);
const char *pp_last_error(void);

/* Create a context on one GPU for a grid (replaces MG_initial,
 * CORE_MG.F90:31-278, and the grid half of init_simulation).  4 to 1024
 * cells per direction (PP_BAD_ARG otherwise).  Levels: a direction is
 * halved while its cell count is at least 4 (an odd count's last coarse cell
 * spans three fine cells), down to at most 4096 cells, which are solved
 * directly. */
int32_t pp_create(pp_ctx **ctx, const pp_grid_desc *g);
int32_t pp_destroy(pp_ctx *ctx);

/* Upload a geometry (call after every set_boundary: MAIN.F90:1549,
 * BOUNDARY_MOVE.F90:210).  Derives ghostCellMark and the outer-face flags
 * (BOUNDARY_SET.F90:54-92, GCM_SET_BOUNDARY.F90:49-118) from iblank and
 * checks them against the ghost list; builds the level hierarchy, ghost
 * tables and the left null vector used by the compatibility projection.
 * Ghost stencil corners on solid non-ghost cells (value 0) leave the
 * operator without an exact null vector; its near-null vector is used
 * then (pp_setup_info reports the residual reached). */
int32_t pp_set_geometry(pp_ctx *ctx, const pp_geometry *geo);

/* solve_poisson (CORE_POISSON.F90:493-676): p_inout = pPrime (initial guess
 * in, solution out), rhs = nlu0, bc_src = Neumann datum per ghost
 * (uNormal, GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:331-343) or NULL for 0.
 * Host pointers, reference layout.  Returns PP_OK or PP_NOT_CONVERGED (the
 * solution is still written, as the reference continues). */
int32_t pp_solve(pp_ctx *ctx, double *p_inout, const double *rhs,
                 const double *bc_src, const pp_solve_opts *o,
                 pp_solve_stats *st);
/* Same with device pointers (reference layout, already in HBM). */
int32_t pp_solve_device(pp_ctx *ctx, double *d_p, const double *d_rhs,
                        const double *d_bc_src, const pp_solve_opts *o,
                        pp_solve_stats *st);

/* Adjoint of the zero-mean solve.  For J = sum(w * p) with p = pp_solve's
 * output (projected RHS, zero fluid mean), returns dJ/d(rhs) in lambda
 * (reference layout) and dJ/d(bc_src) in dJ_dg (n_ghost, may be NULL).
 * No counterpart in the reference (it has no adjoint). */
int32_t pp_adjoint(pp_ctx *ctx, const double *w, double *lambda,
                   double *dJ_dg, const pp_solve_opts *o, pp_solve_stats *st);
/* pp_adjoint on device arrays (wobj, lambda in the reference layout; dJ_dg
 * n_ghost values or NULL); nothing is copied to or from the host. */
int32_t pp_adjoint_device(pp_ctx *ctx, const double *d_wobj, double *d_lambda,
                          double *d_dJ_dg, const pp_solve_opts *opts,
                          pp_solve_stats *stats);

/* ---- single routines of the reference path (parity checks, callers that
 *      keep their own loop) ------------------------------------------- */

/* calc_residual (CORE_POISSON.F90:1111-1222) with nlw0 = 0: res (may be
 * NULL) = masked r - A var; *resm = max|res|, loc = argmax, *l2 = ||res||_2. */
int32_t pp_calc_residual(pp_ctx *ctx, const double *var, const double *r,
                         double *res, double *resm, int32_t *loc, double *l2);
/* GCM_correct_pressure's ghost part: set_outer_ghost_pres (Neumann,
 * CORE_SET_OUTER_GHOST.F90:66-213) + GCM_p_set_bc_internal solved exactly
 * (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:250-384); bc_src may be NULL. */
int32_t pp_ghost_rows(pp_ctx *ctx, double *p, const double *bc_src);
/* subtracting_average_pressure (CORE_POISSON.F90:419-490). */
int32_t pp_mean_remove(pp_ctx *ctx, double *p, double *mean);
/* rhs_poisson, non-cut-cell branch (CORE_POISSON.F90:281-288):
 * nlu0 = dtinv*div(face velocity)*(1-iblank); *sum = sum(nlu0*dV). */
int32_t pp_divergence(pp_ctx *ctx, const double *face_u, const double *face_v,
                      const double *face_w, double dtinv, double *nlu0,
                      double *sum);
/* pp_divergence on device arrays without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight time-step loops, harnesses that forbid hipMalloc).
 * face_u/v/w and nlu0 are device arrays in the reference layout (nlu0's
 * outer layer and blank cells are set to 0).  The reduction's partial sums
 * live in one static device array per GPU (about 4 MB, held until the
 * process ends): grids whose cell launch needs more than 524288 workgroups
 * (beyond about 512x1024x1024 cells) return PP_BAD_ARG.  Calls that overlap
 * on one GPU (any contexts) race on that array.  sum may be NULL: then the
 * call only enqueues work on the context's stream and returns without
 * synchronising; otherwise it waits and writes sum(nlu0*dV). */
int32_t pp_divergence_noalloc(pp_ctx *ctx, const double *d_face_u,
                              const double *d_face_v, const double *d_face_w,
                              double dtinv, double *d_nlu0, double *sum);
/* `sweeps` smoothing sweeps of the finest-level operator on var (inout)
 * with right-hand side r, FP64, host arrays in the reference layout; the
 * ghost rows are not updated (the caller's GCM step does that), as in the
 * smoother routines themselves.  PP_SMOOTH_POINT_JACOBI is
 * MG_itsolv_Point_Jacobi (CORE_MG.F90:3716-3800) called at level 1 of all
 * three directions; PP_SMOOTH_LINE is one zebra x, y, z line pass set per
 * sweep (the line rows of MG_itsolv, CORE_MG.F90:1945-2009, in zebra line
 * order).  omega is the relaxation weight (MAIN.F90:478). */
int32_t pp_smooth(pp_ctx *ctx, int32_t smoother, double omega, int32_t sweeps,
                  double *var, const double *r);
/* ---- fractional step either side of the solve ----------------------
 * Velocity ghost rows for correct_vel: coeffGCMD(8, n_ghost) column-major
 * (GCM_Calc_VanMatrixDN, GCM_SET_BOUNDARY.F90:307-494; pp_geom_dirichlet
 * gives them for the built-in geometry) for the ghost list and stencil
 * cells of the last pp_set_geometry, and the GCM weights
 * bodyInterceptWeight, imagePointWeight (MAIN.F90:913-914: p/(p-1) and
 * 1/(p-1) for probeLengthNormalized p).  Call after every pp_set_geometry.
 * The number of Jacobi sweeps that solves the ghost-ghost coupling to
 * round-off is fixed here (PP_BAD_GEOMETRY if it does not converge). */
int32_t pp_set_vel_ghosts(pp_ctx *ctx, const double *coeffD, double bi_weight,
                          double ip_weight);
/* correct_vel (CORE_FRAC_STEP_ROUTINES.F90:244-452) for non-periodic outer
 * boundaries without cut cells: u,v,w -= dt grad p (face values by the fx
 * interpolation, outer faces by the cell value) on cells with iblank = 0;
 * the ghost velocities by GCM_vel_set_bc_internal
 * (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:126-247) solved to round-off (the
 * reference iterates Gauss-Seidel to restol_ad); face_u,v,w -= dt times
 * the face pressure gradient, outer faces set from bcxu/bcyv/bczw; faces
 * of blank non-ghost cells zeroed (GCM_set_face_vel_body, :387-423).
 * Every field is a host array in the reference layout (0:nx+1,0:ny+1,0:nz+1);
 * vel_bi = (u,v,w) at each body intercept (3*n_ghost, uBodyIntercept..;
 * NULL: bodies at rest). */
int32_t pp_correct_vel(pp_ctx *ctx, double dt, const double *p, double *u,
                       double *v, double *w, double *face_u, double *face_v,
                       double *face_w, const double *bcxu, const double *bcyv,
                       const double *bczw, const double *vel_bi);
/* pp_correct_vel on device arrays (no host copies). */
int32_t pp_correct_vel_device(pp_ctx *ctx, double dt, const double *d_p,
                              double *d_u, double *d_v, double *d_w,
                              double *d_face_u, double *d_face_v,
                              double *d_face_w, const double *d_bcxu,
                              const double *d_bcyv, const double *d_bczw,
                              const double *d_vel_bi);
/* update_pressure (CORE_FRAC_STEP_ROUTINES.F90:949-967): p = p + pPrime
 * (van_kan != 0, frac_step_type == VAN_KAN) or p = pPrime on the whole
 * field, then the pressure ghost rows (GCM_p_set_bc_internal) on p with
 * bc_src as in pp_solve (NULL: homogeneous).  The outer layer is left as
 * the sum (after a solve both hold the Neumann copy). */
int32_t pp_update_pressure(pp_ctx *ctx, int32_t van_kan, const double *pprime,
                           double *p, const double *bc_src);
int32_t pp_update_pressure_device(pp_ctx *ctx, int32_t van_kan,
                                  const double *d_pprime, double *d_p,
                                  const double *d_bc_src);

/* ---- implicit momentum solve of the fractional step ------------------
 * advec_scheme as the reference's (AMODULES.F90:60-62): ADAMS_BASHFORTH2
 * keeps only the viscous terms implicit (the itsolv_ad_*_AB routines),
 * CRANK_NICOLSON1 adds the face-velocity advection terms with half_dt =
 * dt/2.  CRANK_NICOLSON2 recomputes the face velocities inside the loop
 * (face_vel) and is rejected with PP_BAD_ARG. */
#define PP_ADAMS_BASHFORTH2 1
#define PP_CRANK_NICOLSON1 2
#define PP_CRANK_NICOLSON2 3

typedef struct pp_ad_stats {
  int32_t status;
  int32_t iterations;        /* outer iterations of the GCM loop */
  double max_residual;       /* max over u, v, w of calc_residual_ad */
  double residual_uvw[3];    /* last calc_residual_ad of u, v, w */
  int32_t loc[3];            /* argmax (1-based i, j, k) of the largest */
  double ms;                 /* device time of the call (HIP events) */
} pp_ad_stats;

/* set_solve_ad (CORE_SET_SOLVERS.F90:20-223; called by the reference after
 * every set_boundary, MAIN.F90 time loop): the implicit viscous
 * coefficients amx_ad..apz_ad of the momentum equations, kept on the
 * device until the next call or pp_set_geometry.  visc_tot = viscTot
 * (reference layout, outer layer included), bcxvisc..bczvisc the
 * boundary-face viscosities.  With fresh cells in the geometry (its
 * fresh_cell mask) the fresh rows of the GCM branch (:175-218) are set from
 * the tables of pp_set_fresh_cells, which must be called first; pp_solve_ad
 * and pp_calc_residual_ad then add the off-stencil terms of
 * GCM_correct_rhs_ad / GCM_correct_res_ad (GCM_FRESHCELL_AD.F90:2-137).
 * Requires pp_set_geometry. */
int32_t pp_set_solve_ad(pp_ctx *ctx, double dt, const double *visc_tot,
                        const double *bcxvisc, const double *bcyvisc,
                        const double *bczvisc);
int32_t pp_set_solve_ad_device(pp_ctx *ctx, double dt, const double *d_visc_tot,
                               const double *d_bcxvisc, const double *d_bcyvisc,
                               const double *d_bczvisc);
/* The fresh cells of the last pp_set_geometry for the momentum solve
 * (GCM_set_freshcell's tables, as pp_geom_fresh gives them): n_fresh cells
 * (1-based i, j, k), exactly those of the geometry's fresh_cell mask in any
 * order, each with the lower corner of its 2x2x2 stencil (the cell or the
 * one below it in each direction) and coeffGCMFreshD(8, n_fresh)
 * column-major.  The fresh rows are u_f - sum_(stencil nodes but the cell)
 * c u_node = nlu_f with nlu_f = c_self u_BI formed by the caller
 * (rhs_adjust_bc, CORE_RHS_AD.F90:796-815).  pp_set_geometry clears the
 * tables; call pp_set_solve_ad after this. */
int32_t pp_set_fresh_cells(pp_ctx *ctx, int32_t n_fresh, const int32_t *i_fresh,
                           const int32_t *j_fresh, const int32_t *k_fresh,
                           const int32_t *i_cell, const int32_t *j_cell,
                           const int32_t *k_cell, const double *coeffD);

/* solve_ad, GCM branch (CORE_SOLVE_AD.F90:200-267): until max(calc_residual_ad
 * of u, v, w) <= restol_ad or itermax_ad iterations, line sweeps in x, y
 * and z of u, v and w (itsolv_ad, :272-762; here zebra-ordered line passes
 * instead of the reference's lexicographic order, same fixed point) and
 * the velocity ghost update GCM_GhostCell_Vel
 * (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:2-124, Dirichlet outer faces: its
 * mass-flux correction does not act).  u, v, w: initial guess in, solution
 * out; nlu, nlv, nlw: right-hand sides; face_u/v/w: face velocities of the
 * advection terms (read for CRANK_NICOLSON1; CRANK_NICOLSON2 goes through
 * pp_solve_ad_cn2); vel_bi: (u, v, w) per ghost
 * at the body intercept or NULL.  Requires pp_set_solve_ad and, with
 * ghost cells, pp_set_vel_ghosts.  Returns PP_OK or PP_NOT_CONVERGED (the
 * fields are written either way, as the reference continues). */
int32_t pp_solve_ad(pp_ctx *ctx, int32_t advec_scheme, double *u, double *v,
                    double *w, const double *nlu, const double *nlv,
                    const double *nlw, const double *face_u,
                    const double *face_v, const double *face_w,
                    const double *vel_bi, int32_t itermax_ad, double restol_ad,
                    pp_ad_stats *st);
int32_t pp_solve_ad_device(pp_ctx *ctx, int32_t advec_scheme, double *d_u,
                           double *d_v, double *d_w, const double *d_nlu,
                           const double *d_nlv, const double *d_nlw,
                           const double *d_face_u, const double *d_face_v,
                           const double *d_face_w, const double *d_vel_bi,
                           int32_t itermax_ad, double restol_ad,
                           pp_ad_stats *st);

/* solve_ad with advec_scheme CRANK_NICOLSON2 (CORE_SOLVE_AD.F90:218-247):
 * the loop of pp_solve_ad with face_vel after every ghost update, so the
 * residual and the next sweeps use the face velocities of the new fields.
 * face_u/v/w: the face velocities of the first sweeps in, those of the
 * returned fields out; van_kan, p, bcxu..bczw as for pp_face_vel (dt is
 * pp_set_solve_ad's).  pp_solve_ad refuses CRANK_NICOLSON2. */
int32_t pp_solve_ad_cn2(pp_ctx *ctx, double *u, double *v, double *w,
                        const double *nlu, const double *nlv, const double *nlw,
                        double *face_u, double *face_v, double *face_w,
                        int32_t van_kan, const double *p, const double *bcxu,
                        const double *bcyv, const double *bczw,
                        const double *vel_bi, int32_t itermax_ad,
                        double restol_ad, pp_ad_stats *st);
int32_t pp_solve_ad_cn2_device(pp_ctx *ctx, double *d_u, double *d_v,
                               double *d_w, const double *d_nlu,
                               const double *d_nlv, const double *d_nlw,
                               double *d_face_u, double *d_face_v,
                               double *d_face_w, int32_t van_kan,
                               const double *d_p, const double *d_bcxu,
                               const double *d_bcyv, const double *d_bczw,
                               const double *d_vel_bi, int32_t itermax_ad,
                               double restol_ad, pp_ad_stats *st);

/* face_vel (CORE_FRAC_STEP_ROUTINES.F90:6-240), GCM, non-periodic outer
 * boundaries: face velocities from the cell velocities by the fx
 * interpolation of uTilde = u + dt grad p (1 - iblank) (van_kan = 1,
 * frac_step_type VAN_KAN) or of u, the face pressure correction for
 * VAN_KAN, outer faces from bcxu/bcyv/bczw, the faces of ghost cells from
 * the interpolation of u itself (VAN_KAN), the faces of blank non-ghost
 * cells zero.  p may be NULL when van_kan = 0.  Reference layout. */
int32_t pp_face_vel(pp_ctx *ctx, int32_t van_kan, double dt, const double *p,
                    const double *u, const double *v, const double *w,
                    const double *bcxu, const double *bcyv, const double *bczw,
                    double *face_u, double *face_v, double *face_w);
int32_t pp_face_vel_device(pp_ctx *ctx, int32_t van_kan, double dt,
                           const double *d_p, const double *d_u,
                           const double *d_v, const double *d_w,
                           const double *d_bcxu, const double *d_bcyv,
                           const double *d_bczw, double *d_face_u,
                           double *d_face_v, double *d_face_w);

/* calc_residual_ad (CORE_SOLVE_AD.F90:766-907) of one field: *resm = max
 * |res| over the cells, loc its 1-based (i, j, k) (first in the
 * reference's k, j, i loop order), res (may be NULL) the residual field.
 * Host pointers.  Requires pp_set_solve_ad. */
int32_t pp_calc_residual_ad(pp_ctx *ctx, int32_t advec_scheme,
                            const double *var, const double *r,
                            const double *face_u, const double *face_v,
                            const double *face_w, double *res, double *resm,
                            int32_t loc[3]);

/* tdma (CORE_TDMA.F90:10-34) on a batch of independent systems, through
 * the device line solvers (n <= 1024): n_lines systems of length n stored line after
 * line (a: sub-, b: main, c: super-diagonal, r: rhs, u: solution).
 * method 0 = wavefront PCR (lines along the contiguous index),
 * method 1 = partitioned Thomas (lines across it). FP64. */
int32_t pp_tdma_batched(pp_ctx *ctx, const double *a, const double *b,
                        const double *c, const double *r, double *u,
                        int32_t n, int32_t n_lines, int32_t method);
/* ---- single multigrid operators of one level (test and diagnostic access;
 *      each is one step of the V-cycle: no counterpart in the reference as a
 *      separate call) ----------------------------------------------------
 * Host arrays hold the cells of a level densely, i fastest: cx*cy*cz values
 * (pp_level_info dims).  precision: PP_PREC_FP64 or PP_PREC_MIXED (FP32, the
 * mixed solve's V-cycles).  The level's work arrays are overwritten (later
 * solves start from fresh ones). */
#define PP_LOP_RESIDUAL 0     /* out(l) = b(l) - A_l a(l) on fluid cells, 0 on blank
                                 cells (MG_residual, CORE_MG.F90:2274-2387) */
#define PP_LOP_RESTRICT 1     /* out(l+1) = fluid-volume-weighted mean of a(l) over the
                                 children (MG_Restrict_X/Y/Z, CORE_MG.F90:2464-2723) */
#define PP_LOP_RESRESTRICT 2  /* out(l+1) = the restriction of b(l) - A_l a(l), as the
                                 V-cycle forms it (fused where level l halves in x, y, z) */
#define PP_LOP_PROLONG 3      /* out(l) = b(l) + prolongation of a(l+1) on fluid cells
                                 (MG_Prolong_X/Y/Z + correction, CORE_MG.F90:2885-3523,
                                 1354-1355) */
#define PP_LOP_LINE_PASS 4    /* out(l) = a(l) after one zebra x, y, z line pass set with
                                 right-hand side b(l) and weight omega (MG_itsolv,
                                 CORE_MG.F90:1883-2270) */
int32_t pp_level_apply(pp_ctx *ctx, int32_t op, int32_t level, int32_t precision,
                       double omega, const double *a, const double *b, double *out);
/* Level l's flags (cx*cy*cz bytes, i fastest: bit0 blank, bit1 ghost, bits
 * 2-7 face closed on the i-,i+,j-,j+,k-,k+ side; MG_Prepare_Iblank,
 * CORE_MG.F90:699-822, and MG_Prepare, :1716-1879) and per direction the
 * cell widths, centres and face coefficients dxcinv(face)*dxinv(cell)
 * (MG_initial, :186-205), each as cx + cy + cz values (x, then y, then z).
 * Any pointer may be NULL. */
int32_t pp_level_data(pp_ctx *ctx, int32_t level, uint8_t *flags, double *w,
                      double *xc, double *cm, double *cp);
/* Left null vector of the ghost-cell operator (A^T w = 0), normalised so
 * that sum over fluid cells of w equals the fluid volume; reference layout,
 * fluid and ghost entries.  No counterpart in the reference. */
int32_t pp_null_vector(pp_ctx *ctx, double *w);
/* Level hierarchy: *n_levels and (cx,cy,cz) per level into dims[3*l]. */
int32_t pp_level_info(pp_ctx *ctx, int32_t *n_levels, int32_t *dims);
/* Diagnostics of the last pp_set_geometry (no counterpart in the reference):
 * phase_ms[4] = host+device time of the flags/levels, ghost tables, coarsest
 * factorisation and left-null-vector phases; the null vector's iterations
 * and final relative residual.  Any pointer may be NULL. */
int32_t pp_setup_info(pp_ctx *ctx, double *phase_ms, int32_t *null_iterations,
                      double *null_residual);
/* Kernel timing for roofline evidence (no counterpart in the reference):
 * reps back-to-back launches of one multigrid kernel class (PP_K_XLINE,
 * PP_K_YLINE, PP_K_ZLINE, PP_K_RESIDUAL) on one level in the given
 * precision, bracketed by two HIP events on the context's stream;
 * *avg_us = time per launch, *bytes = algorithmic bytes per launch.
 * Level 0 in PP_PREC_MIXED times the line passes on the colour-split copy
 * the mixed solve smooths (where the grid has one).  Overwrites that
 * level's correction/residual work arrays (no effect on later solves,
 * which start from a fresh correction). */
int32_t pp_time_kernel(pp_ctx *ctx, int32_t kernel_class, int32_t level,
                       int32_t precision, int32_t reps, double *avg_us,
                       double *bytes);

/* ---- geometry: body description -> pp_geometry inputs ---------------
 * Replaces set_boundary for canonical solid bodies (BOUNDARY_SET.F90:8-51):
 * iblank (set_iblank_canonical_body_fast, BOUNDARY_IBLANK_FAST.F90:2-501,
 * or the analytic cell-centre-inside rule), find_isolatedIblank(1)
 * (BOUNDARY_SET.F90:440-499), fill_cell (SSM_SET_BC.F90:1425-1511), ghost
 * marking, body intercepts, image points and Neumann coefficients
 * (GCM_set_internal_boundary, GCM_SET_BOUNDARY.F90:2-494).  Host only. */
#define PP_BODY_ELLIPSOID 3       /* canonical_body_type ELLIPSOID */
#define PP_GEOM_ANALYTIC 0        /* exact ellipsoid: centre-inside iblank, closest-point intercepts */
#define PP_GEOM_TRIANGULATED 1    /* the reference's triangulated surface and heuristics */

typedef struct pp_body_desc {
  int32_t shape;             /* PP_BODY_ELLIPSOID */
  double center[3];          /* xcent, ycent, zcent */
  double semi[3];            /* radiusx, radiusy, radiusz */
  double alpha_deg;          /* rotation about z (BOUNDARY_MARKER_INIT.F90:127-142) */
  int32_t n_theta;           /* markers per latitude ring (MAIN.F90:1140-1168) */
} pp_body_desc;

typedef struct pp_geom pp_geom;
/* iblank_prev (reference layout, may be NULL) marks fresh cells. */
int32_t pp_geom_build(pp_geom **out, const pp_grid_desc *g, int32_t n_body,
                      const pp_body_desc *bodies, int32_t mode,
                      double probe_length_normalized, const int8_t *iblank_prev);

/* One caller-given triangulated surface, as unstruc_surface_in.dat holds it
 * (BOUNDARY_MARKER_INIT.F90:209-248): n_markers markers (x, y, z each),
 * n_elems elements (triElemNeig: three 1-based marker ids each) and
 * pointOutsideBody, the point that orients the element normals. */
typedef struct pp_surface_desc {
  int32_t n_markers;
  int32_t n_elems;
  const double *xyz;
  const int32_t *tri;
  double outside[3];
} pp_surface_desc;

/* pp_geom_build in PP_GEOM_TRIANGULATED mode for UNSTRUCTURED_SURFACE bodies
 * (canonical_body_type 4, BOUNDARY_MARKER_INIT.F90:209-248 feeding the same
 * set_boundary pipeline) instead of canonical ellipsoids. */
int32_t pp_geom_build_surfaces(pp_geom **out, const pp_grid_desc *g,
                               int32_t n_surface, const pp_surface_desc *s,
                               double probe_norm, const int8_t *iblank_prev);

/* pp_geom_build (PP_GEOM_TRIANGULATED) and pp_geom_build_surfaces with the
 * per-cell, per-line and per-ghost work on GPU g->device (the band
 * dot-normal classification of set_iblank_canonical_body_fast, the i/j/k ray
 * fills, find_isolatedIblank, fill_cell, the ghost marking and list of
 * GCM_set_internal_boundary, GCM_calc_bodyIntercept_Unstruc, the image points
 * and GCM_Calc_VanMatrixDN's Vandermonde solves, GCM_set_freshcell;
 * BOUNDARY_IBLANK_FAST.F90:2-501, GCM_SET_BOUNDARY.F90:2-494,
 * GCM_SET_FRESHCELL.F90:1-262).  The result is the host build's, bit for bit,
 * behind the same handle and getters.  Grids up to 2^31 - 1 cells including
 * the reference's outer layers. */
int32_t pp_geom_build_device(pp_geom **out, const pp_grid_desc *g,
                             int32_t n_body, const pp_body_desc *bodies,
                             double probe_length_normalized,
                             const int8_t *iblank_prev);
int32_t pp_geom_build_surfaces_device(pp_geom **out, const pp_grid_desc *g,
                                      int32_t n_surface,
                                      const pp_surface_desc *s,
                                      double probe_norm,
                                      const int8_t *iblank_prev);

/* The marker surface PP_GEOM_TRIANGULATED builds for one canonical body
 * (BOUNDARY_MARKER_INIT.F90 ELLIPSOID branch, oriented against the origin),
 * e.g. to write it as unstruc_surface_in.dat.  Call with xyz = tri = NULL
 * for the counts, then with arrays of 3*n_markers and 3*n_elems. */
int32_t pp_geom_canonical_surface(const pp_body_desc *body, int32_t *n_markers,
                                  int32_t *n_elems, double *xyz, int32_t *tri);
int32_t pp_geom_counts(const pp_geom *geo, int32_t *n_ghost, int32_t *n_blank,
                       int32_t *n_fresh);
/* Any output pointer may be NULL.  coeff is coeffGCMN(8,n) column-major;
 * bi/ip/nrm are 3 per ghost (body intercept, image point, inward normal). */
int32_t pp_geom_get(const pp_geom *geo, int8_t *iblank, int8_t *fresh,
                    int32_t *ig, int32_t *jg, int32_t *kg, int32_t *ic,
                    int32_t *jc, int32_t *kc, double *coeff, double *probe,
                    double *bi, double *ip, double *nrm);
/* Per ghost: its body (1-based, bodyNum) and, in PP_GEOM_TRIANGULATED
 * mode, the surface element of its intercept (1-based, closestElementGC,
 * GCM_SET_BOUNDARY.F90:136-137; 0 in analytic mode).  Either may be NULL. */
int32_t pp_geom_ghost_elements(const pp_geom *geo, int32_t *body, int32_t *elem);
/* coeffGCMD(8,n) column-major: the velocity (Dirichlet) coefficients of
 * GCM_Calc_VanMatrixDN for the same ghosts, for pp_set_vel_ghosts. */
int32_t pp_geom_dirichlet(const pp_geom *geo, double *coeffD);
/* Fresh cells of a geometry built with iblank_prev (GCM_set_freshcell,
 * GCM_SET_FRESHCELL.F90:1-262), n_fresh of pp_geom_counts in the reference's
 * order (i fastest, then j, then k): the cell (1-based, iFresh..), the
 * lower corner of its 2x2x2 velocity stencil (iFreshCellIndex..),
 * coeffGCMFreshD(8, n) column-major in the stencil order of the ghost
 * coefficients, the body intercept (3 per cell, xBodyInterceptFresh..), the
 * body (1-based) and, in PP_GEOM_TRIANGULATED mode, the surface element
 * (1-based, closestElementFresh; 0 in analytic mode).  The intercept is taken
 * on the nearest body (the reference uses the body that blanked the cell at
 * the previous geometry).  Any output may be NULL.  PP_BAD_GEOMETRY when a
 * fresh cell's image point leaves the stencil (where the reference stops). */
int32_t pp_geom_fresh(const pp_geom *geo, int32_t *i_fresh, int32_t *j_fresh,
                      int32_t *k_fresh, int32_t *i_cell, int32_t *j_cell,
                      int32_t *k_cell, double *coeffD, double *bi,
                      int32_t *body, int32_t *elem);
int32_t pp_geom_destroy(pp_geom *geo);

#ifdef __cplusplus
}
#endif
#endif
