/* Runs the oracle's restatements once each on a small stretched grid with a
 * moving ellipsoid (set_boundary with fresh cells, GCM_set_freshcell, ghost
 * rows, calc_residual, mean removal, the multigrid solve, the momentum solve,
 * face_vel and the CRANK_NICOLSON2 loop) so that an AddressSanitizer /
 * UndefinedBehaviorSanitizer build reports any invalid access or undefined
 * operation.  Prints "ok" and exits 0 when everything ran. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void nodes(int n, double len, double beta, double *x) {
  for (int i = 0; i <= n; i++) {
    double xi = (double)i / n;
    x[i] = 0.5 * len * (1.0 + atanh((2.0 * xi - 1.0) * tanh(beta)) / beta);
  }
}

int main(void) {
  const int nx = 20, ny = 18, nz = 16;
  double xn[21], yn[19], zn[17];
  nodes(nx, 1.0, 1.2, xn);
  nodes(ny, 1.0, 0.9, yn);
  nodes(nz, 1.0, 1.4, zn);
  or_ctx *c = or_create(nx, ny, nz, xn, yn, zn);
  if (!c) { printf("create failed\n"); return 1; }
  const long n = or_field_size(c);
  int8_t *ib0 = calloc(n, 1), *ib = calloc(n, 1), *fr = calloc(n, 1);
  if (or_geom_ellipsoid(c, 24, 0.47, 0.5, 0.52, 0.22, 0.2, 0.24, 25.0) < 0) return 2;
  if (or_geom_set_boundary(c, NULL, 2.0)) return 3;
  or_geom_get_iblank(c, ib0, NULL);
  or_geom_clear_bodies(c);
  if (or_geom_ellipsoid(c, 24, 0.5, 0.5, 0.51, 0.22, 0.2, 0.24, 25.0) < 0) return 4;
  if (or_geom_set_boundary(c, ib0, 2.0)) return 5;
  or_geom_get_iblank(c, ib, fr);
  int rc = or_geom_set_freshcell(c, 2.0);
  printf("fresh tables %d (rc %d), ghosts %d\n", or_nfresh_tables(c), rc, or_nghost(c));
  double *p = calloc(n, sizeof(double)), *b = calloc(n, sizeof(double)), *res = calloc(n, sizeof(double));
  for (long r = 0; r < n; r++) b[r] = ib[r] ? 0.0 : sin(0.37 * (double)r);
  int32_t loc[3];
  double l2;
  double m = or_calc_residual(c, p, b, res, loc, &l2);
  or_mg_opts o;
  memset(&o, 0, sizeof(o));
  o.iter_max = 3; o.restol = 0.0; o.rel_tol = 0.0; o.iter_finest = 1; o.iter_coarsest = 1; o.omega = 1.0;
  o.itermax_ad = 50; o.restol_ad = 1e-12; o.mean_remove = 1; o.zebra = 0;
  or_mg_stats st;
  or_mg_solve(c, p, b, &o, &st);
  o.zebra = 1;
  or_mg_solve(c, p, b, &o, &st);
  or_mean_remove(c, p);
  or_ghost_rows(c, p, NULL, 50, 1e-12);
  or_point_jacobi(c, p, b, 0.8);
  double *y = calloc(n, sizeof(double));
  or_apply_A(c, p, y);
  or_apply_AT(c, p, y);
  /* fractional step and momentum */
  const int ng = or_nghost(c);
  double *cd = calloc(8 * (ng > 0 ? ng : 1), sizeof(double));
  or_geom_get_dirichlet(c, cd);
  or_set_vel_ghosts(c, cd, 2.0, 1.0);
  double *f[12];
  for (int q = 0; q < 12; q++) {
    f[q] = calloc(n, sizeof(double));
    for (long r = 0; r < n; r++) f[q][r] = 0.1 * cos(0.11 * (double)(r + 7 * q));
  }
  double *vbi = calloc(3 * (ng > 0 ? ng : 1), sizeof(double));
  or_correct_vel(c, 0.01, p, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], vbi, 50, 1e-12);
  or_face_vel(c, 1, 0.01, p, f[0], f[1], f[2], f[6], f[7], f[8], f[3], f[4], f[5]);
  double *coef = calloc(7 * n, sizeof(double));
  or_set_solve_ad(c, 0.01, f[9], f[10], f[10], f[10], coef);
  double mr;
  int iter = or_solve_ad(c, 0.005, coef, f[3], f[4], f[5], f[0], f[1], f[2], f[9], f[10], f[11], vbi, 20, 1e-10, &mr);
  iter += or_solve_ad_cn2(c, 0.01, coef, f[3], f[4], f[5], f[0], f[1], f[2], f[9], f[10], f[11], vbi, 1, p, f[6],
                          f[7], f[8], 5, 1e-10, &mr);
  printf("residual %.3e, mg iterations %d, momentum %d\n", m, st.iterations, iter);
  for (int q = 0; q < 12; q++) free(f[q]);
  free(coef); free(vbi); free(cd); free(y); free(p); free(b); free(res); free(ib0); free(ib); free(fr);
  or_destroy(c);
  printf("ok\n");
  return 0;
}
