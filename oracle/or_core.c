/*
 * or_core.c -- grid metrics, flags, residual, ghost rows, mean removal,
 * outer ghost layer, divergence and Thomas solve of the reference path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Parity pins (tests/test_oracle.py):
 *   make_grid      closed forms on uniform grids (x(i) = (i-1)h, dxc = h)
 *   ghost marking  brute force on small random masks
 *   calc_residual  dense assembly of the 7-point operator vs a scipy.sparse
 *                  matrix built independently in the test; O(h^2) truncation
 *                  of cos(pi x)cos(pi y)cos(pi z) with Neumann walls
 *   ghost rows     trilinear fields with zero normal derivative reproduced
 *                  exactly; sum of non-self coefficients = 1
 *   mean removal   closed form: fluid volume mean of the result is 0
 *   divergence     constant face field -> 0; linear field -> exact constant
 *   tdma           numpy dense solve of random diagonally dominant systems
 */
#include "or_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double *dalloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

/* make_grid (CORE_GRID.F90:29-151) for one direction: nodes x(1..n) given,
 * x(0) = -x(2), x(n+1) = 2x(n)-x(n-1); centres, widths and inverses. */
static void make_dir(int n, const double *xn, double *x, double *xc,
                     double *dx, double *dxinv, double *dxc, double *dxcinv) {
  for (int i = 1; i <= n; i++) x[i] = xn[i - 1];
  x[0] = -x[2];
  x[n + 1] = x[n] + x[n] - x[n - 1];
  for (int i = 0; i <= n; i++) {
    xc[i] = 0.5 * (x[i + 1] + x[i]);
    dx[i] = x[i + 1] - x[i];
    dxinv[i] = 1.0 / dx[i];
  }
  xc[n + 1] = x[n + 1];
  for (int i = 1; i <= n; i++) {
    dxc[i] = xc[i] - xc[i - 1];
    dxcinv[i] = 1.0 / dxc[i];
  }
}

or_ctx *or_create(int nxc, int nyc, int nzc, const double *xn,
                  const double *yn, const double *zn) {
  if (nxc < 2 || nyc < 2 || nzc < 2) return NULL;
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->nxc = nxc; c->nyc = nyc; c->nzc = nzc;
  c->nx = nxc + 1; c->ny = nyc + 1; c->nz = nzc + 1;
  c->ex = c->nx + 2; c->ey = c->ny + 2; c->ez = c->nz + 2;
  c->x = dalloc(c->nx + 2); c->xc = dalloc(c->nx + 2); c->dx = dalloc(c->nx + 2);
  c->dxinv = dalloc(c->nx + 2); c->dxc = dalloc(c->nx + 2); c->dxcinv = dalloc(c->nx + 2);
  c->y = dalloc(c->ny + 2); c->yc = dalloc(c->ny + 2); c->dy = dalloc(c->ny + 2);
  c->dyinv = dalloc(c->ny + 2); c->dyc = dalloc(c->ny + 2); c->dycinv = dalloc(c->ny + 2);
  c->z = dalloc(c->nz + 2); c->zc = dalloc(c->nz + 2); c->dz = dalloc(c->nz + 2);
  c->dzinv = dalloc(c->nz + 2); c->dzc = dalloc(c->nz + 2); c->dzcinv = dalloc(c->nz + 2);
  make_dir(c->nx, xn, c->x, c->xc, c->dx, c->dxinv, c->dxc, c->dxcinv);
  make_dir(c->ny, yn, c->y, c->yc, c->dy, c->dyinv, c->dyc, c->dycinv);
  make_dir(c->nz, zn, c->z, c->zc, c->dz, c->dzinv, c->dzc, c->dzcinv);
  long nf = or_field_size(c);
  c->iblank = (int8_t *)calloc(nf, 1);
  c->ghost = (int8_t *)calloc(nf, 1);
  c->fresh = (int8_t *)calloc(nf, 1);
  c->ium = (int8_t *)calloc(nf, 1); c->iup = (int8_t *)calloc(nf, 1);
  c->jum = (int8_t *)calloc(nf, 1); c->jup = (int8_t *)calloc(nf, 1);
  c->kum = (int8_t *)calloc(nf, 1); c->kup = (int8_t *)calloc(nf, 1);
  c->bodynum = (int32_t *)calloc(nf, sizeof(int32_t));
  or_mark_ghosts(c);
  return c;
}

static void free_ghosts(or_ctx *c) {
  free(c->ig); free(c->jg); free(c->kg);
  free(c->ic); free(c->jc); free(c->kc); free(c->elem);
  free(c->coeff); free(c->probe); free(c->bi); free(c->ipt); free(c->coeffD);
  c->ig = c->jg = c->kg = c->ic = c->jc = c->kc = c->elem = NULL;
  c->coeff = c->probe = c->bi = c->ipt = c->coeffD = NULL;
}

void or_free_fresh(or_ctx *c) {
  free(c->fr_i); free(c->fr_j); free(c->fr_k); free(c->fr_ic); free(c->fr_jc); free(c->fr_kc);
  free(c->fr_body); free(c->fr_elem); free(c->fr_cD); free(c->fr_bi);
  c->fr_i = c->fr_j = c->fr_k = c->fr_ic = c->fr_jc = c->fr_kc = c->fr_body = c->fr_elem = NULL;
  c->fr_cD = c->fr_bi = NULL;
  c->nfr = 0;
}

void or_alloc_fresh(or_ctx *c, int n) {
  or_free_fresh(c);
  c->nfr = n;
  size_t m = n > 0 ? (size_t)n : 1;
  c->fr_i = calloc(m, 4); c->fr_j = calloc(m, 4); c->fr_k = calloc(m, 4);
  c->fr_ic = calloc(m, 4); c->fr_jc = calloc(m, 4); c->fr_kc = calloc(m, 4);
  c->fr_body = calloc(m, 4); c->fr_elem = calloc(m, 4);
  c->fr_cD = calloc(8 * m, sizeof(double)); c->fr_bi = calloc(3 * m, sizeof(double));
}

void or_set_fresh_tables(or_ctx *c, int n, const int32_t *ifr, const int32_t *jfr, const int32_t *kfr,
                         const int32_t *ic, const int32_t *jc, const int32_t *kc, const double *coeffD) {
  or_alloc_fresh(c, n);
  for (int q = 0; q < n; q++) {
    c->fr_i[q] = ifr[q]; c->fr_j[q] = jfr[q]; c->fr_k[q] = kfr[q];
    c->fr_ic[q] = ic[q]; c->fr_jc[q] = jc[q]; c->fr_kc[q] = kc[q];
    memcpy(c->fr_cD + 8 * q, coeffD + 8 * q, 8 * sizeof(double));
  }
}

int or_nfresh_tables(const or_ctx *c) { return c->nfr; }

void or_destroy(or_ctx *c) {
  if (!c) return;
  double *arr[] = {c->x, c->xc, c->dx, c->dxinv, c->dxc, c->dxcinv,
                   c->y, c->yc, c->dy, c->dyinv, c->dyc, c->dycinv,
                   c->z, c->zc, c->dz, c->dzinv, c->dzc, c->dzcinv};
  for (size_t i = 0; i < sizeof(arr) / sizeof(arr[0]); i++) free(arr[i]);
  free(c->iblank); free(c->ghost); free(c->fresh);
  free(c->ium); free(c->iup); free(c->jum); free(c->jup);
  free(c->kum); free(c->kup); free(c->bodynum);
  free_ghosts(c);
  or_free_fresh(c);
  or_geom_clear_bodies(c);
  free(c);
}

long or_field_size(const or_ctx *c) { return c->ex * c->ey * c->ez; }

void or_get_metrics(const or_ctx *c, int dir, double *out) {
  int n = dir == 0 ? c->nx : dir == 1 ? c->ny : c->nz;
  const double *src[6];
  if (dir == 0) {
    src[0] = c->x; src[1] = c->xc; src[2] = c->dx; src[3] = c->dxinv; src[4] = c->dxc; src[5] = c->dxcinv;
  } else if (dir == 1) {
    src[0] = c->y; src[1] = c->yc; src[2] = c->dy; src[3] = c->dyinv; src[4] = c->dyc; src[5] = c->dycinv;
  } else {
    src[0] = c->z; src[1] = c->zc; src[2] = c->dz; src[3] = c->dzinv; src[4] = c->dzc; src[5] = c->dzcinv;
  }
  for (int a = 0; a < 6; a++) memcpy(out + (long)a * (n + 2), src[a], sizeof(double) * (n + 2));
}

/* set_outer_iup (BOUNDARY_SET.F90:54-92) followed by the GCM ghost marking
 * and i-fastest ghost list (GCM_SET_BOUNDARY.F90:49-118).  Under GCM the
 * internal face flags stay 0 (set_internal_iup_GCM is never called,
 * SSM_SET_INTERNAL_IUP_AREA.F90:381-407). */
void or_mark_ghosts(or_ctx *c) {
  long nf = or_field_size(c);
  memset(c->ium, 0, nf); memset(c->iup, 0, nf);
  memset(c->jum, 0, nf); memset(c->jup, 0, nf);
  memset(c->kum, 0, nf); memset(c->kup, 0, nf);
  memset(c->ghost, 0, nf);
  for (int j = 1; j <= c->ny; j++)
    for (int k = 1; k <= c->nz; k++) {
      c->ium[OR_IDX(c, 1, j, k)] = 1;
      c->iup[OR_IDX(c, c->nxc, j, k)] = 1;
    }
  for (int i = 1; i <= c->nx; i++)
    for (int k = 1; k <= c->nz; k++) {
      c->jum[OR_IDX(c, i, 1, k)] = 1;
      c->jup[OR_IDX(c, i, c->nyc, k)] = 1;
    }
  for (int i = 1; i <= c->nx; i++)
    for (int j = 1; j <= c->ny; j++) {
      c->kum[OR_IDX(c, i, j, 1)] = 1;
      c->kup[OR_IDX(c, i, j, c->nzc)] = 1;
    }
  const int8_t *ib = c->iblank;
  int nb = 0;
  for (int k = 1; k <= c->nz - 1; k++)
    for (int j = 1; j <= c->ny - 1; j++)
      for (int i = 1; i <= c->nx - 1; i++) {
        int iM = i - 1 > 1 ? i - 1 : 1, iP = i + 1 < c->nx - 1 ? i + 1 : c->nx - 1;
        int jM = j - 1 > 1 ? j - 1 : 1, jP = j + 1 < c->ny - 1 ? j + 1 : c->ny - 1;
        int kM = k - 1 > 1 ? k - 1 : 1, kP = k + 1 < c->nz - 1 ? k + 1 : c->nz - 1;
        int b = ib[OR_IDX(c, i, j, k)];
        if (b * ib[OR_IDX(c, iM, j, k)] == 0 || b * ib[OR_IDX(c, iP, j, k)] == 0 ||
            b * ib[OR_IDX(c, i, jM, k)] == 0 || b * ib[OR_IDX(c, i, jP, k)] == 0 ||
            b * ib[OR_IDX(c, i, j, kM)] == 0 || b * ib[OR_IDX(c, i, j, kP)] == 0) {
          if (b == 1) { c->ghost[OR_IDX(c, i, j, k)] = 1; nb++; }
        }
      }
  free_ghosts(c);
  c->nghost = nb;
  long ng = nb > 0 ? nb : 1;
  c->ig = (int32_t *)calloc(ng, 4); c->jg = (int32_t *)calloc(ng, 4); c->kg = (int32_t *)calloc(ng, 4);
  c->ic = (int32_t *)calloc(ng, 4); c->jc = (int32_t *)calloc(ng, 4); c->kc = (int32_t *)calloc(ng, 4);
  c->elem = (int32_t *)calloc(ng, 4);
  c->coeff = dalloc(8 * ng); c->probe = dalloc(ng);
  c->bi = dalloc(3 * ng); c->ipt = dalloc(3 * ng);
  c->coeffD = dalloc(8 * ng);
  nb = 0;
  for (int k = 1; k <= c->nz - 1; k++)
    for (int j = 1; j <= c->ny - 1; j++)
      for (int i = 1; i <= c->nx - 1; i++)
        if (c->ghost[OR_IDX(c, i, j, k)] == 1) {
          c->ig[nb] = i; c->jg[nb] = j; c->kg[nb] = k; nb++;
        }
}

int or_set_iblank(or_ctx *c, const int8_t *iblank) {
  memcpy(c->iblank, iblank, or_field_size(c));
  or_mark_ghosts(c);
  return c->nghost;
}

int or_nghost(const or_ctx *c) { return c->nghost; }

void or_get_flags(const or_ctx *c, int8_t *ghost, int8_t *ium, int8_t *iup,
                  int8_t *jum, int8_t *jup, int8_t *kum, int8_t *kup) {
  long nf = or_field_size(c);
  if (ghost) memcpy(ghost, c->ghost, nf);
  if (ium) memcpy(ium, c->ium, nf);
  if (iup) memcpy(iup, c->iup, nf);
  if (jum) memcpy(jum, c->jum, nf);
  if (jup) memcpy(jup, c->jup, nf);
  if (kum) memcpy(kum, c->kum, nf);
  if (kup) memcpy(kup, c->kup, nf);
}

void or_get_ghost_list(const or_ctx *c, int32_t *ig, int32_t *jg, int32_t *kg) {
  memcpy(ig, c->ig, 4 * (size_t)c->nghost);
  memcpy(jg, c->jg, 4 * (size_t)c->nghost);
  memcpy(kg, c->kg, 4 * (size_t)c->nghost);
}

void or_set_ghost_stencils(or_ctx *c, const int32_t *ic, const int32_t *jc,
                           const int32_t *kc, const double *coeff,
                           const double *probe) {
  memcpy(c->ic, ic, 4 * (size_t)c->nghost);
  memcpy(c->jc, jc, 4 * (size_t)c->nghost);
  memcpy(c->kc, kc, 4 * (size_t)c->nghost);
  memcpy(c->coeff, coeff, 8 * 8 * (size_t)c->nghost);
  memcpy(c->probe, probe, 8 * (size_t)c->nghost);
}

void or_set_fresh(or_ctx *c, const int8_t *fresh) {
  memcpy(c->fresh, fresh, or_field_size(c));
  or_free_fresh(c);
}

/* calc_residual (CORE_POISSON.F90:1111-1222), non-cut-cell branch, 3-D,
 * with nlw0 = 0 (pgrad = 0: no ZERO_GRADIENT faces). */
double or_calc_residual(const or_ctx *c, const double *var, const double *r,
                        double *res_out, int32_t *loc, double *l2) {
  double resm = 0.0, s2 = 0.0;
  if (loc) loc[0] = loc[1] = loc[2] = 0;
  if (res_out) memset(res_out, 0, sizeof(double) * or_field_size(c));
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double bmx = c->dxcinv[i] * c->dxinv[i] * (1 - c->ium[id]);
        double bpx = c->dxcinv[i + 1] * c->dxinv[i] * (1 - c->iup[id]);
        double bcx = -(bmx + bpx);
        double bmy = c->dycinv[j] * c->dyinv[j] * (1 - c->jum[id]);
        double bpy = c->dycinv[j + 1] * c->dyinv[j] * (1 - c->jup[id]);
        double bcy = -(bmy + bpy);
        double bmz = c->dzcinv[k] * c->dzinv[k] * (1 - c->kum[id]);
        double bpz = c->dzcinv[k + 1] * c->dzinv[k] * (1 - c->kup[id]);
        double bcz = -(bmz + bpz);
        double fl = (double)(1 - c->iblank[id]);
        bmx *= fl; bpx *= fl; bmy *= fl; bpy *= fl; bmz *= fl; bpz *= fl;
        double bc = (bcx + bcy + bcz) * fl + (double)c->iblank[id];
        double res = r[id] - var[id] * bc - var[id - 1] * bmx - var[id + 1] * bpx -
                     var[id - c->ex] * bmy - var[id + c->ex] * bpy -
                     var[id - c->ex * c->ey] * bmz - var[id + c->ex * c->ey] * bpz;
        res = res * fl;
        if (res_out) res_out[id] = res;
        s2 += res * res;
        if (fabs(res) > resm) {
          resm = fabs(res);
          if (loc) { loc[0] = i; loc[1] = j; loc[2] = k; }
        }
      }
  if (l2) *l2 = sqrt(s2);
  return resm;
}

/* MG_itsolv_Point_Jacobi (CORE_MG.F90:3716-3800) as MG_X/Y/Z call it on the
 * finest level (nLevX = nLevY = nLevZ = 1): the level metrics are the grid's,
 * ium_MG..kup_MG are the fine face flags and ghostcellMark_MG the ghost
 * marks (MG_Prepare_BC at n = 1, CORE_MG.F90:1686-1700), nlw = 0, 3-D.
 * Out of place over the interior cells: var1 = var + omega (rhss - var). */
void or_point_jacobi(const or_ctx *c, double *var, const double *r, double omega) {
  long n = or_field_size(c), sxy = c->ex * c->ey;
  double *var1 = dalloc(n);
  memcpy(var1, var, sizeof(double) * n);
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double bmx = c->dxcinv[i] * c->dxinv[i] * (1.0 - c->ium[id]);
        double bpx = c->dxcinv[i + 1] * c->dxinv[i] * (1.0 - c->iup[id]);
        double bcx = -(bmx + bpx);
        double bmy = c->dycinv[j] * c->dyinv[j] * (1.0 - c->jum[id]);
        double bpy = c->dycinv[j + 1] * c->dyinv[j] * (1.0 - c->jup[id]);
        double bcy = -(bmy + bpy);
        double bmz = c->dzcinv[k] * c->dzinv[k] * (1.0 - c->kum[id]) * 1.0;
        double bpz = c->dzcinv[k + 1] * c->dzinv[k] * (1.0 - c->kup[id]) * 1.0;
        double bcz = -(bmz + bpz);
        double fl = 1.0 - (double)c->iblank[id];
        bmx *= fl; bpx *= fl; bmy *= fl; bpy *= fl; bmz *= fl; bpz *= fl;
        double bc = (bcx + bcy + bcz) * fl + (double)c->iblank[id];
        double rhss = r[id] - var[id - 1] * bmx - var[id + 1] * bpx - var[id - c->ex] * bmy -
                      var[id + c->ex] * bpy - var[id - sxy] * bmz - var[id + sxy] * bpz + 0.0;
        rhss = (rhss / bc) * fl + (double)c->ghost[id] * var[id];
        var1[id] = var[id] + omega * (rhss - var[id]);
      }
  memcpy(var, var1, sizeof(double) * n);
  free(var1);
}

/* subtracting_average_pressure (CORE_POISSON.F90:419-490) */
double or_mean_remove(const or_ctx *c, double *p) {
  double pt = 0.0, vt = 0.0;
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double v = c->dx[i] * c->dy[j] * c->dz[k];
        pt += p[id] * v * (1 - c->iblank[id]);
        vt += v * (1 - c->iblank[id]);
      }
  double pav = pt / vt;
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (c->iblank[id] == 1) p[id] = 0.0;
        else p[id] = p[id] - pav;
      }
  for (int k = 1; k <= c->nz - 1; k++)
    for (int j = 1; j <= c->ny - 1; j++)
      for (int i = 1; i <= c->nx - 1; i++) {
        if (c->fresh[OR_IDX(c, i, j, k)] != 1) continue;
        int i1 = 0;
        double p1 = 0.0;
        for (int kk = k - 1; kk <= k + 1; kk++)
          for (int jj = j - 1; jj <= j + 1; jj++)
            for (int ii = i - 1; ii <= i + 1; ii++) {
              long q = OR_IDX(c, ii, jj, kk);
              if (c->iblank[q] != 1 && c->fresh[q] != 1) { i1++; p1 += p[q]; }
            }
        /* the reference divides by i1 = 0 for a fresh cell with no fluid,
         * non-fresh neighbour (a body that moved more than a cell); such a
         * cell keeps its value here */
        if (i1 > 0) p[OR_IDX(c, i, j, k)] = p1 / (double)i1;
      }
  return pav;
}

/* set_outer_ghost_pres with PBC_NEUMANN on all six faces
 * (CORE_SET_OUTER_GHOST.F90:66-213); index ranges 0..n as in GCM_Pressure. */
void or_outer_ghost(const or_ctx *c, double *p) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++) {
      p[OR_IDX(c, 0, j, k)] = p[OR_IDX(c, 1, j, k)];
      p[OR_IDX(c, nx, j, k)] = p[OR_IDX(c, nx - 1, j, k)];
    }
  for (int k = 0; k <= nz; k++)
    for (int i = 0; i <= nx; i++) {
      p[OR_IDX(c, i, 0, k)] = p[OR_IDX(c, i, 1, k)];
      p[OR_IDX(c, i, ny, k)] = p[OR_IDX(c, i, ny - 1, k)];
    }
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      p[OR_IDX(c, i, j, 0)] = p[OR_IDX(c, i, j, 1)];
      p[OR_IDX(c, i, j, nz)] = p[OR_IDX(c, i, j, nz - 1)];
    }
}

/* GCM_p_set_bc_internal (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:250-384):
 * Gauss-Seidel over the ghost list in list order.  The stencil offsets
 * incI/J/K follow GCM_MEMORY_ALLOCATE.F90:494-518. */
static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

int or_ghost_rows(const or_ctx *c, double *p, const double *g, int itermax,
                  double restol) {
  int it = 0;
  double rmax = 1.0e10;
  while (it < itermax && rmax > restol) {
    rmax = 0.0;
    for (int n = 0; n < c->nghost; n++) {
      int iG = c->ig[n], jG = c->jg[n], kG = c->kg[n];
      double un = g ? g[n] : 0.0;
      double pip = 0.0;
      for (int r = 0; r < 8; r++) {
        int ii = c->ic[n] + incI[r], jj = c->jc[n] + incJ[r], kk = c->kc[n] + incK[r];
        if (ii == iG && jj == jG && kk == kG) pip += c->coeff[8 * n + r] * un;
        else pip += c->coeff[8 * n + r] * p[OR_IDX(c, ii, jj, kk)];
      }
      long gid = OR_IDX(c, iG, jG, kG);
      double prev = p[gid];
      p[gid] = pip + c->probe[n] * un;
      double d = fabs(p[gid] - prev);
      if (d > rmax) rmax = d;
    }
    it++;
  }
  return it;
}

/* rhs_poisson, non-cut-cell branch (CORE_POISSON.F90:281-288) */
double or_divergence(const or_ctx *c, const double *fu, const double *fv,
                     const double *fw, double dtinv, double *nlu0) {
  double sum = 0.0;
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double v = dtinv * ((fu[id + 1] - fu[id]) * c->dxinv[i] +
                            (fv[id + c->ex] - fv[id]) * c->dyinv[j] +
                            (fw[id + c->ex * c->ey] - fw[id]) * c->dzinv[k]);
        v = v * (double)(1 - c->iblank[id]);
        nlu0[id] = v;
        sum += v * c->dx[i] * c->dy[j] * c->dz[k];
      }
  return sum;
}

/* tdma (CORE_TDMA.F90:10-34) */
void or_tdma(const double *a, const double *b, const double *cc,
             const double *r, double *u, int n1, int n2) {
  double *gam = dalloc(n2 + 1);
  double binv = 1.0 / b[n1];
  u[n1] = r[n1] * binv;
  for (int j = n1 + 1; j <= n2; j++) {
    gam[j] = cc[j - 1] * binv;
    binv = 1.0 / (b[j] - a[j] * gam[j]);
    u[j] = (r[j] - a[j] * u[j - 1]) * binv;
  }
  for (int j = n2 - 1; j >= n1; j--) u[j] = u[j] - gam[j + 1] * u[j + 1];
  free(gam);
}

/* Outer-layer indices read by a ghost stencil hold Neumann copies of the
 * adjacent interior cell (set_outer_ghost_pres); clamping the index to the
 * interior reproduces that copy, edges and corners included. */
static long clamp_idx(const or_ctx *c, int i, int j, int k) {
  if (i < 1) i = 1;
  if (i > c->nxc) i = c->nxc;
  if (j < 1) j = 1;
  if (j > c->nyc) j = c->nyc;
  if (k < 1) k = 1;
  if (k > c->nzc) k = c->nzc;
  return OR_IDX(c, i, j, k);
}

/* The linear operator whose residual calc_residual measures: fluid rows are
 * the masked 7-point stretched Laplacian, ghost rows p_g - sum c_r p_r with
 * the ghost's own stencil entry carrying the Neumann datum (not p). */
void or_apply_A(const or_ctx *c, const double *p, double *y) {
  memset(y, 0, sizeof(double) * or_field_size(c));
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (c->iblank[id]) continue;
        double bmx = c->dxcinv[i] * c->dxinv[i] * (1 - c->ium[id]);
        double bpx = c->dxcinv[i + 1] * c->dxinv[i] * (1 - c->iup[id]);
        double bmy = c->dycinv[j] * c->dyinv[j] * (1 - c->jum[id]);
        double bpy = c->dycinv[j + 1] * c->dyinv[j] * (1 - c->jup[id]);
        double bmz = c->dzcinv[k] * c->dzinv[k] * (1 - c->kum[id]);
        double bpz = c->dzcinv[k + 1] * c->dzinv[k] * (1 - c->kup[id]);
        double bc = -(bmx + bpx) - (bmy + bpy) - (bmz + bpz);
        y[id] = p[id] * bc + p[id - 1] * bmx + p[id + 1] * bpx +
                p[id - c->ex] * bmy + p[id + c->ex] * bpy +
                p[id - c->ex * c->ey] * bmz + p[id + c->ex * c->ey] * bpz;
      }
  for (int n = 0; n < c->nghost; n++) {
    int iG = c->ig[n], jG = c->jg[n], kG = c->kg[n];
    long gid = OR_IDX(c, iG, jG, kG);
    double s = p[gid];
    for (int r = 0; r < 8; r++) {
      int ii = c->ic[n] + incI[r], jj = c->jc[n] + incJ[r], kk = c->kc[n] + incK[r];
      if (ii == iG && jj == jG && kk == kG) continue;
      long q = clamp_idx(c, ii, jj, kk);
      if (c->iblank[q] && !c->ghost[q]) continue; /* solid cells hold 0 */
      s -= c->coeff[8 * n + r] * p[q];
    }
    y[gid] = s;
  }
}

void or_apply_AT(const or_ctx *c, const double *x, double *y) {
  memset(y, 0, sizeof(double) * or_field_size(c));
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (c->iblank[id]) continue;
        double bmx = c->dxcinv[i] * c->dxinv[i] * (1 - c->ium[id]);
        double bpx = c->dxcinv[i + 1] * c->dxinv[i] * (1 - c->iup[id]);
        double bmy = c->dycinv[j] * c->dyinv[j] * (1 - c->jum[id]);
        double bpy = c->dycinv[j + 1] * c->dyinv[j] * (1 - c->jup[id]);
        double bmz = c->dzcinv[k] * c->dzinv[k] * (1 - c->kum[id]);
        double bpz = c->dzcinv[k + 1] * c->dzinv[k] * (1 - c->kup[id]);
        double bc = -(bmx + bpx) - (bmy + bpy) - (bmz + bpz);
        double v = x[id];
        y[id] += bc * v;
        y[id - 1] += bmx * v; y[id + 1] += bpx * v;
        y[id - c->ex] += bmy * v; y[id + c->ex] += bpy * v;
        y[id - c->ex * c->ey] += bmz * v; y[id + c->ex * c->ey] += bpz * v;
      }
  for (int n = 0; n < c->nghost; n++) {
    int iG = c->ig[n], jG = c->jg[n], kG = c->kg[n];
    long gid = OR_IDX(c, iG, jG, kG);
    double v = x[gid];
    y[gid] += v;
    for (int r = 0; r < 8; r++) {
      int ii = c->ic[n] + incI[r], jj = c->jc[n] + incJ[r], kk = c->kc[n] + incK[r];
      if (ii == iG && jj == jG && kk == kG) continue;
      long q = clamp_idx(c, ii, jj, kk);
      if (c->iblank[q] && !c->ghost[q]) continue;
      y[q] -= c->coeff[8 * n + r] * v;
    }
  }
  /* only fluid and ghost unknowns exist */
  for (int k = 0; k < c->ez; k++)
    for (int j = 0; j < c->ey; j++)
      for (int i = 0; i < c->ex; i++) {
        long id = OR_IDX(c, i, j, k);
        int interior = i >= 1 && i <= c->nxc && j >= 1 && j <= c->nyc && k >= 1 && k <= c->nzc;
        if (!interior || (c->iblank[id] && !c->ghost[id])) y[id] = 0.0;
      }
}
