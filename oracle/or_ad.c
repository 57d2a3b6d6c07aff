/*
 * or_ad.c -- the implicit momentum (advection-diffusion) solve of the
 * fractional step: set_solve_ad (CORE_SET_SOLVERS.F90:20-223, with the
 * GCM fresh-cell rows of :175-218), the line sweeps itsolv_ad_x/y/z
 * (CORE_SOLVE_AD.F90:314-762; with half_dt = 0 they are the
 * itsolv_ad_*_AB routines of the same file), calc_residual_ad (:766-907)
 * and the GCM loop of solve_ad (:200-267) with the velocity ghost update
 * GCM_GhostCell_Vel (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:2-124; its
 * GCM_enforce_global_mass_consv acts only on ZERO_GRADIENT outer faces and
 * is a no-op for Dirichlet ones).  Test infrastructure only (see
 * oracle.h).  Fresh cells (fresh mask with the tables of
 * or_geom_set_freshcell or or_set_fresh_tables) take the diagonal 1
 * (killForGCMFresh) and the off-stencil terms of GCM_correct_rhs_ad /
 * GCM_correct_res_ad (GCM_FRESHCELL_AD.F90:2-137).  Restated for the
 * configurations of the path: no periodic outer boundaries, advec_scheme
 * ADAMS_BASHFORTH2 (half_dt = 0) or CRANK_NICOLSON1 (half_dt = dt/2).
 *
 * coef holds amx_ad, apx_ad, amy_ad, apy_ad, amz_ad, apz_ad, one field
 * each, in that order.
 *
 * Pins (tests/test_oracle_ad.py):
 *   set_solve_ad     uniform grid: interior coefficients -dt/2 nu / h^2,
 *                    outer-face coefficients twice that (closed form).
 *   calc_residual_ad a field linear in x, y, z, a viscosity linear in
 *                    x, y, z and constant face velocities on a stretched
 *                    grid: the residual of r = var + half_dt (U b_x + V b_y
 *                    + W b_z) - (dt/2) (b_x dnu/dx + b_y dnu/dy + b_z
 *                    dnu/dz) is zero to round-off on every fluid cell off
 *                    the outer faces and the body (both interpolations are
 *                    exact for linear fields).  A quadratic field on a
 *                    uniform grid fixes the diffusion scale.
 *   solve_ad         on a small grid with a body the converged solution
 *                    equals a dense solve (numpy) of the system assembled
 *                    column by column from calc_residual_ad and the
 *                    ghost rows u_g = bw u_BI - iw sum_r cD_r u_r.
 * Fresh-cell pins (tests/test_oracle_fresh_ad.py):
 *   set_freshcell    the weights reproduce every trilinear function at the
 *                    cell centre from its stencil values with the cell's own
 *                    node taken at BI; BI of a sphere is the radial
 *                    projection of the centre within the marker spacing;
 *                    the stencil corner follows a brute-force search.
 *   fresh residual   ADAMS_BASHFORTH2: the residual at a fresh cell is
 *                    r - (u_f - sum over the other seven stencil nodes of
 *                    cD u), the full interpolation row; CRANK_NICOLSON1 adds
 *                    exactly the advection terms an ordinary cell's row
 *                    adds (the same difference with the cell's value 0).
 *   fresh solve_ad   the converged solution equals the dense solve with the
 *                    fresh rows u_f - sum cD u = r_f, and satisfies the
 *                    interpolation at each fresh cell.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "or_internal.h"

/* stencil order of the GCM interpolation (incI, incJ, incK) */
static const int or_incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int or_incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int or_incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

/* fx(i) = (x(i) - xc(i-1)) / (xc(i) - xc(i-1)) (CORE_GRID.F90:172-174) */
static double fint(const double *x, const double *xc, int i) { return (x[i] - xc[i - 1]) / (xc[i] - xc[i - 1]); }

void or_set_solve_ad(const or_ctx *c, double dt, const double *visc, const double *bxv, const double *byv,
                     const double *bzv, double *coef) {
  const long nf = or_field_size(c), ex = c->ex, exy = c->ex * c->ey;
  double *amx = coef, *apx = coef + nf, *amy = coef + 2 * nf, *apy = coef + 3 * nf, *amz = coef + 4 * nf,
         *apz = coef + 5 * nf;
  memset(coef, 0, sizeof(double) * 6 * (size_t)nf);
  const double rnDim = 1.0; /* ndim - DIM_2D for a 3-D grid */
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        const long id = OR_IDX(c, i, j, k);
        const double fxe = fint(c->x, c->xc, i + 1), fxw = fint(c->x, c->xc, i);
        const double fyn = fint(c->y, c->yc, j + 1), fys = fint(c->y, c->yc, j);
        const double fzf = fint(c->z, c->zc, k + 1), fzb = fint(c->z, c->zc, k);
        const int ium = c->ium[id], iup = c->iup[id], jum = c->jum[id], jup = c->jup[id], kum = c->kum[id],
                  kup = c->kup[id];
        const double nuE = (fxe * visc[id + 1] + (1.0 - fxe) * visc[id]) * (1 - iup) + bxv[id] * iup;
        const double nuW = (fxw * visc[id] + (1.0 - fxw) * visc[id - 1]) * (1 - ium) + bxv[id] * ium;
        const double nuN = (fyn * visc[id + ex] + (1.0 - fyn) * visc[id]) * (1 - jup) + byv[id] * jup;
        const double nuS = (fys * visc[id] + (1.0 - fys) * visc[id - ex]) * (1 - jum) + byv[id] * jum;
        const double nuF = (fzf * visc[id + exy] + (1.0 - fzf) * visc[id]) * (1 - kup) + bzv[id] * kup;
        const double nuB = (fzb * visc[id] + (1.0 - fzb) * visc[id - exy]) * (1 - kum) + bzv[id] * kum;
        const double hdt = 0.5 * dt;
        double a;
        a = (c->dxcinv[i] * (1 - ium) + c->dxinv[i] * ium * 2.0) * c->dxinv[i];
        amx[id] = -(hdt * nuW) * a;
        a = (c->dxcinv[i + 1] * (1 - iup) + c->dxinv[i] * iup * 2.0) * c->dxinv[i];
        apx[id] = -(hdt * nuE) * a;
        a = (c->dycinv[j] * (1 - jum) + c->dyinv[j] * jum * 2.0) * c->dyinv[j];
        amy[id] = -(hdt * nuS) * a;
        a = (c->dycinv[j + 1] * (1 - jup) + c->dyinv[j] * jup * 2.0) * c->dyinv[j];
        apy[id] = -(hdt * nuN) * a;
        a = (c->dzcinv[k] * (1 - kum) + c->dzinv[k] * kum * 2.0) * c->dzinv[k];
        amz[id] = -(hdt * nuB) * a * rnDim;
        a = (c->dzcinv[k + 1] * (1 - kup) + c->dzinv[k] * kup * 2.0) * c->dzinv[k];
        apz[id] = -(hdt * nuF) * a * rnDim;
      }
  /* GCM branch (:175-218): the rows of the fresh cells, zero but for the
   * face neighbours in the stencil, which take -coeffGCMFreshD */
  for (int n = 0; n < c->nfr; n++) {
    const int iF = c->fr_i[n], jF = c->fr_j[n], kF = c->fr_k[n];
    const long id = OR_IDX(c, iF, jF, kF);
    amx[id] = apx[id] = amy[id] = apy[id] = amz[id] = apz[id] = 0.0;
    for (int r = 0; r < 8; r++) {
      const int i = c->fr_ic[n] + or_incI[r], j = c->fr_jc[n] + or_incJ[r], k = c->fr_kc[n] + or_incK[r];
      const double cD = c->fr_cD[8 * n + r];
      if (i == iF - 1 && j == jF && k == kF) amx[id] = -cD;
      if (i == iF + 1 && j == jF && k == kF) apx[id] = -cD;
      if (i == iF && j == jF - 1 && k == kF) amy[id] = -cD;
      if (i == iF && j == jF + 1 && k == kF) apy[id] = -cD;
      if (i == iF && j == jF && k == kF - 1) amz[id] = -cD;
      if (i == iF && j == jF && k == kF + 1) apz[id] = -cD;
    }
  }
}

/* the fresh-table entry of cell (i, j, k), -1 if none */
static int fresh_entry(const or_ctx *c, int i, int j, int k) {
  for (int n = 0; n < c->nfr; n++)
    if (c->fr_i[n] == i && c->fr_j[n] == j && c->fr_k[n] == k) return n;
  return -1;
}

/* GCM_correct_rhs_ad / GCM_correct_res_ad (GCM_FRESHCELL_AD.F90:2-137):
 * x + coeffGCMFreshD * var at each stencil node outside the 7-point
 * stencil (two indices off by one), added one by one in stencil order */
static double fresh_correct(const or_ctx *c, int n, const double *var, double x) {
  const int iF = c->fr_i[n], jF = c->fr_j[n], kF = c->fr_k[n];
  for (int r = 0; r < 8; r++) {
    const int i = c->fr_ic[n] + or_incI[r], j = c->fr_jc[n] + or_incJ[r], k = c->fr_kc[n] + or_incK[r];
    const int di = i == iF - 1 || i == iF + 1, dj = j == jF - 1 || j == jF + 1, dk = k == kF - 1 || k == kF + 1;
    if ((di && dj) || (di && dk) || (dj && dk)) x = x + c->fr_cD[8 * n + r] * var[OR_IDX(c, i, j, k)];
  }
  return x;
}

/* The per-cell operator of itsolv_ad_* and calc_residual_ad before any
 * masking: the viscous coefficients plus the Crank-Nicolson advection
 * terms (CORE_SOLVE_AD.F90:391-421), and the unmasked diagonal sum
 * acx + acy + acz. */
typedef struct ad_row {
  double amx, apx, amy, apy, amz, apz, sum;
} ad_row;

static ad_row ad_cell(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                      const double *fw, int i, int j, int k) {
  const long nf = or_field_size(c), ex = c->ex, exy = c->ex * c->ey;
  const long id = OR_IDX(c, i, j, k);
  ad_row r;
  double amx = coef[id], apx = coef[nf + id], amy = coef[2 * nf + id], apy = coef[3 * nf + id],
         amz = coef[4 * nf + id], apz = coef[5 * nf + id];
  double acx = -(amx + apx), acy = -(amy + apy), acz = -(amz + apz);
  const double fxw = fint(c->x, c->xc, i), fxe = fint(c->x, c->xc, i + 1);
  const double fys = fint(c->y, c->yc, j), fyn = fint(c->y, c->yc, j + 1);
  const double fzb = fint(c->z, c->zc, k), fzf = fint(c->z, c->zc, k + 1);
  double t1, t2;
  t1 = 1.0 - fxw; t2 = fxe;
  amx = amx - half_dt * fu[id] * t1 * c->dxinv[i];
  apx = apx + half_dt * fu[id + 1] * t2 * c->dxinv[i];
  t1 = fxw * (1 - c->ium[id]);
  t2 = (1.0 - fxe) * (1 - c->iup[id]);
  acx = acx + half_dt * (fu[id + 1] * t2 - fu[id] * t1) * c->dxinv[i];
  t1 = 1.0 - fys; t2 = fyn;
  amy = amy - half_dt * fv[id] * t1 * c->dyinv[j];
  apy = apy + half_dt * fv[id + ex] * t2 * c->dyinv[j];
  t1 = fys * (1 - c->jum[id]);
  t2 = (1.0 - fyn) * (1 - c->jup[id]);
  acy = acy + half_dt * (fv[id + ex] * t2 - fv[id] * t1) * c->dyinv[j];
  t1 = 1.0 - fzb; t2 = fzf;
  amz = amz - half_dt * fw[id] * t1 * c->dzinv[k];
  apz = apz + half_dt * fw[id + exy] * t2 * c->dzinv[k];
  t1 = fzb * (1 - c->kum[id]);
  t2 = (1.0 - fzf) * (1 - c->kup[id]);
  acz = acz + half_dt * (fw[id + exy] * t2 - fw[id] * t1) * c->dzinv[k];
  r.amx = amx; r.apx = apx; r.amy = amy; r.apy = apy; r.amz = amz; r.apz = apz;
  r.sum = acx + acy + acz;
  return r;
}

double or_calc_residual_ad(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                           const double *fw, const double *var, const double *r, double *res, int32_t *loc) {
  const long ex = c->ex, exy = c->ex * c->ey;
  double resm = 0.0;
  if (loc) loc[0] = loc[1] = loc[2] = 0;
  if (res) memset(res, 0, sizeof(double) * (size_t)or_field_size(c));
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        const long id = OR_IDX(c, i, j, k);
        const ad_row a = ad_cell(c, half_dt, coef, fu, fv, fw, i, j, k);
        const double kill = 1.0 - 1.0 * (double)c->fresh[id];
        double s = r[id] - var[id] * (1.0 + a.sum * kill) - var[id - 1] * a.amx * (1 - c->ium[id]) -
                   var[id + 1] * a.apx * (1 - c->iup[id]) - var[id - ex] * a.amy * (1 - c->jum[id]) -
                   var[id + ex] * a.apy * (1 - c->jup[id]) - var[id - exy] * a.amz * (1 - c->kum[id]) -
                   var[id + exy] * a.apz * (1 - c->kup[id]);
        if (c->fresh[id] == 1) {
          const int n = fresh_entry(c, i, j, k);
          if (n >= 0) s = fresh_correct(c, n, var, s);
        }
        s = s * (double)(1 - c->iblank[id]);
        if (res) res[id] = s;
        if (fabs(s) > resm) {
          resm = fabs(s);
          if (loc) { loc[0] = i; loc[1] = j; loc[2] = k; }
        }
      }
  return resm;
}

/* line arrays of one tdma call, indexed 0..n+1 as the reference's */
typedef struct ad_line {
  double *a, *b, *cc, *r, *x;
} ad_line;

/* one cell's row of a line solve along direction dir (0 x, 1 y, 2 z) */
static void ad_line_row(const or_ctx *c, int dir, double half_dt, const double *coef, const double *fu,
                        const double *fv, const double *fw, const double *var, const double *rr, int i, int j, int k,
                        ad_line *L, int m) {
  const long ex = c->ex, exy = c->ex * c->ey;
  const long id = OR_IDX(c, i, j, k);
  const ad_row a = ad_cell(c, half_dt, coef, fu, fv, fw, i, j, k);
  const double riblank = 1.0 - (double)c->iblank[id];
  const double kill = 1.0 - 1.0 * (double)c->fresh[id];
  const double tJM = var[id - ex] * a.amy * (1 - c->jum[id]), tJP = var[id + ex] * a.apy * (1 - c->jup[id]);
  const double tKM = var[id - exy] * a.amz * (1 - c->kum[id]), tKP = var[id + exy] * a.apz * (1 - c->kup[id]);
  const double tIM = var[id - 1] * a.amx * (1 - c->ium[id]), tIP = var[id + 1] * a.apx * (1 - c->iup[id]);
  double rhs, am, ap;
  if (dir == 0) {
    rhs = rr[id] - (tJM + tJP + tKM + tKP);
    am = a.amx * riblank * (1 - c->ium[id]);
    ap = a.apx * riblank * (1 - c->iup[id]);
  } else if (dir == 1) {
    rhs = rr[id] - (tIM + tIP + tKM + tKP);
    am = a.amy * riblank * (1 - c->jum[id]);
    ap = a.apy * riblank * (1 - c->jup[id]);
  } else {
    rhs = rr[id] - tJM - tJP - tIM - tIP;
    am = a.amz * riblank * (1.0 - c->kum[id]);
    ap = a.apz * riblank * (1.0 - c->kup[id]);
  }
  L->a[m] = am;
  L->cc[m] = ap;
  L->b[m] = 1.0 + a.sum * riblank * kill;
  L->r[m] = rhs * riblank + var[id] * (double)c->ghost[id];
  if (c->fresh[id] == 1) {
    const int n = fresh_entry(c, i, j, k);
    if (n >= 0) L->r[m] = fresh_correct(c, n, var, L->r[m]);
  }
}

void or_itsolv_ad(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                  const double *fw, double *var, const double *r) {
  int nmax = c->nxc > c->nyc ? c->nxc : c->nyc;
  if (c->nzc > nmax) nmax = c->nzc;
  double *buf = (double *)malloc(sizeof(double) * 5 * (size_t)(nmax + 2));
  ad_line L = {buf, buf + (nmax + 2), buf + 2 * (nmax + 2), buf + 3 * (nmax + 2), buf + 4 * (nmax + 2)};
  /* x lines: k, j outer */
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++) {
      for (int i = 1; i <= c->nxc; i++) ad_line_row(c, 0, half_dt, coef, fu, fv, fw, var, r, i, j, k, &L, i);
      or_tdma(L.a, L.b, L.cc, L.r, L.x, 1, c->nxc);
      for (int i = 1; i <= c->nxc; i++) var[OR_IDX(c, i, j, k)] = L.x[i];
    }
  /* y lines: k, i outer */
  for (int k = 1; k <= c->nzc; k++)
    for (int i = 1; i <= c->nxc; i++) {
      for (int j = 1; j <= c->nyc; j++) ad_line_row(c, 1, half_dt, coef, fu, fv, fw, var, r, i, j, k, &L, j);
      or_tdma(L.a, L.b, L.cc, L.r, L.x, 1, c->nyc);
      for (int j = 1; j <= c->nyc; j++) var[OR_IDX(c, i, j, k)] = L.x[j];
    }
  /* z lines: i, j outer */
  for (int i = 1; i <= c->nxc; i++)
    for (int j = 1; j <= c->nyc; j++) {
      for (int k = 1; k <= c->nzc; k++) ad_line_row(c, 2, half_dt, coef, fu, fv, fw, var, r, i, j, k, &L, k);
      or_tdma(L.a, L.b, L.cc, L.r, L.x, 1, c->nzc);
      for (int k = 1; k <= c->nzc; k++) var[OR_IDX(c, i, j, k)] = L.x[k];
    }
  free(buf);
}

int or_solve_ad(const or_ctx *c, double half_dt, const double *coef, const double *fu, const double *fv,
                const double *fw, double *u, double *v, double *w, const double *nlu, const double *nlv,
                const double *nlw, const double *vbi, int itermax, double restol, double *maxres_out) {
  int iter = 0;
  double maxres = 1.0e10;
  while (iter < itermax && maxres > restol) {
    or_itsolv_ad(c, half_dt, coef, fu, fv, fw, u, nlu);
    or_itsolv_ad(c, half_dt, coef, fu, fv, fw, v, nlv);
    or_itsolv_ad(c, half_dt, coef, fu, fv, fw, w, nlw);
    or_vel_ghosts(c, u, v, w, vbi, itermax, restol);
    const double ru = or_calc_residual_ad(c, half_dt, coef, fu, fv, fw, u, nlu, NULL, NULL);
    const double rv = or_calc_residual_ad(c, half_dt, coef, fu, fv, fw, v, nlv, NULL, NULL);
    const double rw = or_calc_residual_ad(c, half_dt, coef, fu, fv, fw, w, nlw, NULL, NULL);
    maxres = ru;
    if (fabs(rv) > fabs(maxres)) maxres = rv;
    if (fabs(rw) > fabs(maxres)) maxres = rw;
    iter++;
  }
  if (maxres_out) *maxres_out = maxres;
  return iter;
}

/* solve_ad with advec_scheme CRANK_NICOLSON2 (CORE_SOLVE_AD.F90:218-247):
 * the GCM loop of or_solve_ad with face_vel after each ghost update, so the
 * residual and the next sweeps use the face velocities of the new fields
 * (fu, fv, fw in/out). */
int or_solve_ad_cn2(const or_ctx *c, double dt, const double *coef, double *fu, double *fv, double *fw, double *u,
                    double *v, double *w, const double *nlu, const double *nlv, const double *nlw, const double *vbi,
                    int van_kan, const double *p, const double *bcxu, const double *bcyv, const double *bczw,
                    int itermax, double restol, double *maxres_out) {
  const double half_dt = 0.5 * dt;
  int iter = 0;
  double maxres = 1.0e10;
  while (iter < itermax && maxres > restol) {
    or_itsolv_ad(c, half_dt, coef, fu, fv, fw, u, nlu);
    or_itsolv_ad(c, half_dt, coef, fu, fv, fw, v, nlv);
    or_itsolv_ad(c, half_dt, coef, fu, fv, fw, w, nlw);
    or_vel_ghosts(c, u, v, w, vbi, itermax, restol);
    or_face_vel(c, van_kan, dt, p, u, v, w, bcxu, bcyv, bczw, fu, fv, fw);
    const double ru = or_calc_residual_ad(c, half_dt, coef, fu, fv, fw, u, nlu, NULL, NULL);
    const double rv = or_calc_residual_ad(c, half_dt, coef, fu, fv, fw, v, nlv, NULL, NULL);
    const double rw = or_calc_residual_ad(c, half_dt, coef, fu, fv, fw, w, nlw, NULL, NULL);
    maxres = ru;
    if (fabs(rv) > fabs(maxres)) maxres = rv;
    if (fabs(rw) > fabs(maxres)) maxres = rw;
    iter++;
  }
  if (maxres_out) *maxres_out = maxres;
  return iter;
}
