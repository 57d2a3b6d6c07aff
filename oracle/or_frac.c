/*
 * or_frac.c -- the fractional-step routines either side of the pressure
 * solve: correct_vel (CORE_FRAC_STEP_ROUTINES.F90:244-452) with the GCM
 * velocity ghost update (GCM_vel_set_bc_internal,
 * GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:126-247) and the solid-face reset
 * (GCM_set_face_vel_body, :387-423), face_vel (CORE_FRAC_STEP_ROUTINES.F90:
 * 6-240) and update_pressure (:949-967).  Test infrastructure only (see
 * oracle.h).  Restated for the configurations of the path: no periodic
 * outer boundaries, no cut-cell (cure_pressure_oscillations) branch.
 *
 * Pins (tests/test_oracle_frac.py):
 *   correct_vel    a pressure field linear in x, y, z on a stretched grid:
 *                  the cell correction is -dt times the exact gradient
 *                  (the fx interpolation is exact for linear fields), and
 *                  every interior face correction is -dt times the exact
 *                  face-normal derivative; outer faces take the bc arrays.
 *   correct_vel    after rhs_poisson + a converged solve, the discrete
 *                  divergence of the corrected face velocities equals
 *                  dt times the solve's residual on every fluid cell
 *                  (the projection property of the fractional step).
 *   solid faces    faces of blank non-ghost cells are zero, brute force.
 *   vel ghosts     the converged ghost values satisfy
 *                  u_g = bw u_BI - iw (sum_r cD_r u_r) with the self
 *                  corner reading u_BI, to round-off.
 *   coeffGCMD      reproduces every trilinear field at the image point
 *                  (closed form), with the self corner at the intercept.
 *   update_pressure  p + p' (or p') on every cell, then the ghost rows.
 *   face_vel       u linear on a stretched grid: every interior face holds
 *                  the exact face value; VAN_KAN with p linear too gives the
 *                  same faces away from the outer layer (uTilde's gradient
 *                  and the face correction cancel; a flipped or dropped
 *                  pressure term leaves dt times the slope); with p
 *                  quadratic the ghost-cell faces keep the exact u value
 *                  while ordinary faces do not; solid faces are zero;
 *                  outer faces take the bc arrays.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "or_internal.h"

static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

void or_set_vel_ghosts(or_ctx *c, const double *coeffD, double bi_weight, double ip_weight) {
  memcpy(c->coeffD, coeffD, sizeof(double) * 8 * (size_t)(c->nghost > 0 ? c->nghost : 0));
  c->bi_weight = bi_weight;
  c->ip_weight = ip_weight;
}

/* GCM_vel_set_bc_internal: Gauss-Seidel over the ghost list in list order
 * until the summed change of (u, v, w) is at most restol.  vbi holds the
 * body-intercept velocity (u, v, w) per ghost, NULL for resting bodies. */
int or_vel_ghosts(const or_ctx *c, double *u, double *v, double *w, const double *vbi, int itermax,
                  double restol) {
  int it = 0;
  double rmax = 1.0e10;
  const double bw = c->bi_weight, iw = c->ip_weight;
  while (it < itermax && rmax > restol) {
    rmax = 0.0;
    for (int n = 0; n < c->nghost; n++) {
      int iG = c->ig[n], jG = c->jg[n], kG = c->kg[n];
      double ub = vbi ? vbi[3 * n] : 0.0, vb = vbi ? vbi[3 * n + 1] : 0.0, wb = vbi ? vbi[3 * n + 2] : 0.0;
      double uip = 0.0, vip = 0.0, wip = 0.0;
      for (int r = 0; r < 8; r++) {
        int ii = c->ic[n] + incI[r], jj = c->jc[n] + incJ[r], kk = c->kc[n] + incK[r];
        double cf = c->coeffD[8 * n + r];
        if (ii != iG || jj != jG || kk != kG) {
          long q = OR_IDX(c, ii, jj, kk);
          uip += cf * u[q];
          vip += cf * v[q];
          wip += cf * w[q];
        } else {
          uip += cf * ub;
          vip += cf * vb;
          wip += cf * wb;
        }
      }
      long g = OR_IDX(c, iG, jG, kG);
      double up = u[g], vp = v[g], wp = w[g];
      u[g] = ub * bw - uip * iw;
      v[g] = vb * bw - vip * iw;
      w[g] = wb * bw - wip * iw;
      double d = fabs(u[g] - up) + fabs(v[g] - vp) + fabs(w[g] - wp);
      if (d > rmax) rmax = d;
    }
    it++;
  }
  return it;
}

/* fx(i) = (x(i) - xc(i-1)) / (xc(i) - xc(i-1)) (CORE_GRID.F90:172-174) */
static double fint(const double *x, const double *xc, int i) { return (x[i] - xc[i - 1]) / (xc[i] - xc[i - 1]); }

/* face_u,v,w -= dt times the face-normal pressure gradient, then the outer
 * faces from bcxu/bcyv/bczw, in the reference's loop and overwrite order
 * (correct_vel :378-444; face_vel :94-159 has the same loops) */
static void face_pressure_correct(const or_ctx *c, double dt, const double *p, double *fu, double *fv, double *fw,
                                  const double *bcxu, const double *bcyv, const double *bczw) {
  const long ex = c->ex, exy = c->ex * c->ey;
  int nmax = c->nx > c->ny ? c->nx : c->ny;
  if (c->nz > nmax) nmax = c->nz;
  double *f1 = (double *)malloc(sizeof(double) * 2 * (size_t)(nmax + 2));
  double *f2 = f1 + nmax + 2;
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++) {
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double gw = (p[id] - p[id - 1]) * c->dxcinv[i];
        double ge = (p[id + 1] - p[id]) * c->dxcinv[i + 1];
        f1[i] = fu[id] - dt * gw * (1 - c->ium[id]);
        f2[i] = fu[id + 1] - dt * ge * (1 - c->iup[id]);
      }
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        fu[id] = f1[i];
        fu[id + 1] = f2[i];
      }
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        fu[id] = fu[id] * (1 - c->ium[id]) + bcxu[id] * c->ium[id];
        fu[id + 1] = fu[id + 1] * (1 - c->iup[id]) + bcxu[id] * c->iup[id];
      }
    }
  for (int k = 1; k <= c->nzc; k++)
    for (int i = 1; i <= c->nxc; i++) {
      for (int j = 1; j <= c->nyc; j++) {
        long id = OR_IDX(c, i, j, k);
        double gs = (p[id] - p[id - ex]) * c->dycinv[j];
        double gn = (p[id + ex] - p[id]) * c->dycinv[j + 1];
        f1[j] = fv[id] - dt * gs * (1 - c->jum[id]);
        f2[j] = fv[id + ex] - dt * gn * (1 - c->jup[id]);
      }
      for (int j = 1; j <= c->nyc; j++) {
        long id = OR_IDX(c, i, j, k);
        fv[id] = f1[j];
        fv[id + ex] = f2[j];
      }
      for (int j = 1; j <= c->nyc; j++) {
        long id = OR_IDX(c, i, j, k);
        fv[id] = fv[id] * (1 - c->jum[id]) + bcyv[id] * c->jum[id];
        fv[id + ex] = fv[id + ex] * (1 - c->jup[id]) + bcyv[id] * c->jup[id];
      }
    }
  for (int j = 1; j <= c->nyc; j++)
    for (int i = 1; i <= c->nxc; i++) {
      for (int k = 1; k <= c->nzc; k++) {
        long id = OR_IDX(c, i, j, k);
        double gb = (p[id] - p[id - exy]) * c->dzcinv[k];
        double gf = (p[id + exy] - p[id]) * c->dzcinv[k + 1];
        f1[k] = fw[id] - dt * gb * (1 - c->kum[id]);
        f2[k] = fw[id + exy] - dt * gf * (1 - c->kup[id]);
      }
      for (int k = 1; k <= c->nzc; k++) {
        long id = OR_IDX(c, i, j, k);
        fw[id] = f1[k];
        fw[id + exy] = f2[k];
      }
      for (int k = 1; k <= c->nzc; k++) {
        long id = OR_IDX(c, i, j, k);
        fw[id] = fw[id] * (1 - c->kum[id]) + bczw[id] * c->kum[id];
        fw[id + exy] = fw[id + exy] * (1 - c->kup[id]) + bczw[id] * c->kup[id];
      }
    }
  free(f1);
}

/* GCM_set_face_vel_body (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:387-423):
 * the faces of blank non-ghost cells are zero */
static void set_face_vel_body(const or_ctx *c, double *fu, double *fv, double *fw) {
  const long ex = c->ex, exy = c->ex * c->ey;
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (c->iblank[id] == 1 && c->ghost[id] == 0) {
          fu[id + 1] = 0.0; fu[id] = 0.0;
          fv[id + ex] = 0.0; fv[id] = 0.0;
          fw[id + exy] = 0.0; fw[id] = 0.0;
        }
      }
}

void or_correct_vel(const or_ctx *c, double dt, const double *p, double *u, double *v, double *w, double *fu,
                    double *fv, double *fw, const double *bcxu, const double *bcyv, const double *bczw,
                    const double *vbi, int itermax_ad, double restol_ad) {
  const long ex = c->ex, exy = c->ex * c->ey;
  /* cell velocities (:270-367) */
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double fxe = fint(c->x, c->xc, i + 1), fxw = fint(c->x, c->xc, i);
        double fyn = fint(c->y, c->yc, j + 1), fys = fint(c->y, c->yc, j);
        double fzf = fint(c->z, c->zc, k + 1), fzb = fint(c->z, c->zc, k);
        double pe = (fxe * p[id + 1] + (1.0 - fxe) * p[id]) * (1 - c->iup[id]) + p[id] * c->iup[id];
        double pw = (fxw * p[id] + (1.0 - fxw) * p[id - 1]) * (1 - c->ium[id]) + p[id] * c->ium[id];
        double pn = (fyn * p[id + ex] + (1.0 - fyn) * p[id]) * (1 - c->jup[id]) + p[id] * c->jup[id];
        double ps = (fys * p[id] + (1.0 - fys) * p[id - ex]) * (1 - c->jum[id]) + p[id] * c->jum[id];
        double pf = (fzf * p[id + exy] + (1.0 - fzf) * p[id]) * (1 - c->kup[id]) + p[id] * c->kup[id];
        double pb = (fzb * p[id] + (1.0 - fzb) * p[id - exy]) * (1 - c->kum[id]) + p[id] * c->kum[id];
        double pgx = (pe - pw) * c->dxinv[i];
        double pgy = (pn - ps) * c->dyinv[j];
        double pgz = (pf - pb) * c->dzinv[k];
        double fl = (double)(1 - c->iblank[id]);
        u[id] = u[id] - dt * pgx * fl;
        v[id] = v[id] - dt * pgy * fl;
        w[id] = w[id] - dt * pgz * fl;
      }
  /* ghost velocities (:375) */
  or_vel_ghosts(c, u, v, w, vbi, itermax_ad, restol_ad);
  /* face velocities, in the reference's loop and overwrite order (:378-444) */
  face_pressure_correct(c, dt, p, fu, fv, fw, bcxu, bcyv, bczw);
  set_face_vel_body(c, fu, fv, fw);
}

/* face_vel (CORE_FRAC_STEP_ROUTINES.F90:6-240), GCM, non-periodic, in the
 * reference's loop and overwrite order: uTilde = u + dt grad p (1 - iblank)
 * for VAN_KAN (:36-70; the cell gradient from fx-interpolated face
 * pressures, the cell's own value on an outer face), else u (:72-76); face
 * values by the fx interpolation of uTilde, outer faces from the bc arrays
 * (:80-107); for VAN_KAN the face pressure correction (:111-175, the loops
 * of correct_vel); the outer faces again (:179-197); for VAN_KAN the faces
 * of ghost cells from the interpolation of u itself (:201-234); the faces of
 * blank non-ghost cells zero (:238). */
void or_face_vel(const or_ctx *c, int van_kan, double dt, const double *p, const double *u, const double *v,
                 const double *w, const double *bcxu, const double *bcyv, const double *bczw, double *fu, double *fv,
                 double *fw) {
  const long nf = or_field_size(c), ex = c->ex, exy = c->ex * c->ey;
  double *ut = (double *)malloc(sizeof(double) * 3 * (size_t)nf);
  double *vt = ut + nf, *wt = ut + 2 * nf;
  memcpy(ut, u, sizeof(double) * nf);
  memcpy(vt, v, sizeof(double) * nf);
  memcpy(wt, w, sizeof(double) * nf);
  if (van_kan)
    for (int k = 1; k <= c->nzc; k++)
      for (int j = 1; j <= c->nyc; j++)
        for (int i = 1; i <= c->nxc; i++) {
          long id = OR_IDX(c, i, j, k);
          double fxe = fint(c->x, c->xc, i + 1), fxw = fint(c->x, c->xc, i);
          double fyn = fint(c->y, c->yc, j + 1), fys = fint(c->y, c->yc, j);
          double fzf = fint(c->z, c->zc, k + 1), fzb = fint(c->z, c->zc, k);
          double pe = (fxe * p[id + 1] + (1.0 - fxe) * p[id]) * (1 - c->iup[id]) + p[id] * c->iup[id];
          double pw = (fxw * p[id] + (1.0 - fxw) * p[id - 1]) * (1 - c->ium[id]) + p[id] * c->ium[id];
          double pn = (fyn * p[id + ex] + (1.0 - fyn) * p[id]) * (1 - c->jup[id]) + p[id] * c->jup[id];
          double ps = (fys * p[id] + (1.0 - fys) * p[id - ex]) * (1 - c->jum[id]) + p[id] * c->jum[id];
          double pf = (fzf * p[id + exy] + (1.0 - fzf) * p[id]) * (1 - c->kup[id]) + p[id] * c->kup[id];
          double pb = (fzb * p[id] + (1.0 - fzb) * p[id - exy]) * (1 - c->kum[id]) + p[id] * c->kum[id];
          double pgx = (pe - pw) * c->dxinv[i];
          double pgy = (pn - ps) * c->dyinv[j];
          double pgz = (pf - pb) * c->dzinv[k];
          double fl = (double)(1 - c->iblank[id]);
          ut[id] = u[id] + dt * pgx * fl;
          vt[id] = v[id] + dt * pgy * fl;
          wt[id] = w[id] + dt * pgz * fl;
        }
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        double fxe = fint(c->x, c->xc, i + 1), fxw = fint(c->x, c->xc, i);
        double fyn = fint(c->y, c->yc, j + 1), fys = fint(c->y, c->yc, j);
        double fzf = fint(c->z, c->zc, k + 1), fzb = fint(c->z, c->zc, k);
        fu[id + 1] = (fxe * ut[id + 1] + (1.0 - fxe) * ut[id]) * (1 - c->iup[id]) + bcxu[id] * c->iup[id];
        fu[id] = (fxw * ut[id] + (1.0 - fxw) * ut[id - 1]) * (1 - c->ium[id]) + bcxu[id] * c->ium[id];
        fv[id + ex] = (fyn * vt[id + ex] + (1.0 - fyn) * vt[id]) * (1 - c->jup[id]) + bcyv[id] * c->jup[id];
        fv[id] = (fys * vt[id] + (1.0 - fys) * vt[id - ex]) * (1 - c->jum[id]) + bcyv[id] * c->jum[id];
        fw[id + exy] = (fzf * wt[id + exy] + (1.0 - fzf) * wt[id]) * (1 - c->kup[id]) + bczw[id] * c->kup[id];
        fw[id] = (fzb * wt[id] + (1.0 - fzb) * wt[id - exy]) * (1 - c->kum[id]) + bczw[id] * c->kum[id];
      }
  free(ut);
  if (van_kan) face_pressure_correct(c, dt, p, fu, fv, fw, bcxu, bcyv, bczw);
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        fu[id] = fu[id] * (1 - c->ium[id]) + bcxu[id] * c->ium[id];
        fu[id + 1] = fu[id + 1] * (1 - c->iup[id]) + bcxu[id] * c->iup[id];
        fv[id] = fv[id] * (1 - c->jum[id]) + bcyv[id] * c->jum[id];
        fv[id + ex] = fv[id + ex] * (1 - c->jup[id]) + bcyv[id] * c->jup[id];
        fw[id] = fw[id] * (1 - c->kum[id]) + bczw[id] * c->kum[id];
        fw[id + exy] = fw[id + exy] * (1 - c->kup[id]) + bczw[id] * c->kup[id];
      }
  if (van_kan)
    for (int k = 1; k <= c->nzc; k++)
      for (int j = 1; j <= c->nyc; j++)
        for (int i = 1; i <= c->nxc; i++) {
          long id = OR_IDX(c, i, j, k);
          double gy = (double)c->ghost[id], gn = 1.0 - gy;
          double fxe = fint(c->x, c->xc, i + 1), fxw = fint(c->x, c->xc, i);
          double fyn = fint(c->y, c->yc, j + 1), fys = fint(c->y, c->yc, j);
          double fzf = fint(c->z, c->zc, k + 1), fzb = fint(c->z, c->zc, k);
          fu[id + 1] = gn * fu[id + 1] + gy * (fxe * u[id + 1] + (1.0 - fxe) * u[id]);
          fu[id] = gn * fu[id] + gy * (fxw * u[id] + (1.0 - fxw) * u[id - 1]);
          fv[id + ex] = gn * fv[id + ex] + gy * (fyn * v[id + ex] + (1.0 - fyn) * v[id]);
          fv[id] = gn * fv[id] + gy * (fys * v[id] + (1.0 - fys) * v[id - ex]);
          fw[id + exy] = gn * fw[id + exy] + gy * (fzf * w[id + exy] + (1.0 - fzf) * w[id]);
          fw[id] = gn * fw[id] + gy * (fzb * w[id] + (1.0 - fzb) * w[id - exy]);
        }
  set_face_vel_body(c, fu, fv, fw);
}

void or_update_pressure(const or_ctx *c, int van_kan, const double *pprime, double *p, const double *g, int itermax,
                        double restol) {
  long nf = or_field_size(c);
  if (van_kan)
    for (long q = 0; q < nf; q++) p[q] = p[q] + pprime[q];
  else
    memcpy(p, pprime, sizeof(double) * nf);
  or_ghost_rows(c, p, g, itermax, restol);
}
