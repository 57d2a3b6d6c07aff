/*
 * or_mg.c -- the reference geometric multigrid (CORE_MG.F90) and the outer
 * solve_poisson loop (CORE_POISSON.F90:493-676), restated in C.
 * TEST INFRASTRUCTURE ONLY (see oracle.h); also the timed CPU baseline.
 *
 * Scope of the restatement (the configuration every benchmark case uses):
 *   pp_solver_type = MG (3), iRedBlack = 0 (lexicographic lines),
 *   mgcycle = 1 (V), Full_Coarsening = F, cure_pressure_oscillations = F,
 *   Conformal_Mapping = F, all outer pressure faces PBC_NEUMANN, GCM bodies,
 *   nlw = nlw0 = 0 (no ZERO_GRADIENT faces, so pgrad = 0).
 * Reference quirks kept: semi-coarsened x, y, z cycles in sequence, no
 * finest-level post-smoothing in the x and y cycles, 1-D masked volume
 * averaging restriction, inverse-distance-squared prolongation, stair-step
 * Neumann faces and no ghosts on coarse levels, ghost rows applied by
 * GCM_Pressure at every finest-level visit.
 * The z-direction level metrics leave zc_MG(0) and zc_MG(m+1) unset
 * (CORE_MG.F90:254-262); every use of them is multiplied by a zero face flag
 * or never reached, so they are computed here with the x/y formulas.
 *
 * Parity pins (tests/test_oracle.py): the converged solution equals a
 * scipy sparse-direct solution of the same linear system (the fixed point
 * does not depend on the cycle); one MG_Solver call reduces the residual
 * (contraction factor < 1) on config-1 geometry.  Per-iterate agreement with
 * the Fortran is parity unpinned (the Fortran does not build here).
 */
#include "or_internal.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* threads the zebra variant runs on (OpenMP's default team size) */
int or_threads(void) { return omp_get_max_threads(); }

#define MAXLEV 12

typedef struct {
  int nlev;
  int m[MAXLEV + 1];                 /* mgrid(n), n = 1..nlev */
  double *x[MAXLEV + 1], *xc[MAXLEV + 1], *dx[MAXLEV + 1], *dxinv[MAXLEV + 1];
  double *dxc[MAXLEV + 1], *dxcinv[MAXLEV + 1];
} mgdir;

typedef struct {
  int cx, cy, cz;                    /* cell counts of the level box */
  long ex, ey;                       /* extents cx+2, cy+2 */
} box;

typedef struct {
  or_ctx *c;
  const or_mg_opts *o;
  mgdir d[3];
} mgctx;

static long bidx(const box *b, int i, int j, int k) {
  return (long)i + b->ex * ((long)j + b->ey * (long)k);
}
static long bsize(const box *b) { return b->ex * b->ey * (long)(b->cz + 2); }

/* MG_initial (CORE_MG.F90:31-278) for one direction */
static void mg_init_dir(mgdir *d, int nc, const double *x, const double *xc,
                        const double *dx, const double *dxinv,
                        const double *dxc, const double *dxcinv) {
  int ii = nc, nlevel;
  for (nlevel = 1; nlevel <= 10; nlevel++) {
    ii = ii / 2;
    if (ii >= 4) continue;
    break;
  }
  d->nlev = nlevel > 10 ? 11 : nlevel;
  if (d->nlev > MAXLEV) d->nlev = MAXLEV;
  d->m[1] = nc;
  for (int n = 2; n <= d->nlev; n++)
    d->m[n] = (d->m[n - 1] % 2 == 0) ? d->m[n - 1] / 2 : (d->m[n - 1] - 1) / 2;
  int sz = nc + 4;
  for (int n = 1; n <= d->nlev; n++) {
    d->x[n] = calloc(sz, 8); d->xc[n] = calloc(sz, 8); d->dx[n] = calloc(sz, 8);
    d->dxinv[n] = calloc(sz, 8); d->dxc[n] = calloc(sz, 8); d->dxcinv[n] = calloc(sz, 8);
  }
  int nn = nc + 1;
  for (int i = 0; i <= nn; i++) {
    d->dxc[1][i] = dxc[i]; d->dx[1][i] = dx[i];
    d->dxcinv[1][i] = dxcinv[i]; d->dxinv[1][i] = dxinv[i]; d->xc[1][i] = xc[i];
  }
  for (int i = 0; i <= nn + 1; i++) d->x[1][i] = x[i];
  for (int n = 2; n <= d->nlev; n++) {
    int m = d->m[n];
    double *X = d->x[n], *Xp = d->x[n - 1];
    for (int i = 1; i <= m + 1; i++) X[i] = Xp[2 * i - 1];
    X[0] = -X[2];
    X[m + 2] = 2.0 * X[m + 1] - X[m];
    for (int i = 0; i <= m + 1; i++) {
      d->dx[n][i] = X[i + 1] - X[i];
      d->dxinv[n][i] = 1.0 / d->dx[n][i];
    }
    for (int i = 0; i <= m + 1; i++) d->xc[n][i] = 0.5 * (X[i] + X[i + 1]);
    for (int i = 1; i <= m + 1; i++) {
      d->dxc[n][i] = d->xc[n][i] - d->xc[n][i - 1];
      d->dxcinv[n][i] = 1.0 / d->dxc[n][i];
    }
  }
}

static void mg_free_dir(mgdir *d) {
  for (int n = 1; n <= d->nlev; n++) {
    free(d->x[n]); free(d->xc[n]); free(d->dx[n]);
    free(d->dxinv[n]); free(d->dxc[n]); free(d->dxcinv[n]);
  }
}

/* Per-level state of one directional cycle. */
typedef struct {
  box b;
  int8_t *ib;                              /* iblank_MG of the level */
  int8_t *gm;                              /* ghostcellMark_MG (level 1) */
  int8_t *f[6];                            /* ium,iup,jum,jup,kum,kup */
  int lv[3];                               /* nLevX, nLevY, nLevZ */
} level;

/* MG_Prepare (CORE_MG.F90:1716-1879): stair-step Neumann faces from the
 * level iblank, then outer faces Neumann (non-periodic, PBC_NEUMANN). */
static void mg_prepare(level *L, int par) {
  box *b = &L->b;
  int mx = b->cx, my = b->cy, mz = b->cz;
  long n = bsize(b);
  for (int a = 0; a < 6; a++) memset(L->f[a], 0, n);
  int8_t *ib = L->ib;
#pragma omp parallel for if (par) schedule(static)
  for (int k = 1; k <= mz; k++) {
    for (int j = 1; j <= my; j++) {
      for (int i = 1; i <= mx - 1; i++) {
        double t = (double)(ib[bidx(b, i + 1, j, k)] - ib[bidx(b, i, j, k)]) / 2.0;
        L->f[1][bidx(b, i, j, k)] = (int8_t)(int)(t + 0.5);
      }
      for (int i = 2; i <= mx; i++) {
        double t = (double)(ib[bidx(b, i - 1, j, k)] - ib[bidx(b, i, j, k)]) / 2.0;
        L->f[0][bidx(b, i, j, k)] = (int8_t)(int)(t + 0.5);
      }
    }
    for (int j = 1; j <= my - 1; j++)
      for (int i = 1; i <= mx; i++) {
        double t = (double)(ib[bidx(b, i, j + 1, k)] - ib[bidx(b, i, j, k)]) / 2.0;
        L->f[3][bidx(b, i, j, k)] = (int8_t)(int)(t + 0.5);
      }
    for (int j = 2; j <= my; j++)
      for (int i = 1; i <= mx; i++) {
        double t = (double)(ib[bidx(b, i, j - 1, k)] - ib[bidx(b, i, j, k)]) / 2.0;
        L->f[2][bidx(b, i, j, k)] = (int8_t)(int)(t + 0.5);
      }
  }
#pragma omp parallel for if (par) schedule(static)
  for (int j = 1; j <= my; j++)
    for (int i = 1; i <= mx; i++) {
      for (int k = 1; k <= mz - 1; k++) {
        double t = (double)(ib[bidx(b, i, j, k + 1)] - ib[bidx(b, i, j, k)]) / 2.0;
        L->f[5][bidx(b, i, j, k)] = (int8_t)(int)(t + 0.5);
      }
      for (int k = 2; k <= mz; k++) {
        double t = (double)(ib[bidx(b, i, j, k - 1)] - ib[bidx(b, i, j, k)]) / 2.0;
        L->f[4][bidx(b, i, j, k)] = (int8_t)(int)(t + 0.5);
      }
    }
  for (int k = 1; k <= mz; k++)
    for (int j = 1; j <= my; j++) {
      L->f[0][bidx(b, 1, j, k)] = 1;
      L->f[1][bidx(b, mx, j, k)] = 1;
    }
  for (int k = 1; k <= mz; k++)
    for (int i = 1; i <= mx; i++) {
      L->f[2][bidx(b, i, 1, k)] = 1;
      L->f[3][bidx(b, i, my, k)] = 1;
    }
  for (int j = 1; j <= my; j++)
    for (int i = 1; i <= mx; i++) {
      L->f[4][bidx(b, i, j, 1)] = 1;
      L->f[5][bidx(b, i, j, mz)] = 1;
    }
}

/* coefficients of cell (i,j,k) at the level (MG_itsolv CORE_MG.F90:1945-1958) */
typedef struct { double amx, apx, acx, amy, apy, acy, amz, apz, acz; } coef7;

static coef7 mg_coef(const mgctx *M, const level *L, int i, int j, int k, long id) {
  const mgdir *X = &M->d[0], *Y = &M->d[1], *Z = &M->d[2];
  int lx = L->lv[0], ly = L->lv[1], lz = L->lv[2];
  coef7 q;
  q.amx = X->dxcinv[lx][i] * X->dxinv[lx][i] * (1.0 - L->f[0][id]);
  q.apx = X->dxcinv[lx][i + 1] * X->dxinv[lx][i] * (1.0 - L->f[1][id]);
  q.acx = -(q.amx + q.apx);
  q.amy = Y->dxcinv[ly][j] * Y->dxinv[ly][j] * (1.0 - L->f[2][id]);
  q.apy = Y->dxcinv[ly][j + 1] * Y->dxinv[ly][j] * (1.0 - L->f[3][id]);
  q.acy = -(q.amy + q.apy);
  q.amz = Z->dxcinv[lz][k] * Z->dxinv[lz][k] * (1.0 - L->f[4][id]);
  q.apz = Z->dxcinv[lz][k + 1] * Z->dxinv[lz][k] * (1.0 - L->f[5][id]);
  q.acz = -(q.amz + q.apz);
  return q;
}

/* One line of MG_itsolv (CORE_MG.F90:1945-2009): rows of the cells along
 * direction dir through (t1, t2) (the two transverse indices in the order
 * x-lines (j,k), y-lines (i,k), z-lines (i,j)), tdma, SOR blend.  buf holds
 * 5 * (nmax + 1) doubles. */
static void mg_line(const mgctx *M, const level *L, double *var, const double *R, int dir, int t1, int t2,
                    double *buf, int nmax) {
  const box *b = &L->b;
  double *a = buf, *bb = buf + (nmax + 1), *cc = buf + 2 * (nmax + 1);
  double *rhs = buf + 3 * (nmax + 1), *sol = buf + 4 * (nmax + 1);
  const double om = M->o->omega;
  const long sy = b->ex, sz = b->ex * b->ey;
  const int n = dir == 0 ? b->cx : dir == 1 ? b->cy : b->cz;
  for (int m = 1; m <= n; m++) {
    int i = dir == 0 ? m : t1, j = dir == 1 ? m : (dir == 0 ? t1 : t2), k = dir == 2 ? m : t2;
    long id = bidx(b, i, j, k);
    coef7 q = mg_coef(M, L, i, j, k, id);
    double r, am, ap;
    if (dir == 0) {
      r = R[id] - var[id - sy] * q.amy - var[id + sy] * q.apy - var[id - sz] * q.amz - var[id + sz] * q.apz;
      am = q.amx; ap = q.apx;
    } else if (dir == 1) {
      r = R[id] - var[id - 1] * q.amx - var[id + 1] * q.apx - var[id - sz] * q.amz - var[id + sz] * q.apz;
      am = q.amy; ap = q.apy;
    } else {
      r = R[id] - var[id - sy] * q.amy - var[id + sy] * q.apy - var[id - 1] * q.amx - var[id + 1] * q.apx;
      am = q.amz; ap = q.apz;
    }
    double fl = 1.0 - (double)L->ib[id];
    a[m] = am * fl;
    cc[m] = ap * fl;
    bb[m] = (q.acx + q.acy + q.acz) * fl + (double)L->ib[id];
    rhs[m] = r * fl + (double)(L->gm ? L->gm[id] : 0) * var[id];
    if (fabs(bb[m]) < 1.0e-12) { a[m] = 0; cc[m] = 0; bb[m] = 1.0; rhs[m] = var[id]; }
  }
  or_tdma(a, bb, cc, rhs, sol, 1, n);
  for (int m = 1; m <= n; m++) {
    int i = dir == 0 ? m : t1, j = dir == 1 ? m : (dir == 0 ? t1 : t2), k = dir == 2 ? m : t2;
    long id = bidx(b, i, j, k);
    var[id] = var[id] * (1.0 - om) + om * sol[m];
  }
}

/* MG_itsolv (CORE_MG.F90:1883-2270).  zebra = 0: iRedBlack = 0, the lines
 * of each direction in the reference's loop order (Gauss-Seidel across
 * lines).  zebra = 1: the lines of each direction in two colours by the
 * parity of their transverse indices, every line of a colour solved in
 * parallel (OpenMP): the same line rows, another order, the same fixed
 * point (the all-core baseline variant). */
static void mg_itsolv(const mgctx *M, const level *L, double *var, const double *R) {
  const box *b = &L->b;
  int nmax = b->cx > b->cy ? b->cx : b->cy;
  if (b->cz > nmax) nmax = b->cz;
  const int T1[3] = {b->cy, b->cx, b->cx}, T2[3] = {b->cz, b->cz, b->cy};
  if (!M->o->zebra) {
    double *buf = calloc(5 * (size_t)(nmax + 1), 8);
    for (int t2 = 1; t2 <= b->cz; t2++)       /* x-lines: k outer, j inner */
      for (int t1 = 1; t1 <= b->cy; t1++) mg_line(M, L, var, R, 0, t1, t2, buf, nmax);
    for (int t2 = 1; t2 <= b->cz; t2++)       /* y-lines: k outer, i inner */
      for (int t1 = 1; t1 <= b->cx; t1++) mg_line(M, L, var, R, 1, t1, t2, buf, nmax);
    for (int t2 = 1; t2 <= b->cy; t2++)       /* z-lines: j outer, i inner */
      for (int t1 = 1; t1 <= b->cx; t1++) mg_line(M, L, var, R, 2, t1, t2, buf, nmax);
    free(buf);
    return;
  }
  for (int dir = 0; dir < 3; dir++)
    for (int col = 0; col < 2; col++) {
      const int n1 = T1[dir], n2 = T2[dir];
#pragma omp parallel
      {
        double *buf = calloc(5 * (size_t)(nmax + 1), 8);
#pragma omp for schedule(static)
        for (long q = 0; q < (long)n1 * n2; q++) {
          int t1 = 1 + (int)(q % n1), t2 = 1 + (int)(q / n1);
          if (((t1 + t2) & 1) == col) mg_line(M, L, var, R, dir, t1, t2, buf, nmax);
        }
        free(buf);
      }
    }
}

/* MG_residual (CORE_MG.F90:2274-2387) */
static void mg_residual(const mgctx *M, const level *L, const double *var,
                        const double *R, double *res) {
  const box *b = &L->b;
  long sy = b->ex, sz = b->ex * b->ey;
#pragma omp parallel for if (M->o->zebra) schedule(static)
  for (int k = 1; k <= b->cz; k++)
    for (int j = 1; j <= b->cy; j++)
      for (int i = 1; i <= b->cx; i++) {
        long id = bidx(b, i, j, k);
        coef7 q = mg_coef(M, L, i, j, k, id);
        double fl = 1.0 - (double)L->ib[id];
        double bc = (q.acx + q.acy + q.acz) * fl + (double)L->ib[id];
        double v = R[id] - var[id] * bc - var[id - 1] * q.amx * fl -
                   var[id + 1] * q.apx * fl - var[id - sy] * q.amy * fl -
                   var[id + sy] * q.apy * fl - var[id - sz] * q.amz * fl -
                   var[id + sz] * q.apz * fl;
        res[id] = v * fl;
      }
}

/* Level-1 ghost handling: GCM_Pressure (CORE_MG.F90:4707-4743) =
 * set_outer_ghost_pres + GCM_p_set_bc_internal on the (0:nx,0:ny,0:nz)
 * working array of MG_Solver. */
static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

static void gcm_pressure(const mgctx *M, const box *b, double *var) {
  const or_ctx *c = M->c;
  int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++) {
      var[bidx(b, 0, j, k)] = var[bidx(b, 1, j, k)];
      var[bidx(b, nx, j, k)] = var[bidx(b, nx - 1, j, k)];
    }
  for (int k = 0; k <= nz; k++)
    for (int i = 0; i <= nx; i++) {
      var[bidx(b, i, 0, k)] = var[bidx(b, i, 1, k)];
      var[bidx(b, i, ny, k)] = var[bidx(b, i, ny - 1, k)];
    }
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      var[bidx(b, i, j, 0)] = var[bidx(b, i, j, 1)];
      var[bidx(b, i, j, nz)] = var[bidx(b, i, j, nz - 1)];
    }
  int it = 0;
  double rmax = 1.0e10;
  while (it < M->o->itermax_ad && rmax > M->o->restol_ad) {
    rmax = 0.0;
    for (int n = 0; n < c->nghost; n++) {
      int iG = c->ig[n], jG = c->jg[n], kG = c->kg[n];
      double pip = 0.0;
      for (int r = 0; r < 8; r++) {
        int ii = c->ic[n] + incI[r], jj = c->jc[n] + incJ[r], kk = c->kc[n] + incK[r];
        if (ii == iG && jj == jG && kk == kG) continue; /* coeff * uNormal, uNormal = 0 */
        if (ii > nx) ii = nx;
        if (jj > ny) jj = ny;
        if (kk > nz) kk = nz;
        pip += c->coeff[8 * n + r] * var[bidx(b, ii, jj, kk)];
      }
      long gid = bidx(b, iG, jG, kG);
      double prev = var[gid];
      var[gid] = pip;
      double d = fabs(pip - prev);
      if (d > rmax) rmax = d;
    }
    it++;
  }
}

/* restriction in direction dir (MG_Restrict_X/Y/Z, CORE_MG.F90:2464-2723) */
static void mg_restrict(const mgctx *M, int dir, int n, const level *Lf,
                        const level *Lc, const double *var, double *r1) {
  const box *bf = &Lf->b, *bc = &Lc->b;
  const double *X = M->d[dir].x[n];
#pragma omp parallel for if (M->o->zebra) schedule(static)
  for (int k = 1; k <= bc->cz; k++)
    for (int j = 1; j <= bc->cy; j++)
      for (int i = 1; i <= bc->cx; i++) {
        int c = dir == 0 ? i : dir == 1 ? j : k;
        int ii = 2 * c, iim = 2 * c - 1, iip = 2 * c + 1;
        double area1 = X[ii] - X[iim], area2 = X[iip] - X[ii];
        double area = area1 + area2;
        double w1 = area1 / area, w2 = area2 / area;
        long f2, f1;
        if (dir == 0) { f2 = bidx(bf, ii, j, k); f1 = bidx(bf, iim, j, k); }
        else if (dir == 1) { f2 = bidx(bf, i, ii, k); f1 = bidx(bf, i, iim, k); }
        else { f2 = bidx(bf, i, j, ii); f1 = bidx(bf, i, j, iim); }
        double d2 = 1.0 - (double)Lf->ib[f2], d1 = 1.0 - (double)Lf->ib[f1];
        double sum1 = d2 * w2 * var[f2] + d1 * w1 * var[f1];
        double temp = d2 * w2 + d1 * w1;
        r1[bidx(bc, i, j, k)] = temp > 0.0 ? sum1 / temp : sum1;
      }
}

/* prolongation in direction dir (MG_Prolong_X/Y/Z, CORE_MG.F90:2885-3523):
 * inverse-distance-squared weights between coarse centres, masked by the
 * coarse iblank; writes phi on the fine level interior. */
static void mg_prolong(const mgctx *M, int dir, int n, const level *Lf,
                       const level *Lc, double *phi, const double *correc) {
  const box *bf = &Lf->b, *bc = &Lc->b;
  const double *xf = M->d[dir].xc[n], *xcc = M->d[dir].xc[n + 1];
  int mc = dir == 0 ? bc->cx : dir == 1 ? bc->cy : bc->cz; /* mx1 - 1 */
  int mfine = dir == 0 ? bf->cx : dir == 1 ? bf->cy : bf->cz;   /* mx - 1 */
  double *w1 = calloc(mc + 2, 8), *w2 = calloc(mc + 2, 8);
  double *w3 = calloc(mc + 2, 8), *w4 = calloc(mc + 2, 8);
  for (int i = 1; i <= mc; i++) {
    int ii = 2 * i, iim = 2 * i - 1;
    double d1 = 1.0 / pow(xf[iim] - xcc[i], 2), d2 = 1.0 / pow(xf[iim] - xcc[i - 1], 2);
    double de = d1 + d2;
    w1[i] = d1 / de; w2[i] = d2 / de;
    d1 = 1.0 / pow(xf[ii] - xcc[i], 2); d2 = 1.0 / pow(xf[ii] - xcc[i + 1], 2);
    de = d1 + d2;
    w3[i] = d1 / de; w4[i] = d2 / de;
  }
  /* loop over the two transverse directions of the fine/coarse box */
  int t1n = dir == 0 ? bf->cy : bf->cx;
  int t2n = dir == 2 ? bf->cy : bf->cz;
#pragma omp parallel for if (M->o->zebra) schedule(static)
  for (int t2 = 1; t2 <= t2n; t2++)
    for (int t1 = 1; t1 <= t1n; t1++) {
#define FI(a) (dir == 0 ? bidx(bf, (a), t1, t2) : dir == 1 ? bidx(bf, t1, (a), t2) : bidx(bf, t1, t2, (a)))
#define CI(a) (dir == 0 ? bidx(bc, (a), t1, t2) : dir == 1 ? bidx(bc, t1, (a), t2) : bidx(bc, t1, t2, (a)))
#define DB(a) (1.0 - (double)Lc->ib[CI(a)])
      phi[FI(1)] = correc[CI(1)];
      double de1 = w3[1] * DB(1), de2 = w4[1] * DB(2), de = de1 + de2;
      phi[FI(2)] = de1 * correc[CI(1)] + de2 * correc[CI(2)];
      if (de > 0.0) phi[FI(2)] = phi[FI(2)] / de;
      for (int i = 2; i <= mc - 1; i++) {
        int ii = 2 * i, iim = 2 * i - 1;
        double v1 = correc[CI(i)], v2 = correc[CI(i - 1)], v3 = correc[CI(i + 1)];
        de1 = w1[i] * DB(i); de2 = w2[i] * DB(i - 1); de = de1 + de2;
        phi[FI(iim)] = de1 * v1 + de2 * v2;
        if (de > 0.0) phi[FI(iim)] = phi[FI(iim)] / de;
        de1 = w3[i] * DB(i); de2 = w4[i] * DB(i + 1); de = de1 + de2;
        phi[FI(ii)] = de1 * v1 + de2 * v3;
        if (de > 0.0) phi[FI(ii)] = phi[FI(ii)] / de;
      }
      int i = mc, ii = 2 * i, iim = 2 * i - 1;
      de1 = w1[i] * DB(i); de2 = w2[i] * DB(i - 1); de = de1 + de2;
      phi[FI(iim)] = de1 * correc[CI(i)] + de2 * correc[CI(i - 1)];
      if (de > 0.0) phi[FI(iim)] = phi[FI(iim)] / de;
      phi[FI(ii)] = correc[CI(i)];
      if (n == 1 && mfine > 2 * i) phi[FI(mfine)] = correc[CI(i)];
#undef FI
#undef CI
#undef DB
    }
  free(w1); free(w2); free(w3); free(w4);
}

/* Directional recursive cycle MG_X / MG_Y / MG_Z (CORE_MG.F90:1239-1663)
 * as a V-cycle (mgcycle = 1, GMM = 1). */
typedef struct {
  int dir, nlev;
  level L[MAXLEV + 1];
} cycle;

static void cycle_level(mgctx *M, cycle *C, int n, double *var, const double *rhs) {
  level *L = &C->L[n];
  int dir = C->dir;
  int totIter = n == 1 ? M->o->iter_finest : (n == C->nlev ? M->o->iter_coarsest : 1);
  if (n != 1) mg_prepare(L, M->o->zebra);
  if (n == 1) gcm_pressure(M, &L->b, var);
  for (int it = 0; it < totIter; it++) mg_itsolv(M, L, var, rhs);
  if (n == C->nlev) return;
  long sz = bsize(&L->b);
  double *r = calloc(sz, 8);
  mg_residual(M, L, var, rhs, r);
  level *Lc = &C->L[n + 1];
  long szc = bsize(&Lc->b);
  double *r1 = calloc(szc, 8), *phi = calloc(szc, 8);
  mg_restrict(M, dir, n, L, Lc, r, r1);
  cycle_level(M, C, n + 1, phi, r1);
  mg_prolong(M, dir, n, L, Lc, r, phi);
  const box *b = &L->b;
#pragma omp parallel for if (M->o->zebra) schedule(static)
  for (int k = 1; k <= b->cz; k++)
    for (int j = 1; j <= b->cy; j++)
      for (int i = 1; i <= b->cx; i++) {
        long id = bidx(b, i, j, k);
        var[id] = var[id] + r[id] * (1.0 - (double)L->ib[id]);
      }
  if (n != 1) mg_prepare(L, M->o->zebra);
  int post;
  if (dir == 2) {
    post = n == 1 ? M->o->iter_finest : 1;
    if (n == 1) gcm_pressure(M, &L->b, var);
  } else {
    post = n == 1 ? 0 : 1;
    if (n == 1) gcm_pressure(M, &L->b, var);
  }
  for (int it = 0; it < post; it++) mg_itsolv(M, L, var, rhs);
  if (n == 1 && dir >= 1) gcm_pressure(M, &L->b, var);
  free(r); free(r1); free(phi);
}

/* MG_Prepare_Iblank (CORE_MG.F90:699-822) and the level boxes of one
 * directional cycle; level 1 copies the fine flags (MG_Prepare_BC(1)). */
static void cycle_build(mgctx *M, cycle *C, int dir) {
  or_ctx *c = M->c;
  C->dir = dir;
  C->nlev = M->d[dir].nlev;
  for (int n = 1; n <= C->nlev; n++) {
    level *L = &C->L[n];
    L->b.cx = c->nxc; L->b.cy = c->nyc; L->b.cz = c->nzc;
    if (dir == 0) L->b.cx = M->d[0].m[n];
    if (dir == 1) L->b.cy = M->d[1].m[n];
    if (dir == 2) L->b.cz = M->d[2].m[n];
    L->b.ex = L->b.cx + 2; L->b.ey = L->b.cy + 2;
    long sz = bsize(&L->b);
    L->ib = calloc(sz, 1);
    for (int a = 0; a < 6; a++) L->f[a] = calloc(sz, 1);
    L->gm = NULL;
    L->lv[0] = L->lv[1] = L->lv[2] = 1;
    L->lv[dir] = n;
    if (n == 1) {
      L->gm = calloc(sz, 1);
      const int8_t *src[6] = {c->ium, c->iup, c->jum, c->jup, c->kum, c->kup};
#pragma omp parallel for if (M->o->zebra) schedule(static)
      for (int k = 0; k <= c->nzc + 1; k++)
        for (int j = 0; j <= c->nyc + 1; j++)
          for (int i = 0; i <= c->nxc + 1; i++) {
            long id = bidx(&L->b, i, j, k), g = OR_IDX(c, i, j, k);
            L->ib[id] = c->iblank[g];
            L->gm[id] = c->ghost[g];
            for (int a = 0; a < 6; a++) L->f[a][id] = src[a][g];
          }
    } else {
      level *P = &C->L[n - 1];
#pragma omp parallel for if (M->o->zebra) schedule(static)
      for (int k = 1; k <= L->b.cz; k++)
        for (int j = 1; j <= L->b.cy; j++)
          for (int i = 1; i <= L->b.cx; i++) {
            long a, bb;
            if (dir == 0) { a = bidx(&P->b, 2 * i - 1, j, k); bb = bidx(&P->b, 2 * i, j, k); }
            else if (dir == 1) { a = bidx(&P->b, i, 2 * j - 1, k); bb = bidx(&P->b, i, 2 * j, k); }
            else { a = bidx(&P->b, i, j, 2 * k - 1); bb = bidx(&P->b, i, j, 2 * k); }
            L->ib[bidx(&L->b, i, j, k)] = P->ib[a] * P->ib[bb];
          }
    }
  }
}

static void cycle_free(cycle *C) {
  for (int n = 1; n <= C->nlev; n++) {
    free(C->L[n].ib); free(C->L[n].gm);
    for (int a = 0; a < 6; a++) free(C->L[n].f[a]);
  }
}

static void mgctx_init(mgctx *M, or_ctx *c, const or_mg_opts *o) {
  M->c = c; M->o = o;
  mg_init_dir(&M->d[0], c->nxc, c->x, c->xc, c->dx, c->dxinv, c->dxc, c->dxcinv);
  mg_init_dir(&M->d[1], c->nyc, c->y, c->yc, c->dy, c->dyinv, c->dyc, c->dycinv);
  mg_init_dir(&M->d[2], c->nzc, c->z, c->zc, c->dz, c->dzinv, c->dzc, c->dzcinv);
}

static void mgctx_free(mgctx *M) {
  for (int d = 0; d < 3; d++) mg_free_dir(&M->d[d]);
}

/* MG_Solver (CORE_MG.F90:281-334) on the module-layout arrays var, rr:
 * copy (0:nx,0:ny,0:nz) into working arrays, run the x, y, z cycles,
 * copying back after each. */
static void mg_solver(mgctx *M, double *var, const double *rr) {
  or_ctx *c = M->c;
  box b1 = {c->nxc, c->nyc, c->nzc, c->nxc + 2, c->nyc + 2};
  long sz = bsize(&b1);
  double *v1 = calloc(sz, 8), *r1 = calloc(sz, 8);
#pragma omp parallel for if (M->o->zebra) schedule(static)
  for (int k = 0; k <= c->nz; k++)
    for (int j = 0; j <= c->ny; j++)
      for (int i = 0; i <= c->nx; i++) {
        v1[bidx(&b1, i, j, k)] = var[OR_IDX(c, i, j, k)];
        r1[bidx(&b1, i, j, k)] = rr[OR_IDX(c, i, j, k)];
      }
  for (int dir = 0; dir < 3; dir++) {
    cycle C;
    memset(&C, 0, sizeof(C));
    cycle_build(M, &C, dir);
    cycle_level(M, &C, 1, v1, r1);
    cycle_free(&C);
#pragma omp parallel for if (M->o->zebra) schedule(static)
    for (int k = 0; k <= c->nz; k++)
      for (int j = 0; j <= c->ny; j++)
        for (int i = 0; i <= c->nx; i++) var[OR_IDX(c, i, j, k)] = v1[bidx(&b1, i, j, k)];
  }
  free(v1); free(r1);
}

/* GCM_correct_pressure (CORE_POISSON.F90:1794-1823) with pgrad = 0 */
static void gcm_correct_pressure(const or_ctx *c, const or_mg_opts *o, double *p) {
  or_outer_ghost(c, p);
  or_ghost_rows(c, p, NULL, o->itermax_ad, o->restol_ad);
}

void or_mg_cycle(or_ctx *c, double *p, const double *rhs, const or_mg_opts *o) {
  mgctx M;
  mgctx_init(&M, c, o);
  mg_solver(&M, p, rhs);
  gcm_correct_pressure(c, o, p);
  mgctx_free(&M);
}

/* solve_poisson, MG branch (CORE_POISSON.F90:493-676) */
int or_mg_solve(or_ctx *c, double *p, const double *rhs, const or_mg_opts *o,
                or_mg_stats *s) {
  mgctx M;
  mgctx_init(&M, c, o);
  if (o->mean_remove) or_mean_remove(c, p);
  double bl2 = 0.0;
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (!c->iblank[id]) bl2 += rhs[id] * rhs[id];
      }
  bl2 = sqrt(bl2);
  int iter = 0;
  double maxres = 1.0e10, l2 = 0.0;
  int32_t loc[3] = {0, 0, 0};
  int conv = 0;
  while (iter < o->iter_max && maxres > o->restol) {
    mg_solver(&M, p, rhs);
    gcm_correct_pressure(c, o, p);
    maxres = or_calc_residual(c, p, rhs, NULL, loc, &l2);
    iter++;
    if (o->rel_tol > 0.0 && l2 <= o->rel_tol * bl2) { conv = 1; break; }
  }
  if (maxres <= o->restol) conv = 1;
  if (s) {
    s->iterations = iter; s->linf = maxres; s->l2 = l2; s->b_l2 = bl2;
    s->loc[0] = loc[0]; s->loc[1] = loc[1]; s->loc[2] = loc[2];
    for (int d = 0; d < 3; d++) s->levels[d] = M.d[d].nlev;
  }
  mgctx_free(&M);
  return conv ? 0 : 1;
}
