/* Internal state shared by the oracle's translation units (test
 * infrastructure only; see oracle.h). */
#ifndef FSI_OR_INTERNAL_H
#define FSI_OR_INTERNAL_H

#include "oracle.h"
#include <stddef.h>

typedef struct or_body {
  int nmarker, nelem;
  double *xm, *ym, *zm;          /* markers 1..nmarker stored 0-based */
  int32_t *tri;                  /* 3*nelem, 1-based marker ids */
  double *cx, *cy, *cz;          /* element centroids */
  double *nxv, *nyv, *nzv;       /* unit element normals (inward) */
  double *area;                  /* element areas */
} or_body;

struct or_ctx {
  int nxc, nyc, nzc;
  int nx, ny, nz;                /* reference node counts */
  long ex, ey, ez;               /* field extents nx+2, ny+2, nz+2 */
  /* 1-D metrics, 0..n+1 (CORE_GRID.F90:97-149) */
  double *x, *xc, *dx, *dxinv, *dxc, *dxcinv;
  double *y, *yc, *dy, *dyinv, *dyc, *dycinv;
  double *z, *zc, *dz, *dzinv, *dzc, *dzcinv;
  /* flags, reference layout */
  int8_t *iblank, *ghost, *fresh;
  int8_t *ium, *iup, *jum, *jup, *kum, *kup;
  /* ghost list and stencils (GCM_arrays, AMODULES.F90:507-579) */
  int nghost;
  int32_t *ig, *jg, *kg, *ic, *jc, *kc, *elem;
  double *coeff, *probe, *bi, *ipt;
  double *coeffD;                /* coeffGCMD(8,n), velocity (Dirichlet) set */
  double bi_weight, ip_weight;   /* bodyInterceptWeight, imagePointWeight */
  /* fresh cells (GCM_set_freshcell, AMODULES.F90:560-570): iFresh..,
   * iFreshCellIndex.., coeffGCMFreshD(8,n), body intercept, body, element */
  int nfr;
  int32_t *fr_i, *fr_j, *fr_k, *fr_ic, *fr_jc, *fr_kc, *fr_body, *fr_elem;
  double *fr_cD, *fr_bi;
  /* bodies */
  int nbody;
  or_body body[64];
  int32_t *bodynum;
};

#define OR_IDX(c, i, j, k) \
  ((long)(i) + (c)->ex * ((long)(j) + (c)->ey * (long)(k)))

void or_mark_ghosts(or_ctx *c);
void or_free_fresh(or_ctx *c);
void or_alloc_fresh(or_ctx *c, int n);

#endif
