/*
 * or_geom.c -- the reference's triangulated-surface geometry for canonical
 * ELLIPSOID bodies under GCM, restated in C.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   surface mesh   initialize_marker, CASE(ELLIPSOID)
 *                  (BOUNDARY_MARKER_INIT.F90:115-205; n_phi MAIN.F90:1156-1166)
 *   orientation    determine_norm_dir_unstruc (BOUNDARY_SET_ARCLENGTH_NORM.F90:200-320)
 *   normals        calculate_arclength_norm_ds (BOUNDARY_SET_ARCLENGTH_NORM.F90:69-167)
 *   iblank         set_iblank_canonical_body_fast (BOUNDARY_IBLANK_FAST.F90:2-501)
 *                  with search_vertex_dotNorm (:504-1024), check_BIInsideTriangle
 *                  (:1028-1159), calc_BIOutsideTriangle (:1163-1435), get_inter and
 *                  get_distance_face (:3513-3555), check_if_edge_marker (:3580-3703),
 *                  pointInTriangle (:3862-3910), vector_angle/mo/dot/cross
 *                  (AMODULES.F90:176-182,905-988)
 *   cleanup        find_isolatedIblank(1) (BOUNDARY_SET.F90:440-499), fill_cell
 *                  (SSM_SET_BC.F90:1425-1511), in the order of set_internal_boundary
 *                  (BOUNDARY_SET.F90:264-410)
 *   ghosts         GCM_set_internal_boundary (GCM_SET_BOUNDARY.F90:2-304) with
 *                  GCM_calc_bodyIntercept_Unstruc (GCM_INTERCEPT_KERNELS_UNSTRUC.F90:26-398)
 *                  and GCM_Calc_VanMatrixDN (GCM_SET_BOUNDARY.F90:307-494)
 *
 * Expressions keep the reference's operation order (Fortran evaluates
 * a*b*c as (a*b)*c and a+b+c as (a+b)+c); contraction is off.  REAL(4)
 * literals of the reference (1e-20 in pointInTriangle, 1e-10 in get_inter)
 * are used at single precision, as the Fortran compares against them.
 *
 * Decisions where the reference leaves a value undefined:
 *   - pointOutsideBody is never assigned for ELLIPSOID bodies; the origin is
 *     used (a zero-initialised allocation) and a body whose bounding box
 *     contains the origin is rejected.
 *   - vector_angle of a zero vector (0/0) returns NaN (the Fortran function
 *     result is unassigned there).
 *   - check_if_edge_marker with fewer than two elements on the marker
 *     reports "not an edge" (the Fortran leaves edgeM unassigned).
 *   - get_inter's closed form is the line-plane intersection
 *     p1 + lhs1/(lhs1-lhs2) (p2-p1), algebraically the reference's expanded
 *     polynomial; it rounds differently.
 *   - bodyNum of the k = 0 layer (read by the k-direction fill, never
 *     written) is 0.
 *   - the 8x8 Vandermonde inverse is Gauss-Jordan elimination with partial
 *     pivoting instead of LAPACK DGETRF/DGETRI (Lapack_VICAR.F90).
 *
 * Parity pins (tests/test_geometry.py): the mesh has Euler characteristic
 * 2, every edge is shared by exactly two elements with opposite
 * orientation, all normals point into the body and the areas sum to the
 * area of the inscribed polyhedron (closed forms on a sphere); iblank
 * agrees with "centre inside the faceted body" away from the surface band
 * and with brute-force point-in-polyhedron tests on the band; intercepts lie
 * on their element's plane; the coefficients reproduce trilinear fields with
 * zero normal derivative at the intercept.  Agreement with the Fortran
 * itself is parity unpinned (it does not build here).
 */
#include "or_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define UNDECIDED 1
#define NSAVE 7
#define NCHECK 3

static double g_pi(void) { return 4.0 * atan(1.0); }

/* ---- operation module (AMODULES.F90:176-182, 905-988) ------------------ */
static double mo3(const double a[3]) {
  double s = 0.0;
  for (int i = 0; i < 3; i++) s = s + a[i] * a[i];
  return sqrt(s);
}
static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double vector_angle(const double a[3], const double b[3]) {
  double pi = 4.0 * atan(1.0);
  double cr[3];
  double c = dot3(a, b) / (mo3(a) * mo3(b));
  cross3(a, b, cr);
  double d = mo3(cr) / (mo3(a) * mo3(b));
  double ang = NAN;
  if (c < 1.0 && c > -1.0) ang = acos(c);
  else if (c > 0.0) ang = 0.0;
  else if (c < 0.0) ang = pi;
  if (d >= -pi / 2.0 && d < 0.0) ang = 2.0 * pi - ang;
  return ang;
}

/* ---- body surfaces ---------------------------------------------------- */
void or_geom_clear_bodies(or_ctx *c) {
  for (int b = 0; b < c->nbody; b++) {
    or_body *B = &c->body[b];
    free(B->xm); free(B->ym); free(B->zm); free(B->tri);
    free(B->cx); free(B->cy); free(B->cz);
    free(B->nxv); free(B->nyv); free(B->nzv); free(B->area);
    memset(B, 0, sizeof(*B));
  }
  c->nbody = 0;
}

int or_geom_nbodies(const or_ctx *c) { return c->nbody; }

static void elem_vectors(const or_body *B, int e, double *nx, double *ny, double *nz) {
  int n1 = B->tri[3 * e] - 1, n2 = B->tri[3 * e + 1] - 1, n3 = B->tri[3 * e + 2] - 1;
  double v1x = B->xm[n2] - B->xm[n1], v1y = B->ym[n2] - B->ym[n1], v1z = B->zm[n2] - B->zm[n1];
  double v2x = B->xm[n3] - B->xm[n2], v2y = B->ym[n3] - B->ym[n2], v2z = B->zm[n3] - B->zm[n2];
  *nx = v1y * v2z - v1z * v2y;
  *ny = v1z * v2x - v1x * v2z;
  *nz = v1x * v2y - v1y * v2x;
}

int or_geom_ellipsoid(or_ctx *c, int n_theta, double xc0, double yc0, double zc0, double ra, double rb, double rc,
                      double alpha_deg) {
  if (c->nbody >= 64 || n_theta < 3 || ra <= 0 || rb <= 0 || rc <= 0) return -1;
  double PI = g_pi();
  double aratio = rc / ra;  /* GCM: Aratio = radiusz/radiusx (MAIN.F90:1156-1160) */
  int n_phi = (n_theta % 2 == 0) ? (int)lround(0.5 * (double)n_theta * aratio)
                                  : (int)lround(0.5 * (double)(n_theta - 1) * aratio) + 1;
  if (n_phi < 3) return -1;
  double ca = cos(alpha_deg * PI / 180.0), sa = sin(alpha_deg * PI / 180.0);
  or_body *B = &c->body[c->nbody];
  memset(B, 0, sizeof(*B));
  int cap = n_theta * n_phi;
  B->xm = calloc(cap, 8); B->ym = calloc(cap, 8); B->zm = calloc(cap, 8);
  int m = 0;
  for (int i = 1; i <= n_phi; i++) {
    double phi = ((double)i - 1.0) * PI / ((double)n_phi - 1.0);
    int nring = (i == 1 || i == n_phi) ? 1 : n_theta;
    for (int j = 1; j <= nring; j++) {
      double theta = nring == 1 ? 0.0 : ((double)j - 1.0) * 2.0 * PI / (double)n_theta;
      double xT = ra * cos(theta) * sin(phi);
      double yT = rb * sin(theta) * sin(phi);
      B->xm[m] = xc0 + xT * ca - yT * sa;
      B->ym[m] = yc0 + xT * sa + yT * ca;
      B->zm[m] = zc0 + rc * cos(phi);
      m++;
    }
  }
  B->nmarker = m;
  int ne = 2 * n_theta + 2 * (m - n_theta - 2);
  B->tri = calloc(3 * (size_t)ne, sizeof(int32_t));
  int j = 0;
  for (int k = 1; k <= n_theta; k++, j++) {
    B->tri[3 * j] = 1;
    B->tri[3 * j + 1] = 1 + k;
    B->tri[3 * j + 2] = (1 + k == n_theta + 1) ? 2 : 1 + k + 1;
  }
  int nnt = 1;
  for (int i = 2; i <= m - n_theta - 1; i++) {
    B->tri[3 * j] = i;
    B->tri[3 * j + 1] = i + n_theta;
    B->tri[3 * j + 2] = (i == nnt * n_theta + 1) ? i + 1 : i + n_theta + 1;
    j++;
    B->tri[3 * j] = i;
    if (i == nnt * n_theta + 1) {
      B->tri[3 * j + 1] = i + 1;
      B->tri[3 * j + 2] = i + 1 - n_theta;
      nnt++;
    } else {
      B->tri[3 * j + 1] = i + n_theta + 1;
      B->tri[3 * j + 2] = i + 1;
    }
    j++;
  }
  for (int i = m - n_theta; i <= m - 1; i++, j++) {
    B->tri[3 * j] = i;
    B->tri[3 * j + 1] = m;
    B->tri[3 * j + 2] = (i + 1 == m) ? m - n_theta : i + 1;
  }
  B->nelem = j;
  /* orientation from the element at the marker closest to the outside
   * point (the origin): its normal must point away from that point */
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300, zmin = 1e300, zmax = -1e300;
  for (int q = 0; q < m; q++) {
    xmin = fmin(xmin, B->xm[q]); xmax = fmax(xmax, B->xm[q]);
    ymin = fmin(ymin, B->ym[q]); ymax = fmax(ymax, B->ym[q]);
    zmin = fmin(zmin, B->zm[q]); zmax = fmax(zmax, B->zm[q]);
  }
  if (xmin <= 0.0 && xmax >= 0.0 && ymin <= 0.0 && ymax >= 0.0 && zmin <= 0.0 && zmax >= 0.0) {
    free(B->xm); free(B->ym); free(B->zm); free(B->tri);
    memset(B, 0, sizeof(*B));
    return -2;
  }
  const double po[3] = {0.0, 0.0, 0.0};
  double dmin = 1.0e8;
  int cnode = 0;
  for (int q = 1; q <= m; q++) {
    double d = sqrt((B->xm[q - 1] - po[0]) * (B->xm[q - 1] - po[0]) + (B->ym[q - 1] - po[1]) * (B->ym[q - 1] - po[1]) +
                    (B->zm[q - 1] - po[2]) * (B->zm[q - 1] - po[2]));
    if (d <= dmin) { dmin = d; cnode = q; }
  }
  int celem = 0;
  for (int e = 0; e < B->nelem; e++)
    if (B->tri[3 * e] == cnode || B->tri[3 * e + 1] == cnode || B->tri[3 * e + 2] == cnode) celem = e;
  double nx, ny, nz;
  elem_vectors(B, celem, &nx, &ny, &nz);
  int a1 = B->tri[3 * celem] - 1, a2 = B->tri[3 * celem + 1] - 1, a3 = B->tri[3 * celem + 2] - 1;
  double cxe = (B->xm[a1] + B->xm[a2] + B->xm[a3]) / 3.0;
  double cye = (B->ym[a1] + B->ym[a2] + B->ym[a3]) / 3.0;
  double cze = (B->zm[a1] + B->zm[a2] + B->zm[a3]) / 3.0;
  double vp = (cxe - po[0]) * nx + (cye - po[1]) * ny + (cze - po[2]) * nz;
  if (vp < 0.0)
    for (int e = 0; e < B->nelem; e++) {
      int t = B->tri[3 * e + 1];
      B->tri[3 * e + 1] = B->tri[3 * e + 2];
      B->tri[3 * e + 2] = t;
    }
  /* unit normals, areas and centroids */
  B->cx = calloc(B->nelem, 8); B->cy = calloc(B->nelem, 8); B->cz = calloc(B->nelem, 8);
  B->nxv = calloc(B->nelem, 8); B->nyv = calloc(B->nelem, 8); B->nzv = calloc(B->nelem, 8);
  B->area = calloc(B->nelem, 8);
  for (int e = 0; e < B->nelem; e++) {
    elem_vectors(B, e, &nx, &ny, &nz);
    double mag = sqrt(nx * nx + ny * ny + nz * nz);
    B->area[e] = 0.5 * mag;
    B->nxv[e] = nx / mag; B->nyv[e] = ny / mag; B->nzv[e] = nz / mag;
    int n1 = B->tri[3 * e] - 1, n2 = B->tri[3 * e + 1] - 1, n3 = B->tri[3 * e + 2] - 1;
    B->cx[e] = (B->xm[n1] + B->xm[n2] + B->xm[n3]) / 3.0;
    B->cy[e] = (B->ym[n1] + B->ym[n2] + B->ym[n3]) / 3.0;
    B->cz[e] = (B->zm[n1] + B->zm[n2] + B->zm[n3]) / 3.0;
  }
  c->nbody++;
  return m;
}

void or_geom_surface_sizes(const or_ctx *c, int body, int *nmarker, int *nelem) {
  *nmarker = c->body[body].nmarker;
  *nelem = c->body[body].nelem;
}

void or_geom_get_surface(const or_ctx *c, int body, double *xyz, int32_t *tri, double *cent, double *norm) {
  const or_body *B = &c->body[body];
  for (int q = 0; q < B->nmarker; q++) {
    xyz[3 * q] = B->xm[q]; xyz[3 * q + 1] = B->ym[q]; xyz[3 * q + 2] = B->zm[q];
  }
  memcpy(tri, B->tri, sizeof(int32_t) * 3 * (size_t)B->nelem);
  for (int e = 0; e < B->nelem; e++) {
    cent[3 * e] = B->cx[e]; cent[3 * e + 1] = B->cy[e]; cent[3 * e + 2] = B->cz[e];
    norm[3 * e] = B->nxv[e]; norm[3 * e + 1] = B->nyv[e]; norm[3 * e + 2] = B->nzv[e];
  }
}

/* ---- geometric kernels of BOUNDARY_IBLANK_FAST.F90 -------------------- */
static void marker(const or_body *B, int m1, double p[3]) {
  p[0] = B->xm[m1 - 1]; p[1] = B->ym[m1 - 1]; p[2] = B->zm[m1 - 1];
}
static void enorm(const or_body *B, int e, double n[3]) {
  n[0] = B->nxv[e]; n[1] = B->nyv[e]; n[2] = B->nzv[e];
}
static void ecent(const or_body *B, int e, double r[3]) {
  r[0] = B->cx[e]; r[1] = B->cy[e]; r[2] = B->cz[e];
}

static double get_distance_face(const double n[3], const double ref[3], const double p[3]) {
  if (p[0] == ref[0] && p[1] == ref[1] && p[2] == ref[2]) return 0.0;
  double v[3] = {p[0] - ref[0], p[1] - ref[1], p[2] - ref[2]};
  double ang = vector_angle(n, v);
  return mo3(v) * cos(ang);
}

static void get_inter(const double n[3], const double ref[3], const double p1[3], const double p2[3], double out[3]) {
  double lhs1 = n[0] * (p1[0] - ref[0]) + n[1] * (p1[1] - ref[1]) + n[2] * (p1[2] - ref[2]);
  double lhs2 = n[0] * (p2[0] - ref[0]) + n[1] * (p2[1] - ref[1]) + n[2] * (p2[2] - ref[2]);
  const double eps = (double)1e-10f;
  if (fabs(lhs1) < eps) {
    memcpy(out, p1, 3 * sizeof(double));
  } else if (fabs(lhs2) < eps) {
    memcpy(out, p2, 3 * sizeof(double));
  } else {
    double t = lhs1 / (lhs1 - lhs2);
    for (int d = 0; d < 3; d++) out[d] = p1[d] + t * (p2[d] - p1[d]);
  }
}

static int point_in_triangle(const double a[3], const double b[3], const double cc[3], const double p[3]) {
  double pi = 4.0 * atan(1.0);
  double v1[3], v2[3], v3[3], w1[3], w2[3], w3[3], c1[3], c2[3], c3[3];
  for (int d = 0; d < 3; d++) {
    v1[d] = b[d] - a[d]; v2[d] = cc[d] - b[d]; v3[d] = a[d] - cc[d];
    w1[d] = p[d] - a[d]; w2[d] = p[d] - b[d]; w3[d] = p[d] - cc[d];
  }
  cross3(v1, w1, c1); cross3(v2, w2, c2); cross3(v3, w3, c3);
  double m1 = mo3(c1), m2 = mo3(c2), m3 = mo3(c3);
  const double eps = (double)1e-20f;
  if (m1 < eps || m2 < eps || m3 < eps) return 0;
  for (int d = 0; d < 3; d++) { c1[d] = c1[d] / m1; c2[d] = c2[d] / m1; c3[d] = c3[d] / m1; }
  double a12 = vector_angle(c1, c2), a23 = vector_angle(c2, c3);
  return a12 < pi / 10.0 && a23 < pi / 10.0;
}

/* check_BIInsideTriangle: normal projection of P on the element plane and
 * the Heron area test quantities */
static void bi_inside(const or_body *B, int e, const double P[3], double bi[3], double *area123, double *areadiff,
                      double *dist_inside) {
  int n1 = B->tri[3 * e], n2 = B->tri[3 * e + 1], n3 = B->tri[3 * e + 2];
  double p1[3], p2[3], p3[3], n[3];
  marker(B, n1, p1); marker(B, n2, p2); marker(B, n3, p3);
  enorm(B, e, n);
  double pc = -n[0] * p1[0] - n[1] * p1[1] - n[2] * p1[2];
  double dtp = -(n[0] * P[0] + n[1] * P[1] + n[2] * P[2] + pc);
  *dist_inside = fabs(dtp);
  for (int d = 0; d < 3; d++) bi[d] = P[d] + n[d] * dtp;
#define SQD(u, v) ((u[0] - v[0]) * (u[0] - v[0]) + (u[1] - v[1]) * (u[1] - v[1]) + (u[2] - v[2]) * (u[2] - v[2]))
  double s12 = sqrt(SQD(p2, p1)), s23 = sqrt(SQD(p3, p2)), s31 = sqrt(SQD(p1, p3));
  double s14 = sqrt(SQD(bi, p1)), s24 = sqrt(SQD(bi, p2)), s34 = sqrt(SQD(bi, p3));
#undef SQD
  double h123 = 0.5 * (s12 + s23 + s31), h124 = 0.5 * (s12 + s24 + s14);
  double h234 = 0.5 * (s23 + s24 + s34), h314 = 0.5 * (s31 + s34 + s14);
  double a123 = sqrt(h123 * (h123 - s12) * (h123 - s23) * (h123 - s31));
  double a124 = sqrt(h124 * (h124 - s12) * (h124 - s24) * (h124 - s14));
  double a234 = sqrt(h234 * (h234 - s23) * (h234 - s24) * (h234 - s34));
  double a314 = sqrt(h314 * (h314 - s31) * (h314 - s34) * (h314 - s14));
  *area123 = a123;
  *areadiff = a124 + a234 + a314 - a123;
}

/* calc_BIOutsideTriangle: the angle between (P - closest marker) and the
 * outward normal of the element */
static double bi_outside(const or_body *B, int e, const double P[3], int cmarker) {
  double cm[3], v[3], n[3];
  marker(B, cmarker, cm);
  for (int d = 0; d < 3; d++) v[d] = P[d] - cm[d];
  enorm(B, e, n);
  for (int d = 0; d < 3; d++) n[d] = -n[d];
  return vector_angle(v, n);
}

/* indices (1-based) of the k markers closest to P: repeated MINLOC
 * (first index among equal distances) */
static void closest_markers(const or_body *B, const double P[3], int k, int *out) {
  double *d2 = malloc(sizeof(double) * B->nmarker);
  for (int m = 0; m < B->nmarker; m++) {
    double dx = B->xm[m] - P[0], dy = B->ym[m] - P[1], dz = B->zm[m] - P[2];
    d2[m] = dx * dx + dy * dy + dz * dz;
  }
  for (int s = 0; s < k; s++) {
    int best = 0;
    for (int m = 1; m < B->nmarker; m++)
      if (d2[m] < d2[best]) best = m;
    out[s] = best + 1;
    d2[best] = 1.0e20;
  }
  free(d2);
}

static int elem_has(const or_body *B, int e, int m1) {
  return B->tri[3 * e] == m1 || B->tri[3 * e + 1] == m1 || B->tri[3 * e + 2] == m1;
}

/* check_if_edge_marker: two elements at the marker whose normals are at
 * least 90 degrees apart */
static int edge_marker(const or_body *B, int m1) {
  double pi = g_pi();
  int *lst = malloc(sizeof(int) * B->nelem), n = 0;
  for (int e = 0; e < B->nelem; e++)
    if (elem_has(B, e, m1)) lst[n++] = e;
  int edge = 0;
  for (int a = 0; a < n - 1 && !edge; a++)
    for (int b = a + 1; b < n; b++) {
      double r[3], q[3];
      enorm(B, lst[a], r); enorm(B, lst[b], q);
      for (int d = 0; d < 3; d++) { r[d] = -r[d]; q[d] = -q[d]; }
      if (vector_angle(r, q) >= pi / 2.0) { edge = 1; break; }
    }
  free(lst);
  return edge;
}

/* search_vertex_dotNorm for one band cell and one body: 0 and *dotnorm, or
 * 4 when every saved marker is blocked */
static int search_vertex_dotnorm(const or_body *B, const double P[3], double *dotnorm) {
  int cv[NSAVE];
  *dotnorm = -1.0;
  closest_markers(B, P, NSAVE, cv);
  int *nb = malloc(sizeof(int) * B->nelem), nn = 0;
  for (int e = 0; e < B->nelem; e++) {
    int hit = 0;
    for (int s = 0; s < NSAVE; s++) hit |= elem_has(B, e, cv[s]);
    if (hit) nb[nn++] = e;
  }
  int choice = -1;
  for (int s = 0; s < NSAVE; s++) {
    double pm[3];
    marker(B, cv[s], pm);
    int change = 0;
    for (int q = 0; q < nn; q++) {
      int e = nb[q];
      if (elem_has(B, e, cv[s])) continue;
      double n[3], r[3];
      enorm(B, e, n); ecent(B, e, r);
      double dc = get_distance_face(n, r, P), dm = get_distance_face(n, r, pm);
      if (dc * dm > 0.0) continue;
      double po[3], a[3], b[3], cc[3];
      get_inter(n, r, P, pm, po);
      marker(B, B->tri[3 * e], a); marker(B, B->tri[3 * e + 1], b); marker(B, B->tri[3 * e + 2], cc);
      if (point_in_triangle(a, b, cc, po)) { change = 1; break; }
    }
    if (!change) { choice = s; break; }
  }
  if (choice < 0) { free(nb); return 4; }
  int cm = cv[choice];
  if (edge_marker(B, cm)) { free(nb); *dotnorm = -1.0; return 0; }
  nn = 0;
  for (int e = 0; e < B->nelem; e++)
    if (elem_has(B, e, cm)) nb[nn++] = e;
  double dmin = 1.0e16, dist_in = 1.0e16;
  const double eps_area = 1.0e-4;
  int closest = -1, inside = 0;
  for (int q = 0; q < nn; q++) {
    int e = nb[q];
    double bi[3], a123, adiff, din;
    bi_inside(B, e, P, bi, &a123, &adiff, &din);
    if (fabs(adiff) < eps_area * a123) {
      inside = 1;
      if (din <= dist_in) { dist_in = din; closest = e; }
    } else if (!inside) {
      double dbe = bi_outside(B, e, P, cm);
      if (dbe <= dmin) { dmin = dbe; closest = e; }
    }
  }
  free(nb);
  if (closest < 0) return 5;
  *dotnorm = (P[0] - B->cx[closest]) * B->nxv[closest] + (P[1] - B->cy[closest]) * B->nyv[closest] +
             (P[2] - B->cz[closest]) * B->nzv[closest];
  return 0;
}

/* GCM_calc_bodyIntercept_Unstruc: intercept and element for a ghost */
static int body_intercept(const or_body *B, const double P[3], double bi[3], int *elem) {
  int cn[NCHECK];
  closest_markers(B, P, NCHECK, cn);
  double bit[NCHECK][3];
  int cel[NCHECK];
  int *nb = malloc(sizeof(int) * B->nelem);
  for (int s = 0; s < NCHECK; s++) {
    int nn = 0;
    for (int e = 0; e < B->nelem; e++)
      if (elem_has(B, e, cn[s])) nb[nn++] = e;
    int closest = -1, done = 0;
    double dmin = 1.0e16, xb[3] = {0, 0, 0};
    const double eps_area = 1.0e-4;
    for (int q = 0; q < nn; q++) {
      int e = nb[q];
      double t[3], a123, adiff, din;
      bi_inside(B, e, P, t, &a123, &adiff, &din);
      if (fabs(adiff) < eps_area * a123) {
        memcpy(xb, t, sizeof(t));
        closest = e;
        done = 1;
        break;
      }
      double dbe = bi_outside(B, e, P, cn[s]);
      if (dbe <= dmin) { dmin = dbe; closest = e; memcpy(xb, t, sizeof(t)); }
    }
    if (closest < 0) { free(nb); return 5; }
    if (!done) {
      /* the closest point on the two element edges through the marker */
      double cvp[3];
      marker(B, cn[s], cvp);
      int t1 = B->tri[3 * closest], t2 = B->tri[3 * closest + 1], t3 = B->tri[3 * closest + 2];
      int na, nbn;
      if (t1 == cn[s]) { na = t2; nbn = t3; }
      else if (t2 == cn[s]) { na = t3; nbn = t1; }
      else { na = t1; nbn = t2; }
      double v01[3] = {P[0] - cvp[0], P[1] - cvp[1], P[2] - cvp[2]};
      double best = 1.0e16, ei[3] = {cvp[0], cvp[1], cvp[2]};
      for (int ed = 0; ed < 2; ed++) {
        double q[3];
        marker(B, ed == 0 ? na : nbn, q);
        double v12[3] = {q[0] - cvp[0], q[1] - cvp[1], q[2] - cvp[2]};
        double mag = sqrt(v12[0] * v12[0] + v12[1] * v12[1] + v12[2] * v12[2]);
        double inv = 1.0 / mag;
        for (int d = 0; d < 3; d++) v12[d] = v12[d] * inv;
        double proj = v01[0] * v12[0] + v01[1] * v12[1] + v01[2] * v12[2];
        double et[3];
        if (proj < 0.0 || proj > mag) {
          memcpy(et, cvp, sizeof(et));
        } else {
          for (int d = 0; d < 3; d++) et[d] = cvp[d] + proj * v12[d];
        }
        double dd = (et[0] - P[0]) * (et[0] - P[0]) + (et[1] - P[1]) * (et[1] - P[1]) + (et[2] - P[2]) * (et[2] - P[2]);
        if (dd < best) { best = dd; memcpy(ei, et, sizeof(et)); }
      }
      memcpy(xb, ei, sizeof(ei));
    }
    memcpy(bit[s], xb, sizeof(xb));
    cel[s] = closest;
  }
  free(nb);
  int sel = 0;
  double dsel = 0.0;
  for (int s = 0; s < NCHECK; s++) {
    double d = (bit[s][0] - P[0]) * (bit[s][0] - P[0]) + (bit[s][1] - P[1]) * (bit[s][1] - P[1]) +
               (bit[s][2] - P[2]) * (bit[s][2] - P[2]);
    if (s == 0 || d < dsel) { dsel = d; sel = s; }
  }
  memcpy(bi, bit[sel], 3 * sizeof(double));
  *elem = cel[sel];
  return 0;
}

/* ---- set_iblank_canonical_body_fast ----------------------------------- */
static int set_iblank_fast(or_ctx *c, int8_t *temp, int32_t *bnum) {
  long nf = or_field_size(c);
  int nx = c->nx, ny = c->ny, nz = c->nz, nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  int8_t *und = calloc(nf, 1);
  memset(temp, 0, nf);
  memset(bnum, 0, sizeof(int32_t) * nf);
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) und[OR_IDX(c, i, j, k)] = UNDECIDED;
  /* the band: cells in the bounding box of any element */
  for (int b = 0; b < c->nbody; b++) {
    const or_body *B = &c->body[b];
    for (int e = 0; e < B->nelem; e++) {
      double xv[3], yv[3], zv[3];
      for (int v = 0; v < 3; v++) {
        int m1 = B->tri[3 * e + v] - 1;
        xv[v] = B->xm[m1]; yv[v] = B->ym[m1]; zv[v] = B->zm[m1];
      }
      if ((xv[0] <= c->x[1] && xv[1] <= c->x[1] && xv[2] <= c->x[1]) ||
          (xv[0] >= c->x[nx] && xv[1] >= c->x[nx] && xv[2] >= c->x[nx]) ||
          (yv[0] <= c->y[1] && yv[1] <= c->y[1] && yv[2] <= c->y[1]) ||
          (yv[0] >= c->y[ny] && yv[1] >= c->y[ny] && yv[2] >= c->y[ny]) ||
          (zv[0] <= c->z[1] && zv[1] <= c->z[1] && zv[2] <= c->z[1]) ||
          (zv[0] >= c->z[nz] && zv[1] >= c->z[nz] && zv[2] >= c->z[nz]))
        continue;
      const double *cen[3] = {c->xc, c->yc, c->zc};
      const double *vv[3] = {xv, yv, zv};
      int nn[3] = {nx, ny, nz}, ncc[3] = {nxc, nyc, nzc};
      int lo[3], hi[3];
      for (int d = 0; d < 3; d++) {
        double bmin = vv[d][0], bmax = vv[d][0];
        for (int v = 1; v < 3; v++) {
          if (vv[d][v] < bmin) bmin = vv[d][v];
          if (vv[d][v] > bmax) bmax = vv[d][v];
        }
        int cmin = 0, cmax = 0;
        for (int i = 0; i <= nn[d] - 1; i++) {
          if ((cen[d][i] - bmin) * (cen[d][i + 1] - bmin) <= 0.0) cmin = i;
          if ((cen[d][i] - bmax) * (cen[d][i + 1] - bmax) <= 0.0) cmax = i + 1;
        }
        if (cmin == 0) cmin = 1;
        if (cmax == 0 || cmax == nn[d]) cmax = ncc[d];
        lo[d] = cmin;
        hi[d] = cmax;
      }
      for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
          for (int i = lo[0]; i <= hi[0]; i++) und[OR_IDX(c, i, j, k)] = -UNDECIDED;
    }
  }
  /* dot-normal classification of the band cells, bodies in order */
  int rc = 0;
  for (int k = 1; k <= nzc && !rc; k++)
    for (int j = 1; j <= nyc && !rc; j++)
      for (int i = 1; i <= nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (und[id] != -UNDECIDED) continue;
        double P[3] = {c->xc[i], c->yc[j], c->zc[k]};
        temp[id] = 0;
        for (int b = 0; b < c->nbody; b++) {
          double dn;
          rc = search_vertex_dotnorm(&c->body[b], P, &dn);
          if (rc) break;
          if (dn >= 0.0) { temp[id] = 1; bnum[id] = b + 1; break; }
        }
        if (rc) break;
      }
  if (rc) { free(und); return rc; }
  /* ray fills along i, then j, then k (each overwrites the undecided cells) */
  for (int k = 1; k <= nzc; k++)
    for (int j = 1; j <= nyc; j++)
      for (int ii = 1; ii <= nx; ii++)
        if (und[OR_IDX(c, ii, j, k)] != UNDECIDED) { temp[OR_IDX(c, 0, j, k)] = temp[OR_IDX(c, ii, j, k)]; break; }
  for (int i = 1; i <= nx; i++)
    for (int k = 1; k <= nzc; k++)
      for (int j = 1; j <= nyc; j++) {
        long id = OR_IDX(c, i, j, k);
        if (und[id] != UNDECIDED) continue;
        temp[id] = temp[id - 1];
        if (temp[id] == 1) bnum[id] = bnum[id - 1];
      }
  for (int k = 1; k <= nzc; k++)
    for (int i = 1; i <= nxc; i++)
      for (int jj = 1; jj <= ny; jj++)
        if (und[OR_IDX(c, i, jj, k)] != UNDECIDED) { temp[OR_IDX(c, i, 0, k)] = temp[OR_IDX(c, i, jj, k)]; break; }
  for (int j = 1; j <= ny; j++)
    for (int k = 1; k <= nzc; k++)
      for (int i = 1; i <= nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (und[id] != UNDECIDED) continue;
        temp[id] = temp[id - c->ex];
        if (temp[id] == 1) bnum[id] = bnum[id - c->ex];
      }
  long sxy = c->ex * c->ey;
  for (int j = 1; j <= nyc; j++)
    for (int i = 1; i <= nxc; i++)
      for (int kk = 1; kk <= nzc + 1; kk++)
        if (und[OR_IDX(c, i, j, kk)] != UNDECIDED) { temp[OR_IDX(c, i, j, 0)] = temp[OR_IDX(c, i, j, kk)]; break; }
  for (int k = 1; k <= nzc + 1; k++)
    for (int j = 1; j <= nyc; j++)
      for (int i = 1; i <= nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (und[id] != UNDECIDED) continue;
        temp[id] = temp[id - sxy];
        if (temp[id] == 1) bnum[id] = bnum[id - sxy];
      }
  free(und);
  return 0;
}

/* find_isolatedIblank(1) */
static void isolated_iblank(or_ctx *c, int8_t *ib, int32_t *bnum) {
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        if (ib[OR_IDX(c, i, j, k)] != 1) continue;
        int s = 0;
        for (int kk = (k - 1 > 1 ? k - 1 : 1); kk <= (k + 1 < c->nzc ? k + 1 : c->nzc); kk++)
          for (int jj = (j - 1 > 1 ? j - 1 : 1); jj <= (j + 1 < c->nyc ? j + 1 : c->nyc); jj++)
            for (int ii = (i - 1 > 1 ? i - 1 : 1); ii <= (i + 1 < c->nxc ? i + 1 : c->nxc); ii++)
              s += ib[OR_IDX(c, ii, jj, kk)];
        if (s <= 1) { ib[OR_IDX(c, i, j, k)] = 0; bnum[OR_IDX(c, i, j, k)] = 0; }
      }
}

/* fill_cell with the outer-face flags of set_outer_iup */
static void fill_cell(or_ctx *c, int8_t *ib, int8_t *fresh) {
  long st[3] = {1, c->ex, c->ex * c->ey};
  const int8_t *up[3] = {c->iup, c->jup, c->kup}, *um[3] = {c->ium, c->jum, c->kum};
  for (int d = 0; d < 3; d++)
    for (int k = 1; k <= c->nzc; k++)
      for (int j = 1; j <= c->nyc; j++)
        for (int i = 1; i <= c->nxc; i++) {
          long id = OR_IDX(c, i, j, k);
          if (ib[id] != 0) continue;
          if ((ib[id - st[d]] == 1 && up[d][id] == 1) || (ib[id + st[d]] == 1 && um[d][id] == 1)) {
            ib[id] = 1;
            if (fresh[id] == 1) fresh[id] = 0;
          }
        }
}

static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

/* inverse of an 8x8 matrix by Gauss-Jordan elimination with partial
 * pivoting; returns 0 when singular */
static int invert8(double A[8][8], double Ainv[8][8]) {
  double M[8][16];
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) { M[i][j] = A[i][j]; M[i][8 + j] = i == j ? 1.0 : 0.0; }
  for (int k = 0; k < 8; k++) {
    int p = k;
    for (int i = k + 1; i < 8; i++)
      if (fabs(M[i][k]) > fabs(M[p][k])) p = i;
    if (M[p][k] == 0.0) return 0;
    if (p != k)
      for (int j = 0; j < 16; j++) { double t = M[k][j]; M[k][j] = M[p][j]; M[p][j] = t; }
    double piv = M[k][k];
    for (int j = 0; j < 16; j++) M[k][j] /= piv;
    for (int i = 0; i < 8; i++) {
      if (i == k) continue;
      double f = M[i][k];
      if (f == 0.0) continue;
      for (int j = 0; j < 16; j++) M[i][j] -= f * M[k][j];
    }
  }
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) Ainv[i][j] = M[i][8 + j];
  return 1;
}

int or_geom_set_boundary(or_ctx *c, const int8_t *iblank_prev, double probe_norm) {
  long nf = or_field_size(c);
  int8_t *temp = calloc(nf, 1);
  or_free_fresh(c);
  int rc = set_iblank_fast(c, temp, c->bodynum);
  if (rc) { free(temp); return rc; }
  memset(c->fresh, 0, nf);
  for (int k = 1; k <= c->nzc; k++)
    for (int j = 1; j <= c->nyc; j++)
      for (int i = 1; i <= c->nxc; i++) {
        long id = OR_IDX(c, i, j, k);
        if (iblank_prev && temp[id] == 0 && iblank_prev[id] == 1) c->fresh[id] = 1;
        c->iblank[id] = temp[id];
      }
  free(temp);
  isolated_iblank(c, c->iblank, c->bodynum);
  fill_cell(c, c->iblank, c->fresh);
  or_mark_ghosts(c);
  /* MAIN.F90:913-914 */
  c->bi_weight = probe_norm / (probe_norm - 1.0);
  c->ip_weight = 1.0 / (probe_norm - 1.0);
  for (int n = 0; n < c->nghost; n++) {
    int iG = c->ig[n], jG = c->jg[n], kG = c->kg[n];
    long id = OR_IDX(c, iG, jG, kG);
    int b = c->bodynum[id];
    if (b < 1) return 7;
    const or_body *B = &c->body[b - 1];
    double GC[3] = {c->xc[iG], c->yc[jG], c->zc[kG]}, bi[3];
    int elem;
    rc = body_intercept(B, GC, bi, &elem);
    if (rc) return rc;
    double ds = sqrt((GC[0] - bi[0]) * (GC[0] - bi[0]) + (GC[1] - bi[1]) * (GC[1] - bi[1]) +
                     (GC[2] - bi[2]) * (GC[2] - bi[2]));
    double diag = sqrt(c->dxc[iG] * c->dxc[iG] + c->dyc[jG] * c->dyc[jG] + c->dzc[kG] * c->dzc[kG]);
    if (ds > diag && ds / diag > 2.0) return 2;
    double sl[3] = {(bi[0] - GC[0]) / ds, (bi[1] - GC[1]) / ds, (bi[2] - GC[2]) / ds};
    double probe = ds * probe_norm;
    double ip[3] = {GC[0] + sl[0] * probe, GC[1] + sl[1] * probe, GC[2] + sl[2] * probe};
    const double *cen[3] = {c->xc, c->yc, c->zc}, *wid[3] = {c->dx, c->dy, c->dz};
    int g3[3] = {iG, jG, kG}, nn[3] = {c->nx, c->ny, c->nz}, cell[3];
    for (int d = 0; d < 3; d++) {
      int range = (int)lround(probe / wid[d][g3[d]]) + 1;
      int lo = g3[d] - range > 0 ? g3[d] - range : 0, hi = g3[d] + range < nn[d] ? g3[d] + range : nn[d];
      cell[d] = -1;
      for (int i = lo; i <= hi; i++)
        if (cen[d][i] <= ip[d] && cen[d][i + 1] > ip[d]) cell[d] = i;
      if (cell[d] < 0) return 1;
    }
    double V[8][8], Vi[8][8], VD[8][8], VDi[8][8];
    double nrm[3] = {B->nxv[elem], B->nyv[elem], B->nzv[elem]};
    for (int r = 0; r < 8; r++) {
      int i = cell[0] + incI[r], j = cell[1] + incJ[r], k = cell[2] + incK[r];
      double x1 = c->xc[i], x2 = c->yc[j], x3 = c->zc[k];
      if (i == iG && j == jG && k == kG) {
        x1 = bi[0]; x2 = bi[1]; x3 = bi[2];
        /* Dirichlet row: the value at the body intercept */
        VD[r][0] = x1 * x2 * x3; VD[r][1] = x1 * x2; VD[r][2] = x1 * x3; VD[r][3] = x2 * x3;
        VD[r][4] = x1; VD[r][5] = x2; VD[r][6] = x3; VD[r][7] = 1.0;
        V[r][0] = nrm[0] * x2 * x3 + nrm[1] * x1 * x3 + nrm[2] * x1 * x2;
        V[r][1] = nrm[0] * x2 + nrm[1] * x1;
        V[r][2] = nrm[0] * x3 + nrm[2] * x1;
        V[r][3] = nrm[1] * x3 + nrm[2] * x2;
        V[r][4] = nrm[0]; V[r][5] = nrm[1]; V[r][6] = nrm[2]; V[r][7] = 0.0;
      } else {
        V[r][0] = x1 * x2 * x3; V[r][1] = x1 * x2; V[r][2] = x1 * x3; V[r][3] = x2 * x3;
        V[r][4] = x1; V[r][5] = x2; V[r][6] = x3; V[r][7] = 1.0;
        for (int q = 0; q < 8; q++) VD[r][q] = V[r][q];
      }
    }
    if (!invert8(V, Vi)) return 3;
    if (!invert8(VD, VDi)) return 3;
    for (int r = 0; r < 8; r++)
      c->coeffD[8 * n + r] = VDi[0][r] * ip[0] * ip[1] * ip[2] + VDi[1][r] * ip[0] * ip[1] +
                             VDi[2][r] * ip[0] * ip[2] + VDi[3][r] * ip[1] * ip[2] + VDi[4][r] * ip[0] +
                             VDi[5][r] * ip[1] + VDi[6][r] * ip[2] + VDi[7][r];
    for (int r = 0; r < 8; r++)
      c->coeff[8 * n + r] = Vi[0][r] * ip[0] * ip[1] * ip[2] + Vi[1][r] * ip[0] * ip[1] + Vi[2][r] * ip[0] * ip[2] +
                            Vi[3][r] * ip[1] * ip[2] + Vi[4][r] * ip[0] + Vi[5][r] * ip[1] + Vi[6][r] * ip[2] +
                            Vi[7][r];
    c->ic[n] = cell[0]; c->jc[n] = cell[1]; c->kc[n] = cell[2];
    c->probe[n] = probe;
    c->elem[n] = elem + 1;
    for (int d = 0; d < 3; d++) { c->bi[3 * n + d] = bi[d]; c->ipt[3 * n + d] = ip[d]; }
  }
  return 0;
}

/* GCM_set_freshcell (GCM_SET_FRESHCELL.F90:1-262) for the fresh cells of
 * the last or_geom_set_boundary: the list in i-fastest order (:56-72), the
 * body intercept of each cell centre (GCM_calc_bodyIntercept_Unstruc on
 * the body whose closest marker is nearest, the lower index on ties; the
 * reference reads bodyNum, which for a fresh cell still holds the body that
 * blanked it), the image point FIP = BI + (FC - BI)/|FC - BI| * |FC - BI| *
 * probeLengthNormalized (:101-121), the stencil search over the cell and
 * its two neighbours clamped to 1..n-1 (:123-145) with the reference's stop
 * when the corner is neither the cell nor the one below it (:149-166), and
 * coeffGCMFreshD from the inverse of the Dirichlet Vandermonde matrix whose
 * row at the cell itself is taken at BI, evaluated at the cell centre
 * (:196-260).  Returns 0, or 10 (image point outside the stencil), 11 (BI
 * at the cell centre), 3 (singular) or an intercept error (4, 5). */
int or_geom_set_freshcell(or_ctx *c, double probe_norm) {
  int n = 0;
  for (int k = 1; k <= c->nz - 1; k++)
    for (int j = 1; j <= c->ny - 1; j++)
      for (int i = 1; i <= c->nx - 1; i++) n += c->fresh[OR_IDX(c, i, j, k)] == 1;
  or_alloc_fresh(c, n);
  n = 0;
  for (int k = 1; k <= c->nz - 1; k++)
    for (int j = 1; j <= c->ny - 1; j++)
      for (int i = 1; i <= c->nx - 1; i++)
        if (c->fresh[OR_IDX(c, i, j, k)] == 1) { c->fr_i[n] = i; c->fr_j[n] = j; c->fr_k[n] = k; n++; }
  for (int q = 0; q < n; q++) {
    const int iF = c->fr_i[q], jF = c->fr_j[q], kF = c->fr_k[q];
    const double FC[3] = {c->xc[iF], c->yc[jF], c->zc[kF]};
    /* nearest body by its closest marker */
    int bsel = -1;
    double dbest = 0.0;
    for (int b = 0; b < c->nbody; b++) {
      const or_body *B = &c->body[b];
      for (int m = 0; m < B->nmarker; m++) {
        const double dx = B->xm[m] - FC[0], dy = B->ym[m] - FC[1], dz = B->zm[m] - FC[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (bsel < 0 || d2 < dbest) { dbest = d2; bsel = b; }
      }
    }
    if (bsel < 0) return 7;
    double bi[3];
    int elem;
    int rc = body_intercept(&c->body[bsel], FC, bi, &elem);
    if (rc) return rc;
    c->fr_body[q] = bsel + 1;
    c->fr_elem[q] = elem + 1;
    for (int d = 0; d < 3; d++) c->fr_bi[3 * q + d] = bi[d];
    const double ds = sqrt((FC[1] - bi[1]) * (FC[1] - bi[1]) + (FC[0] - bi[0]) * (FC[0] - bi[0]) +
                           (FC[2] - bi[2]) * (FC[2] - bi[2]));
    if (!(ds > 0.0)) return 11;
    const double dsinv = 1.0 / ds, plen = ds * probe_norm;
    const double *cen[3] = {c->xc, c->yc, c->zc};
    const int F3[3] = {iF, jF, kF}, nmax[3] = {c->nx - 1, c->ny - 1, c->nz - 1};
    int cell[3];
    for (int d = 0; d < 3; d++) {
      const double fip = bi[d] + (FC[d] - bi[d]) * dsinv * plen;
      const int lo = F3[d] - 1 > 1 ? F3[d] - 1 : 1, hi = F3[d] + 1 < nmax[d] ? F3[d] + 1 : nmax[d];
      cell[d] = -1;
      for (int i = lo; i <= hi; i++)
        if (cen[d][i] <= fip && cen[d][i + 1] > fip) cell[d] = i;
      if (cell[d] - F3[d] > 0 || cell[d] - F3[d] < -1) return 10;
    }
    double VD[8][8], VDi[8][8];
    for (int r = 0; r < 8; r++) {
      const int i = cell[0] + incI[r], j = cell[1] + incJ[r], k = cell[2] + incK[r];
      double x1 = c->xc[i], x2 = c->yc[j], x3 = c->zc[k];
      if (i == iF && j == jF && k == kF) { x1 = bi[0]; x2 = bi[1]; x3 = bi[2]; }
      VD[r][0] = x1 * x2 * x3; VD[r][1] = x1 * x2; VD[r][2] = x1 * x3; VD[r][3] = x2 * x3;
      VD[r][4] = x1; VD[r][5] = x2; VD[r][6] = x3; VD[r][7] = 1.0;
    }
    if (!invert8(VD, VDi)) return 3;
    const double x = FC[0], y = FC[1], z = FC[2];
    for (int r = 0; r < 8; r++)
      c->fr_cD[8 * q + r] = VDi[0][r] * x * y * z + VDi[1][r] * x * y + VDi[2][r] * x * z + VDi[3][r] * y * z +
                            VDi[4][r] * x + VDi[5][r] * y + VDi[6][r] * z + VDi[7][r];
    c->fr_ic[q] = cell[0]; c->fr_jc[q] = cell[1]; c->fr_kc[q] = cell[2];
  }
  return 0;
}

void or_geom_get_fresh(const or_ctx *c, int32_t *ifr, int32_t *jfr, int32_t *kfr, int32_t *ic, int32_t *jc,
                       int32_t *kc, double *coeffD, double *bi, int32_t *body, int32_t *elem) {
  size_t n = (size_t)c->nfr;
  if (ifr) memcpy(ifr, c->fr_i, 4 * n);
  if (jfr) memcpy(jfr, c->fr_j, 4 * n);
  if (kfr) memcpy(kfr, c->fr_k, 4 * n);
  if (ic) memcpy(ic, c->fr_ic, 4 * n);
  if (jc) memcpy(jc, c->fr_jc, 4 * n);
  if (kc) memcpy(kc, c->fr_kc, 4 * n);
  if (coeffD) memcpy(coeffD, c->fr_cD, 64 * n);
  if (bi) memcpy(bi, c->fr_bi, 24 * n);
  if (body) memcpy(body, c->fr_body, 4 * n);
  if (elem) memcpy(elem, c->fr_elem, 4 * n);
}

void or_geom_get_dirichlet(const or_ctx *c, double *coeffD) {
  memcpy(coeffD, c->coeffD, sizeof(double) * 8 * (size_t)c->nghost);
}

void or_geom_get_iblank(const or_ctx *c, int8_t *iblank, int8_t *fresh) {
  long nf = or_field_size(c);
  if (iblank) memcpy(iblank, c->iblank, nf);
  if (fresh) memcpy(fresh, c->fresh, nf);
}

void or_geom_get_ghost_geometry(const or_ctx *c, int32_t *ic, int32_t *jc, int32_t *kc, double *coeff, double *probe,
                                double *bi, double *ip, int32_t *elem, int32_t *body) {
  size_t n = (size_t)c->nghost;
  if (ic) memcpy(ic, c->ic, 4 * n);
  if (jc) memcpy(jc, c->jc, 4 * n);
  if (kc) memcpy(kc, c->kc, 4 * n);
  if (coeff) memcpy(coeff, c->coeff, 64 * n);
  if (probe) memcpy(probe, c->probe, 8 * n);
  if (bi) memcpy(bi, c->bi, 24 * n);
  if (ip) memcpy(ip, c->ipt, 24 * n);
  if (elem) memcpy(elem, c->elem, 4 * n);
  if (body)
    for (size_t q = 0; q < n; q++) body[q] = c->bodynum[OR_IDX(c, c->ig[q], c->jg[q], c->kg[q])];
}
