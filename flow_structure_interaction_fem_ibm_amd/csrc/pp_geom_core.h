// Per-point routines of the geometry pipeline, shared by the host build
// (pp_geom.cpp, pp_geom_tri.cpp) and the device build (pp_geom_dev.hip):
// each is __host__ __device__, reads plain-pointer views of the grid and of
// the triangulated bodies, and evaluates every value that reaches a
// decision or an output in the reference's operation order with
// contraction off, so the two builds give the same bits wherever the math
// library calls they make (sqrt, and acos / cos near the angle thresholds
// only) agree.
//
//   dot_norm          search_vertex_dotNorm (BOUNDARY_IBLANK_FAST.F90:504-1024)
//   tri_intercept     GCM_calc_bodyIntercept_Unstruc (GCM_INTERCEPT_KERNELS_UNSTRUC.F90:26-398)
//   ghost_stencil     GCM_set_internal_boundary image points and GCM_Calc_VanMatrixDN
//                     (GCM_SET_BOUNDARY.F90:141-257, 307-494)
//   fresh_stencil     GCM_set_freshcell (GCM_SET_FRESHCELL.F90:101-262)
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIP__
#include <hip/hip_runtime.h>
#define PPG_HD __host__ __device__
#else
#define PPG_HD
#endif
#define PPG_EXACT _Pragma("clang fp contract(off)")

namespace ppg {

// the angle constants of vector_angle's tests, evaluated once on the host
// (pi = 4 atan(1) as the reference forms it) and passed to device code
struct GeomConsts {
  double pi, tenth_pi, cos_tenth_pi, half_pi;
};

// Reference grid arrays x(0:n+1), xc(0:n+1), dx(0:n+1) per direction and
// the reference field indexing
struct GridV {
  int nxc, nyc, nzc;
  const double *X[3], *XC[3], *DX[3];
  long ex, exy;
  PPG_HD long r(int i, int j, int k) const { return (long)i + ex * j + exy * k; }
  PPG_HD double dxc(int d, int i) const { return XC[d][i] - XC[d][i - 1]; }
};

struct TreeNode {
  double lo[3], hi[3];
  int a, b, left, right;
};

// one triangulated body: markers (0-based arrays), elements (three 1-based
// marker ids each), element centroids and unit normals, the elements at each
// marker (CSR, ascending), and the marker kd-tree (nodes, permutation)
struct TriV {
  int nm, ne;
  const double *xm, *ym, *zm;
  const int *tri;
  const double *cx, *cy, *cz, *nx, *ny, *nz;
  const int *ve_ptr, *ve_idx;
  const TreeNode *nodes;
  const int *perm;
};

struct GhostOut {
  int ic = -1, jc = -1, kc = -1;
  int elem = 0;   // closest surface element (1-based), triangulated mode
  int body = -1;  // body of the ghost (0-based)
  double coeff[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // coeffGCMN (pressure, Neumann)
  double coeffD[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // coeffGCMD (velocity, Dirichlet)
  double probe = 0, bi[3] = {0, 0, 0}, ip[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
};

// One fresh cell's velocity interpolation (GCM_set_freshcell,
// GCM_SET_FRESHCELL.F90:86-262): the 2x2x2 stencil below the image point
// FIP = BI + (FC - BI) probeLengthNormalized and coeffGCMFreshD, the weights
// that give the value at the cell centre FC from the stencil values with the
// cell's own node replaced by the body intercept BI.
struct FreshOut {
  int i = 0, j = 0, k = 0;         // the fresh cell (1-based)
  int ic = -1, jc = -1, kc = -1;   // iFreshCellIndex..: the stencil's lower corner
  int body = -1, elem = 0;         // body (0-based) and surface element (1-based, triangulated mode)
  double bi[3] = {0, 0, 0};
  double coeffD[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// error codes (0 = success)
enum {
  TRI_ERR_NO_MARKER = 4,   // no unblocked marker among the 7 closest
  TRI_ERR_NO_ELEMENT = 5,  // no element selected for a cell or ghost
  TRI_ERR_SMALL = 6,       // too few markers/elements or rings, or grid centres not increasing
  TRI_ERR_MARKER_ID = 8,   // an element names a marker out of range
  TRI_ERR_ORIENT = 9,      // the origin (canonical bodies' orientation point) inside a body's bounding box
};
enum {
  GHOST_ERR_IP_CELL = 1,    // image point cell not found
  GHOST_ERR_LONG = 2,       // intercept longer than twice the cell diagonal
  GHOST_ERR_SINGULAR = 3,   // singular Vandermonde
  GHOST_ERR_NO_BODY = 7,    // ghost (or fresh) cell without a body
  FRESH_ERR_STENCIL = 10,   // the image point's cell is not the fresh cell or the one below it
  FRESH_ERR_AT_BI = 11,     // the body intercept coincides with the cell centre
};

namespace core {

constexpr int kSave = 7;   // markers saved by search_vertex_dotNorm
constexpr int kCheck = 3;  // markers tried by GCM_calc_bodyIntercept_Unstruc

PPG_HD inline int imax(int a, int b) { return a > b ? a : b; }
PPG_HD inline int imin(int a, int b) { return a < b ? a : b; }

// mo / dot / cross of the operation module (AMODULES.F90:176-182, 905-934)
PPG_HD inline double mo3(const double *a) {
  PPG_EXACT
  double s = 0.0;
  s = s + a[0] * a[0];
  s = s + a[1] * a[1];
  s = s + a[2] * a[2];
  return sqrt(s);
}
PPG_HD inline double dot3(const double *a, const double *b) {
  PPG_EXACT
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
PPG_HD inline void cross3(const double *a, const double *b, double *o) {
  PPG_EXACT
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
// cosine term of vector_angle (AMODULES.F90:968-988)
PPG_HD inline double cos_term(const double *a, const double *b) {
  PPG_EXACT
  return dot3(a, b) / (mo3(a) * mo3(b));
}
// vector_angle itself; its 2*pi - angle branch needs |a x b| < 0 and never
// applies
PPG_HD inline double angle_of(double c, const GeomConsts &K) {
  if (c < 1.0 && c > -1.0) return acos(c);
  if (c > 0.0) return 0.0;
  if (c < 0.0) return K.pi;
  return NAN;
}

// sign of get_distance_face (BOUNDARY_IBLANK_FAST.F90:3531-3550): +1, -1, or
// 0 for p == ref.  The sign of mo(v) cos(acos(c)) is read from c when
// |c| >= 1e-14 (far above acos's error), the full expression below.
PPG_HD inline int face_side(const double *n, const double *ref, const double *p, const GeomConsts &K) {
  PPG_EXACT
  if (p[0] == ref[0] && p[1] == ref[1] && p[2] == ref[2]) return 0;
  double v[3] = {p[0] - ref[0], p[1] - ref[1], p[2] - ref[2]};
  double c = cos_term(n, v);
  if (c >= 1e-14) return 1;
  if (c <= -1e-14) return -1;
  double d = mo3(v) * cos(angle_of(c, K));
  return d > 0.0 ? 1 : (d < 0.0 ? -1 : 0);
}

// vector_angle(a, b) < pi/10, read from the cosine away from the threshold
PPG_HD inline bool angle_below_tenth_pi(const double *a, const double *b, const GeomConsts &K) {
  PPG_EXACT
  const double c = cos_term(a, b);
  if (c > K.cos_tenth_pi + 1e-12) return true;
  if (c < K.cos_tenth_pi - 1e-12) return false;
  return angle_of(c, K) < K.tenth_pi;
}

// vector_angle(a, b) >= pi/2
PPG_HD inline bool angle_at_least_half_pi(const double *a, const double *b, const GeomConsts &K) {
  PPG_EXACT
  const double c = cos_term(a, b);
  if (c < -1e-14) return true;
  if (c > 1e-14) return false;
  return angle_of(c, K) >= K.half_pi;
}

// get_inter (BOUNDARY_IBLANK_FAST.F90:3513-3529): the line-plane
// intersection p1 + lhs1/(lhs1-lhs2)(p2-p1)
PPG_HD inline void line_plane(const double *n, const double *ref, const double *p1, const double *p2, double *out) {
  PPG_EXACT
  double lhs1 = n[0] * (p1[0] - ref[0]) + n[1] * (p1[1] - ref[1]) + n[2] * (p1[2] - ref[2]);
  double lhs2 = n[0] * (p2[0] - ref[0]) + n[1] * (p2[1] - ref[1]) + n[2] * (p2[2] - ref[2]);
  const double eps = (double)1e-10f;
  if (fabs(lhs1) < eps) {
    for (int d = 0; d < 3; d++) out[d] = p1[d];
  } else if (fabs(lhs2) < eps) {
    for (int d = 0; d < 3; d++) out[d] = p2[d];
  } else {
    double t = lhs1 / (lhs1 - lhs2);
    for (int d = 0; d < 3; d++) out[d] = p1[d] + t * (p2[d] - p1[d]);
  }
}

// pointInTriangle (BOUNDARY_IBLANK_FAST.F90:3862-3910); the second and third
// cross products are scaled by the first one's norm, as there
PPG_HD inline bool in_triangle(const double *a, const double *b, const double *cc, const double *p,
                               const GeomConsts &K) {
  PPG_EXACT
  double v1[3], v2[3], v3[3], w1[3], w2[3], w3[3], c1[3], c2[3], c3[3];
  for (int d = 0; d < 3; d++) {
    v1[d] = b[d] - a[d]; v2[d] = cc[d] - b[d]; v3[d] = a[d] - cc[d];
    w1[d] = p[d] - a[d]; w2[d] = p[d] - b[d]; w3[d] = p[d] - cc[d];
  }
  cross3(v1, w1, c1); cross3(v2, w2, c2); cross3(v3, w3, c3);
  const double m1 = mo3(c1), m2 = mo3(c2), m3 = mo3(c3);
  const double eps = (double)1e-20f;
  if (m1 < eps || m2 < eps || m3 < eps) return false;
  for (int d = 0; d < 3; d++) { c1[d] = c1[d] / m1; c2[d] = c2[d] / m1; c3[d] = c3[d] / m1; }
  return angle_below_tenth_pi(c1, c2, K) && angle_below_tenth_pi(c2, c3, K);
}

PPG_HD inline void marker(const TriV &B, int m1, double *p) {
  p[0] = B.xm[m1 - 1]; p[1] = B.ym[m1 - 1]; p[2] = B.zm[m1 - 1];
}
PPG_HD inline void normal(const TriV &B, int e, double *n) { n[0] = B.nx[e]; n[1] = B.ny[e]; n[2] = B.nz[e]; }
PPG_HD inline void centroid(const TriV &B, int e, double *r) { r[0] = B.cx[e]; r[1] = B.cy[e]; r[2] = B.cz[e]; }
PPG_HD inline bool has_marker(const TriV &B, int e, int m1) {
  return B.tri[3 * e] == m1 || B.tri[3 * e + 1] == m1 || B.tri[3 * e + 2] == m1;
}

// check_BIInsideTriangle (BOUNDARY_IBLANK_FAST.F90:1028-1159): the foot of
// P on the element's plane and the Heron areas of the element and of the
// three triangles the foot makes with its edges
PPG_HD inline double seg_len(const double *u, const double *v) {
  PPG_EXACT
  return sqrt((u[0] - v[0]) * (u[0] - v[0]) + (u[1] - v[1]) * (u[1] - v[1]) + (u[2] - v[2]) * (u[2] - v[2]));
}
PPG_HD inline void bi_inside(const TriV &B, int e, const double *P, double *bi, double &area123, double &areadiff,
                             double &dist_in) {
  PPG_EXACT
  double p1[3], p2[3], p3[3], n[3];
  marker(B, B.tri[3 * e], p1); marker(B, B.tri[3 * e + 1], p2); marker(B, B.tri[3 * e + 2], p3);
  normal(B, e, n);
  const double pc = -n[0] * p1[0] - n[1] * p1[1] - n[2] * p1[2];
  const double dtp = -(n[0] * P[0] + n[1] * P[1] + n[2] * P[2] + pc);
  dist_in = fabs(dtp);
  for (int d = 0; d < 3; d++) bi[d] = P[d] + n[d] * dtp;
  const double s12 = seg_len(p2, p1), s23 = seg_len(p3, p2), s31 = seg_len(p1, p3);
  const double s14 = seg_len(bi, p1), s24 = seg_len(bi, p2), s34 = seg_len(bi, p3);
  const double h123 = 0.5 * (s12 + s23 + s31), h124 = 0.5 * (s12 + s24 + s14);
  const double h234 = 0.5 * (s23 + s24 + s34), h314 = 0.5 * (s31 + s34 + s14);
  const double a123 = sqrt(h123 * (h123 - s12) * (h123 - s23) * (h123 - s31));
  const double a124 = sqrt(h124 * (h124 - s12) * (h124 - s24) * (h124 - s14));
  const double a234 = sqrt(h234 * (h234 - s23) * (h234 - s24) * (h234 - s34));
  const double a314 = sqrt(h314 * (h314 - s31) * (h314 - s34) * (h314 - s14));
  area123 = a123;
  areadiff = a124 + a234 + a314 - a123;
}

// calc_BIOutsideTriangle (BOUNDARY_IBLANK_FAST.F90:1163-1435): angle between
// (P - closest marker) and the element's outward normal
PPG_HD inline double bi_outside(const TriV &B, int e, const double *P, int cmarker, const GeomConsts &K) {
  PPG_EXACT
  double cm[3], v[3], n[3];
  marker(B, cmarker, cm);
  for (int d = 0; d < 3; d++) v[d] = P[d] - cm[d];
  normal(B, e, n);
  for (int d = 0; d < 3; d++) n[d] = -n[d];
  return angle_of(cos_term(v, n), K);
}

// the k best (squared distance, index) pairs seen, ascending, lower index
// first on equal distances
struct KBest {
  int k, n;
  double d[kSave];
  int id[kSave];
  PPG_HD bool full() const { return n == k; }
  PPG_HD bool better(double d2, int m) const {
    return !full() || d2 < d[n - 1] || (d2 == d[n - 1] && m < id[n - 1]);
  }
  PPG_HD void insert(double d2, int m) {
    int p = full() ? k - 1 : n++;
    while (p > 0 && (d2 < d[p - 1] || (d2 == d[p - 1] && m < id[p - 1]))) {
      d[p] = d[p - 1]; id[p] = id[p - 1];
      p--;
    }
    d[p] = d2; id[p] = m;
  }
};

// squared distance of marker m (0-based) to P, as the reference forms it
PPG_HD inline double marker_d2(const TriV &B, int m, const double *P) {
  PPG_EXACT
  double dx = B.xm[m] - P[0], dy = B.ym[m] - P[1], dz = B.zm[m] - P[2];
  return dx * dx + dy * dy + dz * dz;
}

// a rounded squared distance no larger than marker_d2 of any marker in the
// box (rounding is monotone)
PPG_HD inline double box_d2(const TreeNode &nd, const double *P) {
  PPG_EXACT
  double g[3];
  for (int d = 0; d < 3; d++) {
    if (P[d] < nd.lo[d]) g[d] = nd.lo[d] - P[d];
    else if (P[d] > nd.hi[d]) g[d] = nd.hi[d] - P[d];
    else g[d] = 0.0;
  }
  return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
}

// 1-based ids of the k <= kSave markers closest to P (repeated MINLOC order:
// smallest squared distance first, lowest index on ties), by the kd-tree
PPG_HD inline void closest_markers(const TriV &B, const double *P, int k, int *out) {
  KBest kb;
  kb.k = k;
  kb.n = 0;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const TreeNode &nd = B.nodes[stack[--sp]];
    const double lb = box_d2(nd, P);
    if (kb.full() && lb > kb.d[k - 1]) continue;
    if (nd.left < 0) {
      for (int q = nd.a; q < nd.b; q++) {
        int m = B.perm[q];
        double d2 = marker_d2(B, m, P);
        if (kb.better(d2, m)) kb.insert(d2, m);
      }
      continue;
    }
    const double dl = box_d2(B.nodes[nd.left], P), dr = box_d2(B.nodes[nd.right], P);
    // push the farther child first so the nearer one is visited first
    if (dl <= dr) { stack[sp++] = nd.right; stack[sp++] = nd.left; }
    else { stack[sp++] = nd.left; stack[sp++] = nd.right; }
  }
  for (int s = 0; s < k; s++) out[s] = kb.id[s] + 1;
}

// check_if_edge_marker (BOUNDARY_IBLANK_FAST.F90:3580-3703)
PPG_HD inline bool edge_marker(const TriV &B, int m1, const GeomConsts &K) {
  const int a0 = B.ve_ptr[m1 - 1], a1 = B.ve_ptr[m1];
  for (int a = a0; a < a1 - 1; a++)
    for (int b = a + 1; b < a1; b++) {
      double r[3], q[3];
      normal(B, B.ve_idx[a], r); normal(B, B.ve_idx[b], q);
      for (int d = 0; d < 3; d++) { r[d] = -r[d]; q[d] = -q[d]; }
      if (angle_at_least_half_pi(r, q, K)) return true;
    }
  return false;
}

// Does an element at one of the markers cv[0..kSave) other than those at
// marker cv[s] block the segment from P to marker cv[s]?  The reference
// scans the union of the elements at the saved markers and stops at the
// first blocking one; whether one blocks does not depend on the order or on
// repeats, so the lists are scanned marker by marker.
PPG_HD inline bool blocked(const TriV &B, const int *cv, int s, const double *P, const double *pm,
                           const GeomConsts &K) {
  PPG_EXACT
  for (int t = 0; t < kSave; t++)
    for (int q = B.ve_ptr[cv[t] - 1]; q < B.ve_ptr[cv[t]]; q++) {
      const int e = B.ve_idx[q];
      if (has_marker(B, e, cv[s])) continue;
      double n[3], r[3];
      normal(B, e, n); centroid(B, e, r);
      const int sc = face_side(n, r, P, K), sm = face_side(n, r, pm, K);
      if (sc * sm > 0) continue;
      double po[3], a[3], b[3], cc[3];
      line_plane(n, r, P, pm, po);
      marker(B, B.tri[3 * e], a); marker(B, B.tri[3 * e + 1], b); marker(B, B.tri[3 * e + 2], cc);
      if (in_triangle(a, b, cc, po, K)) return true;
    }
  return false;
}

// search_vertex_dotNorm (BOUNDARY_IBLANK_FAST.F90:504-1024) for one cell and
// one body: 0 and dotnorm, or TRI_ERR_NO_MARKER / TRI_ERR_NO_ELEMENT
PPG_HD inline int dot_norm(const TriV &B, const double *P, double &dotnorm, const GeomConsts &K) {
  PPG_EXACT
  dotnorm = -1.0;
  int cv[kSave];
  closest_markers(B, P, kSave, cv);
  int choice = -1;
  for (int s = 0; s < kSave && choice < 0; s++) {
    double pm[3];
    marker(B, cv[s], pm);
    if (!blocked(B, cv, s, P, pm, K)) choice = s;
  }
  if (choice < 0) return TRI_ERR_NO_MARKER;
  const int cm = cv[choice];
  if (edge_marker(B, cm, K)) return 0;
  double dmin = 1.0e16, dist_in = 1.0e16;
  int closest = -1;
  bool inside = false;
  for (int q = B.ve_ptr[cm - 1]; q < B.ve_ptr[cm]; q++) {
    const int e = B.ve_idx[q];
    double bi[3], a123, adiff, din;
    bi_inside(B, e, P, bi, a123, adiff, din);
    if (fabs(adiff) < 1.0e-4 * a123) {
      inside = true;
      if (din <= dist_in) { dist_in = din; closest = e; }
    } else if (!inside) {
      const double dbe = bi_outside(B, e, P, cm, K);
      if (dbe <= dmin) { dmin = dbe; closest = e; }
    }
  }
  if (closest < 0) return TRI_ERR_NO_ELEMENT;
  dotnorm = (P[0] - B.cx[closest]) * B.nx[closest] + (P[1] - B.cy[closest]) * B.ny[closest] +
            (P[2] - B.cz[closest]) * B.nz[closest];
  return 0;
}

// squared distance from P to the body's closest marker
PPG_HD inline double nearest_marker_d2(const TriV &B, const double *P) {
  int m;
  closest_markers(B, P, 1, &m);
  return marker_d2(B, m - 1, P);
}

// GCM_calc_bodyIntercept_Unstruc for the point P: body intercept, inward
// element normal and element (1-based); 0 or TRI_ERR_NO_ELEMENT
PPG_HD inline int tri_intercept(const TriV &B, const double *P, double *bi, double *nrm, int &elem,
                                const GeomConsts &K) {
  PPG_EXACT
  int cn[kCheck];
  closest_markers(B, P, kCheck, cn);
  double bit[kCheck][3];
  int cel[kCheck];
  for (int s = 0; s < kCheck; s++) {
    int closest = -1;
    bool done = false;
    double dmin = 1.0e16, xb[3] = {0.0, 0.0, 0.0};
    for (int q = B.ve_ptr[cn[s] - 1]; q < B.ve_ptr[cn[s]]; q++) {
      const int e = B.ve_idx[q];
      double t[3], a123, adiff, din;
      bi_inside(B, e, P, t, a123, adiff, din);
      if (fabs(adiff) < 1.0e-4 * a123) {
        for (int d = 0; d < 3; d++) xb[d] = t[d];
        closest = e;
        done = true;
        break;
      }
      const double dbe = bi_outside(B, e, P, cn[s], K);
      if (dbe <= dmin) {
        dmin = dbe;
        closest = e;
        for (int d = 0; d < 3; d++) xb[d] = t[d];
      }
    }
    if (closest < 0) return TRI_ERR_NO_ELEMENT;
    if (!done) {
      // the closest point on the element's two edges through the marker
      double cvp[3];
      marker(B, cn[s], cvp);
      const int t1 = B.tri[3 * closest], t2 = B.tri[3 * closest + 1], t3 = B.tri[3 * closest + 2];
      int na, nb;
      if (t1 == cn[s]) { na = t2; nb = t3; }
      else if (t2 == cn[s]) { na = t3; nb = t1; }
      else { na = t1; nb = t2; }
      const double v01[3] = {P[0] - cvp[0], P[1] - cvp[1], P[2] - cvp[2]};
      double best = 1.0e16, ei[3] = {cvp[0], cvp[1], cvp[2]};
      for (int ed = 0; ed < 2; ed++) {
        double q[3];
        marker(B, ed == 0 ? na : nb, q);
        double v12[3] = {q[0] - cvp[0], q[1] - cvp[1], q[2] - cvp[2]};
        const double mag = sqrt(v12[0] * v12[0] + v12[1] * v12[1] + v12[2] * v12[2]);
        const double inv = 1.0 / mag;
        for (int d = 0; d < 3; d++) v12[d] = v12[d] * inv;
        const double proj = v01[0] * v12[0] + v01[1] * v12[1] + v01[2] * v12[2];
        double et[3];
        if (proj < 0.0 || proj > mag) {
          for (int d = 0; d < 3; d++) et[d] = cvp[d];
        } else {
          for (int d = 0; d < 3; d++) et[d] = cvp[d] + proj * v12[d];
        }
        const double dd =
            (et[0] - P[0]) * (et[0] - P[0]) + (et[1] - P[1]) * (et[1] - P[1]) + (et[2] - P[2]) * (et[2] - P[2]);
        if (dd < best) {
          best = dd;
          for (int d = 0; d < 3; d++) ei[d] = et[d];
        }
      }
      for (int d = 0; d < 3; d++) xb[d] = ei[d];
    }
    for (int d = 0; d < 3; d++) bit[s][d] = xb[d];
    cel[s] = closest;
  }
  int sel = 0;
  double dsel = 0.0;
  for (int s = 0; s < kCheck; s++) {
    const double d = (bit[s][0] - P[0]) * (bit[s][0] - P[0]) + (bit[s][1] - P[1]) * (bit[s][1] - P[1]) +
                     (bit[s][2] - P[2]) * (bit[s][2] - P[2]);
    if (s == 0 || d < dsel) { dsel = d; sel = s; }
  }
  for (int d = 0; d < 3; d++) bi[d] = bit[sel][d];
  const int e = cel[sel];
  nrm[0] = B.nx[e]; nrm[1] = B.ny[e]; nrm[2] = B.nz[e];
  elem = e + 1;
  return 0;
}

// last i in [0, n-1] with (c[i]-b)(c[i+1]-b) <= 0, or -1 (c strictly
// increasing, n + 1 entries)
PPG_HD inline int last_bracket(const double *c, int n, double b) {
  // p: last index in [0, n] with c[p] <= b (binary search)
  int lo = 0, hi = n + 1;  // first index with c > b lies in [lo, hi]
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (c[mid] <= b) lo = mid + 1;
    else hi = mid;
  }
  const int p = lo - 1;
  if (p < 0) return -1;
  if (p == n) return c[n] == b ? n - 1 : -1;
  return p;
}

// Solve M^T c = beta for an 8x8 Vandermonde matrix M (partial pivoting): c
// are the coefficients of GCM_Calc_VanMatrixDN, the row beta of M^-1
PPG_HD inline bool solve8T(const double M[8][8], const double beta[8], double c[8]) {
  PPG_EXACT
  double A[8][9];
  for (int i = 0; i < 8; i++) {
    for (int j = 0; j < 8; j++) A[i][j] = M[j][i];
    A[i][8] = beta[i];
  }
  for (int k = 0; k < 8; k++) {
    int p = k;
    for (int i = k + 1; i < 8; i++)
      if (fabs(A[i][k]) > fabs(A[p][k])) p = i;
    if (A[p][k] == 0.0) return false;
    if (p != k)
      for (int j = 0; j < 9; j++) {
        const double t = A[k][j];
        A[k][j] = A[p][j];
        A[p][j] = t;
      }
    for (int i = k + 1; i < 8; i++) {
      double f = A[i][k] / A[k][k];
      for (int j = k; j < 9; j++) A[i][j] -= f * A[k][j];
    }
  }
  for (int i = 7; i >= 0; i--) {
    double s = A[i][8];
    for (int j = i + 1; j < 8; j++) s -= A[i][j] * c[j];
    c[i] = s / A[i][i];
  }
  return true;
}

// corner r of a 2x2x2 stencil: offsets in i, j, k
PPG_HD inline int inc_i(int r) { return (r == 2 || r == 3 || r == 6 || r == 7) ? 1 : 0; }
PPG_HD inline int inc_j(int r) { return (r == 1 || r == 2 || r == 5 || r == 6) ? 1 : 0; }
PPG_HD inline int inc_k(int r) { return r >= 4 ? 1 : 0; }

// Image point, its interpolation cell and the Neumann and Dirichlet
// coefficients of one ghost (GCM_SET_BOUNDARY.F90:141-257, 399-491).
// Returns 0, GHOST_ERR_IP_CELL, GHOST_ERR_LONG or GHOST_ERR_SINGULAR.
PPG_HD inline int ghost_stencil(const GridV &G, int iG, int jG, int kG, const double *bi, const double *nrm,
                                double probe_norm, GhostOut &o) {
  PPG_EXACT
  double xg = G.XC[0][iG], yg = G.XC[1][jG], zg = G.XC[2][kG];
  double ds = sqrt((xg - bi[0]) * (xg - bi[0]) + (yg - bi[1]) * (yg - bi[1]) + (zg - bi[2]) * (zg - bi[2]));
  double diag = sqrt(G.dxc(0, iG) * G.dxc(0, iG) + G.dxc(1, jG) * G.dxc(1, jG) + G.dxc(2, kG) * G.dxc(2, kG));
  if (ds > diag && ds / diag > 2.0) return GHOST_ERR_LONG;
  double sx = (bi[0] - xg) / ds, sy = (bi[1] - yg) / ds, sz = (bi[2] - zg) / ds;
  double probe = ds * probe_norm;
  double ip[3] = {xg + sx * probe, yg + sy * probe, zg + sz * probe};
  int idx[3] = {iG, jG, kG};
  int cell[3] = {-1, -1, -1};
  int nn[3] = {G.nxc + 1, G.nyc + 1, G.nzc + 1};
  for (int d = 0; d < 3; d++) {
    int range = (int)lround(probe / G.DX[d][idx[d]]) + 1;
    int lo = imax(idx[d] - range, 0), hi = imin(idx[d] + range, nn[d]);
    for (int i = lo; i <= hi; i++)
      if (G.XC[d][i] <= ip[d] && G.XC[d][i + 1] > ip[d]) cell[d] = i;
  }
  if (cell[0] < 0 || cell[1] < 0 || cell[2] < 0) return GHOST_ERR_IP_CELL;
  double M[8][8], MD[8][8];
  for (int r = 0; r < 8; r++) {
    int i = cell[0] + inc_i(r), j = cell[1] + inc_j(r), k = cell[2] + inc_k(r);
    double x1 = G.XC[0][i], x2 = G.XC[1][j], x3 = G.XC[2][k];
    if (i == iG && j == jG && k == kG) {
      x1 = bi[0]; x2 = bi[1]; x3 = bi[2];
      // velocity (Dirichlet) row: the value at the body intercept
      MD[r][0] = x1 * x2 * x3; MD[r][1] = x1 * x2; MD[r][2] = x1 * x3; MD[r][3] = x2 * x3;
      MD[r][4] = x1; MD[r][5] = x2; MD[r][6] = x3; MD[r][7] = 1.0;
      double n1 = nrm[0], n2 = nrm[1], n3 = nrm[2];
      M[r][0] = n1 * x2 * x3 + n2 * x1 * x3 + n3 * x1 * x2;
      M[r][1] = n1 * x2 + n2 * x1;
      M[r][2] = n1 * x3 + n3 * x1;
      M[r][3] = n2 * x3 + n3 * x2;
      M[r][4] = n1; M[r][5] = n2; M[r][6] = n3; M[r][7] = 0.0;
    } else {
      M[r][0] = x1 * x2 * x3; M[r][1] = x1 * x2; M[r][2] = x1 * x3; M[r][3] = x2 * x3;
      M[r][4] = x1; M[r][5] = x2; M[r][6] = x3; M[r][7] = 1.0;
      for (int q = 0; q < 8; q++) MD[r][q] = M[r][q];
    }
  }
  double beta[8] = {ip[0] * ip[1] * ip[2], ip[0] * ip[1], ip[0] * ip[2], ip[1] * ip[2], ip[0], ip[1], ip[2], 1.0};
  if (!solve8T(M, beta, o.coeff)) return GHOST_ERR_SINGULAR;
  if (!solve8T(MD, beta, o.coeffD)) return GHOST_ERR_SINGULAR;
  o.ic = cell[0]; o.jc = cell[1]; o.kc = cell[2];
  o.probe = probe;
  for (int d = 0; d < 3; d++) { o.bi[d] = bi[d]; o.ip[d] = ip[d]; o.nrm[d] = nrm[d]; }
  return 0;
}

// GCM_set_freshcell for one cell o.i, o.j, o.k whose body intercept is bi
// (GCM_SET_FRESHCELL.F90:101-262): the image point, the stencil search over
// the cell and its two neighbours in each direction (clamped to 1..n-1 of
// the cell centres) with the reference's test that the stencil corner is the
// cell or the one below it, and the Dirichlet Vandermonde weights at the
// cell centre.  Returns 0, GHOST_ERR_SINGULAR, FRESH_ERR_STENCIL or
// FRESH_ERR_AT_BI.
PPG_HD inline int fresh_stencil(const GridV &G, const double *bi, double probe_norm, FreshOut &o) {
  PPG_EXACT
  const int idx[3] = {o.i, o.j, o.k}, nn[3] = {G.nxc, G.nyc, G.nzc};
  const double fc[3] = {G.XC[0][o.i], G.XC[1][o.j], G.XC[2][o.k]};
  const double ds = sqrt((fc[1] - bi[1]) * (fc[1] - bi[1]) + (fc[0] - bi[0]) * (fc[0] - bi[0]) +
                         (fc[2] - bi[2]) * (fc[2] - bi[2]));
  if (!(ds > 0.0)) return FRESH_ERR_AT_BI;
  const double dsinv = 1.0 / ds, probe = ds * probe_norm;
  int cell[3];
  for (int d = 0; d < 3; d++) {
    const double slope = (fc[d] - bi[d]) * dsinv;
    const double fip = bi[d] + slope * probe;
    const int lo = imax(idx[d] - 1, 1), hi = imin(idx[d] + 1, nn[d]);
    cell[d] = -1;
    for (int i = lo; i <= hi; i++)
      if (G.XC[d][i] <= fip && G.XC[d][i + 1] > fip) cell[d] = i;
    if (cell[d] < idx[d] - 1 || cell[d] > idx[d]) return FRESH_ERR_STENCIL;
  }
  double MD[8][8];
  for (int r = 0; r < 8; r++) {
    const int i = cell[0] + inc_i(r), j = cell[1] + inc_j(r), k = cell[2] + inc_k(r);
    double x1 = G.XC[0][i], x2 = G.XC[1][j], x3 = G.XC[2][k];
    if (i == o.i && j == o.j && k == o.k) { x1 = bi[0]; x2 = bi[1]; x3 = bi[2]; }
    MD[r][0] = x1 * x2 * x3; MD[r][1] = x1 * x2; MD[r][2] = x1 * x3; MD[r][3] = x2 * x3;
    MD[r][4] = x1; MD[r][5] = x2; MD[r][6] = x3; MD[r][7] = 1.0;
  }
  const double beta[8] = {fc[0] * fc[1] * fc[2], fc[0] * fc[1], fc[0] * fc[2], fc[1] * fc[2],
                          fc[0], fc[1], fc[2], 1.0};
  if (!solve8T(MD, beta, o.coeffD)) return GHOST_ERR_SINGULAR;
  o.ic = cell[0]; o.jc = cell[1]; o.kc = cell[2];
  for (int d = 0; d < 3; d++) o.bi[d] = bi[d];
  return 0;
}

}  // namespace core
}  // namespace ppg
