// Triangulated-surface geometry of canonical ELLIPSOID bodies under GCM: the
// reference's marker mesh, its dot-normal iblank and its closest-element body
// intercepts, host C++ (threaded).
//
//   surface mesh  initialize_marker CASE(ELLIPSOID) (BOUNDARY_MARKER_INIT.F90:115-205,
//                 n_phi MAIN.F90:1156-1166), orientation determine_norm_dir_unstruc
//                 (BOUNDARY_SET_ARCLENGTH_NORM.F90:200-320), unit normals, areas and
//                 centroids calculate_arclength_norm_ds (:69-167)
//   iblank        set_iblank_canonical_body_fast (BOUNDARY_IBLANK_FAST.F90:2-501):
//                 band = cells in any element's bounding box; a band cell is blank when
//                 search_vertex_dotNorm (:504-1024) gives dot >= 0 for some body (bodies
//                 in order); the other cells by ray fills along i, then j, then k, each
//                 overwriting the previous one
//   intercepts    GCM_calc_bodyIntercept_Unstruc (GCM_INTERCEPT_KERNELS_UNSTRUC.F90:26-398)
//
// Every value that reaches a decision or an output is computed with the
// reference's operation order (contraction off).  Three things make it fast
// without changing a result:
//   - the k closest markers come from a kd-tree whose pruning bound is a
//     rounded squared distance that can only be smaller than the rounded
//     squared distance of any marker in the pruned box (rounding is
//     monotone), with ties kept in index order;
//   - the sign of get_distance_face (mo(v) cos(acos(c))) is read from c when
//     |c| >= 1e-14, and the angle tests acos(c) < pi/10 and acos(c) >= pi/2
//     from c away from their thresholds (margins far above acos's error); the
//     full expressions are evaluated near the thresholds;
//   - band cells, ray-fill lines and ghosts run on host threads.
// Decisions where the reference leaves a value undefined (orientation point,
// vector_angle of a zero vector, an unassigned edge flag, bodyNum of the
// k = 0 layer) are taken as follows: the orientation point is the origin and
// a body whose bounding box holds it is rejected; a zero vector gives NaN; a
// marker with fewer than two elements is not an edge; the k = 0 layer has no
// body.  get_inter is the line-plane intersection p1 + lhs1/(lhs1-lhs2)(p2-p1)
// (algebraically the reference's expanded form).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "pp_geom.h"

#pragma STDC FP_CONTRACT OFF

namespace ppg {

namespace {

constexpr int kUndecided = 1;
constexpr int kSave = 7;   // markers saved by search_vertex_dotNorm
constexpr int kCheck = 3;  // markers tried by GCM_calc_bodyIntercept_Unstruc

inline double pi_ref() { return 4.0 * std::atan(1.0); }

// mo / dot / cross of the operation module (AMODULES.F90:176-182, 905-934)
inline double mo3(const double *a) {
  double s = 0.0;
  s = s + a[0] * a[0];
  s = s + a[1] * a[1];
  s = s + a[2] * a[2];
  return std::sqrt(s);
}
inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
// cosine term of vector_angle (AMODULES.F90:968-988)
inline double cos_term(const double *a, const double *b) { return dot3(a, b) / (mo3(a) * mo3(b)); }
// vector_angle itself; its 2*pi - angle branch needs |a x b| < 0 and never
// applies
inline double angle_of(double c) {
  if (c < 1.0 && c > -1.0) return std::acos(c);
  if (c > 0.0) return 0.0;
  if (c < 0.0) return pi_ref();
  return NAN;
}

// sign of get_distance_face (BOUNDARY_IBLANK_FAST.F90:3531-3550): +1, -1, or
// 0 for p == ref
inline int face_side(const double *n, const double *ref, const double *p) {
  if (p[0] == ref[0] && p[1] == ref[1] && p[2] == ref[2]) return 0;
  double v[3] = {p[0] - ref[0], p[1] - ref[1], p[2] - ref[2]};
  double c = cos_term(n, v);
  if (c >= 1e-14) return 1;
  if (c <= -1e-14) return -1;
  double d = mo3(v) * std::cos(angle_of(c));
  return d > 0.0 ? 1 : (d < 0.0 ? -1 : 0);
}

// vector_angle(a, b) < pi/10
inline bool angle_below_tenth_pi(const double *a, const double *b) {
  const double c = cos_term(a, b);
  static const double thr = pi_ref() / 10.0, cthr = std::cos(pi_ref() / 10.0);
  if (c > cthr + 1e-12) return true;
  if (c < cthr - 1e-12) return false;
  return angle_of(c) < thr;
}

// vector_angle(a, b) >= pi/2
inline bool angle_at_least_half_pi(const double *a, const double *b) {
  const double c = cos_term(a, b);
  if (c < -1e-14) return true;
  if (c > 1e-14) return false;
  return angle_of(c) >= pi_ref() / 2.0;
}

// get_inter (BOUNDARY_IBLANK_FAST.F90:3513-3529)
inline void line_plane(const double *n, const double *ref, const double *p1, const double *p2, double *out) {
  double lhs1 = n[0] * (p1[0] - ref[0]) + n[1] * (p1[1] - ref[1]) + n[2] * (p1[2] - ref[2]);
  double lhs2 = n[0] * (p2[0] - ref[0]) + n[1] * (p2[1] - ref[1]) + n[2] * (p2[2] - ref[2]);
  const double eps = (double)1e-10f;
  if (std::fabs(lhs1) < eps) {
    for (int d = 0; d < 3; d++) out[d] = p1[d];
  } else if (std::fabs(lhs2) < eps) {
    for (int d = 0; d < 3; d++) out[d] = p2[d];
  } else {
    double t = lhs1 / (lhs1 - lhs2);
    for (int d = 0; d < 3; d++) out[d] = p1[d] + t * (p2[d] - p1[d]);
  }
}

// pointInTriangle (BOUNDARY_IBLANK_FAST.F90:3862-3910); the second and third
// cross products are scaled by the first one's norm, as there
bool in_triangle(const double *a, const double *b, const double *cc, const double *p) {
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
  return angle_below_tenth_pi(c1, c2) && angle_below_tenth_pi(c2, c3);
}

struct Surf {
  const TriBody &B;
  void marker(int m1, double *p) const { p[0] = B.xm[m1 - 1]; p[1] = B.ym[m1 - 1]; p[2] = B.zm[m1 - 1]; }
  void normal(int e, double *n) const { n[0] = B.nx[e]; n[1] = B.ny[e]; n[2] = B.nz[e]; }
  void centroid(int e, double *r) const { r[0] = B.cx[e]; r[1] = B.cy[e]; r[2] = B.cz[e]; }
  bool has(int e, int m1) const { return B.tri[3 * e] == m1 || B.tri[3 * e + 1] == m1 || B.tri[3 * e + 2] == m1; }
};

// check_BIInsideTriangle (BOUNDARY_IBLANK_FAST.F90:1028-1159)
void bi_inside(const Surf &S, int e, const double *P, double *bi, double &area123, double &areadiff,
               double &dist_in) {
  double p1[3], p2[3], p3[3], n[3];
  S.marker(S.B.tri[3 * e], p1); S.marker(S.B.tri[3 * e + 1], p2); S.marker(S.B.tri[3 * e + 2], p3);
  S.normal(e, n);
  const double pc = -n[0] * p1[0] - n[1] * p1[1] - n[2] * p1[2];
  const double dtp = -(n[0] * P[0] + n[1] * P[1] + n[2] * P[2] + pc);
  dist_in = std::fabs(dtp);
  for (int d = 0; d < 3; d++) bi[d] = P[d] + n[d] * dtp;
  auto len = [](const double *u, const double *v) {
    return std::sqrt((u[0] - v[0]) * (u[0] - v[0]) + (u[1] - v[1]) * (u[1] - v[1]) + (u[2] - v[2]) * (u[2] - v[2]));
  };
  const double s12 = len(p2, p1), s23 = len(p3, p2), s31 = len(p1, p3);
  const double s14 = len(bi, p1), s24 = len(bi, p2), s34 = len(bi, p3);
  const double h123 = 0.5 * (s12 + s23 + s31), h124 = 0.5 * (s12 + s24 + s14);
  const double h234 = 0.5 * (s23 + s24 + s34), h314 = 0.5 * (s31 + s34 + s14);
  const double a123 = std::sqrt(h123 * (h123 - s12) * (h123 - s23) * (h123 - s31));
  const double a124 = std::sqrt(h124 * (h124 - s12) * (h124 - s24) * (h124 - s14));
  const double a234 = std::sqrt(h234 * (h234 - s23) * (h234 - s24) * (h234 - s34));
  const double a314 = std::sqrt(h314 * (h314 - s31) * (h314 - s34) * (h314 - s14));
  area123 = a123;
  areadiff = a124 + a234 + a314 - a123;
}

// calc_BIOutsideTriangle (BOUNDARY_IBLANK_FAST.F90:1163-1435): angle between
// (P - closest marker) and the element's outward normal
double bi_outside(const Surf &S, int e, const double *P, int cmarker) {
  double cm[3], v[3], n[3];
  S.marker(cmarker, cm);
  for (int d = 0; d < 3; d++) v[d] = P[d] - cm[d];
  S.normal(e, n);
  for (int d = 0; d < 3; d++) n[d] = -n[d];
  return angle_of(cos_term(v, n));
}

// ---- kd-tree of markers ------------------------------------------------
void tree_build(TriBody &B) {
  MarkerTree &T = B.tree;
  T.perm.resize(B.nm);
  for (int m = 0; m < B.nm; m++) T.perm[m] = m;
  T.nodes.clear();
  T.nodes.reserve(2 * (B.nm / 4 + 1));
  const double *X[3] = {B.xm.data(), B.ym.data(), B.zm.data()};
  // iterative build with an explicit stack of (node, a, b)
  struct Item { int node, a, b; };
  std::vector<Item> st;
  T.nodes.push_back(MarkerTree::Node());
  st.push_back({0, 0, B.nm});
  while (!st.empty()) {
    Item it = st.back();
    st.pop_back();
    MarkerTree::Node nd;
    nd.a = it.a; nd.b = it.b; nd.left = nd.right = -1;
    for (int d = 0; d < 3; d++) { nd.lo[d] = INFINITY; nd.hi[d] = -INFINITY; }
    for (int q = it.a; q < it.b; q++)
      for (int d = 0; d < 3; d++) {
        nd.lo[d] = std::min(nd.lo[d], X[d][T.perm[q]]);
        nd.hi[d] = std::max(nd.hi[d], X[d][T.perm[q]]);
      }
    if (it.b - it.a > 8) {
      int ax = 0;
      for (int d = 1; d < 3; d++)
        if (nd.hi[d] - nd.lo[d] > nd.hi[ax] - nd.lo[ax]) ax = d;
      int mid = (it.a + it.b) / 2;
      std::nth_element(T.perm.begin() + it.a, T.perm.begin() + mid, T.perm.begin() + it.b,
                       [&](int u, int v) { return X[ax][u] < X[ax][v]; });
      nd.left = (int)T.nodes.size();
      T.nodes.push_back(MarkerTree::Node());
      nd.right = (int)T.nodes.size();
      T.nodes.push_back(MarkerTree::Node());
      st.push_back({nd.left, it.a, mid});
      st.push_back({nd.right, mid, it.b});
    }
    T.nodes[it.node] = nd;
  }
}

struct KBest {
  int k, n = 0;
  double d[kSave];
  int id[kSave];
  bool full() const { return n == k; }
  // (d2, m) precedes the current k-th entry
  bool better(double d2, int m) const { return !full() || d2 < d[n - 1] || (d2 == d[n - 1] && m < id[n - 1]); }
  void insert(double d2, int m) {
    int p = full() ? k - 1 : n++;
    while (p > 0 && (d2 < d[p - 1] || (d2 == d[p - 1] && m < id[p - 1]))) {
      d[p] = d[p - 1]; id[p] = id[p - 1];
      p--;
    }
    d[p] = d2; id[p] = m;
  }
};

// squared distance of marker m to P, as the reference forms it
inline double marker_d2(const TriBody &B, int m, const double *P) {
  double dx = B.xm[m] - P[0], dy = B.ym[m] - P[1], dz = B.zm[m] - P[2];
  return dx * dx + dy * dy + dz * dz;
}

// a rounded squared distance no larger than marker_d2 of any marker in the box
inline double box_d2(const MarkerTree::Node &nd, const double *P) {
  double g[3];
  for (int d = 0; d < 3; d++) {
    if (P[d] < nd.lo[d]) g[d] = nd.lo[d] - P[d];
    else if (P[d] > nd.hi[d]) g[d] = nd.hi[d] - P[d];
    else g[d] = 0.0;
  }
  return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
}

// 1-based ids of the k markers closest to P (repeated MINLOC order)
void closest_markers(const TriBody &B, const double *P, int k, int *out) {
  KBest kb;
  kb.k = k;
  const MarkerTree &T = B.tree;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const MarkerTree::Node &nd = T.nodes[stack[--sp]];
    const double lb = box_d2(nd, P);
    if (kb.full() && lb > kb.d[k - 1]) continue;
    if (nd.left < 0) {
      for (int q = nd.a; q < nd.b; q++) {
        int m = T.perm[q];
        double d2 = marker_d2(B, m, P);
        if (kb.better(d2, m)) kb.insert(d2, m);
      }
      continue;
    }
    const double dl = box_d2(T.nodes[nd.left], P), dr = box_d2(T.nodes[nd.right], P);
    // push the farther child first so the nearer one is visited first
    if (dl <= dr) { stack[sp++] = nd.right; stack[sp++] = nd.left; }
    else { stack[sp++] = nd.left; stack[sp++] = nd.right; }
  }
  for (int s = 0; s < k; s++) out[s] = kb.id[s] + 1;
}

// elements at any of the given markers, ascending
void elements_at(const TriBody &B, const int *m1, int nm, std::vector<int> &out) {
  out.clear();
  for (int s = 0; s < nm; s++)
    for (int q = B.ve_ptr[m1[s] - 1]; q < B.ve_ptr[m1[s]]; q++) out.push_back(B.ve_idx[q]);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
}

// check_if_edge_marker (BOUNDARY_IBLANK_FAST.F90:3580-3703)
bool edge_marker(const Surf &S, int m1) {
  const TriBody &B = S.B;
  const int a0 = B.ve_ptr[m1 - 1], a1 = B.ve_ptr[m1];
  for (int a = a0; a < a1 - 1; a++)
    for (int b = a + 1; b < a1; b++) {
      double r[3], q[3];
      S.normal(B.ve_idx[a], r); S.normal(B.ve_idx[b], q);
      for (int d = 0; d < 3; d++) { r[d] = -r[d]; q[d] = -q[d]; }
      if (angle_at_least_half_pi(r, q)) return true;
    }
  return false;
}

// search_vertex_dotNorm (BOUNDARY_IBLANK_FAST.F90:504-1024) for one cell and
// one body: 0 and dotnorm, or 4 / 5
int dot_norm(const TriBody &B, const double *P, double &dotnorm, std::vector<int> &nb) {
  Surf S{B};
  dotnorm = -1.0;
  int cv[kSave];
  closest_markers(B, P, kSave, cv);
  elements_at(B, cv, kSave, nb);
  int choice = -1;
  for (int s = 0; s < kSave && choice < 0; s++) {
    double pm[3];
    S.marker(cv[s], pm);
    bool change = false;
    for (int e : nb) {
      if (S.has(e, cv[s])) continue;
      double n[3], r[3];
      S.normal(e, n); S.centroid(e, r);
      const int sc = face_side(n, r, P), sm = face_side(n, r, pm);
      if (sc * sm > 0) continue;
      double po[3], a[3], b[3], cc[3];
      line_plane(n, r, P, pm, po);
      S.marker(B.tri[3 * e], a); S.marker(B.tri[3 * e + 1], b); S.marker(B.tri[3 * e + 2], cc);
      if (in_triangle(a, b, cc, po)) { change = true; break; }
    }
    if (!change) choice = s;
  }
  if (choice < 0) return TRI_ERR_NO_MARKER;
  const int cm = cv[choice];
  if (edge_marker(S, cm)) return 0;
  double dmin = 1.0e16, dist_in = 1.0e16;
  int closest = -1;
  bool inside = false;
  for (int q = B.ve_ptr[cm - 1]; q < B.ve_ptr[cm]; q++) {
    const int e = B.ve_idx[q];
    double bi[3], a123, adiff, din;
    bi_inside(S, e, P, bi, a123, adiff, din);
    if (std::fabs(adiff) < 1.0e-4 * a123) {
      inside = true;
      if (din <= dist_in) { dist_in = din; closest = e; }
    } else if (!inside) {
      const double dbe = bi_outside(S, e, P, cm);
      if (dbe <= dmin) { dmin = dbe; closest = e; }
    }
  }
  if (closest < 0) return TRI_ERR_NO_ELEMENT;
  dotnorm = (P[0] - B.cx[closest]) * B.nx[closest] + (P[1] - B.cy[closest]) * B.ny[closest] +
            (P[2] - B.cz[closest]) * B.nz[closest];
  return 0;
}

// last i in [0, n-1] with (c[i]-b)(c[i+1]-b) <= 0, or -1 (c strictly increasing)
int last_bracket(const std::vector<double> &c, int n, double b) {
  // p: last index in [0, n] with c[p] <= b
  int p = (int)(std::upper_bound(c.begin(), c.begin() + n + 1, b) - c.begin()) - 1;
  if (p < 0) return -1;
  if (p == n) return c[n] == b ? n - 1 : -1;
  return p;
}

}  // namespace

// ---- meshes ---------------------------------------------------------------
int tri_build(const Grid &G, const std::vector<Body> &bodies, TriBodies &T) {
  (void)G;
  const double PI = pi_ref();
  T.b.assign(bodies.size(), TriBody());
  for (size_t bi = 0; bi < bodies.size(); bi++) {
    const Body &Bd = bodies[bi];
    TriBody &B = T.b[bi];
    const int nt = Bd.n_theta;
    if (nt < 3) return TRI_ERR_SMALL;
    const double aratio = Bd.r[2] / Bd.r[0];
    const int nphi = (nt % 2 == 0) ? (int)std::lround(0.5 * (double)nt * aratio)
                                   : (int)std::lround(0.5 * (double)(nt - 1) * aratio) + 1;
    if (nphi < 3) return TRI_ERR_SMALL;
    const double ca = std::cos(Bd.alpha_deg * PI / 180.0), sa = std::sin(Bd.alpha_deg * PI / 180.0);
    for (int i = 1; i <= nphi; i++) {
      const double phi = ((double)i - 1.0) * PI / ((double)nphi - 1.0);
      const int nring = (i == 1 || i == nphi) ? 1 : nt;
      for (int j = 1; j <= nring; j++) {
        const double theta = nring == 1 ? 0.0 : ((double)j - 1.0) * 2.0 * PI / (double)nt;
        const double xT = Bd.r[0] * std::cos(theta) * std::sin(phi);
        const double yT = Bd.r[1] * std::sin(theta) * std::sin(phi);
        B.xm.push_back(Bd.c[0] + xT * ca - yT * sa);
        B.ym.push_back(Bd.c[1] + xT * sa + yT * ca);
        B.zm.push_back(Bd.c[2] + Bd.r[2] * std::cos(phi));
      }
    }
    const int m = (int)B.xm.size();
    B.nm = m;
    auto add = [&](int a, int b, int c) { B.tri.push_back(a); B.tri.push_back(b); B.tri.push_back(c); };
    for (int k = 1; k <= nt; k++) add(1, 1 + k, (1 + k == nt + 1) ? 2 : k + 2);
    int nnt = 1;
    for (int i = 2; i <= m - nt - 1; i++) {
      const bool wrap = (i == nnt * nt + 1);
      add(i, i + nt, wrap ? i + 1 : i + nt + 1);
      if (wrap) {
        add(i, i + 1, i + 1 - nt);
        nnt++;
      } else {
        add(i, i + nt + 1, i + 1);
      }
    }
    for (int i = m - nt; i <= m - 1; i++) add(i, m, (i + 1 == m) ? m - nt : i + 1);
    B.ne = (int)B.tri.size() / 3;
    const double origin[3] = {0.0, 0.0, 0.0};
    const int rc = tri_finish(B, origin, true);
    if (rc) return rc;
  }
  return 0;
}

// The rest of a body's surface set-up once its markers and elements are
// known (the orientation step of determine_norm_dir_unstruc with the point
// outside the body P): the element at the marker closest to P gets its
// normal pointing away from P (every element flipped otherwise); element
// normals, areas and centroids; the elements at each marker; the marker
// search tree.  0, or TRI_ERR_ORIENT when P lies inside the markers'
// bounding box and check_box is set (the canonical bodies, whose orientation
// point is the origin: one inside the box cannot be relied on to see the
// outside of the surface).  A caller-given pointOutsideBody is used wherever
// it lies, as determine_norm_dir_unstruc does
// (BOUNDARY_SET_ARCLENGTH_NORM.F90:229-317).
int tri_finish(TriBody &B, const double *P, bool check_box) {
  {
    const int m = B.nm;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int q = 0; q < m; q++) {
      const double v[3] = {B.xm[q], B.ym[q], B.zm[q]};
      for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], v[d]); hi[d] = std::max(hi[d], v[d]); }
    }
    if (check_box && lo[0] <= P[0] && hi[0] >= P[0] && lo[1] <= P[1] && hi[1] >= P[1] && lo[2] <= P[2] &&
        hi[2] >= P[2])
      return TRI_ERR_ORIENT;
    double dmin = 1.0e8;
    int cnode = 0;
    for (int q = 1; q <= m; q++) {
      const double dx = B.xm[q - 1] - P[0], dy = B.ym[q - 1] - P[1], dz = B.zm[q - 1] - P[2];
      const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (d <= dmin) { dmin = d; cnode = q; }
    }
    auto raw_normal = [&](int e, double *n) {
      const int a = B.tri[3 * e] - 1, b = B.tri[3 * e + 1] - 1, c = B.tri[3 * e + 2] - 1;
      const double v1[3] = {B.xm[b] - B.xm[a], B.ym[b] - B.ym[a], B.zm[b] - B.zm[a]};
      const double v2[3] = {B.xm[c] - B.xm[b], B.ym[c] - B.ym[b], B.zm[c] - B.zm[b]};
      n[0] = v1[1] * v2[2] - v1[2] * v2[1];
      n[1] = v1[2] * v2[0] - v1[0] * v2[2];
      n[2] = v1[0] * v2[1] - v1[1] * v2[0];
    };
    int celem = 0;
    for (int e = 0; e < B.ne; e++)
      if (B.tri[3 * e] == cnode || B.tri[3 * e + 1] == cnode || B.tri[3 * e + 2] == cnode) celem = e;
    double n0[3];
    raw_normal(celem, n0);
    {
      const int a = B.tri[3 * celem] - 1, b = B.tri[3 * celem + 1] - 1, c = B.tri[3 * celem + 2] - 1;
      const double gx = (B.xm[a] + B.xm[b] + B.xm[c]) / 3.0, gy = (B.ym[a] + B.ym[b] + B.ym[c]) / 3.0,
                   gz = (B.zm[a] + B.zm[b] + B.zm[c]) / 3.0;
      const double vp = (gx - P[0]) * n0[0] + (gy - P[1]) * n0[1] + (gz - P[2]) * n0[2];
      if (vp < 0.0)
        for (int e = 0; e < B.ne; e++) std::swap(B.tri[3 * e + 1], B.tri[3 * e + 2]);
    }
    B.cx.resize(B.ne); B.cy.resize(B.ne); B.cz.resize(B.ne);
    B.nx.resize(B.ne); B.ny.resize(B.ne); B.nz.resize(B.ne); B.area.resize(B.ne);
    for (int e = 0; e < B.ne; e++) {
      double n[3];
      raw_normal(e, n);
      const double mag = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      B.area[e] = 0.5 * mag;
      B.nx[e] = n[0] / mag; B.ny[e] = n[1] / mag; B.nz[e] = n[2] / mag;
      const int a = B.tri[3 * e] - 1, b = B.tri[3 * e + 1] - 1, c = B.tri[3 * e + 2] - 1;
      B.cx[e] = (B.xm[a] + B.xm[b] + B.xm[c]) / 3.0;
      B.cy[e] = (B.ym[a] + B.ym[b] + B.ym[c]) / 3.0;
      B.cz[e] = (B.zm[a] + B.zm[b] + B.zm[c]) / 3.0;
    }
    // elements at each marker (ascending element index)
    B.ve_ptr.assign(m + 1, 0);
    for (int e = 0; e < B.ne; e++)
      for (int v = 0; v < 3; v++) B.ve_ptr[B.tri[3 * e + v]]++;
    for (int q = 0; q < m; q++) B.ve_ptr[q + 1] += B.ve_ptr[q];
    B.ve_idx.assign(B.ve_ptr[m], 0);
    std::vector<int> fill(B.ve_ptr.begin(), B.ve_ptr.end() - 1);
    for (int e = 0; e < B.ne; e++)
      for (int v = 0; v < 3; v++) {
        const int q = B.tri[3 * e + v] - 1;
        B.ve_idx[fill[q]++] = e;
      }
    if (m < kSave) return TRI_ERR_SMALL;
    tree_build(B);
  }
  return 0;
}

const char *tri_error_text(int rc) {
  switch (rc) {
    case TRI_ERR_NO_MARKER: return "no unblocked marker among the 7 closest to a band cell";
    case TRI_ERR_NO_ELEMENT: return "no surface element selected for a cell or ghost";
    case TRI_ERR_SMALL: return "too few markers, elements or rings, or grid centres not increasing";
    case TRI_ERR_MARKER_ID: return "an element names a marker out of range";
    case TRI_ERR_ORIENT:
      return "the origin, which orients a canonical body's normals, lies inside the body's bounding box";
    default: return "unknown error";
  }
}

int tri_build_surfaces(const std::vector<Surface> &surf, TriBodies &T) {
  T.b.assign(surf.size(), TriBody());
  for (size_t bi = 0; bi < surf.size(); bi++) {
    const Surface &S = surf[bi];
    TriBody &B = T.b[bi];
    if (S.nm < 3 || S.ne < 1) return TRI_ERR_SMALL;
    B.nm = S.nm;
    B.ne = S.ne;
    B.xm.resize(S.nm); B.ym.resize(S.nm); B.zm.resize(S.nm);
    for (int q = 0; q < S.nm; q++) { B.xm[q] = S.xyz[3 * q]; B.ym[q] = S.xyz[3 * q + 1]; B.zm[q] = S.xyz[3 * q + 2]; }
    B.tri.assign(S.tri, S.tri + 3L * S.ne);
    for (int v : B.tri)
      if (v < 1 || v > S.nm) return TRI_ERR_MARKER_ID;
    const int rc = tri_finish(B, S.out, false);
    if (rc) return rc;
  }
  return 0;
}

// ---- iblank ---------------------------------------------------------------
int tri_iblank(const Grid &G, const TriBodies &T, ByteField &ib, IntField &bodynum) {
  const int nxc = G.nxc, nyc = G.nyc, nzc = G.nzc, nx = nxc + 1, ny = nyc + 1, nz = nzc + 1;
  const long ex = G.ex, exy = G.exy;
  for (int d = 0; d < 3; d++)
    for (size_t i = 1; i < G.XC[d].size(); i++)
      if (!(G.XC[d][i] > G.XC[d][i - 1])) return TRI_ERR_SMALL;
  ByteField und, temp;
  par_fill(temp, G.size, (int8_t)0);
  par_fill(und, G.size, (int8_t)0);
  parallel_for(nz + 1, [&](long k0, long k1) {
    for (long k = k0; k < k1; k++)
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++) und[G.r(i, j, (int)k)] = kUndecided;
  }, kPlaneGrain);
  // the band: cells in the bounding box of any element
  const int nn[3] = {nx, ny, nz}, ncc[3] = {nxc, nyc, nzc};
  for (const TriBody &B : T.b)
    for (int e = 0; e < B.ne; e++) {
      double v[3][3];
      for (int q = 0; q < 3; q++) {
        const int m = B.tri[3 * e + q] - 1;
        v[0][q] = B.xm[m]; v[1][q] = B.ym[m]; v[2][q] = B.zm[m];
      }
      bool out = false;
      for (int d = 0; d < 3 && !out; d++) {
        const double first = G.X[d][1], last = G.X[d][nn[d]];
        out = (v[d][0] <= first && v[d][1] <= first && v[d][2] <= first) ||
              (v[d][0] >= last && v[d][1] >= last && v[d][2] >= last);
      }
      if (out) continue;
      int lo[3], hi[3];
      for (int d = 0; d < 3; d++) {
        const double bmin = std::min(std::min(v[d][0], v[d][1]), v[d][2]);
        const double bmax = std::max(std::max(v[d][0], v[d][1]), v[d][2]);
        const int a = last_bracket(G.XC[d], nn[d], bmin), b = last_bracket(G.XC[d], nn[d], bmax);
        int cmin = a < 0 ? 0 : a, cmax = b < 0 ? 0 : b + 1;
        if (cmin == 0) cmin = 1;
        if (cmax == 0 || cmax == nn[d]) cmax = ncc[d];
        lo[d] = cmin; hi[d] = cmax;
      }
      for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
          for (int i = lo[0]; i <= hi[0]; i++) und[G.r(i, j, k)] = -kUndecided;
    }
  // dot-normal classification of the band cells (listed in index order,
  // plane by plane in parallel)
  std::vector<std::vector<long>> bandk(nzc + 1);
  parallel_for(nzc, [&](long k0, long k1) {
    for (int k = (int)k0 + 1; k <= (int)k1; k++)
      for (int j = 1; j <= nyc; j++)
        for (int i = 1; i <= nxc; i++)
          if (und[G.r(i, j, k)] == -kUndecided) bandk[k].push_back(G.r(i, j, k));
  }, kPlaneGrain);
  std::vector<long> band;
  for (int k = 1; k <= nzc; k++) band.insert(band.end(), bandk[k].begin(), bandk[k].end());
  std::atomic<int> err(0);
  parallel_for((long)band.size(), [&](long a, long b) {
    std::vector<int> scratch;
    for (long q = a; q < b && !err.load(std::memory_order_relaxed); q++) {
      const long id = band[q];
      const int k = (int)(id / exy), j = (int)((id % exy) / ex), i = (int)(id % ex);
      const double P[3] = {G.XC[0][i], G.XC[1][j], G.XC[2][k]};
      temp[id] = 0;
      for (size_t bb = 0; bb < T.b.size(); bb++) {
        double dn;
        const int rc = dot_norm(T.b[bb], P, dn, scratch);
        if (rc) { err.store(rc); break; }
        if (dn >= 0.0) {
          temp[id] = 1;
          bodynum[id] = (int)bb;
          break;
        }
      }
    }
  });
  if (err.load()) return err.load();
  // ray fills along i, j, k; each line is independent
  parallel_for(nzc, [&](long k0, long k1) {
    for (int k = (int)k0 + 1; k <= (int)k1; k++)
      for (int j = 1; j <= nyc; j++) {
        for (int ii = 1; ii <= nx; ii++)
          if (und[G.r(ii, j, k)] != kUndecided) { temp[G.r(0, j, k)] = temp[G.r(ii, j, k)]; break; }
        for (int i = 1; i <= nx; i++) {
          const long id = G.r(i, j, k);
          if (und[id] != kUndecided) continue;
          temp[id] = temp[id - 1];
          if (temp[id] == 1) bodynum[id] = bodynum[id - 1];
        }
      }
  }, kPlaneGrain);
  parallel_for(nzc, [&](long k0, long k1) {
    for (int k = (int)k0 + 1; k <= (int)k1; k++)
      for (int i = 1; i <= nxc; i++) {
        for (int jj = 1; jj <= ny; jj++)
          if (und[G.r(i, jj, k)] != kUndecided) { temp[G.r(i, 0, k)] = temp[G.r(i, jj, k)]; break; }
        for (int j = 1; j <= ny; j++) {
          const long id = G.r(i, j, k);
          if (und[id] != kUndecided) continue;
          temp[id] = temp[id - ex];
          if (temp[id] == 1) bodynum[id] = bodynum[id - ex];
        }
      }
  }, kPlaneGrain);
  parallel_for(nyc, [&](long j0, long j1) {
    for (int j = (int)j0 + 1; j <= (int)j1; j++)
      for (int i = 1; i <= nxc; i++) {
        for (int kk = 1; kk <= nzc + 1; kk++)
          if (und[G.r(i, j, kk)] != kUndecided) { temp[G.r(i, j, 0)] = temp[G.r(i, j, kk)]; break; }
        for (int k = 1; k <= nzc + 1; k++) {
          const long id = G.r(i, j, k);
          if (und[id] != kUndecided) continue;
          temp[id] = temp[id - exy];
          if (temp[id] == 1) bodynum[id] = bodynum[id - exy];
        }
      }
  }, kPlaneGrain);
  parallel_for(nzc, [&](long k0, long k1) {
    for (int k = (int)k0 + 1; k <= (int)k1; k++)
      for (int j = 1; j <= nyc; j++)
        for (int i = 1; i <= nxc; i++) ib[G.r(i, j, k)] = temp[G.r(i, j, k)];
  }, kPlaneGrain);
  return 0;
}

// ---- body intercepts ----------------------------------------------------------
double tri_nearest_marker_d2(const TriBodies &T, int body, const double *P) {
  int m;
  closest_markers(T.b[body], P, 1, &m);
  return marker_d2(T.b[body], m - 1, P);
}

int tri_intercept(const Grid &G, const TriBodies &T, int body, int iG, int jG, int kG, double *bi, double *nrm,
                  int &elem) {
  const TriBody &B = T.b[body];
  Surf S{B};
  const double P[3] = {G.XC[0][iG], G.XC[1][jG], G.XC[2][kG]};
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
      bi_inside(S, e, P, t, a123, adiff, din);
      if (std::fabs(adiff) < 1.0e-4 * a123) {
        for (int d = 0; d < 3; d++) xb[d] = t[d];
        closest = e;
        done = true;
        break;
      }
      const double dbe = bi_outside(S, e, P, cn[s]);
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
      S.marker(cn[s], cvp);
      const int t1 = B.tri[3 * closest], t2 = B.tri[3 * closest + 1], t3 = B.tri[3 * closest + 2];
      int na, nb;
      if (t1 == cn[s]) { na = t2; nb = t3; }
      else if (t2 == cn[s]) { na = t3; nb = t1; }
      else { na = t1; nb = t2; }
      const double v01[3] = {P[0] - cvp[0], P[1] - cvp[1], P[2] - cvp[2]};
      double best = 1.0e16, ei[3] = {cvp[0], cvp[1], cvp[2]};
      for (int ed = 0; ed < 2; ed++) {
        double q[3];
        S.marker(ed == 0 ? na : nb, q);
        double v12[3] = {q[0] - cvp[0], q[1] - cvp[1], q[2] - cvp[2]};
        const double mag = std::sqrt(v12[0] * v12[0] + v12[1] * v12[1] + v12[2] * v12[2]);
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

}  // namespace ppg
