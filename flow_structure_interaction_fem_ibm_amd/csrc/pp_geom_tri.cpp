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

inline double pi_ref() { return 4.0 * std::atan(1.0); }

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
    if (m < core::kSave) return TRI_ERR_SMALL;
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
        const int a = core::last_bracket(G.XC[d].data(), nn[d], bmin),
                  b = core::last_bracket(G.XC[d].data(), nn[d], bmax);
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
  const GeomConsts K = geom_consts();
  std::vector<TriV> views(T.b.size());
  for (size_t bb = 0; bb < T.b.size(); bb++) views[bb] = T.b[bb].view();
  parallel_for((long)band.size(), [&](long a, long b) {
    for (long q = a; q < b && !err.load(std::memory_order_relaxed); q++) {
      const long id = band[q];
      const int k = (int)(id / exy), j = (int)((id % exy) / ex), i = (int)(id % ex);
      const double P[3] = {G.XC[0][i], G.XC[1][j], G.XC[2][k]};
      temp[id] = 0;
      for (size_t bb = 0; bb < T.b.size(); bb++) {
        double dn;
        const int rc = core::dot_norm(views[bb], P, dn, K);
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
  return core::nearest_marker_d2(T.b[body].view(), P);
}

int tri_intercept(const Grid &G, const TriBodies &T, int body, int iG, int jG, int kG, double *bi, double *nrm,
                  int &elem) {
  const double P[3] = {G.XC[0][iG], G.XC[1][jG], G.XC[2][kG]};
  return core::tri_intercept(T.b[body].view(), P, bi, nrm, elem, geom_consts());
}

GeomConsts geom_consts() {
  GeomConsts K;
  K.pi = pi_ref();
  K.tenth_pi = K.pi / 10.0;
  K.cos_tenth_pi = std::cos(K.pi / 10.0);
  K.half_pi = K.pi / 2.0;
  return K;
}

}  // namespace ppg
