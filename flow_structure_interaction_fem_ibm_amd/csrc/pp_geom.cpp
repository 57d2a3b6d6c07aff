// Body geometry -> the GCM inputs of the pressure solve (iblank, ghost list,
// image-point cells, Neumann interpolation coefficients), host C++.
//
// Replaces set_boundary for canonical solid bodies (BOUNDARY_SET.F90:8-51):
//   iblank        set_iblank_canonical_body_fast (BOUNDARY_IBLANK_FAST.F90:2-501)
//                 or, in analytic mode, "cell centre inside the body"
//   cleanup       find_isolatedIblank(1) (BOUNDARY_SET.F90:440-499),
//                 fill_cell (SSM_SET_BC.F90:1425-1511)
//   ghosts        GCM_set_internal_boundary (GCM_SET_BOUNDARY.F90:2-304)
//   coefficients  GCM_Calc_VanMatrixDN (GCM_SET_BOUNDARY.F90:307-494)
// The triangulated mode lives in pp_geom_tri.cpp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "pp_geom.h"

namespace ppg {

void parallel_for(long n, const std::function<void(long, long)> &fn, long grain) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < grain || n < 2 || nt == 1) { fn(0, n); return; }
  std::vector<std::thread> th;
  long chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    long a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([&, a, b] { fn(a, b); });
  }
  for (auto &t : th) t.join();
}

Grid::Grid(int nxc_, int nyc_, int nzc_, const double *xn, const double *yn, const double *zn)
    : nxc(nxc_), nyc(nyc_), nzc(nzc_) {
  const double *src[3] = {xn, yn, zn};
  int ns[3] = {nxc, nyc, nzc};
  for (int d = 0; d < 3; d++) {
    int n = ns[d] + 1;  // reference node count
    auto &x = X[d];
    x.assign(n + 2, 0.0);
    XC[d].assign(n + 2, 0.0);
    DX[d].assign(n + 2, 0.0);
    for (int i = 1; i <= n; i++) x[i] = src[d][i - 1];
    x[0] = -x[2];
    x[n + 1] = x[n] + x[n] - x[n - 1];
    for (int i = 0; i <= n; i++) {
      XC[d][i] = 0.5 * (x[i + 1] + x[i]);
      DX[d][i] = x[i + 1] - x[i];
    }
    XC[d][n + 1] = x[n + 1];
  }
  ex = nxc + 3;
  exy = (long)ex * (nyc + 3);
  size = exy * (nzc + 3);
}

// ---------------------------------------------------------------------
// analytic bodies
// ---------------------------------------------------------------------
static inline void to_body(const Body &b, const double *p, double *q) {
  double ca = std::cos(b.alpha_deg * M_PI / 180.0), sa = std::sin(b.alpha_deg * M_PI / 180.0);
  double dx = p[0] - b.c[0], dy = p[1] - b.c[1], dz = p[2] - b.c[2];
  q[0] = ca * dx + sa * dy;
  q[1] = -sa * dx + ca * dy;
  q[2] = dz;
}
static inline void from_body_vec(const Body &b, const double *q, double *v) {
  double ca = std::cos(b.alpha_deg * M_PI / 180.0), sa = std::sin(b.alpha_deg * M_PI / 180.0);
  v[0] = ca * q[0] - sa * q[1];
  v[1] = sa * q[0] + ca * q[1];
  v[2] = q[2];
}

bool inside_analytic(const Body &b, const double *p) {
  double q[3];
  to_body(b, p, q);
  double s = 0.0;
  for (int d = 0; d < 3; d++) s += (q[d] / b.r[d]) * (q[d] / b.r[d]);
  return s <= 1.0;
}

// Closest point on the ellipsoid surface to p (body intercept) and the
// inward unit normal there.  Distance-minimising point x = a^2 q/(a^2+t)
// with t the root of sum (a q/(a^2+t))^2 = 1 (bisection, then Newton).
void closest_point_analytic(const Body &b, const double *p, double *bi, double *nin) {
  double q[3];
  to_body(b, p, q);
  const double *a = b.r;
  double x[3];
  if (a[0] == a[1] && a[1] == a[2]) {
    double r = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if (r == 0.0) { x[0] = a[0]; x[1] = 0; x[2] = 0; }
    else for (int d = 0; d < 3; d++) x[d] = q[d] * a[0] / r;
  } else {
    auto f = [&](double t) {
      double s = 0.0;
      for (int d = 0; d < 3; d++) {
        double v = a[d] * q[d] / (a[d] * a[d] + t);
        s += v * v;
      }
      return s - 1.0;
    };
    double amin2 = std::min(a[0] * a[0], std::min(a[1] * a[1], a[2] * a[2]));
    double lo = -amin2 + 1e-300, hi = 1.0;
    // bracket: f decreasing on (-amin^2, inf)
    double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    double amax = std::max(a[0], std::max(a[1], a[2]));
    hi = std::max(1.0, amax * qn);
    while (f(hi) > 0) hi *= 2.0;
    // move lo up until f(lo) > 0 (it tends to +inf as t -> -amin^2 unless the
    // matching q component is 0)
    double l = -amin2;
    double step = amin2;
    lo = l + step;
    for (int it = 0; it < 200 && f(lo) <= 0; it++) { step *= 0.5; lo = l + step; }
    double t = 0.5 * (lo + hi);
    for (int it = 0; it < 200; it++) {
      double fv = f(t);
      if (fv > 0) lo = t; else hi = t;
      double dfv = 0.0;
      for (int d = 0; d < 3; d++) {
        double den = a[d] * a[d] + t;
        dfv += -2.0 * a[d] * a[d] * q[d] * q[d] / (den * den * den);
      }
      double tn = t - fv / dfv;
      if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
      if (std::fabs(tn - t) <= 1e-15 * std::max(1.0, std::fabs(t))) { t = tn; break; }
      t = tn;
    }
    for (int d = 0; d < 3; d++) x[d] = a[d] * a[d] * q[d] / (a[d] * a[d] + t);
  }
  double g[3] = {-x[0] / (a[0] * a[0]), -x[1] / (a[1] * a[1]), -x[2] / (a[2] * a[2])};
  double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  for (int d = 0; d < 3; d++) g[d] /= gn;
  double w[3];
  from_body_vec(b, x, w);
  for (int d = 0; d < 3; d++) bi[d] = b.c[d] + w[d];
  from_body_vec(b, g, nin);
}

// ---------------------------------------------------------------------
// shared pipeline
// ---------------------------------------------------------------------

// find_isolatedIblank(1) (BOUNDARY_SET.F90:440-499): a blank cell with no
// other blank cell in its clamped 3x3x3 neighbourhood is cleared and loses
// its body number.  The reference clears in place in loop order; that equals
// testing every cell against the original field: a neighbour visited earlier
// saw this (still blank) cell and so was not cleared, and later neighbours
// are unchanged at this cell's visit.  Hence a parallel two-pass form.
void isolated_iblank(const Grid &G, ByteField &ib, IntField &bodynum) {
  std::vector<std::vector<long>> clear(G.nzc + 1);
  parallel_for(G.nzc, [&](long k0, long k1) {
    for (int k = (int)k0 + 1; k <= (int)k1; k++)
      for (int j = 1; j <= G.nyc; j++)
        for (int i = 1; i <= G.nxc; i++) {
          if (ib[G.r(i, j, k)] != 1) continue;
          int s = 0;
          for (int kk = std::max(1, k - 1); kk <= std::min(G.nzc, k + 1); kk++)
            for (int jj = std::max(1, j - 1); jj <= std::min(G.nyc, j + 1); jj++)
              for (int ii = std::max(1, i - 1); ii <= std::min(G.nxc, i + 1); ii++) s += ib[G.r(ii, jj, kk)];
          if (s <= 1) clear[k].push_back(G.r(i, j, k));
        }
  }, kPlaneGrain);
  for (auto &v : clear)
    for (long r : v) {
      ib[r] = 0;
      bodynum[r] = -1;
    }
}

// fill_cell (SSM_SET_BC.F90:1425-1511) with the outer-face flags of
// set_outer_iup: a fluid cell on an outer face whose inward neighbour
// along that axis is blank becomes blank (sweeps in x, then y, then z).
// Only cells on the two faces normal to the sweep's axis can change, and
// (at least 3 cells per direction) none of them is another's neighbour, so
// each sweep visits just those faces.
void fill_cell(const Grid &G, ByteField &ib, ByteField &fresh) {
  const int n[3] = {G.nxc, G.nyc, G.nzc};
  const long st[3] = {1, G.ex, G.exy};
  for (int axis = 0; axis < 3; axis++) {
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    for (int face = 0; face < 2; face++) {
      const int c = face == 0 ? 1 : n[axis];
      for (int v = 1; v <= n[a2]; v++)
        for (int u = 1; u <= n[a1]; u++) {
          int idx[3];
          idx[axis] = c; idx[a1] = u; idx[a2] = v;
          const long r = G.r(idx[0], idx[1], idx[2]);
          if (ib[r] != 0) continue;
          const bool up = (c == n[axis]), um = (c == 1);
          if ((ib[r - st[axis]] == 1 && up) || (ib[r + st[axis]] == 1 && um)) {
            ib[r] = 1;
            if (fresh[r] == 1) fresh[r] = 0;
          }
        }
    }
  }
}

// GCM ghost marking and list (GCM_SET_BOUNDARY.F90:49-118); per-plane lists
// concatenated in plane order keep the reference's i-fastest list order
void mark_ghosts(const Grid &G, const ByteField &ib, ByteField &gm, std::vector<int> &list) {
  par_fill(gm, G.size, (int8_t)0);
  list.clear();
  const int nx = G.nxc + 1, ny = G.nyc + 1, nz = G.nzc + 1;
  std::vector<std::vector<int>> plane(nz);
  parallel_for(nz - 1, [&](long k0, long k1) {
    for (int k = (int)k0 + 1; k <= (int)k1; k++)
      for (int j = 1; j <= ny - 1; j++)
        for (int i = 1; i <= nx - 1; i++) {
          long r = G.r(i, j, k);
          if (ib[r] != 1) continue;
          int iM = std::max(i - 1, 1), iP = std::min(i + 1, nx - 1);
          int jM = std::max(j - 1, 1), jP = std::min(j + 1, ny - 1);
          int kM = std::max(k - 1, 1), kP = std::min(k + 1, nz - 1);
          if (ib[G.r(iM, j, k)] == 0 || ib[G.r(iP, j, k)] == 0 || ib[G.r(i, jM, k)] == 0 ||
              ib[G.r(i, jP, k)] == 0 || ib[G.r(i, j, kM)] == 0 || ib[G.r(i, j, kP)] == 0) {
            gm[r] = 1;
            plane[k].push_back(i);
            plane[k].push_back(j);
            plane[k].push_back(k);
          }
        }
  }, kPlaneGrain);
  for (auto &v : plane) list.insert(list.end(), v.begin(), v.end());
}

// Image point, interpolation cell and coefficients of one ghost, and one
// fresh cell's interpolation: the shared routines of pp_geom_core.h.
int ghost_stencil(const Grid &G, int iG, int jG, int kG, const double *bi, const double *nrm, double probe_norm,
                  GhostOut &o) {
  return core::ghost_stencil(G.view(), iG, jG, kG, bi, nrm, probe_norm, o);
}

int fresh_stencil(const Grid &G, const double *bi, double probe_norm, FreshOut &o) {
  return core::fresh_stencil(G.view(), bi, probe_norm, o);
}

}  // namespace ppg

// ======================================================================
// C-ABI
// ======================================================================
using namespace ppg;

namespace ppg {
const char *ghost_error_text(int e) {
  return e == GHOST_ERR_IP_CELL    ? "image point cell not found"
         : e == GHOST_ERR_LONG     ? "intercept too long"
         : e == GHOST_ERR_SINGULAR ? "singular Vandermonde"
         : e == TRI_ERR_NO_ELEMENT ? "no element selected for the intercept"
                                   : "ghost cell without a body";
}

void count_cells(const Grid &G, pp_geom *o) {
  std::vector<long> nb(G.nzc + 3, 0), nf(G.nzc + 3, 0);
  parallel_for(G.nzc + 3, [&](long k0, long k1) {
    for (long k = k0; k < k1; k++)
      for (long r = k * G.exy; r < (k + 1) * G.exy; r++) { nb[k] += o->iblank[r]; nf[k] += o->fresh[r]; }
  }, kPlaneGrain);
  o->nblank = o->nfresh = 0;
  for (int k = 0; k < G.nzc + 3; k++) { o->nblank += (int)nb[k]; o->nfresh += (int)nf[k]; }
}
}  // namespace ppg

namespace {
// the part of pp_geom_build after its bodies are known: canonical bodies
// (bs) or, in the triangulated mode, caller-given surfaces (surf)
int32_t build_geometry(pp_geom **out, const Grid &G, int32_t mode, const std::vector<Body> &bs,
                       const std::vector<Surface> *surf, double probe_norm, const int8_t *iblank_prev) {
  const int nbody = surf ? (int)surf->size() : (int)bs.size();
  pp_geom *o = new pp_geom();
  par_fill(o->iblank, G.size, (int8_t)0);
  par_fill(o->fresh, G.size, (int8_t)0);
  IntField bodynum;
  par_fill(bodynum, G.size, -1);
  int rc = 0;
  TriBodies tri;
  if (mode == PP_GEOM_ANALYTIC) {
    // cell centres inside a body; bodies in order (the first one containing
    // a centre owns it), each scanned over the cells of its bounding box
    for (int b = 0; b < nbody; b++) {
      const Body &B = bs[b];
      double ca = std::cos(B.alpha_deg * M_PI / 180.0), sa = std::sin(B.alpha_deg * M_PI / 180.0);
      double ext[3] = {std::sqrt(B.r[0] * B.r[0] * ca * ca + B.r[1] * B.r[1] * sa * sa),
                       std::sqrt(B.r[0] * B.r[0] * sa * sa + B.r[1] * B.r[1] * ca * ca), B.r[2]};
      int lo[3], hi[3];
      int nn[3] = {G.nxc, G.nyc, G.nzc};
      for (int d = 0; d < 3; d++) {
        double a = B.c[d] - ext[d] * (1.0 + 1e-9) - 1e-12, z = B.c[d] + ext[d] * (1.0 + 1e-9) + 1e-12;
        lo[d] = 1;
        while (lo[d] <= nn[d] && G.XC[d][lo[d]] < a) lo[d]++;
        hi[d] = nn[d];
        while (hi[d] >= 1 && G.XC[d][hi[d]] > z) hi[d]--;
      }
      if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) continue;
      parallel_for((long)(hi[2] - lo[2] + 1), [&](long k0, long k1) {
        for (long kk = k0; kk < k1; kk++) {
          int k = lo[2] + (int)kk;
          for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
              long r = G.r(i, j, k);
              if (o->iblank[r]) continue;
              double p[3] = {G.XC[0][i], G.XC[1][j], G.XC[2][k]};
              if (inside_analytic(B, p)) {
                o->iblank[r] = 1;
                bodynum[r] = b;
              }
            }
        }
      }, kPlaneGrain);
    }
  } else {
    rc = surf ? tri_build_surfaces(*surf, tri) : tri_build(G, bs, tri);
    if (rc == 0) rc = tri_iblank(G, tri, o->iblank, bodynum);
    if (rc) {
      delete o;
      pp::set_error("pp_geom_build: triangulated set_boundary failed (%d: %s)", rc, tri_error_text(rc));
      return PP_BAD_GEOMETRY;
    }
  }
  // fresh cells: fluid now, blank before (BOUNDARY_IBLANK_FAST.F90:451-457)
  if (iblank_prev)
    parallel_for(G.nzc, [&](long k0, long k1) {
      for (int k = (int)k0 + 1; k <= (int)k1; k++)
        for (int j = 1; j <= G.nyc; j++)
          for (int i = 1; i <= G.nxc; i++) {
            long r = G.r(i, j, k);
            if (o->iblank[r] == 0 && iblank_prev[r] == 1) o->fresh[r] = 1;
          }
    }, kPlaneGrain);
  isolated_iblank(G, o->iblank, bodynum);
  fill_cell(G, o->iblank, o->fresh);
  mark_ghosts(G, o->iblank, o->ghost, o->list);
  int ng = (int)o->list.size() / 3;
  o->gout.resize(ng);
  std::vector<int> err(ng, 0);
  parallel_for(ng, [&](long a, long b) {
    for (long q = a; q < b; q++) {
      int iG = o->list[3 * q], jG = o->list[3 * q + 1], kG = o->list[3 * q + 2];
      double bi[3], nrm[3];
      double p[3] = {G.XC[0][iG], G.XC[1][jG], G.XC[2][kG]};
      if (mode == PP_GEOM_ANALYTIC) {
        int b0 = bodynum[G.r(iG, jG, kG)];
        if (b0 < 0) b0 = 0;
        // closest body surface among all bodies
        double best = 1e300;
        for (int bb = 0; bb < nbody; bb++) {
          double tb[3], tn[3];
          closest_point_analytic(bs[bb], p, tb, tn);
          double d2 = (tb[0] - p[0]) * (tb[0] - p[0]) + (tb[1] - p[1]) * (tb[1] - p[1]) + (tb[2] - p[2]) * (tb[2] - p[2]);
          if (d2 < best || (bb == b0 && d2 == best)) {
            best = d2;
            o->gout[q].body = bb;
            for (int d = 0; d < 3; d++) { bi[d] = tb[d]; nrm[d] = tn[d]; }
          }
        }
      } else {
        const int b0 = bodynum[G.r(iG, jG, kG)];
        if (b0 < 0) { err[q] = 7; continue; }
        o->gout[q].body = b0;
        err[q] = tri_intercept(G, tri, b0, iG, jG, kG, bi, nrm, o->gout[q].elem);
        if (err[q]) continue;
      }
      err[q] = ghost_stencil(G, iG, jG, kG, bi, nrm, probe_norm, o->gout[q]);
    }
  });
  for (int q = 0; q < ng; q++)
    if (err[q]) {
      pp::set_error("pp_geom_build: ghost %d at (%d,%d,%d): %s", q, o->list[3 * q], o->list[3 * q + 1],
                    o->list[3 * q + 2], ghost_error_text(err[q]));
      delete o;
      return PP_BAD_GEOMETRY;
    }
  count_cells(G, o);
  // fresh-cell interpolation (GCM_set_freshcell).  The reference takes the
  // body intercept on bodyNum of the cell, the body that blanked it at the
  // previous geometry; that is not known here, so the body is the nearest
  // one (closest surface point, analytic mode; closest marker, triangulated
  // mode; the lower index on ties), which is the same body whenever the one
  // that left the cell is still the nearest.
  if (o->nfresh > 0) {
    for (int k = 1; k <= G.nzc; k++)
      for (int j = 1; j <= G.nyc; j++)
        for (int i = 1; i <= G.nxc; i++)
          if (o->fresh[G.r(i, j, k)] == 1) {
            FreshOut f;
            f.i = i; f.j = j; f.k = k;
            o->fout.push_back(f);
          }
    const int nfr = (int)o->fout.size();
    std::vector<int> ferr(nfr, 0);
    parallel_for(nfr, [&](long a, long b) {
      for (long q = a; q < b; q++) {
        FreshOut &f = o->fout[q];
        const double p[3] = {G.XC[0][f.i], G.XC[1][f.j], G.XC[2][f.k]};
        double bi[3] = {0.0, 0.0, 0.0}, nrm[3];
        double best = 1e300;
        if (mode == PP_GEOM_ANALYTIC) {
          for (int bb = 0; bb < nbody; bb++) {
            double tb[3], tn[3];
            closest_point_analytic(bs[bb], p, tb, tn);
            const double d2 = (tb[0] - p[0]) * (tb[0] - p[0]) + (tb[1] - p[1]) * (tb[1] - p[1]) +
                              (tb[2] - p[2]) * (tb[2] - p[2]);
            if (d2 < best) {
              best = d2;
              f.body = bb;
              for (int d = 0; d < 3; d++) bi[d] = tb[d];
            }
          }
        } else {
          for (int bb = 0; bb < nbody; bb++) {
            const double d2 = tri_nearest_marker_d2(tri, bb, p);
            if (d2 < best) { best = d2; f.body = bb; }
          }
          if (f.body >= 0) ferr[q] = tri_intercept(G, tri, f.body, f.i, f.j, f.k, bi, nrm, f.elem);
        }
        if (f.body < 0) ferr[q] = 7;
        if (!ferr[q]) ferr[q] = fresh_stencil(G, bi, probe_norm, f);
      }
    });
    for (int q = 0; q < nfr; q++)
      if (ferr[q]) {
        o->fresh_err = ferr[q];
        o->fresh_err_cell[0] = o->fout[q].i;
        o->fresh_err_cell[1] = o->fout[q].j;
        o->fresh_err_cell[2] = o->fout[q].k;
        break;
      }
  }
  *out = o;
  return PP_OK;
}
}  // namespace

extern "C" {

int32_t pp_geom_build(pp_geom **out, const pp_grid_desc *g, int32_t nbody, const pp_body_desc *bodies, int32_t mode,
                      double probe_norm, const int8_t *iblank_prev) {
  if (!out || !g || nbody < 0 || nbody > 256 || (nbody > 0 && !bodies) || probe_norm <= 0.0) {
    pp::set_error("pp_geom_build: bad argument");
    return PP_BAD_ARG;
  }
  *out = nullptr;
  if (mode != PP_GEOM_ANALYTIC && mode != PP_GEOM_TRIANGULATED) {
    pp::set_error("pp_geom_build: unknown mode %d", mode);
    return PP_BAD_ARG;
  }
  Grid G(g->nxc, g->nyc, g->nzc, g->x, g->y, g->z);
  std::vector<Body> bs(nbody);
  for (int b = 0; b < nbody; b++) {
    const pp_body_desc &d = bodies[b];
    if (d.shape != PP_BODY_ELLIPSOID || d.semi[0] <= 0 || d.semi[1] <= 0 || d.semi[2] <= 0) {
      pp::set_error("pp_geom_build: body %d is not a valid ellipsoid", b);
      return PP_BAD_ARG;
    }
    for (int q = 0; q < 3; q++) { bs[b].c[q] = d.center[q]; bs[b].r[q] = d.semi[q]; }
    bs[b].alpha_deg = d.alpha_deg;
    bs[b].n_theta = d.n_theta;
  }
  return build_geometry(out, G, mode, bs, nullptr, probe_norm, iblank_prev);
}

#ifndef PP_HOST_ONLY  // the host-only sanitizer build has no device code
// the triangulated bodies of pp_geom_build's arguments, or a status
static int32_t tri_bodies_of(const char *who, const Grid &G, int32_t nbody, const pp_body_desc *bodies,
                             TriBodies &tri) {
  std::vector<Body> bs(nbody);
  for (int b = 0; b < nbody; b++) {
    const pp_body_desc &d = bodies[b];
    if (d.shape != PP_BODY_ELLIPSOID || d.semi[0] <= 0 || d.semi[1] <= 0 || d.semi[2] <= 0) {
      pp::set_error("%s: body %d is not a valid ellipsoid", who, b);
      return PP_BAD_ARG;
    }
    for (int q = 0; q < 3; q++) { bs[b].c[q] = d.center[q]; bs[b].r[q] = d.semi[q]; }
    bs[b].alpha_deg = d.alpha_deg;
    bs[b].n_theta = d.n_theta;
  }
  const int rc = tri_build(G, bs, tri);
  if (rc) {
    pp::set_error("%s: triangulated set_boundary failed (%d: %s)", who, rc, tri_error_text(rc));
    return PP_BAD_GEOMETRY;
  }
  return PP_OK;
}

int32_t pp_geom_build_device(pp_geom **out, const pp_grid_desc *g, int32_t nbody, const pp_body_desc *bodies,
                             double probe_norm, const int8_t *iblank_prev) {
  if (!out || !g || nbody < 0 || nbody > 256 || (nbody > 0 && !bodies) || probe_norm <= 0.0) {
    pp::set_error("pp_geom_build_device: bad argument");
    return PP_BAD_ARG;
  }
  *out = nullptr;
  Grid G(g->nxc, g->nyc, g->nzc, g->x, g->y, g->z);
  TriBodies tri;
  const int32_t rc = tri_bodies_of("pp_geom_build_device", G, nbody, bodies, tri);
  if (rc) return rc;
  return build_geometry_device(out, G, tri, probe_norm, iblank_prev, g->device);
}

int32_t pp_geom_build_surfaces_device(pp_geom **out, const pp_grid_desc *g, int32_t n_surface,
                                      const pp_surface_desc *s, double probe_norm, const int8_t *iblank_prev) {
  if (!out || !g || n_surface < 0 || n_surface > 256 || (n_surface > 0 && !s) || probe_norm <= 0.0) {
    pp::set_error("pp_geom_build_surfaces_device: bad argument");
    return PP_BAD_ARG;
  }
  *out = nullptr;
  std::vector<Surface> sv(n_surface);
  for (int b = 0; b < n_surface; b++) {
    if (s[b].n_markers < 3 || s[b].n_elems < 1 || !s[b].xyz || !s[b].tri) {
      pp::set_error("pp_geom_build_surfaces_device: surface %d has no markers or elements", b);
      return PP_BAD_ARG;
    }
    sv[b].nm = s[b].n_markers;
    sv[b].ne = s[b].n_elems;
    sv[b].xyz = s[b].xyz;
    sv[b].tri = s[b].tri;
    for (int d = 0; d < 3; d++) sv[b].out[d] = s[b].outside[d];
  }
  Grid G(g->nxc, g->nyc, g->nzc, g->x, g->y, g->z);
  TriBodies tri;
  const int rc = tri_build_surfaces(sv, tri);
  if (rc) {
    pp::set_error("pp_geom_build_surfaces_device: triangulated set_boundary failed (%d: %s)", rc,
                  tri_error_text(rc));
    return PP_BAD_GEOMETRY;
  }
  return build_geometry_device(out, G, tri, probe_norm, iblank_prev, g->device);
}

#endif

int32_t pp_geom_build_surfaces(pp_geom **out, const pp_grid_desc *g, int32_t n_surface, const pp_surface_desc *s,
                               double probe_norm, const int8_t *iblank_prev) {
  if (!out || !g || n_surface < 0 || n_surface > 256 || (n_surface > 0 && !s) || probe_norm <= 0.0) {
    pp::set_error("pp_geom_build_surfaces: bad argument");
    return PP_BAD_ARG;
  }
  *out = nullptr;
  std::vector<Surface> sv(n_surface);
  for (int b = 0; b < n_surface; b++) {
    if (s[b].n_markers < 3 || s[b].n_elems < 1 || !s[b].xyz || !s[b].tri) {
      pp::set_error("pp_geom_build_surfaces: surface %d has no markers or elements", b);
      return PP_BAD_ARG;
    }
    sv[b].nm = s[b].n_markers;
    sv[b].ne = s[b].n_elems;
    sv[b].xyz = s[b].xyz;
    sv[b].tri = s[b].tri;
    for (int d = 0; d < 3; d++) sv[b].out[d] = s[b].outside[d];
  }
  Grid G(g->nxc, g->nyc, g->nzc, g->x, g->y, g->z);
  return build_geometry(out, G, PP_GEOM_TRIANGULATED, std::vector<Body>(), &sv, probe_norm, iblank_prev);
}

int32_t pp_geom_canonical_surface(const pp_body_desc *d, int32_t *n_markers, int32_t *n_elems, double *xyz,
                                  int32_t *tri) {
  if (!d || !n_markers || !n_elems || d->shape != PP_BODY_ELLIPSOID || d->semi[0] <= 0 || d->semi[1] <= 0 ||
      d->semi[2] <= 0) {
    pp::set_error("pp_geom_canonical_surface: bad argument");
    return PP_BAD_ARG;
  }
  std::vector<Body> bs(1);
  for (int q = 0; q < 3; q++) { bs[0].c[q] = d->center[q]; bs[0].r[q] = d->semi[q]; }
  bs[0].alpha_deg = d->alpha_deg;
  bs[0].n_theta = d->n_theta;
  const double x1[2] = {0.0, 1.0};
  Grid G(1, 1, 1, x1, x1, x1);
  TriBodies T;
  if (tri_build(G, bs, T)) {
    pp::set_error("pp_geom_canonical_surface: the body's surface cannot be built");
    return PP_BAD_GEOMETRY;
  }
  const TriBody &B = T.b[0];
  if (xyz && tri) {
    if (*n_markers < B.nm || *n_elems < B.ne) {
      pp::set_error("pp_geom_canonical_surface: arrays hold %d markers and %d elements, need %d and %d",
                    *n_markers, *n_elems, B.nm, B.ne);
      return PP_BAD_ARG;
    }
    for (int q = 0; q < B.nm; q++) { xyz[3 * q] = B.xm[q]; xyz[3 * q + 1] = B.ym[q]; xyz[3 * q + 2] = B.zm[q]; }
    for (int q = 0; q < 3 * B.ne; q++) tri[q] = B.tri[q];
  }
  *n_markers = B.nm;
  *n_elems = B.ne;
  return PP_OK;
}

int32_t pp_geom_counts(const pp_geom *o, int32_t *n_ghost, int32_t *n_blank, int32_t *n_fresh) {
  if (!o) { pp::set_error("pp_geom_counts: null handle"); return PP_BAD_ARG; }
  if (n_ghost) *n_ghost = (int32_t)o->gout.size();
  if (n_blank) *n_blank = o->nblank;
  if (n_fresh) *n_fresh = o->nfresh;
  return PP_OK;
}

int32_t pp_geom_get(const pp_geom *o, int8_t *iblank, int8_t *fresh, int32_t *ig, int32_t *jg, int32_t *kg,
                    int32_t *ic, int32_t *jc, int32_t *kc, double *coeff, double *probe, double *bi, double *ip,
                    double *nrm) {
  if (!o) { pp::set_error("pp_geom_get: null handle"); return PP_BAD_ARG; }
  if (iblank) memcpy(iblank, o->iblank.data(), o->iblank.size());
  if (fresh) memcpy(fresh, o->fresh.data(), o->fresh.size());
  for (size_t q = 0; q < o->gout.size(); q++) {
    const GhostOut &g = o->gout[q];
    if (ig) ig[q] = o->list[3 * q];
    if (jg) jg[q] = o->list[3 * q + 1];
    if (kg) kg[q] = o->list[3 * q + 2];
    if (ic) ic[q] = g.ic;
    if (jc) jc[q] = g.jc;
    if (kc) kc[q] = g.kc;
    if (coeff) memcpy(coeff + 8 * q, g.coeff, 8 * sizeof(double));
    if (probe) probe[q] = g.probe;
    for (int d = 0; d < 3; d++) {
      if (bi) bi[3 * q + d] = g.bi[d];
      if (ip) ip[3 * q + d] = g.ip[d];
      if (nrm) nrm[3 * q + d] = g.nrm[d];
    }
  }
  return PP_OK;
}

int32_t pp_geom_ghost_elements(const pp_geom *o, int32_t *body, int32_t *elem) {
  if (!o) { pp::set_error("pp_geom_ghost_elements: null handle"); return PP_BAD_ARG; }
  for (size_t q = 0; q < o->gout.size(); q++) {
    if (body) body[q] = o->gout[q].body + 1;
    if (elem) elem[q] = o->gout[q].elem;
  }
  return PP_OK;
}

int32_t pp_geom_dirichlet(const pp_geom *o, double *coeffD) {
  if (!o || !coeffD) { pp::set_error("pp_geom_dirichlet: null argument"); return PP_BAD_ARG; }
  for (size_t q = 0; q < o->gout.size(); q++) memcpy(coeffD + 8 * q, o->gout[q].coeffD, 8 * sizeof(double));
  return PP_OK;
}

int32_t pp_geom_fresh(const pp_geom *o, int32_t *ifr, int32_t *jfr, int32_t *kfr, int32_t *ic, int32_t *jc,
                      int32_t *kc, double *coeffD, double *bi, int32_t *body, int32_t *elem) {
  if (!o) { pp::set_error("pp_geom_fresh: null handle"); return PP_BAD_ARG; }
  if (o->fresh_err) {
    const int e = o->fresh_err;
    pp::set_error("pp_geom_fresh: fresh cell (%d,%d,%d): %s", o->fresh_err_cell[0], o->fresh_err_cell[1],
                  o->fresh_err_cell[2],
                  e == FRESH_ERR_STENCIL ? "image point for fresh cell is not inside the 2x2x2 stencil"
                  : e == FRESH_ERR_AT_BI ? "body intercept at the cell centre"
                  : e == 3               ? "singular Vandermonde"
                  : e == 7               ? "no body"
                                         : tri_error_text(e));
    return PP_BAD_GEOMETRY;
  }
  for (size_t q = 0; q < o->fout.size(); q++) {
    const FreshOut &f = o->fout[q];
    if (ifr) ifr[q] = f.i;
    if (jfr) jfr[q] = f.j;
    if (kfr) kfr[q] = f.k;
    if (ic) ic[q] = f.ic;
    if (jc) jc[q] = f.jc;
    if (kc) kc[q] = f.kc;
    if (coeffD) memcpy(coeffD + 8 * q, f.coeffD, 8 * sizeof(double));
    if (bi) memcpy(bi + 3 * q, f.bi, 3 * sizeof(double));
    if (body) body[q] = f.body + 1;
    if (elem) elem[q] = f.elem;
  }
  return PP_OK;
}

int32_t pp_geom_destroy(pp_geom *o) {
  delete o;
  return PP_OK;
}

}  // extern "C"
