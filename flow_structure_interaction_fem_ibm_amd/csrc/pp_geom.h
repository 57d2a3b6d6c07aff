// Internal declarations of the host geometry module.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <functional>
#include <vector>

#include "pp_geom_core.h"
#include "pp_hip.h"

namespace pp {
void set_error(const char *fmt, ...);
}

namespace ppg {

// std::vector whose resize leaves the elements uninitialised: the field
// arrays (tens of millions of cells) are filled in parallel right after
// (par_fill) instead of value-initialised on one thread
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind { using other = NoInitAlloc<U>; };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <typename U>
  void construct(U *p) noexcept { ::new ((void *)p) U; }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
};
using ByteField = std::vector<int8_t, NoInitAlloc<int8_t>>;
using IntField = std::vector<int, NoInitAlloc<int>>;

// fn over [0, n) split into contiguous ranges on up to 16 threads; serial
// below `grain` items (per-cell and per-ghost loops use the default, loops
// over planes of the grid pass a small grain)
constexpr long kPlaneGrain = 8;
void parallel_for(long n, const std::function<void(long, long)> &fn, long grain = 4096);
// v = n copies of x, written by all threads
template <typename F, typename T>
void par_fill(F &v, long n, T x) {
  v.resize(n);
  parallel_for((n + 65535) / 65536, [&](long a, long b) {
    std::fill(v.data() + a * 65536, v.data() + std::min(n, b * 65536), x);
  }, 2);
}

// Reference grid arrays x(0:n+1), xc(0:n+1), dx(0:n+1) per direction
// (make_grid, CORE_GRID.F90:29-151) and the reference field indexing.
struct Grid {
  int nxc, nyc, nzc;
  std::vector<double> X[3], XC[3], DX[3];
  long ex, exy, size;
  Grid(int nxc, int nyc, int nzc, const double *xn, const double *yn, const double *zn);
  long r(int i, int j, int k) const { return (long)i + ex * j + exy * k; }
  double dxc(int d, int i) const { return XC[d][i] - XC[d][i - 1]; }
  GridV view() const {
    GridV v;
    v.nxc = nxc; v.nyc = nyc; v.nzc = nzc;
    for (int d = 0; d < 3; d++) { v.X[d] = X[d].data(); v.XC[d] = XC[d].data(); v.DX[d] = DX[d].data(); }
    v.ex = ex; v.exy = exy;
    return v;
  }
};
// pi = 4 atan(1), pi/10, cos(pi/10), pi/2 on the host's math library
GeomConsts geom_consts();

struct Body {
  double c[3], r[3], alpha_deg;
  int n_theta;
};

// triangulated bodies (pp_geom_tri.cpp)
// k markers closest to a point: a kd-tree over a body's markers whose
// pruning bound never exceeds the computed squared distance of a marker in
// the pruned box, so the result is exactly that of repeated MINLOC over
// all markers (smallest squared distance first, lowest index on ties).
struct MarkerTree {
  using Node = TreeNode;
  std::vector<Node> nodes;
  std::vector<int> perm;
};
struct TriBody {
  int nm = 0, ne = 0;
  std::vector<double> xm, ym, zm;            // markers (0-based)
  std::vector<int> tri;                      // 3 per element, 1-based marker ids
  std::vector<double> cx, cy, cz, nx, ny, nz, area;
  std::vector<int> ve_ptr, ve_idx;           // elements at each marker, ascending (CSR)
  MarkerTree tree;
  TriV view() const {
    TriV v;
    v.nm = nm; v.ne = ne;
    v.xm = xm.data(); v.ym = ym.data(); v.zm = zm.data();
    v.tri = tri.data();
    v.cx = cx.data(); v.cy = cy.data(); v.cz = cz.data();
    v.nx = nx.data(); v.ny = ny.data(); v.nz = nz.data();
    v.ve_ptr = ve_ptr.data(); v.ve_idx = ve_idx.data();
    v.nodes = tree.nodes.data(); v.perm = tree.perm.data();
    return v;
  }
};
struct TriBodies {
  std::vector<TriBody> b;
};
// A triangulated surface given by the caller (unstruc_surface_in.dat):
// nm markers (x, y, z each), ne elements (three 1-based marker ids each),
// and the point outside the body that orients the normals.
struct Surface {
  int nm = 0, ne = 0;
  const double *xyz = nullptr;
  const int32_t *tri = nullptr;
  double out[3] = {0.0, 0.0, 0.0};
};
const char *tri_error_text(int rc);
// 0 or an error code: TRI_ERR_SMALL or TRI_ERR_ORIENT
int tri_build(const Grid &G, const std::vector<Body> &bodies, TriBodies &T);
// caller-given surfaces instead of canonical bodies; 0, TRI_ERR_SMALL or
// TRI_ERR_MARKER_ID (pointOutsideBody may lie anywhere)
int tri_build_surfaces(const std::vector<Surface> &surf, TriBodies &T);
// orientation, normals, element lists and search tree of one surface
int tri_finish(TriBody &B, const double *P, bool check_box);
// iblank of set_iblank_canonical_body_fast and bodynum (0-based body, -1
// none); 0 or an error code: 4 no unblocked marker among the 7 closest,
// 5 no element selected
// bodynum must hold -1 everywhere on entry
int tri_iblank(const Grid &G, const TriBodies &T, ByteField &iblank, IntField &bodynum);
// body intercept, inward element normal and element (1-based) of a ghost;
// 0 or 5
int tri_intercept(const Grid &G, const TriBodies &T, int body, int iG, int jG, int kG, double *bi, double *nrm,
                  int &elem);
// squared distance from P to the body's closest marker
double tri_nearest_marker_d2(const TriBodies &T, int body, const double *P);

bool inside_analytic(const Body &b, const double *p);
void closest_point_analytic(const Body &b, const double *p, double *bi, double *nin);
void isolated_iblank(const Grid &G, ByteField &ib, IntField &bodynum);
void fill_cell(const Grid &G, ByteField &ib, ByteField &fresh);
void mark_ghosts(const Grid &G, const ByteField &ib, ByteField &gm, std::vector<int> &list);
int ghost_stencil(const Grid &G, int iG, int jG, int kG, const double *bi, const double *nrm, double probe_norm,
                  GhostOut &o);

int fresh_stencil(const Grid &G, const double *bi, double probe_norm, FreshOut &o);
const char *ghost_error_text(int e);

}  // namespace ppg

// the built geometry behind the C-ABI handle
struct pp_geom {
  ppg::ByteField iblank, fresh, ghost;
  std::vector<int> list;       // ig, jg, kg triples (1-based)
  std::vector<ppg::GhostOut> gout;
  int nblank = 0, nfresh = 0;
  // fresh cells in k, j, i order and the first failure of their stencils
  std::vector<ppg::FreshOut> fout;
  int fresh_err = 0, fresh_err_cell[3] = {0, 0, 0};
};

namespace ppg {
// nblank and nfresh of o from its iblank and fresh fields
void count_cells(const Grid &G, pp_geom *o);
// the triangulated pipeline of pp_geom_build on device dev: bodies built,
// the rest on the GPU; PP_OK or a status with the error text set
int32_t build_geometry_device(pp_geom **out, const Grid &G, const TriBodies &tri, double probe_norm,
                              const int8_t *iblank_prev, int dev);
}  // namespace ppg
