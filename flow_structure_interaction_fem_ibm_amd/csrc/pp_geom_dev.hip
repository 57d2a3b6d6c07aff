// The triangulated set_boundary on the GPU: the band classification, ray
// fills, cleanup, ghost marking, body intercepts and Vandermonde solves of
// pp_geom_build (PP_GEOM_TRIANGULATED), with the bodies' surfaces built on
// the host (small) and everything per cell, per line and per ghost on the
// device.  The per-point decisions are the shared routines of
// pp_geom_core.h, so every output is the host build's, bit for bit.
//
//   band          the cells in each element's bounding box (one thread per element)
//   classify      search_vertex_dotNorm per band cell, bodies in order
//                 (BOUNDARY_IBLANK_FAST.F90:2-501; one thread per band cell)
//   ray fills     along i, then j, then k (one thread per grid line)
//   cleanup       find_isolatedIblank(1) (BOUNDARY_SET.F90:440-499) in two
//                 passes, fill_cell (SSM_SET_BC.F90:1425-1511) face by face
//   ghosts        GCM_set_internal_boundary (GCM_SET_BOUNDARY.F90:49-118) and
//                 its list in i-fastest order (ordered compaction)
//   intercepts    GCM_calc_bodyIntercept_Unstruc + GCM_Calc_VanMatrixDN per
//                 ghost; GCM_set_freshcell per fresh cell
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <vector>

#include "pp_geom.h"

namespace ppg {
namespace {

constexpr int kUndecided = 1;
constexpr int kThreads = 256;

#define GCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      pp::set_error("pp_geom_build_device: %s (%s:%d)", hipGetErrorString(e_),   \
                    __FILE__, __LINE__);                                          \
      return PP_HIP_ERROR;                                                        \
    }                                                                             \
  } while (0)

// every device allocation of one build, freed together
struct Arena {
  std::vector<void *> p;
  ~Arena() {
    for (void *q : p) (void)hipFree(q);
  }
  template <typename T>
  hipError_t alloc(T **out, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (n > 0 ? n : 1));
    if (e == hipSuccess) p.push_back(q);
    *out = (T *)q;
    return e;
  }
  template <typename T>
  hipError_t upload(T **out, const T *src, size_t n) {
    hipError_t e = alloc(out, n);
    if (e == hipSuccess && n > 0) e = hipMemcpy(*out, src, sizeof(T) * n, hipMemcpyHostToDevice);
    return e;
  }
};

inline int blocks_for(long n) {
  long b = (n + kThreads - 1) / kThreads;
  return (int)(b < 1 ? 1 : (b > 65535L * 16 ? 65535L * 16 : b));
}

__device__ inline void cell_of(const GridV &G, long r, int &i, int &j, int &k) {
  k = (int)(r / G.exy);
  const long rem = r - (long)k * G.exy;
  j = (int)(rem / G.ex);
  i = (int)(rem - (long)j * G.ex);
}

__device__ inline bool interior(const GridV &G, int i, int j, int k) {
  return i >= 1 && i <= G.nxc && j >= 1 && j <= G.nyc && k >= 1 && k <= G.nzc;
}

// und = kUndecided on cells 0..n+1 of each direction (n + 1 nodes), temp 0,
// bodynum -1, iblank, fresh and the work masks 0
__global__ void k_geo_init(GridV G, long size, int8_t *und, int8_t *temp, int *bodynum, int8_t *ib, int8_t *fresh,
                           int8_t *mask) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(G, r, i, j, k);
    und[r] = (i <= G.nxc + 1 && j <= G.nyc + 1 && k <= G.nzc + 1) ? kUndecided : 0;
    temp[r] = 0;
    bodynum[r] = -1;
    ib[r] = 0;
    fresh[r] = 0;
    mask[r] = 0;
  }
}

// the band: the cells of each element's bounding box (cell centres bracketed
// by the box, BOUNDARY_IBLANK_FAST.F90:120-200); concurrent writes store the
// same value
__global__ void k_geo_band(GridV G, TriV B, int8_t *und) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B.ne) return;
  const int nn[3] = {G.nxc + 1, G.nyc + 1, G.nzc + 1}, ncc[3] = {G.nxc, G.nyc, G.nzc};
  double v[3][3];
  for (int q = 0; q < 3; q++) {
    const int m = B.tri[3 * e + q] - 1;
    v[0][q] = B.xm[m]; v[1][q] = B.ym[m]; v[2][q] = B.zm[m];
  }
  for (int d = 0; d < 3; d++) {
    const double first = G.X[d][1], last = G.X[d][nn[d]];
    if ((v[d][0] <= first && v[d][1] <= first && v[d][2] <= first) ||
        (v[d][0] >= last && v[d][1] >= last && v[d][2] >= last))
      return;
  }
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    const double bmin = fmin(fmin(v[d][0], v[d][1]), v[d][2]);
    const double bmax = fmax(fmax(v[d][0], v[d][1]), v[d][2]);
    const int a = core::last_bracket(G.XC[d], nn[d], bmin), b = core::last_bracket(G.XC[d], nn[d], bmax);
    int cmin = a < 0 ? 0 : a, cmax = b < 0 ? 0 : b + 1;
    if (cmin == 0) cmin = 1;
    if (cmax == 0 || cmax == nn[d]) cmax = ncc[d];
    lo[d] = cmin; hi[d] = cmax;
  }
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) und[G.r(i, j, k)] = -kUndecided;
}

// interior band cells (und = -kUndecided) as a flag for the compaction
__global__ void k_geo_band_flags(GridV G, long size, const int8_t *und, int8_t *flag) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(G, r, i, j, k);
    flag[r] = (interior(G, i, j, k) && und[r] == -kUndecided) ? 1 : 0;
  }
}

// search_vertex_dotNorm per band cell: blank (temp 1) by the first body, in
// order, whose closest element has the cell on its inner side
__global__ void k_geo_classify(GridV G, const TriV *bodies, int nbody, GeomConsts K, long nband, const long *band,
                               int8_t *temp, int *bodynum, int *err) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nband) return;
  const long id = band[q];
  int i, j, k;
  cell_of(G, id, i, j, k);
  const double P[3] = {G.XC[0][i], G.XC[1][j], G.XC[2][k]};
  temp[id] = 0;
  for (int bb = 0; bb < nbody; bb++) {
    double dn;
    const int rc = core::dot_norm(bodies[bb], P, dn, K);
    if (rc) {
      atomicCAS(err, 0, rc);
      return;
    }
    if (dn >= 0.0) {
      temp[id] = 1;
      bodynum[id] = bb;
      return;
    }
  }
}

// ray fills: an undecided cell takes its predecessor's value (and body)
// along the line; the cell before the line takes the first decided value
__global__ void k_geo_fill_i(GridV G, const int8_t *und, int8_t *temp, int *bodynum) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.nyc * G.nzc) return;
  const int j = 1 + (int)(t % G.nyc), k = 1 + (int)(t / G.nyc);
  const int nx = G.nxc + 1;
  for (int ii = 1; ii <= nx; ii++)
    if (und[G.r(ii, j, k)] != kUndecided) { temp[G.r(0, j, k)] = temp[G.r(ii, j, k)]; break; }
  for (int i = 1; i <= nx; i++) {
    const long id = G.r(i, j, k);
    if (und[id] != kUndecided) continue;
    temp[id] = temp[id - 1];
    if (temp[id] == 1) bodynum[id] = bodynum[id - 1];
  }
}
__global__ void k_geo_fill_j(GridV G, const int8_t *und, int8_t *temp, int *bodynum) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.nxc * G.nzc) return;
  const int i = 1 + (int)(t % G.nxc), k = 1 + (int)(t / G.nxc);
  const int ny = G.nyc + 1;
  for (int jj = 1; jj <= ny; jj++)
    if (und[G.r(i, jj, k)] != kUndecided) { temp[G.r(i, 0, k)] = temp[G.r(i, jj, k)]; break; }
  for (int j = 1; j <= ny; j++) {
    const long id = G.r(i, j, k);
    if (und[id] != kUndecided) continue;
    temp[id] = temp[id - G.ex];
    if (temp[id] == 1) bodynum[id] = bodynum[id - G.ex];
  }
}
__global__ void k_geo_fill_k(GridV G, const int8_t *und, int8_t *temp, int *bodynum) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.nxc * G.nyc) return;
  const int i = 1 + (int)(t % G.nxc), j = 1 + (int)(t / G.nxc);
  for (int kk = 1; kk <= G.nzc + 1; kk++)
    if (und[G.r(i, j, kk)] != kUndecided) { temp[G.r(i, j, 0)] = temp[G.r(i, j, kk)]; break; }
  for (int k = 1; k <= G.nzc + 1; k++) {
    const long id = G.r(i, j, k);
    if (und[id] != kUndecided) continue;
    temp[id] = temp[id - G.exy];
    if (temp[id] == 1) bodynum[id] = bodynum[id - G.exy];
  }
}

// iblank of the interior from temp; fresh cells (fluid now, blank before,
// BOUNDARY_IBLANK_FAST.F90:451-457) when the previous iblank is given
__global__ void k_geo_iblank(GridV G, long size, const int8_t *temp, const int8_t *prev, int8_t *ib, int8_t *fresh) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(G, r, i, j, k);
    if (!interior(G, i, j, k)) continue;
    ib[r] = temp[r];
    if (prev && temp[r] == 0 && prev[r] == 1) fresh[r] = 1;
  }
}

// find_isolatedIblank(1): a blank cell with no blank neighbour in its
// (clamped) 3x3x3 block is cleared; every cell is tested against the field
// as it was, so the marks come first and the clearing after
__global__ void k_geo_isolated_mark(GridV G, long size, const int8_t *ib, int8_t *clr) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(G, r, i, j, k);
    clr[r] = 0;
    if (!interior(G, i, j, k) || ib[r] != 1) continue;
    int s = 0;
    for (int kk = core::imax(1, k - 1); kk <= core::imin(G.nzc, k + 1); kk++)
      for (int jj = core::imax(1, j - 1); jj <= core::imin(G.nyc, j + 1); jj++)
        for (int ii = core::imax(1, i - 1); ii <= core::imin(G.nxc, i + 1); ii++) s += ib[G.r(ii, jj, kk)];
    if (s <= 1) clr[r] = 1;
  }
}
__global__ void k_geo_isolated_clear(long size, const int8_t *clr, int8_t *ib, int *bodynum) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x)
    if (clr[r]) {
      ib[r] = 0;
      bodynum[r] = -1;
    }
}

// fill_cell on one outer face normal to an axis: a fluid cell whose inward
// neighbour is blank becomes blank (the faces are launched in the host
// build's order, first then last)
__global__ void k_geo_fill_face(GridV G, int axis, int face, int8_t *ib, int8_t *fresh) {
  const int n[3] = {G.nxc, G.nyc, G.nzc};
  const long st[3] = {1, G.ex, G.exy};
  const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
  const long per_face = (long)n[a1] * n[a2];
  const long uv = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (uv >= per_face) return;
  const int c = face == 0 ? 1 : n[axis];
  int idx[3];
  idx[axis] = c;
  idx[a1] = 1 + (int)(uv % n[a1]);
  idx[a2] = 1 + (int)(uv / n[a1]);
  const long r = G.r(idx[0], idx[1], idx[2]);
  if (ib[r] != 0) return;
  const bool up = (c == n[axis]), um = (c == 1);
  if ((ib[r - st[axis]] == 1 && up) || (ib[r + st[axis]] == 1 && um)) {
    ib[r] = 1;
    if (fresh[r] == 1) fresh[r] = 0;
  }
}

// GCM ghost marking (GCM_SET_BOUNDARY.F90:49-118): a blank interior cell with
// a fluid face neighbour (clamped to the interior)
__global__ void k_geo_ghosts(GridV G, long size, const int8_t *ib, int8_t *gm) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(G, r, i, j, k);
    int8_t g = 0;
    if (interior(G, i, j, k) && ib[r] == 1) {
      const int iM = core::imax(i - 1, 1), iP = core::imin(i + 1, G.nxc);
      const int jM = core::imax(j - 1, 1), jP = core::imin(j + 1, G.nyc);
      const int kM = core::imax(k - 1, 1), kP = core::imin(k + 1, G.nzc);
      g = (ib[G.r(iM, j, k)] == 0 || ib[G.r(iP, j, k)] == 0 || ib[G.r(i, jM, k)] == 0 ||
           ib[G.r(i, jP, k)] == 0 || ib[G.r(i, j, kM)] == 0 || ib[G.r(i, j, kP)] == 0)
              ? 1
              : 0;
    }
    gm[r] = g;
  }
}

// interior fresh cells as a flag for the compaction
__global__ void k_geo_fresh_flags(GridV G, long size, const int8_t *fresh, int8_t *flag) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(G, r, i, j, k);
    flag[r] = (interior(G, i, j, k) && fresh[r] == 1) ? 1 : 0;
  }
}

// per ghost: its body (the one that blanked it), the intercept on that
// body's surface, the image point and the coefficients
__global__ void k_geo_ghost_stencils(GridV G, const TriV *bodies, GeomConsts K, long ng, const long *cells,
                                     const int *bodynum, double probe_norm, GhostOut *out, int *err) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= ng) return;
  int i, j, k;
  cell_of(G, cells[q], i, j, k);
  GhostOut o;
  int rc = 0;
  const int b0 = bodynum[cells[q]];
  if (b0 < 0) {
    rc = GHOST_ERR_NO_BODY;
  } else {
    o.body = b0;
    const double P[3] = {G.XC[0][i], G.XC[1][j], G.XC[2][k]};
    double bi[3], nrm[3];
    rc = core::tri_intercept(bodies[b0], P, bi, nrm, o.elem, K);
    if (!rc) rc = core::ghost_stencil(G, i, j, k, bi, nrm, probe_norm, o);
  }
  out[q] = o;
  err[q] = rc;
}

// per fresh cell: the nearest body (closest marker, lower index on ties), its
// intercept and GCM_set_freshcell's stencil
__global__ void k_geo_fresh_stencils(GridV G, const TriV *bodies, int nbody, GeomConsts K, long nf,
                                     const long *cells, double probe_norm, FreshOut *out, int *err) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nf) return;
  FreshOut f;
  cell_of(G, cells[q], f.i, f.j, f.k);
  const double p[3] = {G.XC[0][f.i], G.XC[1][f.j], G.XC[2][f.k]};
  double bi[3] = {0.0, 0.0, 0.0}, nrm[3];
  double best = 1e300;
  int rc = 0;
  for (int bb = 0; bb < nbody; bb++) {
    const double d2 = core::nearest_marker_d2(bodies[bb], p);
    if (d2 < best) { best = d2; f.body = bb; }
  }
  if (f.body >= 0) rc = core::tri_intercept(bodies[f.body], p, bi, nrm, f.elem, K);
  if (f.body < 0) rc = GHOST_ERR_NO_BODY;
  if (!rc) rc = core::fresh_stencil(G, bi, probe_norm, f);
  out[q] = f;
  err[q] = rc;
}

// number of nonzero flags: one atomic add per block
__global__ void k_geo_count(long size, const int8_t *flag, unsigned long long *count) {
  __shared__ unsigned int part[kThreads];
  unsigned int c = 0;
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < size; r += (long)gridDim.x * blockDim.x)
    c += flag[r] != 0;
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(count, (unsigned long long)part[0]);
}

// the indices r in [0, size) with flag[r] != 0, ascending, on the device
// (*dev_out) and on the host
int32_t compact(Arena &A, long size, const int8_t *flag, std::vector<long> &host_out, long **dev_out) {
  unsigned long long *d_cnt = nullptr;
  GCHK(A.alloc(&d_cnt, 1));
  GCHK(hipMemset(d_cnt, 0, sizeof(unsigned long long)));
  k_geo_count<<<blocks_for(size) < 4096 ? blocks_for(size) : 4096, kThreads>>>(size, flag, d_cnt);
  GCHK(hipGetLastError());
  unsigned long long n = 0;
  GCHK(hipMemcpy(&n, d_cnt, sizeof(n), hipMemcpyDeviceToHost));
  long *d_out = nullptr, *d_num = nullptr;
  GCHK(A.alloc(&d_out, (size_t)n));
  GCHK(A.alloc(&d_num, 1));
  hipcub::CountingInputIterator<long> it(0);
  size_t bytes = 0;
  GCHK(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flag, d_out, d_num, (int)size));
  char *tmp = nullptr;
  GCHK(A.alloc(&tmp, bytes));
  GCHK(hipcub::DeviceSelect::Flagged(tmp, bytes, it, flag, d_out, d_num, (int)size));
  long got = 0;
  GCHK(hipMemcpy(&got, d_num, sizeof(long), hipMemcpyDeviceToHost));
  if ((unsigned long long)got != n) {
    pp::set_error("pp_geom_build_device: compaction selected %ld of %llu cells", got, n);
    return PP_HIP_ERROR;
  }
  host_out.resize(n);
  if (n > 0) GCHK(hipMemcpy(host_out.data(), d_out, sizeof(long) * n, hipMemcpyDeviceToHost));
  *dev_out = d_out;
  return PP_OK;
}

}  // namespace

int32_t build_geometry_device(pp_geom **out, const Grid &Gh, const TriBodies &tri, double probe_norm,
                              const int8_t *iblank_prev, int dev) {
  *out = nullptr;
  int prev_dev = 0;
  GCHK(hipGetDevice(&prev_dev));
  GCHK(hipSetDevice(dev));
  struct Restore {
    int d;
    ~Restore() { (void)hipSetDevice(d); }
  } restore{prev_dev};
  for (int d = 0; d < 3; d++)
    for (size_t i = 1; i < Gh.XC[d].size(); i++)
      if (!(Gh.XC[d][i] > Gh.XC[d][i - 1])) {
        pp::set_error("pp_geom_build_device: triangulated set_boundary failed (%d: %s)", TRI_ERR_SMALL,
                      tri_error_text(TRI_ERR_SMALL));
        return PP_BAD_GEOMETRY;
      }
  if (Gh.size >= 0x7fffffffL) {
    pp::set_error("pp_geom_build_device: %ld cells exceed the compaction's int range", Gh.size);
    return PP_BAD_ARG;
  }
  Arena A;
  const long size = Gh.size;
  const int nbody = (int)tri.b.size();
  const GeomConsts K = geom_consts();
  // grid and bodies
  GridV G = Gh.view();
  for (int d = 0; d < 3; d++) {
    double *x, *xc, *dx;
    GCHK(A.upload(&x, Gh.X[d].data(), Gh.X[d].size()));
    GCHK(A.upload(&xc, Gh.XC[d].data(), Gh.XC[d].size()));
    GCHK(A.upload(&dx, Gh.DX[d].data(), Gh.DX[d].size()));
    G.X[d] = x; G.XC[d] = xc; G.DX[d] = dx;
  }
  std::vector<TriV> hv(nbody);
  for (int b = 0; b < nbody; b++) {
    const TriBody &B = tri.b[b];
    TriV v = B.view();
    double *p;
    int *ip;
    TreeNode *tn;
    GCHK(A.upload(&p, B.xm.data(), B.xm.size())); v.xm = p;
    GCHK(A.upload(&p, B.ym.data(), B.ym.size())); v.ym = p;
    GCHK(A.upload(&p, B.zm.data(), B.zm.size())); v.zm = p;
    GCHK(A.upload(&ip, B.tri.data(), B.tri.size())); v.tri = ip;
    GCHK(A.upload(&p, B.cx.data(), B.cx.size())); v.cx = p;
    GCHK(A.upload(&p, B.cy.data(), B.cy.size())); v.cy = p;
    GCHK(A.upload(&p, B.cz.data(), B.cz.size())); v.cz = p;
    GCHK(A.upload(&p, B.nx.data(), B.nx.size())); v.nx = p;
    GCHK(A.upload(&p, B.ny.data(), B.ny.size())); v.ny = p;
    GCHK(A.upload(&p, B.nz.data(), B.nz.size())); v.nz = p;
    GCHK(A.upload(&ip, B.ve_ptr.data(), B.ve_ptr.size())); v.ve_ptr = ip;
    GCHK(A.upload(&ip, B.ve_idx.data(), B.ve_idx.size())); v.ve_idx = ip;
    GCHK(A.upload(&tn, B.tree.nodes.data(), B.tree.nodes.size())); v.nodes = tn;
    GCHK(A.upload(&ip, B.tree.perm.data(), B.tree.perm.size())); v.perm = ip;
    hv[b] = v;
  }
  TriV *d_bodies;
  GCHK(A.upload(&d_bodies, hv.data(), hv.size()));
  // fields
  int8_t *und, *temp, *ib, *fresh, *mask, *prev = nullptr;
  int *bodynum, *d_err;
  GCHK(A.alloc(&und, size));
  GCHK(A.alloc(&temp, size));
  GCHK(A.alloc(&ib, size));
  GCHK(A.alloc(&fresh, size));
  GCHK(A.alloc(&mask, size));
  GCHK(A.alloc(&bodynum, size));
  GCHK(A.alloc(&d_err, 1));
  GCHK(hipMemset(d_err, 0, sizeof(int)));
  if (iblank_prev) GCHK(A.upload(&prev, iblank_prev, size));
  const int gb = blocks_for(size);
  k_geo_init<<<gb, kThreads>>>(G, size, und, temp, bodynum, ib, fresh, mask);
  for (int b = 0; b < nbody; b++)
    if (hv[b].ne > 0) k_geo_band<<<(hv[b].ne + kThreads - 1) / kThreads, kThreads>>>(G, hv[b], und);
  k_geo_band_flags<<<gb, kThreads>>>(G, size, und, mask);
  GCHK(hipGetLastError());
  std::vector<long> band;
  long *d_band = nullptr;
  int32_t rc = compact(A, size, mask, band, &d_band);
  if (rc) return rc;
  const long nband = (long)band.size();
  if (nband > 0)
    k_geo_classify<<<(int)((nband + kThreads - 1) / kThreads), kThreads>>>(G, d_bodies, nbody, K, nband, d_band,
                                                                              temp, bodynum, d_err);
  int herr = 0;
  GCHK(hipGetLastError());
  GCHK(hipMemcpy(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr) {
    pp::set_error("pp_geom_build_device: triangulated set_boundary failed (%d: %s)", herr, tri_error_text(herr));
    return PP_BAD_GEOMETRY;
  }
  k_geo_fill_i<<<blocks_for((long)Gh.nyc * Gh.nzc), kThreads>>>(G, und, temp, bodynum);
  k_geo_fill_j<<<blocks_for((long)Gh.nxc * Gh.nzc), kThreads>>>(G, und, temp, bodynum);
  k_geo_fill_k<<<blocks_for((long)Gh.nxc * Gh.nyc), kThreads>>>(G, und, temp, bodynum);
  k_geo_iblank<<<gb, kThreads>>>(G, size, temp, prev, ib, fresh);
  k_geo_isolated_mark<<<gb, kThreads>>>(G, size, ib, mask);
  k_geo_isolated_clear<<<gb, kThreads>>>(size, mask, ib, bodynum);
  const long nface[3] = {(long)Gh.nyc * Gh.nzc, (long)Gh.nzc * Gh.nxc, (long)Gh.nxc * Gh.nyc};
  for (int axis = 0; axis < 3; axis++)
    for (int face = 0; face < 2; face++)
      k_geo_fill_face<<<blocks_for(nface[axis]), kThreads>>>(G, axis, face, ib, fresh);
  k_geo_ghosts<<<gb, kThreads>>>(G, size, ib, mask);
  GCHK(hipGetLastError());
  pp_geom *o = new pp_geom();
  struct Guard {
    pp_geom *&o;
    ~Guard() { delete o; }
  } guard{o};
  std::vector<long> gcells;
  long *d_gcells = nullptr;
  rc = compact(A, size, mask, gcells, &d_gcells);
  if (rc) return rc;
  const long ng = (long)gcells.size();
  // the ghost mask itself (the host build keeps it)
  o->ghost.resize(size);
  GCHK(hipMemcpy(o->ghost.data(), mask, size, hipMemcpyDeviceToHost));
  o->list.resize(3 * ng);
  for (long q = 0; q < ng; q++) {
    const long r = gcells[q];
    const int k = (int)(r / Gh.exy), j = (int)((r % Gh.exy) / Gh.ex), i = (int)(r % Gh.ex);
    o->list[3 * q] = i; o->list[3 * q + 1] = j; o->list[3 * q + 2] = k;
  }
  GhostOut *d_gout;
  int *d_gerr;
  GCHK(A.alloc(&d_gout, ng));
  GCHK(A.alloc(&d_gerr, ng));
  if (ng > 0)
    k_geo_ghost_stencils<<<(int)((ng + kThreads - 1) / kThreads), kThreads>>>(G, d_bodies, K, ng, d_gcells, bodynum,
                                                                               probe_norm, d_gout, d_gerr);
  GCHK(hipGetLastError());
  o->gout.resize(ng);
  std::vector<int> gerr(ng);
  if (ng > 0) {
    GCHK(hipMemcpy(o->gout.data(), d_gout, sizeof(GhostOut) * ng, hipMemcpyDeviceToHost));
    GCHK(hipMemcpy(gerr.data(), d_gerr, sizeof(int) * ng, hipMemcpyDeviceToHost));
  }
  for (long q = 0; q < ng; q++)
    if (gerr[q]) {
      pp::set_error("pp_geom_build_device: ghost %ld at (%d,%d,%d): %s", q, o->list[3 * q], o->list[3 * q + 1],
                    o->list[3 * q + 2], ghost_error_text(gerr[q]));
      return PP_BAD_GEOMETRY;
    }
  o->iblank.resize(size);
  o->fresh.resize(size);
  GCHK(hipMemcpy(o->iblank.data(), ib, size, hipMemcpyDeviceToHost));
  GCHK(hipMemcpy(o->fresh.data(), fresh, size, hipMemcpyDeviceToHost));
  count_cells(Gh, o);
  if (o->nfresh > 0) {
    k_geo_fresh_flags<<<gb, kThreads>>>(G, size, fresh, mask);
    GCHK(hipGetLastError());
    std::vector<long> fcells;
    long *d_fcells = nullptr;
    rc = compact(A, size, mask, fcells, &d_fcells);
    if (rc) return rc;
    const long nf = (long)fcells.size();
    FreshOut *d_fout;
    int *d_ferr;
    GCHK(A.alloc(&d_fout, nf));
    GCHK(A.alloc(&d_ferr, nf));
    k_geo_fresh_stencils<<<(int)((nf + kThreads - 1) / kThreads), kThreads>>>(G, d_bodies, nbody, K, nf, d_fcells,
                                                                               probe_norm, d_fout, d_ferr);
    GCHK(hipGetLastError());
    o->fout.resize(nf);
    std::vector<int> ferr(nf);
    GCHK(hipMemcpy(o->fout.data(), d_fout, sizeof(FreshOut) * nf, hipMemcpyDeviceToHost));
    GCHK(hipMemcpy(ferr.data(), d_ferr, sizeof(int) * nf, hipMemcpyDeviceToHost));
    for (long q = 0; q < nf; q++)
      if (ferr[q]) {
        o->fresh_err = ferr[q];
        o->fresh_err_cell[0] = o->fout[q].i;
        o->fresh_err_cell[1] = o->fout[q].j;
        o->fresh_err_cell[2] = o->fout[q].k;
        break;
      }
  }
  *out = o;
  o = nullptr;  // the guard frees nothing now
  return PP_OK;
}

}  // namespace ppg
