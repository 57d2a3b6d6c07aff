// Host-side set-up of a context: grid metrics, level hierarchy, flags,
// ghost tables and the coarsest-level pseudo-inverse.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "pp_internal.h"
#include "pp_pool.h"
#include "pp_setup.h"


namespace pp {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
const char *last_error() { return g_err.c_str(); }

// fn(a, b) over [0, n) split across up to 16 host threads (contiguous ranges)
void host_parallel(long n, const std::function<void(long, long)> &fn) { pool_parallel(n, fn); }

// make_grid (CORE_GRID.F90:29-151) for one direction.  xn = x(1..n),
// outputs indexed 0..n+1 like the reference arrays.
void make_grid_dir(int n, const double *xn, std::vector<double> &x, std::vector<double> &xc,
                   std::vector<double> &dx, std::vector<double> &dxinv, std::vector<double> &dxc,
                   std::vector<double> &dxcinv) {
  x.assign(n + 2, 0.0); xc.assign(n + 2, 0.0); dx.assign(n + 2, 0.0);
  dxinv.assign(n + 2, 0.0); dxc.assign(n + 2, 0.0); dxcinv.assign(n + 2, 0.0);
  for (int i = 1; i <= n; i++) x[i] = xn[i - 1];
  x[0] = -x[2];
  x[n + 1] = x[n] + x[n] - x[n - 1];
  for (int i = 0; i <= n; i++) {
    xc[i] = 0.5 * (x[i + 1] + x[i]);
    dx[i] = x[i + 1] - x[i];
    dxinv[i] = 1.0 / dx[i];
  }
  xc[n + 1] = x[n + 1];
  for (int i = 1; i <= n; i++) {
    dxc[i] = xc[i] - xc[i - 1];
    dxcinv[i] = 1.0 / dxc[i];
  }
}

// Metrics of one direction of a level from its nodes X[0..n] (0-based,
// X[q] = x(q+1) of the reference): cell widths, centres and the two face
// coefficients of every cell, dxcinv(face)*dxinv(cell), formed in the same
// order as the reference (x(0) = -x(2), x(n+1) = 2x(n)-x(n-1); MG_initial
// CORE_MG.F90:186-205 builds the coarse levels the same way).
static void level_metrics(const std::vector<double> &X, std::vector<double> &w, std::vector<double> &xc,
                          std::vector<double> &cm, std::vector<double> &cp) {
  int n = (int)X.size() - 1;
  std::vector<double> x(n + 3);  // reference indexing 0..n+2 (x(q) = X[q-1])
  for (int q = 1; q <= n + 1; q++) x[q] = X[q - 1];
  x[0] = -x[2];
  x[n + 2] = x[n + 1] + x[n + 1] - x[n];
  std::vector<double> rxc(n + 3), rdx(n + 3), rdxinv(n + 3), rdxcinv(n + 3);
  for (int i = 0; i <= n + 1; i++) {
    rxc[i] = 0.5 * (x[i + 1] + x[i]);
    rdx[i] = x[i + 1] - x[i];
    rdxinv[i] = 1.0 / rdx[i];
  }
  for (int i = 1; i <= n + 1; i++) rdxcinv[i] = 1.0 / (rxc[i] - rxc[i - 1]);
  w.assign(n, 0.0); xc.assign(n, 0.0); cm.assign(n, 0.0); cp.assign(n, 0.0);
  for (int c = 0; c < n; c++) {
    int i = c + 1;
    w[c] = rdx[i];
    xc[c] = rxc[i];
    cm[c] = rdxcinv[i] * rdxinv[i];
    cp[c] = rdxcinv[i + 1] * rdxinv[i];
  }
}

template <typename T>
static T *dev_upload(const std::vector<T> &h) {
  T *d = nullptr;
  if (h.empty()) return nullptr;
  if (hipMalloc(&d, sizeof(T) * h.size()) != hipSuccess) return nullptr;
  hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
  return d;
}
template <typename T>
static std::vector<float> to_f32(const std::vector<T> &h) {
  std::vector<float> r(h.size());
  for (size_t i = 0; i < h.size(); i++) r[i] = (float)h[i];
  return r;
}

void free_levels(pp_ctx *c) {
  for (auto &L : c->lv) {
    hipFree(L.flags);
    for (int d = 0; d < 3; d++) {
      hipFree(L.cm64[d]); hipFree(L.cp64[d]); hipFree(L.w64[d]); hipFree(L.wpar64[d]);
      hipFree(L.cm32[d]); hipFree(L.cp32[d]); hipFree(L.w32[d]); hipFree(L.wpar32[d]);
      hipFree(L.par_d[d]); hipFree(L.oth_d[d]);
    }
    hipFree(L.e32); hipFree(L.rhs32); hipFree(L.res32);
    hipFree(L.e64); hipFree(L.rhs64); hipFree(L.res64);
  }
  c->lv.clear();
  hipFree(c->sp.e); hipFree(c->sp.rhs); hipFree(c->sp.fl); hipFree(c->sp.nshape);
  c->sp = Split0();
}

void free_ghosts(pp_ctx *c) {
  GhostTables &g = c->gh;
  void *ptrs[] = {g.g_ref, g.g_al, g.bself, g.bnd,
                  g.c_ptr, g.c_ref, g.c_al, g.c_w64, g.c_w32, g.bnd_eff,
                  g.nb_ref, g.nb_al, g.nb_w64, g.nb_w32, g.t_ref, g.t_al, g.t_ptr, g.t_ghost, g.t_coef64,
                  g.t_coef32, g.t_vinv64, g.t_vinv32, g.g_tidx, g.lam64[0], g.lam64[1], g.rhsg64,
                  g.tbase64, g.lam32[0], g.lam32[1], g.rhsg32, g.tbase32, g.olam64[0], g.olam64[1],
                  g.orhsg64, g.otbase64, g.v_cD, g.v_buf[0], g.v_buf[1], g.v_kind, g.v_off,
                  g.ct_ref, g.ct_al, g.ct_ptr, g.ct_ghost, g.ct_w64, g.ct_vinv64, g.ctbase64,
                  g.ct_w32, g.ct_vinv32, g.ctbase32, g.g_as, g.c_as, g.nb_as, g.ct_as};
  for (void *p : ptrs) hipFree(p);
  g = GhostTables();
  hipFree(c->cs.G); hipFree(c->cs.cell);
  c->cs = CoarseSolve();
}

// Level hierarchy: level 0 is the grid; a direction is halved while its
// cell count is at least 4 (n -> n/2 coarse cells); coarsening stops when
// the level has at most kCoarseMax cells or no direction can be halved.
// An odd count keeps every other fine node and the last one, so the last
// coarse cell agglomerates three fine cells and the coarse level spans the
// whole domain (the reference's MG_initial halves odd counts by dropping the
// last cell instead, CORE_MG.F90:98-126; the coarse levels only precondition,
// so the converged solution is the same).
static constexpr long kCoarseMax = 4096;
static constexpr int kMaxLevels = 12;

static int coarsen_factor(int n) { return n >= 4 ? 2 : 1; }

long coarsest_cells(int nxc, int nyc, int nzc) {
  int n[3] = {nxc, nyc, nzc};
  for (int l = 1; l < kMaxLevels; l++) {
    if ((long)n[0] * n[1] * n[2] <= kCoarseMax) break;
    int cf[3];
    bool any = false;
    for (int d = 0; d < 3; d++) {
      cf[d] = coarsen_factor(n[d]);
      any |= cf[d] == 2;
    }
    if (!any) break;
    for (int d = 0; d < 3; d++) n[d] /= cf[d];
  }
  return (long)n[0] * n[1] * n[2];
}

int build_levels(pp_ctx *c) {
  free_levels(c);
  std::vector<Level> lv;
  Level L0;
  L0.cx = c->nxc; L0.cy = c->nyc; L0.cz = c->nzc;
  for (int d = 0; d < 3; d++) {
    int n = d == 0 ? c->nxc : d == 1 ? c->nyc : c->nzc;
    L0.node[d].assign(c->hx[d].begin() + 1, c->hx[d].begin() + n + 2);  // x(1..n+1)
  }
  lv.push_back(L0);
  while ((int)lv.size() < kMaxLevels) {
    Level &F = lv.back();
    long cells = (long)F.cx * F.cy * F.cz;
    if (cells <= kCoarseMax) break;
    int ncell[3] = {F.cx, F.cy, F.cz};
    int cf[3];
    bool any = false;
    for (int d = 0; d < 3; d++) {
      cf[d] = coarsen_factor(ncell[d]);
      any |= cf[d] == 2;
    }
    if (!any) break;
    Level C;
    C.cx = F.cx / cf[0]; C.cy = F.cy / cf[1]; C.cz = F.cz / cf[2];
    for (int d = 0; d < 3; d++) {
      C.cf[d] = cf[d];
      if (cf[d] == 1) {
        C.node[d] = F.node[d];
      } else {
        // every other fine node; the last coarse node is the last fine node
        // (for an odd count the last coarse cell spans three fine cells)
        int nc = ncell[d] / 2;
        C.node[d].resize(nc + 1);
        for (int q = 0; q < nc; q++) C.node[d][q] = F.node[d][2 * q];
        C.node[d][nc] = F.node[d][ncell[d]];
      }
    }
    lv.push_back(C);
  }
  for (auto &L : lv) {
    L.sx = ((L.cx + 8 + 15) / 16) * 16;
    L.sy = L.cy + 2;
    L.sxy = (long)L.sx * L.sy;
    L.total = L.sxy * (L.cz + 2) + 64;
    for (int d = 0; d < 3; d++) level_metrics(L.node[d], L.w[d], L.xc[d], L.cm[d], L.cp[d]);
  }
  // prolongation maps fine l -> coarse l+1
  for (size_t l = 0; l + 1 < lv.size(); l++) {
    Level &F = lv[l], &C = lv[l + 1];
    for (int d = 0; d < 3; d++) {
      int nf = (int)F.w[d].size(), nc = (int)C.w[d].size();
      F.par[d].assign(nf, 0); F.oth[d].assign(nf, 0); F.wpar[d].assign(nf, 1.0);
      for (int i = 0; i < nf; i++) {
        if (C.cf[d] == 1) { F.par[d][i] = i; F.oth[d][i] = i; F.wpar[d][i] = 1.0; continue; }
        // parent, and the coarse neighbour on the child's side of the
        // parent's centre (none for a child centred on its parent)
        const int P = std::min(i / 2, nc - 1);
        const double xf = F.xc[d][i], xp = C.xc[d][P];
        const int O = xf < xp ? P - 1 : (xf > xp ? P + 1 : P);
        F.par[d][i] = P;
        if (O < 0 || O >= nc || O == P) { F.oth[d][i] = P; F.wpar[d][i] = 1.0; continue; }
        F.oth[d][i] = O;
        F.wpar[d][i] = (F.xc[d][i] - C.xc[d][O]) / (C.xc[d][P] - C.xc[d][O]);
      }
    }
  }
  c->lv = std::move(lv);
  return PP_OK;
}

// Flags of every level, built on the device.  Level 0: blank/ghost from the
// geometry (a ghost is a blank cell with a fluid face neighbour, neighbours
// clamped to the interior, GCM_SET_BOUNDARY.F90:49-73), closed faces on the
// outer boundary only (set_outer_iup, BOUNDARY_SET.F90:54-92; under GCM the
// fine internal faces stay open, so fluid cells couple to the ghost values).
// Coarse levels: blank if every child is blank, no ghosts, a face closed
// where a fluid cell meets a blank neighbour (MG_Prepare,
// CORE_MG.F90:1736-1781) or the outer boundary.  Padding and halo cells
// are blank.
__global__ void k_flags_level0(RefArgs R, int sx, long sxy, const int8_t *__restrict__ ib, uint8_t *__restrict__ fl,
                               unsigned long long *__restrict__ nghost) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= R.nxc) return;
  // reference (1-based) indices of the cell and its clamped neighbours
  const int I = i + 1, J = j + 1, K = k + 1;
  auto at = [&](int a, int b, int c) { return ib[(long)a + R.ex * b + R.exy * c]; };
  const bool bl = at(I, J, K) != 0;
  uint8_t f = 0;
  if (bl) {
    f |= FL_BLANK;
    const int iM = max(I - 1, 1), iP = min(I + 1, R.nxc), jM = max(J - 1, 1), jP = min(J + 1, R.nyc);
    const int kM = max(K - 1, 1), kP = min(K + 1, R.nzc);
    if (at(iM, J, K) == 0 || at(iP, J, K) == 0 || at(I, jM, K) == 0 || at(I, jP, K) == 0 || at(I, J, kM) == 0 ||
        at(I, J, kP) == 0) {
      f |= FL_GHOST;
      atomicAdd(nghost, 1ull);
    }
  }
  if (i == 0) f |= FL_IM;
  if (i == R.nxc - 1) f |= FL_IP;
  if (j == 0) f |= FL_JM;
  if (j == R.nyc - 1) f |= FL_JP;
  if (k == 0) f |= FL_KM;
  if (k == R.nzc - 1) f |= FL_KP;
  fl[lidx(sx, sxy, i, j, k)] = f;
}

// a coarse cell is blank when every child is (the last coarse cell of a
// direction also takes the fine cell an odd count leaves over)
__global__ void k_flags_coarse_blank(int fsx, long fsxy, int fcx, int fcy, int fcz, const uint8_t *__restrict__ ffl,
                                     int cx, int cy, int cz, int cf0, int cf1, int cf2, uint8_t *__restrict__ blank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= cx) return;
  const int i0 = i * cf0, i1 = i == cx - 1 ? fcx : i0 + cf0;
  const int j0 = j * cf1, j1 = j == cy - 1 ? fcy : j0 + cf1;
  const int k0 = k * cf2, k1 = k == cz - 1 ? fcz : k0 + cf2;
  int all = 1;
  for (int kk = k0; kk < k1; kk++)
    for (int jj = j0; jj < j1; jj++)
      for (int ii = i0; ii < i1; ii++) all &= (ffl[lidx(fsx, fsxy, ii, jj, kk)] & FL_BLANK) ? 1 : 0;
  blank[i + (long)cx * (j + (long)cy * k)] = (uint8_t)all;
}

__global__ void k_flags_coarse(int sx, long sxy, int cx, int cy, int cz, const uint8_t *__restrict__ blank,
                               uint8_t *__restrict__ fl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= cx) return;
  auto b = [&](int a, int c, int d) { return blank[a + (long)cx * (c + (long)cy * d)] != 0; };
  const bool bl = b(i, j, k);
  uint8_t f = bl ? FL_BLANK : 0;
  if (i == 0 || (!bl && b(i - 1, j, k))) f |= FL_IM;
  if (i == cx - 1 || (!bl && b(i + 1, j, k))) f |= FL_IP;
  if (j == 0 || (!bl && b(i, j - 1, k))) f |= FL_JM;
  if (j == cy - 1 || (!bl && b(i, j + 1, k))) f |= FL_JP;
  if (k == 0 || (!bl && b(i, j, k - 1))) f |= FL_KM;
  if (k == cz - 1 || (!bl && b(i, j, k + 1))) f |= FL_KP;
  fl[lidx(sx, sxy, i, j, k)] = f;
}

// lane-group width G and rows per lane S of the x-line pass for a line of
// n <= 1024 cells (FP64 holds 8 rows per lane in registers up to 512 cells)
void line_shape(int n, int &G, int &S, int elem) {
  if (elem == 8) G = n <= 256 ? 32 : 64;
  else G = n <= 128 ? 16 : (n <= 512 ? 32 : 64);  // FP32: 16-lane groups measured faster up to 128 cells
  int need = (n + G - 1) / G;
  static const int opts[] = {1, 2, 3, 4, 6, 8, 12, 16};
  S = 16;
  for (int o : opts)
    if (o >= need) { S = o; break; }
}

// level-0 flags (halo included) into the colour-split copy
__global__ void k_flags_split(int cx, int cy, int cz, int sx, long sxy, const uint8_t *__restrict__ fl, int ssx,
                              long ssxy, int hp, uint8_t *__restrict__ fs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x - 1;
  const int j = (int)blockIdx.y - 1, k = (int)blockIdx.z - 1;
  if (i > cx) return;
  fs[sidx(ssx, ssxy, hp, i, j, k)] = fl[lidx(sx, sxy, i, j, k)];
}

// The colour-split copy of level 0 (Split0) where it applies: the level
// halves by 2 (the fused transfer kernels) and the FP32 x-line chunks are
// even.  Rows: SPLIT_E0 + two halves of hp cells (hp a multiple of 32, at
// least cx/2 + 2, so the halo cells fit and both halves start on a
// 128-byte line).
int setup_split0(pp_ctx *c) {
  Split0 &sp = c->sp;
  hipFree(sp.e); hipFree(sp.rhs); hipFree(sp.fl); hipFree(sp.nshape);
  sp = Split0();
  if (c->lv.size() < 2) return PP_OK;
  const Level &L = c->lv[0], &C = c->lv[1];
  const bool by2 = C.cf[0] == 2 && C.cf[1] == 2 && C.cf[2] == 2 && L.cx == 2 * C.cx && L.cy == 2 * C.cy &&
                   L.cz == 2 * C.cz;
  int G, S;
  line_shape(L.cx, G, S, 4);
  if (!by2 || S % 2 != 0) return PP_OK;
  sp.hp = ((L.cx / 2 + 2 + 31) / 32) * 32;
  sp.sx = SPLIT_E0 + 2 * sp.hp;
  sp.sxy = (long)sp.sx * (L.cy + 2);
  sp.total = sp.sxy * (L.cz + 2);
  if (hipMalloc(&sp.e, sp.total * 4) != hipSuccess || hipMalloc(&sp.rhs, sp.total * 4) != hipSuccess ||
      hipMalloc(&sp.nshape, sp.total * 4) != hipSuccess || hipMalloc(&sp.fl, sp.total) != hipSuccess) {
    set_error("hipMalloc failed for the colour-split level 0 (%ld cells)", sp.total);
    return PP_HIP_ERROR;
  }
  // halo cells of the split correction stay zero (only interior cells are written)
  if (hipMemset(sp.e, 0, sp.total * 4) != hipSuccess || hipMemset(sp.rhs, 0, sp.total * 4) != hipSuccess ||
      hipMemset(sp.nshape, 0, sp.total * 4) != hipSuccess || hipMemset(sp.fl, FL_BLANK, sp.total) != hipSuccess) {
    set_error("colour-split level 0 initialisation failed");
    return PP_HIP_ERROR;
  }
  sp.ok = true;
  return PP_OK;
}

// The flags of every level from the reference-layout iblank on the device
// (d_iblank); the coarsest level's are copied back for build_coarse, and
// *n_ghost is the number of ghost cells the iblank has.
int build_flags(pp_ctx *c, const int8_t *d_iblank, long *n_ghost) {
  hipStream_t s = c->stream;
  if (!c->d_count && hipMalloc(&c->d_count, sizeof(unsigned long long)) != hipSuccess) {
    set_error("counter allocation failed");
    return PP_HIP_ERROR;
  }
  if (hipMemsetAsync(c->d_count, 0, sizeof(unsigned long long), s) != hipSuccess) {
    set_error("counter reset failed");
    return PP_HIP_ERROR;
  }
  for (size_t l = 0; l < c->lv.size(); l++) {
    Level &L = c->lv[l];
    if (hipMemsetAsync(L.flags, FL_BLANK, L.total, s) != hipSuccess) {
      set_error("flag fill failed");
      return PP_HIP_ERROR;
    }
    dim3 grid((L.cx + 63) / 64, L.cy, L.cz), block(64);
    if (l == 0) {
      k_flags_level0<<<grid, block, 0, s>>>(c->ref, L.sx, L.sxy, d_iblank, L.flags, c->d_count);
    } else {
      const Level &F = c->lv[l - 1];
      // the coarse level's FP32 residual buffer holds the blank mask (free at set-up)
      uint8_t *blank = (uint8_t *)L.res32;
      k_flags_coarse_blank<<<grid, block, 0, s>>>(F.sx, F.sxy, F.cx, F.cy, F.cz, F.flags, L.cx, L.cy, L.cz, L.cf[0],
                                                 L.cf[1], L.cf[2], blank);
      k_flags_coarse<<<grid, block, 0, s>>>(L.sx, L.sxy, L.cx, L.cy, L.cz, blank, L.flags);
      if (hipMemsetAsync(blank, 0, (size_t)L.cx * L.cy * L.cz, s) != hipSuccess) {
        set_error("residual buffer reset failed");
        return PP_HIP_ERROR;
      }
    }
  }
  if (c->sp.ok) {
    const Level &L = c->lv[0];
    dim3 grid((L.cx + 2 + 63) / 64, L.cy + 2, L.cz + 2), block(64);
    k_flags_split<<<grid, block, 0, s>>>(L.cx, L.cy, L.cz, L.sx, L.sxy, L.flags, c->sp.sx, c->sp.sxy, c->sp.hp,
                                          c->sp.fl);
  }
  if (hipGetLastError() != hipSuccess) {
    set_error("flag kernels failed to launch");
    return PP_HIP_ERROR;
  }
  Level &C = c->lv.back();
  C.hflags.resize(C.total);
  unsigned long long ng = 0;
  if (hipMemcpyAsync(C.hflags.data(), C.flags, C.total, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&ng, c->d_count, sizeof(ng), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("flag copy failed");
    return PP_HIP_ERROR;
  }
  *n_ghost = (long)ng;
  return PP_OK;
}

int upload_levels(pp_ctx *c) {
  for (size_t l = 0; l < c->lv.size(); l++) {
    Level &L = c->lv[l];
    if (hipMalloc(&L.flags, L.total) != hipSuccess) L.flags = nullptr;
    for (int d = 0; d < 3; d++) {
      L.cm64[d] = dev_upload(L.cm[d]); L.cp64[d] = dev_upload(L.cp[d]); L.w64[d] = dev_upload(L.w[d]);
      L.cm32[d] = dev_upload(to_f32(L.cm[d])); L.cp32[d] = dev_upload(to_f32(L.cp[d]));
      L.w32[d] = dev_upload(to_f32(L.w[d]));
      if (l + 1 < c->lv.size()) {
        L.wpar64[d] = dev_upload(L.wpar[d]); L.wpar32[d] = dev_upload(to_f32(L.wpar[d]));
        L.par_d[d] = dev_upload(L.par[d]); L.oth_d[d] = dev_upload(L.oth[d]);
      }
    }
    size_t n = L.total;
    if (hipMalloc(&L.e32, n * 4) || hipMalloc(&L.rhs32, n * 4) || hipMalloc(&L.res32, n * 4) ||
        hipMalloc(&L.e64, n * 8) || hipMalloc(&L.rhs64, n * 8) || hipMalloc(&L.res64, n * 8)) {
      set_error("hipMalloc failed for level %zu (%zu cells)", l, n);
      return PP_HIP_ERROR;
    }
    hipMemset(L.e32, 0, n * 4); hipMemset(L.rhs32, 0, n * 4); hipMemset(L.res32, 0, n * 4);
    hipMemset(L.e64, 0, n * 8); hipMemset(L.rhs64, 0, n * 8); hipMemset(L.res64, 0, n * 8);
    if (!L.flags) { set_error("hipMalloc failed for level flags"); return PP_HIP_ERROR; }
  }
  return setup_split0(c);
}

// Ghost marking exactly as GCM_set_internal_boundary (GCM_SET_BOUNDARY.F90:49-73):
// reference offset r is a ghost cell when it is an interior blank cell with
// a fluid 6-neighbour, neighbours clamped to the interior (the rule of
// k_flags_level0, for the cells the host set-up looks at)
bool ref_is_ghost(const pp_ctx *c, long r) {
  const auto &ib = c->h_iblank;
  const int K = (int)(r / c->ref.exy), J = (int)((r % c->ref.exy) / c->ref.ex), I = (int)(r % c->ref.ex);
  if (I < 1 || I > c->nxc || J < 1 || J > c->nyc || K < 1 || K > c->nzc || ib[r] != 1) return false;
  auto R = [&](int i, int j, int k) { return (long)i + c->ref.ex * j + c->ref.exy * k; };
  const int iM = std::max(I - 1, 1), iP = std::min(I + 1, c->nxc), jM = std::max(J - 1, 1),
            jP = std::min(J + 1, c->nyc), kM = std::max(K - 1, 1), kP = std::min(K + 1, c->nzc);
  return ib[R(iM, J, K)] == 0 || ib[R(iP, J, K)] == 0 || ib[R(I, jM, K)] == 0 || ib[R(I, jP, K)] == 0 ||
         ib[R(I, J, kM)] == 0 || ib[R(I, J, kP)] == 0;
}

// std::stable_sort of v by cmp on the host threads: contiguous chunks
// sorted in parallel, then merged pairwise (std::merge keeps the first
// range's element first among equals, so the result is the stable order)
template <typename T, typename Cmp>
static void par_stable_sort(std::vector<T> &v, Cmp cmp) {
  const long n = (long)v.size();
  const long parts = n < 8192 ? 1 : 16;
  if (parts == 1) { std::stable_sort(v.begin(), v.end(), cmp); return; }
  std::vector<long> cut(parts + 1);
  for (long p = 0; p <= parts; p++) cut[p] = n * p / parts;
  auto each = [](long m, const std::function<void(long)> &fn) { HostPool::get().run(m, fn); };
  each(parts, [&](long p) { std::stable_sort(v.begin() + cut[p], v.begin() + cut[p + 1], cmp); });
  std::vector<T> buf(n);
  for (long width = 1; width < parts; width *= 2) {
    const long pairs = (parts + 2 * width - 1) / (2 * width);
    each(pairs, [&](long p) {
      const long lo = cut[std::min(parts, 2 * width * p)], mid = cut[std::min(parts, 2 * width * p + width)];
      const long hi = cut[std::min(parts, 2 * width * (p + 1))];
      std::merge(v.begin() + lo, v.begin() + mid, v.begin() + mid, v.begin() + hi, buf.begin() + lo, cmp);
    });
    v.swap(buf);
  }
}

static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};

int build_ghosts(pp_ctx *c, const pp_geometry *g) {
  GhostTables &G = c->gh;
  Level &L0 = c->lv[0];
  int n = g->n_ghost;
  G.n = n;
  G.h_ig.resize(n); G.h_jg.resize(n); G.h_kg.resize(n);
  G.h_g_ref.resize(n); G.h_g_al.resize(n);
  G.h_s_ref.resize(8L * n); G.h_s_al.resize(8L * n); G.h_coef.resize(8L * n); G.h_s_raw.resize(8L * n);
  G.h_bself.resize(n);
  G.solid_in_stencil = 0;
  auto inb = [&](int i, int j, int k) {
    return i >= 1 && i <= c->nxc && j >= 1 && j <= c->nyc && k >= 1 && k <= c->nzc;
  };
  for (int q = 0; q < n; q++) {
    int iG = g->i_ghost[q], jG = g->j_ghost[q], kG = g->k_ghost[q];
    if (!inb(iG, jG, kG)) { set_error("ghost %d index (%d,%d,%d) outside the grid", q, iG, jG, kG); return PP_BAD_GEOMETRY; }
    long rg = ridx(c->ref, iG - 1, jG - 1, kG - 1);
    if (!ref_is_ghost(c, rg)) {
      set_error("ghost %d at (%d,%d,%d) is not a ghost cell of iblank", q, iG, jG, kG);
      return PP_BAD_GEOMETRY;
    }
    G.h_ig[q] = iG; G.h_jg[q] = jG; G.h_kg[q] = kG;
    G.h_g_ref[q] = rg;
    G.h_g_al[q] = lidx(L0.sx, L0.sxy, iG - 1, jG - 1, kG - 1);
    double cself = 0.0;
    for (int r = 0; r < 8; r++) {
      int ii = g->i_cell[q] + incI[r], jj = g->j_cell[q] + incJ[r], kk = g->k_cell[q] + incK[r];
      double cf = g->coeff[8L * q + r];
      if (!std::isfinite(cf)) { set_error("ghost %d coefficient %d is not finite", q, r); return PP_BAD_GEOMETRY; }
      if (ii < 0 || jj < 0 || kk < 0 || ii > c->nxc + 1 || jj > c->nyc + 1 || kk > c->nzc + 1) {
        set_error("ghost %d stencil cell (%d,%d,%d) outside the grid", q, ii, jj, kk);
        return PP_BAD_GEOMETRY;
      }
      G.h_s_raw[8L * q + r] = ridx(c->ref, ii - 1, jj - 1, kk - 1);
      if (ii == iG && jj == jG && kk == kG) {
        cself += cf;
        G.h_coef[8L * q + r] = 0.0;
        G.h_s_ref[8L * q + r] = rg;
        G.h_s_al[8L * q + r] = G.h_g_al[q];
        continue;
      }
      // outer ghost layer holds the Neumann copy of the adjacent interior cell
      int ci = std::min(std::max(ii, 1), c->nxc), cj = std::min(std::max(jj, 1), c->nyc),
          ck = std::min(std::max(kk, 1), c->nzc);
      long rs = ridx(c->ref, ci - 1, cj - 1, ck - 1);
      if (c->h_iblank[rs] == 1 && !ref_is_ghost(c, rs)) G.solid_in_stencil++;
      G.h_coef[8L * q + r] = cf;
      G.h_s_ref[8L * q + r] = rs;
      G.h_s_al[8L * q + r] = lidx(L0.sx, L0.sxy, ci - 1, cj - 1, ck - 1);
    }
    G.h_bself[q] = cself + g->probe_length[q];
  }
  // Jacobi sweeps that solve the ghost-ghost coupling to round-off
  // (the reference iterates Gauss-Seidel to restol_ad).
  // ghost index of a reference offset (-1: not a ghost); the ghost list is
  // in loop order (k, j, i), i.e. sorted by offset
  std::vector<long> gsorted(G.h_g_ref.begin(), G.h_g_ref.end());
  std::vector<int> gorder(n);
  for (int q = 0; q < n; q++) gorder[q] = q;
  if (!std::is_sorted(gsorted.begin(), gsorted.end())) {
    std::sort(gorder.begin(), gorder.end(), [&](int a, int b) { return G.h_g_ref[a] < G.h_g_ref[b]; });
    for (int q = 0; q < n; q++) gsorted[q] = G.h_g_ref[gorder[q]];
  }
  auto ghost_of = [&](long r) -> int {
    auto it = std::lower_bound(gsorted.begin(), gsorted.end(), r);
    return (it != gsorted.end() && *it == r) ? gorder[it - gsorted.begin()] : -1;
  };
  {
    // stencil values: ghosts iterate, fluid cells hold fixed pseudo-random
    // values (a hash of the offset), solid cells 0
    std::vector<int> sg(8L * n);
    std::vector<double> sv(8L * n), gv(n, 0.0), nw(n);
    host_parallel(8L * n, [&](long e0, long e1) {
      for (long e = e0; e < e1; e++) {
        long r = G.h_s_ref[e];
        sg[e] = ghost_of(r);
        uint64_t h = (uint64_t)r * 0x9E3779B97F4A7C15ull + 12345u;
        h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 27; h *= 0x94D049BB133111EBull; h ^= h >> 33;
        sv[e] = c->h_iblank[r] ? 0.0 : (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
      }
    });
    int it = 0;
    double dmax = 1.0;
    while (it < 200 && dmax > 1e-15) {
      dmax = 0.0;
      host_parallel(n, [&](long a, long b) {
        for (long q = a; q < b; q++) {
          double s = 0.0;
          for (int r = 0; r < 8; r++) {
            long e = 8L * q + r;
            s += G.h_coef[e] * (sg[e] >= 0 ? gv[sg[e]] : sv[e]);
          }
          nw[q] = s;
        }
      });
      for (int q = 0; q < n; q++) {
        dmax = std::max(dmax, std::fabs(nw[q] - gv[q]));
        gv[q] = nw[q];
      }
      it++;
    }
    if (dmax > 1e-15) {
      set_error("ghost rows do not converge under Jacobi iteration (change %g after %d sweeps)", dmax, it);
      return PP_BAD_GEOMETRY;
    }
    G.kexact = std::max(it, 1);
  }
  // ---- composite ghost rows ----
  // Row g of M = sum_{t < kexact} C_gg^t C_gf (the kexact Jacobi sweeps
  // above from zero ghost values, expanded once): every solved ghost value
  // becomes one gather over fluid cells.  Stencil entries on solid
  // non-ghost cells read 0 (blank non-ghost cells hold 0 throughout).
  {
    using Row = std::vector<std::pair<long, double>>;
    auto merge = [](Row &r) {
      std::sort(r.begin(), r.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
      Row o;
      for (auto &e : r) {
        if (!o.empty() && o.back().first == e.first) o.back().second += e.second;
        else o.push_back(e);
      }
      r.swap(o);
    };
    std::vector<Row> cgf(n);
    G.h_gg_ptr.assign(1, 0); G.h_gg_idx.clear(); G.h_gg_w.clear();
    for (int q = 0; q < n; q++) {
      for (int r = 0; r < 8; r++) {
        double cf = G.h_coef[8L * q + r];
        long rs = G.h_s_ref[8L * q + r];
        if (cf == 0.0 || rs == G.h_g_ref[q]) continue;
        const int gq = ghost_of(rs);
        if (gq >= 0) {
          G.h_gg_idx.push_back(gq);
          G.h_gg_w.push_back(cf);
        } else if (c->h_iblank[rs] == 0) {
          cgf[q].push_back({rs, cf});
        }
      }
      merge(cgf[q]);
      G.h_gg_ptr.push_back((int)G.h_gg_idx.size());
    }
    // M = S C_gf with S = sum_{t < kexact} C_gg^t: S's rows (over ghosts,
    // a few hops of the ghost layer) by the same recursion, S_t = I + C_gg
    // S_{t-1}, then every row of M once from the fluid rows of C_gf
    std::vector<Row> cur(n);
    bool coupled = !G.h_gg_idx.empty();
    if (coupled) {
      // each thread gathers into one reused buffer, sorts it by offset and
      // sums equal offsets in place (merge's order), then copies the row
      // into its own storage (the capacity the step before last left)
      auto merge_into = [](Row &buf, Row &out) {
        std::sort(buf.begin(), buf.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        size_t m = 0;
        for (size_t i = 0; i < buf.size(); i++) {
          if (m > 0 && buf[m - 1].first == buf[i].first) buf[m - 1].second += buf[i].second;
          else buf[m++] = buf[i];
        }
        out.assign(buf.begin(), buf.begin() + m);
      };
      std::vector<Row> gcur(n), gnxt(n);
      for (int q = 0; q < n; q++) gcur[q] = {{(long)q, 1.0}};
      for (int t = 1; t < G.kexact; t++) {
        // rows are independent within a step: ghosts in parallel
        host_parallel(n, [&](long a, long b) {
          Row buf;
          for (long q = a; q < b; q++) {
            buf.assign(1, {q, 1.0});
            for (int e = G.h_gg_ptr[q]; e < G.h_gg_ptr[q + 1]; e++)
              for (auto &x : gcur[G.h_gg_idx[e]]) buf.push_back({x.first, G.h_gg_w[e] * x.second});
            merge_into(buf, gnxt[q]);
          }
        });
        gcur.swap(gnxt);
      }
      host_parallel(n, [&](long a, long b) {
        Row buf;
        for (long q = a; q < b; q++) {
          buf.clear();
          for (auto &gx : gcur[q])
            for (auto &x : cgf[gx.first]) buf.push_back({x.first, gx.second * x.second});
          merge_into(buf, cur[q]);
        }
      });
    } else {
      cur = cgf;
    }
    G.h_c_ptr.assign(1, 0); G.h_c_ref.clear(); G.h_c_al.clear(); G.h_c_w.clear();
    for (int q = 0; q < n; q++) {
      for (auto &x : cur[q]) {
        long rs = x.first;
        long z = rs / c->ref.exy, rem = rs % c->ref.exy, y = rem / c->ref.ex, xx = rem % c->ref.ex;
        G.h_c_ref.push_back(rs);
        G.h_c_al.push_back(lidx(L0.sx, L0.sxy, (int)xx - 1, (int)y - 1, (int)z - 1));
        G.h_c_w.push_back(x.second);
      }
      G.h_c_ptr.push_back((int)G.h_c_ref.size());
    }
  }
  // ---- adjoint tables ----
  // fluid 6-neighbours of each ghost with the face conductance a*V of the
  // neighbour's row entry for the ghost (fine-level coefficients
  // dxcinv(face)*dxinv(cell) as in calc_residual)
  {
    auto V = [&](int i, int j, int k) { return c->rdx[0][i] * c->rdx[1][j] * c->rdx[2][k]; };
    G.h_nb_ref.assign(6L * n, -1); G.h_nb_al.assign(6L * n, -1); G.h_nb_w.assign(6L * n, 0.0);
    // each ghost fills its own six slots: ghosts in parallel
    host_parallel(n, [&](long qa, long qb) {
      for (int q = (int)qa; q < (int)qb; q++) {
        int gi = G.h_ig[q], gj = G.h_jg[q], gk = G.h_kg[q];   // 1-based
        int nbs[6][3] = {{gi - 1, gj, gk}, {gi + 1, gj, gk}, {gi, gj - 1, gk},
                         {gi, gj + 1, gk}, {gi, gj, gk - 1}, {gi, gj, gk + 1}};
        for (int m = 0; m < 6; m++) {
          int i = nbs[m][0], j = nbs[m][1], k = nbs[m][2];
          if (!inb(i, j, k)) continue;
          long r = ridx(c->ref, i - 1, j - 1, k - 1);
          if (c->h_iblank[r]) continue;
          // the neighbour sits on side m of the ghost; the ghost is on the
          // opposite side of the neighbour
          double coef;
          switch (m) {
            case 0: coef = c->rdxcinv[0][i + 1] * c->rdxinv[0][i]; break;   // ghost is i+1 of f
            case 1: coef = c->rdxcinv[0][i] * c->rdxinv[0][i]; break;       // ghost is i-1 of f
            case 2: coef = c->rdxcinv[1][j + 1] * c->rdxinv[1][j]; break;
            case 3: coef = c->rdxcinv[1][j] * c->rdxinv[1][j]; break;
            case 4: coef = c->rdxcinv[2][k + 1] * c->rdxinv[2][k]; break;
            default: coef = c->rdxcinv[2][k] * c->rdxinv[2][k]; break;
          }
          G.h_nb_ref[6L * q + m] = r;
          G.h_nb_al[6L * q + m] = lidx(L0.sx, L0.sxy, i - 1, j - 1, k - 1);
          G.h_nb_w[6L * q + m] = coef * V(i, j, k);
        }
      }
    });
    // touched cells: CSR of (ghost, coefficient) by stencil cell; the
    // entries in ghost-major order (which stencil entries a ghost keeps is
    // marked in parallel, then each ghost writes from its prefix offset)
    std::vector<uint8_t> keep(8L * n, 0);
    host_parallel(n, [&](long qa, long qb) {
      for (long q = qa; q < qb; q++)
        for (int r = 0; r < 8; r++) {
          const double cf = G.h_coef[8L * q + r];
          const long rs = G.h_s_ref[8L * q + r];
          if (cf == 0.0 || rs == G.h_g_ref[q]) continue;
          if (c->h_iblank[rs] == 1 && !ref_is_ghost(c, rs)) continue;  // solid cells are not unknowns
          keep[8L * q + r] = 1;
        }
    });
    std::vector<long> eoff(n + 1, 0);
    for (int q = 0; q < n; q++) {
      int kq = 0;
      for (int r = 0; r < 8; r++) kq += keep[8L * q + r];
      eoff[q + 1] = eoff[q] + kq;
    }
    std::vector<std::pair<long, std::pair<int, double>>> ent(eoff[n]);
    host_parallel(n, [&](long qa, long qb) {
      for (long q = qa; q < qb; q++) {
        long o = eoff[q];
        for (int r = 0; r < 8; r++)
          if (keep[8L * q + r]) ent[o++] = {G.h_s_ref[8L * q + r], {(int)q, G.h_coef[8L * q + r]}};
      }
    });
    par_stable_sort(ent, [](const auto &a, const auto &b) { return a.first < b.first; });
    G.h_t_ref.clear(); G.h_t_al.clear(); G.h_t_ptr.assign(1, 0); G.h_t_ghost.clear(); G.h_t_coef.clear();
    G.h_t_vinv.clear();
    G.h_g_tidx.assign(n, -1);
    for (size_t e = 0; e < ent.size(); e++) {
      long rs = ent[e].first;
      if (e == 0 || rs != ent[e - 1].first) {
        if (e > 0) G.h_t_ptr.push_back((int)G.h_t_ghost.size());
        int t = (int)G.h_t_ref.size();
        G.h_t_ref.push_back(rs);
        long z = rs / c->ref.exy, rem = rs % c->ref.exy, y = rem / c->ref.ex, x = rem % c->ref.ex;
        G.h_t_al.push_back(lidx(L0.sx, L0.sxy, (int)x - 1, (int)y - 1, (int)z - 1));
        bool fl = c->h_iblank[rs] == 0;
        G.h_t_vinv.push_back(fl ? 1.0 / V((int)x, (int)y, (int)z) : 0.0);
        const int gq = ghost_of(rs);
        if (gq >= 0) G.h_g_tidx[gq] = t;
      }
      G.h_t_ghost.push_back(ent[e].second.first);
      G.h_t_coef.push_back(ent[e].second.second);
    }
    if (!ent.empty()) G.h_t_ptr.push_back((int)G.h_t_ghost.size());
    G.nt = (int)G.h_t_ref.size();
    // transpose of the composite rows by fluid cell
    // (entry e of ghost q's composite row goes to slot e: ghosts in parallel)
    std::vector<std::pair<long, std::pair<int, double>>> ct(G.h_c_ptr.empty() ? 0 : G.h_c_ptr[n]);
    host_parallel(n, [&](long qa, long qb) {
      for (long q = qa; q < qb; q++)
        for (int e = G.h_c_ptr[q]; e < G.h_c_ptr[q + 1]; e++) ct[e] = {G.h_c_ref[e], {(int)q, G.h_c_w[e]}};
    });
    par_stable_sort(ct, [](const auto &a, const auto &b) { return a.first < b.first; });
    G.h_ct_ref.clear(); G.h_ct_al.clear(); G.h_ct_ptr.assign(1, 0); G.h_ct_ghost.clear(); G.h_ct_w.clear();
    G.h_ct_vinv.clear();
    for (size_t e = 0; e < ct.size(); e++) {
      const long rs = ct[e].first;
      if (e == 0 || rs != ct[e - 1].first) {
        if (e > 0) G.h_ct_ptr.push_back((int)G.h_ct_ghost.size());
        long z = rs / c->ref.exy, rem = rs % c->ref.exy, y = rem / c->ref.ex, x = rem % c->ref.ex;
        G.h_ct_ref.push_back(rs);
        G.h_ct_al.push_back(lidx(L0.sx, L0.sxy, (int)x - 1, (int)y - 1, (int)z - 1));
        G.h_ct_vinv.push_back(1.0 / V((int)x, (int)y, (int)z));
      }
      G.h_ct_ghost.push_back(ct[e].second.first);
      G.h_ct_w.push_back(ct[e].second.second);
    }
    if (!ct.empty()) G.h_ct_ptr.push_back((int)G.h_ct_ghost.size());
    G.nct = (int)G.h_ct_ref.size();
  }
  if (c->sp.ok) {
    // the gather's offsets in the colour-split level 0
    const Split0 &sp = c->sp;
    auto split_of = [&](long r) {
      const long z = r / c->ref.exy, rem = r % c->ref.exy, y = rem / c->ref.ex, x = rem % c->ref.ex;
      return sidx(sp.sx, sp.sxy, sp.hp, (int)x - 1, (int)y - 1, (int)z - 1);
    };
    std::vector<long> gs(G.h_g_ref.size()), cs(G.h_c_ref.size()), nb(G.h_nb_ref.size()), ct(G.h_ct_ref.size());
    for (size_t q = 0; q < gs.size(); q++) gs[q] = split_of(G.h_g_ref[q]);
    for (size_t q = 0; q < cs.size(); q++) cs[q] = split_of(G.h_c_ref[q]);
    for (size_t q = 0; q < nb.size(); q++) nb[q] = G.h_nb_ref[q] >= 0 ? split_of(G.h_nb_ref[q]) : -1;
    for (size_t q = 0; q < ct.size(); q++) ct[q] = split_of(G.h_ct_ref[q]);
    G.g_as = dev_upload(gs); G.c_as = dev_upload(cs);
    G.nb_as = dev_upload(nb); G.ct_as = dev_upload(ct);
  }
  G.g_ref = dev_upload(G.h_g_ref); G.g_al = dev_upload(G.h_g_al);
  G.c_ptr = dev_upload(G.h_c_ptr); G.c_ref = dev_upload(G.h_c_ref); G.c_al = dev_upload(G.h_c_al);
  G.c_w64 = dev_upload(G.h_c_w); G.c_w32 = dev_upload(to_f32(G.h_c_w));
  G.nb_ref = dev_upload(G.h_nb_ref); G.nb_al = dev_upload(G.h_nb_al);
  G.nb_w64 = dev_upload(G.h_nb_w); G.nb_w32 = dev_upload(to_f32(G.h_nb_w));
  G.t_ref = dev_upload(G.h_t_ref); G.t_al = dev_upload(G.h_t_al); G.t_ptr = dev_upload(G.h_t_ptr);
  G.t_ghost = dev_upload(G.h_t_ghost); G.t_coef64 = dev_upload(G.h_t_coef);
  G.t_coef32 = dev_upload(to_f32(G.h_t_coef)); G.t_vinv64 = dev_upload(G.h_t_vinv);
  G.t_vinv32 = dev_upload(to_f32(G.h_t_vinv)); G.g_tidx = dev_upload(G.h_g_tidx);
  G.ct_ref = dev_upload(G.h_ct_ref); G.ct_al = dev_upload(G.h_ct_al); G.ct_ptr = dev_upload(G.h_ct_ptr);
  G.ct_ghost = dev_upload(G.h_ct_ghost); G.ct_w64 = dev_upload(G.h_ct_w); G.ct_w32 = dev_upload(to_f32(G.h_ct_w));
  G.ct_vinv64 = dev_upload(G.h_ct_vinv); G.ct_vinv32 = dev_upload(to_f32(G.h_ct_vinv));
  if (G.nct > 0) {
    hipMalloc(&G.ctbase64, sizeof(double) * G.nct);
    hipMalloc(&G.ctbase32, sizeof(float) * G.nct);
  }
  if (n > 0) {
    for (int b = 0; b < 2; b++) {
      hipMalloc(&G.lam64[b], sizeof(double) * n);
      hipMalloc(&G.olam64[b], sizeof(double) * n);
      hipMalloc(&G.lam32[b], sizeof(float) * n);
    }
    hipMalloc(&G.rhsg64, sizeof(double) * n);
    hipMalloc(&G.orhsg64, sizeof(double) * n);
    hipMalloc(&G.rhsg32, sizeof(float) * n);
  }
  if (G.nt > 0) {
    hipMalloc(&G.tbase64, sizeof(double) * G.nt);
    hipMalloc(&G.otbase64, sizeof(double) * G.nt);
    hipMalloc(&G.tbase32, sizeof(float) * G.nt);
  }
  G.bself = dev_upload(G.h_bself);
  if (n > 0) {
    hipMalloc(&G.bnd, sizeof(double) * n);
    hipMemset(G.bnd, 0, sizeof(double) * n);
    hipMalloc(&G.bnd_eff, sizeof(double) * n);
    hipMemset(G.bnd_eff, 0, sizeof(double) * n);
  }
  return PP_OK;
}

// set-up kernels of the coarsest level's dense operator: the bordered
// matrix's nonzeros into a zeroed dense matrix (unique offsets), the unit
// right-hand sides of the fluid cells (column p holds a 1 at row p), and
// the solution columns transposed into the by-output-cell layout
// G[q*n + p] = X[p*m + q] (32x32 tiles through LDS)
__global__ void k_coarse_scatter(long nn, const long *__restrict__ idx, const double *__restrict__ val,
                                 double *__restrict__ a) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < nn) a[idx[t]] = val[t];
}
__global__ void k_coarse_unit_rhs(long nf, const int *__restrict__ cells, int m, double *__restrict__ b) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < nf) {
    const long p = cells[t];
    b[p * m + p] = 1.0;
  }
}
__global__ void k_coarse_transpose(int n, int m, const double *__restrict__ x, double *__restrict__ g) {
  __shared__ double tile[32][33];
  const int p0 = blockIdx.x * 32, q0 = blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int p = p0 + r, q = q0 + threadIdx.x;
    if (p < n && q < n) tile[r][threadIdx.x] = x[(long)p * m + q];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int q = q0 + r, p = p0 + threadIdx.x;
    if (p < n && q < n) g[(long)q * n + p] = tile[threadIdx.x][r];
  }
}

// Coarsest level: G maps the coarse rhs to the correction that solves the
// coarse equations after removing the volume-weighted mean of the rhs on
// every connected fluid component, with zero volume-weighted mean of the
// correction (bordered system, one constraint per component).
int build_coarse(pp_ctx *c) {
  Level &L = c->lv.back();
  int n = L.cx * L.cy * L.cz;
  if (n > kCoarseDenseMax) {
    set_error("coarsest level of %d cells (%dx%dx%d): at most %ld are solved directly", n, L.cx, L.cy, L.cz,
              kCoarseDenseMax);
    return PP_BAD_ARG;
  }
  auto cid = [&](int i, int j, int k) { return i + L.cx * (j + L.cy * k); };
  std::vector<int> comp(n, -1);
  std::vector<char> fluid(n, 0);
  for (int k = 0; k < L.cz; k++)
    for (int j = 0; j < L.cy; j++)
      for (int i = 0; i < L.cx; i++) fluid[cid(i, j, k)] = !(L.hflags[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK);
  int ncomp = 0;
  std::vector<int> stack;
  for (int s = 0; s < n; s++) {
    if (!fluid[s] || comp[s] >= 0) continue;
    comp[s] = ncomp;
    stack.push_back(s);
    while (!stack.empty()) {
      int q = stack.back();
      stack.pop_back();
      int i = q % L.cx, t = q / L.cx, j = t % L.cy, k = t / L.cy;
      uint8_t f = L.hflags[lidx(L.sx, L.sxy, i, j, k)];
      int nb[6][4] = {{i - 1, j, k, FL_IM}, {i + 1, j, k, FL_IP}, {i, j - 1, k, FL_JM},
                      {i, j + 1, k, FL_JP}, {i, j, k - 1, FL_KM}, {i, j, k + 1, FL_KP}};
      for (auto &e : nb) {
        if (f & e[3]) continue;
        int r = cid(e[0], e[1], e[2]);
        if (fluid[r] && comp[r] < 0) { comp[r] = ncomp; stack.push_back(r); }
      }
    }
    ncomp++;
  }
  int m = n + ncomp;
  // the bordered matrix's nonzeros (row-major offsets q*m + column), summed
  // per offset; the dense matrix is formed on the device
  std::vector<std::pair<long, double>> nz;
  nz.reserve((size_t)8 * n + 8);
  for (int q = 0; q < n; q++) {
    if (!fluid[q]) { nz.push_back({(long)q * m + q, 1.0}); continue; }
    int i = q % L.cx, t = q / L.cx, j = t % L.cy, k = t / L.cy;
    uint8_t f = L.hflags[lidx(L.sx, L.sxy, i, j, k)];
    double amx = (f & FL_IM) ? 0 : L.cm[0][i], apx = (f & FL_IP) ? 0 : L.cp[0][i];
    double amy = (f & FL_JM) ? 0 : L.cm[1][j], apy = (f & FL_JP) ? 0 : L.cp[1][j];
    double amz = (f & FL_KM) ? 0 : L.cm[2][k], apz = (f & FL_KP) ? 0 : L.cp[2][k];
    nz.push_back({(long)q * m + q, -(amx + apx) - (amy + apy) - (amz + apz)});
    if (amx != 0) nz.push_back({(long)q * m + cid(i - 1, j, k), amx});
    if (apx != 0) nz.push_back({(long)q * m + cid(i + 1, j, k), apx});
    if (amy != 0) nz.push_back({(long)q * m + cid(i, j - 1, k), amy});
    if (apy != 0) nz.push_back({(long)q * m + cid(i, j + 1, k), apy});
    if (amz != 0) nz.push_back({(long)q * m + cid(i, j, k - 1), amz});
    if (apz != 0) nz.push_back({(long)q * m + cid(i, j, k + 1), apz});
    nz.push_back({(long)q * m + n + comp[q], 1.0});
    double vol = L.w[0][i] * L.w[1][j] * L.w[2][k];
    nz.push_back({(long)(n + comp[q]) * m + q, vol});
  }
  std::stable_sort(nz.begin(), nz.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  std::vector<long> nzi;
  std::vector<double> nzv;
  for (auto &e : nz) {
    if (!nzi.empty() && nzi.back() == e.first) nzv.back() += e.second;
    else { nzi.push_back(e.first); nzv.push_back(e.second); }
  }
  std::vector<int> fluid_cells;
  for (int p = 0; p < n; p++)
    if (fluid[p]) fluid_cells.push_back(p);
  // Explicit solution operator on the GPU: LU (rocSOLVER getrf) of the
  // bordered matrix and solves for the unit right-hand sides of the fluid
  // cells (getrs).  A is assembled row-major, i.e. column-major A^T, hence
  // the transposed solve.  Column p of the solution holds the response of
  // every cell q to a unit rhs at p; the operator is kept by output cell,
  // G[q*n + p], transposed on the device.
  if (!c->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) {
      set_error("rocblas_create_handle failed");
      return PP_HIP_ERROR;
    }
    c->blas = h;
  }
  rocblas_handle h = (rocblas_handle)c->blas;
  rocblas_set_stream(h, c->stream);
  hipStream_t st = c->stream;
  double *dA = nullptr, *dB = nullptr, *dG = nullptr;
  long *dnzi = dev_upload(nzi);
  double *dnzv = dev_upload(nzv);
  int *dfl = dev_upload(fluid_cells);
  rocblas_int *dpiv = nullptr, *dinfo = nullptr;
  hipMalloc(&dA, sizeof(double) * m * m);
  hipMalloc(&dB, sizeof(double) * m * n);
  hipMalloc(&dG, sizeof(double) * n * n);
  hipMalloc(&dpiv, sizeof(rocblas_int) * m);
  hipMalloc(&dinfo, sizeof(rocblas_int));
  int rc = PP_OK;
  if (!dA || !dB || !dG || !dpiv || !dinfo || !dnzi || !dnzv || (!dfl && !fluid_cells.empty())) {
    set_error("coarse solve: device allocation failed");
    rc = PP_HIP_ERROR;
  } else if (hipMemsetAsync(dA, 0, sizeof(double) * m * m, st) != hipSuccess ||
             hipMemsetAsync(dB, 0, sizeof(double) * m * n, st) != hipSuccess) {
    set_error("coarse solve: initialisation failed");
    rc = PP_HIP_ERROR;
  } else {
    const long nn = (long)nzi.size(), nf = (long)fluid_cells.size();
    k_coarse_scatter<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(nn, dnzi, dnzv, dA);
    if (nf > 0) k_coarse_unit_rhs<<<(unsigned)((nf + 255) / 256), 256, 0, st>>>(nf, dfl, m, dB);
    if (hipGetLastError() != hipSuccess) {
      set_error("coarse solve: assembly kernels failed to launch");
      rc = PP_HIP_ERROR;
    } else if (rocsolver_dgetrf(h, m, m, dA, m, dpiv, dinfo) != rocblas_status_success) {
      set_error("rocsolver_dgetrf failed");
      rc = PP_HIP_ERROR;
    } else {
      rocblas_int info = 0;
      hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st);
      hipStreamSynchronize(st);
      if (info != 0) {
        set_error("coarsest-level bordered matrix is singular (getrf info %d)", (int)info);
        rc = PP_BAD_GEOMETRY;
      } else if (rocsolver_dgetrs(h, rocblas_operation_transpose, m, n, dA, m, dpiv, dB, m) !=
                 rocblas_status_success) {
        set_error("rocsolver_dgetrs failed");
        rc = PP_HIP_ERROR;
      }
    }
  }
  if (rc == PP_OK) {
    // solution column p (response to a unit rhs at p) -> G[q*n + p]
    dim3 tg((n + 31) / 32, (n + 31) / 32), tb(32, 8);
    k_coarse_transpose<<<tg, tb, 0, st>>>(n, m, dB, dG);
    std::vector<int> off(n);
    for (int q = 0; q < n; q++) off[q] = (int)lidx(L.sx, L.sxy, q % L.cx, (q / L.cx) % L.cy, q / (L.cx * L.cy));
    int *dcell = dev_upload(off);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess || !dcell) {
      set_error("coarse solve: transpose failed");
      hipFree(dcell);
      rc = PP_HIP_ERROR;
    } else {
      hipFree(c->cs.G); hipFree(c->cs.cell);
      c->cs = CoarseSolve();
      c->cs.n = n;
      c->cs.G = dG;
      c->cs.cell = dcell;
      dG = nullptr;
    }
  }
  hipFree(dnzi); hipFree(dnzv); hipFree(dfl); hipFree(dG);
  hipFree(dA); hipFree(dB); hipFree(dpiv); hipFree(dinfo);
  return rc;
}

}  // namespace pp
