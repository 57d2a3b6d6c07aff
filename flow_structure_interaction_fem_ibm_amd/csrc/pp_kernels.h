// CDNA4 (gfx950) kernels of the pressure-Poisson multigrid.
//
// Every kernel is HBM-bound (a 7-point stencil is not a dense contraction,
// so there is no MFMA work).  Algorithmic bytes per cell of the level the
// kernel runs on (FP32 / FP64 element size s):
//   residual            3s + 1   (e, rhs, flags in; res out)
//   line pass (1 dir)   3s + 1   (e, rhs, flags in; e out) over both colours
//   restrict            s + 1 + s/8 per fine cell
//   prolong + correct   2s + 1 + s/8 per fine cell
//   outer FP64 residual 8+8+1+s  (p, b, flags in; rhs0 out)
//   FP64 update         8+s+1+8  (p, e, flags in; p out)
#pragma once

#include "pp_internal.h"

namespace pp {

// ---------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------
template <typename T>
__device__ inline T tabs(T v) { return v < T(0) ? -v : v; }

__device__ inline double shfl_up_d(double v, int d) { return __shfl_up(v, d, 64); }
__device__ inline float shfl_up_d(float v, int d) { return __shfl_up(v, d, 64); }
__device__ inline double shfl_down_d(double v, int d) { return __shfl_down(v, d, 64); }
__device__ inline float shfl_down_d(float v, int d) { return __shfl_down(v, d, 64); }

// One row of the level operator for cell id with flag f.  Fluid rows are
// the stretched 7-point Laplacian (MG_itsolv CORE_MG.F90:1945-1968); blank
// rows are identity holding the ghost value (rhs = ghostMark*var,
// CORE_MG.F90:1970-1975).
template <typename T>
struct Row {
  T amx, apx, amy, apy, amz, apz, diag;
};

template <typename T>
__device__ inline Row<T> make_row(const LevelArgs<T> &L, uint8_t f, int i, int j, int k) {
  Row<T> r;
  r.amx = (f & FL_IM) ? T(0) : L.cmx[i];
  r.apx = (f & FL_IP) ? T(0) : L.cpx[i];
  r.amy = (f & FL_JM) ? T(0) : L.cmy[j];
  r.apy = (f & FL_JP) ? T(0) : L.cpy[j];
  r.amz = (f & FL_KM) ? T(0) : L.cmz[k];
  r.apz = (f & FL_KP) ? T(0) : L.cpz[k];
  // (acx + acy + acz) in the reference's order
  r.diag = (-(r.amx + r.apx) + -(r.amy + r.apy)) + -(r.amz + r.apz);
  return r;
}

// ---------------------------------------------------------------------
// block reductions (deterministic: fixed tree order)
// ---------------------------------------------------------------------
struct SumMax {
  double sum;
  double mx;
  long arg;
};

__device__ inline SumMax summax_combine(SumMax a, SumMax b) {
  SumMax r;
  r.sum = a.sum + b.sum;
  if (b.mx > a.mx || (b.mx == a.mx && b.arg >= 0 && (a.arg < 0 || b.arg < a.arg))) {
    r.mx = b.mx; r.arg = b.arg;
  } else {
    r.mx = a.mx; r.arg = a.arg;
  }
  return r;
}

template <int NT>
__device__ inline SumMax block_summax(SumMax v) {
  __shared__ double s_sum[NT / 64];
  __shared__ double s_mx[NT / 64];
  __shared__ long s_arg[NT / 64];
  for (int d = 32; d >= 1; d >>= 1) {
    SumMax o;
    o.sum = __shfl_down(v.sum, d, 64);
    o.mx = __shfl_down(v.mx, d, 64);
    o.arg = __shfl_down(v.arg, d, 64);
    v = summax_combine(v, o);
  }
  int tid = threadIdx.x + threadIdx.y * blockDim.x;
  int w = tid >> 6, lane = tid & 63;
  if (lane == 0) { s_sum[w] = v.sum; s_mx[w] = v.mx; s_arg[w] = v.arg; }
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < NT / 64; q++) {
      SumMax o{s_sum[q], s_mx[q], s_arg[q]};
      v = summax_combine(v, o);
    }
  }
  return v;  // valid in thread 0
}

// partial[blk*3 + {0,1,2}] = {sum, max, arg}
__global__ void k_reduce_summax(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  SumMax v{0.0, -1.0, -1};
  for (int b = threadIdx.x; b < nblk; b += 256) {
    SumMax o{part[3 * b], part[3 * b + 1], (long)part[3 * b + 2]};
    v = summax_combine(v, o);
  }
  // per-thread partials combined in thread order -> deterministic
  v = block_summax<256>(v);
  if (threadIdx.x == 0) { out[0] = v.sum; out[1] = v.mx; out[2] = (double)v.arg; }
}

// multiple independent sums: part[blk*nq + q]
__global__ void k_reduce_sums(const double *__restrict__ part, int nblk, int nq, double *__restrict__ out) {
  __shared__ double s[256];
  for (int q = 0; q < nq; q++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) v += part[(long)b * nq + q];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
      if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = s[0];
    __syncthreads();
  }
}

template <int NT>
__device__ inline double block_sum(double v, double *s) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
  int tid = threadIdx.x + threadIdx.y * blockDim.x;
  int w = tid >> 6, lane = tid & 63;
  if (lane == 0) s[w] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int q = 0; q < NT / 64; q++) r += s[q];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------
// fills and copies
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *__restrict__ p, long n, T v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}

// ---------------------------------------------------------------------
// FP64 outer residual on the reference layout (calc_residual,
// CORE_POISSON.F90:1111-1222): rhs0 = (b - A p) on fluid cells, 0 elsewhere;
// e0 = 0; per-block partial sum of squares, max and argmax.
// Grid: x = row tiles of 256 cells, y = j, z = k.
// ---------------------------------------------------------------------
template <typename TE, int NT>
__global__ void __launch_bounds__(NT) k_outer_residual(RefArgs R, LevelArgs<double> L, const double *__restrict__ p,
                                                       const double *__restrict__ b, TE *__restrict__ rhs0,
                                                       TE *__restrict__ e0, double *__restrict__ part,
                                                       double *__restrict__ rout) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  SumMax v{0.0, 0.0, -1};
  if (i < R.nxc && j < R.nyc) {
    long a = lidx(L.sx, L.sxy, i, j, k);
    long r = ridx(R, i, j, k);
    uint8_t f = L.fl[a];
    double res = 0.0;
    if (!(f & FL_BLANK)) {
      Row<double> w = make_row(L, f, i, j, k);
      res = b[r] - p[r] * w.diag - p[r - 1] * w.amx - p[r + 1] * w.apx - p[r - R.ex] * w.amy -
            p[r + R.ex] * w.apy - p[r - R.exy] * w.amz - p[r + R.exy] * w.apz;
    }
    rhs0[a] = (TE)res;
    e0[a] = TE(0);
    if (rout) rout[r] = res;
    v.sum = res * res;
    v.mx = fabs(res);
    v.arg = (long)r;
  }
  v = block_summax<NT>(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    part[3 * blk] = v.sum;
    part[3 * blk + 1] = v.mx;
    part[3 * blk + 2] = (double)v.arg;
  }
}

// Same residual written as a full FP64 field in the reference layout
// (pp_calc_residual), masked by (1-iblank) like calc_residual.
template <int NT>
__global__ void __launch_bounds__(NT) k_ref_residual(RefArgs R, LevelArgs<double> L, const double *__restrict__ p,
                                                     const double *__restrict__ b, double *__restrict__ res_out,
                                                     double *__restrict__ part) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  SumMax v{0.0, 0.0, -1};
  if (i < R.nxc && j < R.nyc) {
    long a = lidx(L.sx, L.sxy, i, j, k);
    long r = ridx(R, i, j, k);
    uint8_t f = L.fl[a];
    double res = 0.0;
    if (!(f & FL_BLANK)) {
      Row<double> w = make_row(L, f, i, j, k);
      res = b[r] - p[r] * w.diag - p[r - 1] * w.amx - p[r + 1] * w.apx - p[r - R.ex] * w.amy -
            p[r + R.ex] * w.apy - p[r - R.exy] * w.amz - p[r + R.exy] * w.apz;
    }
    if (res_out) res_out[r] = res;
    v.sum = res * res;
    v.mx = fabs(res);
    v.arg = (long)r;
  }
  v = block_summax<NT>(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    part[3 * blk] = v.sum;
    part[3 * blk + 1] = v.mx;
    part[3 * blk + 2] = (double)v.arg;
  }
}

// p += e on fluid cells (refinement update, FP64 reference layout)
template <typename TE>
__global__ void __launch_bounds__(256) k_update(RefArgs R, LevelArgs<double> L, double *__restrict__ p,
                                                const TE *__restrict__ e) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  if (L.fl[a] & FL_BLANK) return;
  long r = ridx(R, i, j, k);
  p[r] = p[r] + (double)e[a];
}

// b_work = rhs on fluid cells, 0 elsewhere (halo included: zeroed beforehand)
__global__ void __launch_bounds__(256) k_mask_rhs(RefArgs R, LevelArgs<double> L, const double *__restrict__ rhs,
                                                  double *__restrict__ bw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  long r = ridx(R, i, j, k);
  bw[r] = (L.fl[a] & FL_BLANK) ? 0.0 : rhs[r];
}

// per-block partial sums over fluid cells: q0 = sum(w*b), q1 = sum(w),
// q2 = sum(b*b), q3 = sum(V*p), q4 = sum(V)
template <int NT>
__global__ void __launch_bounds__(NT) k_fluid_sums(RefArgs R, LevelArgs<double> L, const double *__restrict__ w,
                                                   const double *__restrict__ b, const double *__restrict__ p,
                                                   double *__restrict__ part) {
  __shared__ double s[NT / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double q[5] = {0, 0, 0, 0, 0};
  if (i < R.nxc && j < R.nyc) {
    long a = lidx(L.sx, L.sxy, i, j, k);
    if (!(L.fl[a] & FL_BLANK)) {
      long r = ridx(R, i, j, k);
      double vol = L.wx[i] * L.wy[j] * L.wz[k];
      double wv = w ? w[r] : vol;
      double bv = b ? b[r] : 0.0;
      q[0] = wv * bv;
      q[1] = wv;
      q[2] = bv * bv;
      q[3] = p ? p[r] * vol : 0.0;
      q[4] = vol;
    }
  }
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  for (int t = 0; t < 5; t++) {
    double v = block_sum<NT>(q[t], s);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blk * 5 + t] = v;
  }
}

// x[fluid] += alpha (alpha read from device memory: scale * src[0] (+ scale2*src[1]))
__global__ void __launch_bounds__(256) k_fluid_shift(RefArgs R, LevelArgs<double> L, double *__restrict__ x,
                                                     const double *__restrict__ num, const double *__restrict__ den,
                                                     double sign) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  if (L.fl[a] & FL_BLANK) return;
  long r = ridx(R, i, j, k);
  x[r] = x[r] + sign * (num[0] / den[0]);
}

// subtracting_average_pressure body (CORE_POISSON.F90:447-457):
// p = p - mean on fluid, 0 on every blank cell.
__global__ void __launch_bounds__(256) k_mean_sub(RefArgs R, LevelArgs<double> L, double *__restrict__ p,
                                                  const double *__restrict__ sums) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  long r = ridx(R, i, j, k);
  double mean = sums[3] / sums[4];
  if (L.fl[a] & FL_BLANK) p[r] = 0.0;
  else p[r] = p[r] - mean;
}

// fresh cells: mean of non-blank, non-fresh 3x3x3 neighbours
// (CORE_POISSON.F90:459-488); reads a snapshot so the result does not depend
// on thread order.  The reference walks the cells in order and a fresh cell
// can read a fresh neighbour only through the fresh_cell test, which
// excludes them, so the snapshot gives the same values.
__global__ void __launch_bounds__(256) k_fresh_fill(RefArgs R, const int8_t *__restrict__ iblank_ref,
                                                    const int8_t *__restrict__ fresh, const double *__restrict__ src,
                                                    double *__restrict__ p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long r = ridx(R, i, j, k);
  if (fresh[r] != 1) return;
  int n = 0;
  double s = 0.0;
  for (int dk = -1; dk <= 1; dk++)
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        long q = r + di + dj * R.ex + dk * R.exy;
        if (iblank_ref[q] != 1 && fresh[q] != 1) { n++; s += src[q]; }
      }
  p[r] = s / (double)n;
}

// ---------------------------------------------------------------------
// ghost rows: p_g = sum_r c_r p_r + bnd_g (GCM_p_set_bc_internal), Jacobi
// with a buffer so the result does not depend on thread order.
// ---------------------------------------------------------------------
// adjoint (transposed) ghost rows.  With lambda_f = V_f mu_f on fluid cells
// the transposed fluid block is the same Laplacian acting on mu; the ghost
// columns give
//   lambda_g = s_g - sum_{fluid f adj g} (a_fg V_f) mu_f + sum_g' c_g'g lambda_g'
// and the fluid rows pick up t_f = sum_g c_gf lambda_g as a source:
//   L mu = (s_f + t_f) / V_f.
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_adj_ghost_rhs(int n, const long *__restrict__ nb_idx,
                                                       const T *__restrict__ nb_w, const double *__restrict__ sg,
                                                       const T *__restrict__ mu, T *__restrict__ rhsg) {
  int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  T s = sg ? (T)sg[g] : T(0);
#pragma unroll
  for (int m = 0; m < 6; m++) {
    long q = nb_idx[6 * g + m];
    if (q >= 0) s -= nb_w[6 * g + m] * mu[q];
  }
  rhsg[g] = s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_ghost_jacobi(int n, const int *__restrict__ g_tidx,
                                                          const int *__restrict__ t_ptr,
                                                          const int *__restrict__ t_ghost,
                                                          const T *__restrict__ t_coef, const T *__restrict__ rhsg,
                                                          const T *__restrict__ lam_old, T *__restrict__ lam_new) {
  int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  T s = rhsg[g];
  int t = g_tidx[g];
  if (t >= 0)
    for (int q = t_ptr[t]; q < t_ptr[t + 1]; q++) s += t_coef[q] * lam_old[t_ghost[q]];
  lam_new[g] = s;
}

// target[t_idx[t]] = base[t] + vinv[t] * sum c lambda  for touched fluid cells
template <typename T>
__global__ void __launch_bounds__(256) k_adj_touch(int nt, const long *__restrict__ t_idx,
                                                   const int *__restrict__ t_ptr, const int *__restrict__ t_ghost,
                                                   const T *__restrict__ t_coef, const T *__restrict__ vinv,
                                                   const T *__restrict__ lam, const T *__restrict__ base,
                                                   T *__restrict__ target) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nt) return;
  T vi = vinv[t];
  if (vi == T(0)) return;  // ghost cell: handled by the Jacobi rows
  T s = T(0);
  for (int q = t_ptr[t]; q < t_ptr[t + 1]; q++) s += t_coef[q] * lam[t_ghost[q]];
  target[t_idx[t]] = base[t] + vi * s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_gather_idx(int n, const long *__restrict__ idx, const T *__restrict__ src,
                                                    T *__restrict__ dst) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  dst[t] = src[idx[t]];
}

// ---------------------------------------------------------------------
template <typename T, typename TC>
__global__ void __launch_bounds__(256) k_ghost_gather(int n, const long *__restrict__ sidx, const TC *__restrict__ coef,
                                                      const double *__restrict__ bnd, const T *__restrict__ p,
                                                      T *__restrict__ buf) {
  int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  TC s = bnd ? (TC)bnd[g] : TC(0);
#pragma unroll
  for (int r = 0; r < 8; r++) s += coef[8 * g + r] * (TC)p[sidx[8 * g + r]];
  buf[g] = (T)s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_ghost_scatter(int n, const long *__restrict__ gidx, const T *__restrict__ buf,
                                                       T *__restrict__ p) {
  int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  p[gidx[g]] = buf[g];
}

// outer ghost layer, Neumann copy (set_outer_ghost_pres,
// CORE_SET_OUTER_GHOST.F90:66-213) on the reference layout: faces in x,
// then y, then z, each reading the layer written before (edges and corners
// end up as the nearest interior value).  One kernel per axis.
__global__ void k_outer_ghost_x(RefArgs R, double *__restrict__ p) {
  int j = blockIdx.x * 256 + threadIdx.x - 0;  // 0..ny(ref)
  int k = blockIdx.y;
  int nx = R.nxc + 1;
  if (j > R.nyc + 1 || k > R.nzc + 1) return;
  long row = R.ex * j + R.exy * k;
  p[row + 0] = p[row + 1];
  p[row + nx] = p[row + nx - 1];
}
__global__ void k_outer_ghost_y(RefArgs R, double *__restrict__ p) {
  int i = blockIdx.x * 256 + threadIdx.x;
  int k = blockIdx.y;
  int ny = R.nyc + 1;
  if (i > R.nxc + 1 || k > R.nzc + 1) return;
  long b = i + R.exy * k;
  p[b] = p[b + R.ex];
  p[b + R.ex * ny] = p[b + R.ex * (ny - 1)];
}
__global__ void k_outer_ghost_z(RefArgs R, double *__restrict__ p) {
  int i = blockIdx.x * 256 + threadIdx.x;
  int j = blockIdx.y;
  int nz = R.nzc + 1;
  if (i > R.nxc + 1 || j > R.nyc + 1) return;
  long b = i + R.ex * j;
  p[b] = p[b + R.exy];
  p[b + R.exy * nz] = p[b + R.exy * (nz - 1)];
}

// ---------------------------------------------------------------------
// level residual: res = rhs - L e on fluid cells, 0 on blank cells
// (MG_residual, CORE_MG.F90:2274-2387)
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_residual(LevelArgs<T> L, const T *__restrict__ e, const T *__restrict__ rhs,
                                                  T *__restrict__ res) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= L.cx || j >= L.cy) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  uint8_t f = L.fl[a];
  T v = T(0);
  if (!(f & FL_BLANK)) {
    Row<T> w = make_row(L, f, i, j, k);
    v = rhs[a] - e[a] * w.diag - e[a - 1] * w.amx - e[a + 1] * w.apx - e[a - L.sx] * w.amy - e[a + L.sx] * w.apy -
        e[a - L.sxy] * w.amz - e[a + L.sxy] * w.apz;
  }
  res[a] = v;
}

// ---------------------------------------------------------------------
// restriction: coarse rhs = fluid-volume-weighted mean of the children's
// residuals (the reference's masked averaging, MG_Restrict_X/Y/Z
// CORE_MG.F90:2511-2538, applied over all coarsened directions at once);
// 0 where no child is fluid.  Also zeroes the coarse correction.
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(LevelArgs<T> F, LevelArgs<T> C, int fx, int fy, int fz,
                                                  const T *__restrict__ res, T *__restrict__ rhs_c,
                                                  T *__restrict__ e_c) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  int J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
  if (I >= C.cx || J >= C.cy) return;
  T num = T(0), den = T(0);
  for (int c = 0; c < fz; c++) {
    int k = K * fz + c;
    for (int b = 0; b < fy; b++) {
      int j = J * fy + b;
      T wyz = F.wy[j] * F.wz[k];
      for (int a = 0; a < fx; a++) {
        int i = I * fx + a;
        long id = lidx(F.sx, F.sxy, i, j, k);
        if (!(F.fl[id] & FL_BLANK)) {
          T v = F.wx[i] * wyz;
          num += v * res[id];
          den += v;
        }
      }
    }
  }
  long cid = lidx(C.sx, C.sxy, I, J, K);
  rhs_c[cid] = den > T(0) ? num / den : T(0);
  e_c[cid] = T(0);
}

// ---------------------------------------------------------------------
// prolongation + correction: e_f += trilinear interpolation of the coarse
// correction over the fluid coarse cells (weights renormalised over fluid
// coarse neighbours as the reference's masked prolongation does,
// CORE_MG.F90:3004-3022), on fluid fine cells only (CORE_MG.F90:1354-1355).
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_prolong(LevelArgs<T> F, LevelArgs<T> C, ProlongDir<T> px, ProlongDir<T> py,
                                                 ProlongDir<T> pz, const T *__restrict__ e_c, T *__restrict__ e_f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= F.cx || j >= F.cy) return;
  long id = lidx(F.sx, F.sxy, i, j, k);
  if (F.fl[id] & FL_BLANK) return;
  int ci[2] = {px.par[i], px.oth[i]};
  int cj[2] = {py.par[j], py.oth[j]};
  int ck[2] = {pz.par[k], pz.oth[k]};
  T wi[2] = {px.wpar[i], T(1) - px.wpar[i]};
  T wj[2] = {py.wpar[j], T(1) - py.wpar[j]};
  T wk[2] = {pz.wpar[k], T(1) - pz.wpar[k]};
  T num = T(0), den = T(0);
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        T w = wi[a] * wj[b] * wk[c];
        if (w == T(0)) continue;
        long q = lidx(C.sx, C.sxy, ci[a], cj[b], ck[c]);
        if (C.fl[q] & FL_BLANK) continue;
        num += w * e_c[q];
        den += w;
      }
  if (den > T(0)) e_f[id] = e_f[id] + num / den;
}

// ---------------------------------------------------------------------
// x-line zebra pass (lines along the contiguous index): one wavefront per
// line (j,k) with (j+k)%2 == color.  Lane l owns S consecutive cells; each
// lane eliminates its segment down to one reduced unknown (its last cell),
// the 64 reduced equations are solved by parallel cyclic reduction across
// the wavefront with lane shuffles, and each lane back-substitutes.
// ---------------------------------------------------------------------
template <typename T>
struct Tri {
  T a, b, c, d;
};

// Local elimination of a segment of S rows (x_m in terms of the previous
// segment's last unknown XL and this segment's last unknown XS):
//   x_m = D_m + F_m XL + G_m XS   (m < S-1),   x_{S-1} = XS.
// Returns the reduced equation pieces (row S-1 after forward elimination).
template <typename T, int S>
__device__ inline void seg_eliminate(const T (&a)[S], const T (&b)[S], const T (&c)[S], const T (&d)[S], T (&D)[S],
                                     T (&F)[S], T (&G)[S], T &cl, T &dl, T &fl) {
  T cp[S], dp[S], fp[S];
  {
    T inv = T(1) / b[0];
    cp[0] = c[0] * inv;
    dp[0] = d[0] * inv;
    fp[0] = -a[0] * inv;
  }
#pragma unroll
  for (int m = 1; m < S; m++) {
    T inv = T(1) / (b[m] - a[m] * cp[m - 1]);
    cp[m] = c[m] * inv;
    dp[m] = (d[m] - a[m] * dp[m - 1]) * inv;
    fp[m] = -a[m] * fp[m - 1] * inv;
  }
  D[S - 1] = T(0);
  F[S - 1] = T(0);
  G[S - 1] = T(1);
#pragma unroll
  for (int m = S - 2; m >= 0; m--) {
    D[m] = dp[m] - cp[m] * D[m + 1];
    F[m] = fp[m] - cp[m] * F[m + 1];
    G[m] = -cp[m] * G[m + 1];
  }
  cl = cp[S - 1];
  dl = dp[S - 1];
  fl = fp[S - 1];
}

// PCR over the 64 lanes of a wavefront for one reduced equation per lane:
// A X_{l-1} + B X_l + C X_{l+1} = R.  Returns X_l.
template <typename T>
__device__ inline T wave_pcr(T A, T B, T C, T R) {
  int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T Am = shfl_up_d(A, d), Bm = shfl_up_d(B, d), Cm = shfl_up_d(C, d), Rm = shfl_up_d(R, d);
    T Ap = shfl_down_d(A, d), Bp = shfl_down_d(B, d), Cp = shfl_down_d(C, d), Rp = shfl_down_d(R, d);
    if (lane < d) { Am = T(0); Bm = T(1); Cm = T(0); Rm = T(0); }
    if (lane + d >= 64) { Ap = T(0); Bp = T(1); Cp = T(0); Rp = T(0); }
    T k1 = A / Bm, k2 = C / Bp;
    T nA = -Am * k1;
    T nC = -Cp * k2;
    T nB = B - Cm * k1 - Ap * k2;
    T nR = R - Rm * k1 - Rp * k2;
    A = nA; B = nB; C = nC; R = nR;
  }
  return R / B;
}

// Solve one tridiagonal line held as S rows per lane across a wavefront.
template <typename T, int S>
__device__ inline void wave_line_solve(const T (&a)[S], const T (&b)[S], const T (&c)[S], const T (&d)[S],
                                       T (&x)[S]) {
  T D[S], F[S], G[S], cl, dl, fl;
  seg_eliminate<T, S>(a, b, c, d, D, F, G, cl, dl, fl);
  T D0n = shfl_down_d(D[0], 1), F0n = shfl_down_d(F[0], 1), G0n = shfl_down_d(G[0], 1);
  int lane = threadIdx.x & 63;
  if (lane == 63) { D0n = T(0); F0n = T(0); G0n = T(0); }
  T A = -fl;
  T B = T(1) + cl * F0n;
  T C = cl * G0n;
  T R = dl - cl * D0n;
  T X = wave_pcr<T>(A, B, C, R);
  T XL = shfl_up_d(X, 1);
  if (lane == 0) XL = T(0);
#pragma unroll
  for (int m = 0; m < S - 1; m++) x[m] = D[m] + F[m] * XL + G[m] * X;
  x[S - 1] = X;
}

template <typename T, int S>
__global__ void __launch_bounds__(256) k_xline(LevelArgs<T> L, const T *__restrict__ rhs, T *__restrict__ e,
                                               int color, T omega) {
  int lane = threadIdx.x & 63;
  int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  int nyh = (L.cy + 1) >> 1;
  int k = wave / nyh;
  int jj = wave - k * nyh;
  if (k >= L.cz) return;  // uniform per wavefront
  int j = 2 * jj + ((k + color) & 1);
  if (j >= L.cy) return;  // uniform per wavefront
  long row = lidx(L.sx, L.sxy, 0, j, k);
  T a[S], b[S], c[S], d[S], eo[S], x[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    int i = lane * S + s;
    a[s] = T(0); b[s] = T(1); c[s] = T(0); d[s] = T(0); eo[s] = T(0);
    if (i < L.cx) {
      long id = row + i;
      uint8_t f = L.fl[id];
      T ev = e[id];
      eo[s] = ev;
      if (f & FL_BLANK) {
        d[s] = (f & FL_GHOST) ? ev : T(0);
      } else {
        Row<T> w = make_row(L, f, i, j, k);
        T r = rhs[id] - e[id - L.sx] * w.amy - e[id + L.sx] * w.apy - e[id - L.sxy] * w.amz -
              e[id + L.sxy] * w.apz;
        if (tabs(w.diag) < T(1.0e-12)) {
          d[s] = ev;
        } else {
          a[s] = w.amx; b[s] = w.diag; c[s] = w.apx; d[s] = r;
        }
      }
    }
  }
  wave_line_solve<T, S>(a, b, c, d, x);
#pragma unroll
  for (int s = 0; s < S; s++) {
    int i = lane * S + s;
    if (i < L.cx) e[row + i] = (T(1) - omega) * eo[s] + omega * x[s];
  }
}

// ---------------------------------------------------------------------
// y- and z-line zebra passes (lines across the contiguous index): lines
// (i, plane) with (i + plane) % 2 == colour, plane = k for DIR = 1 (lines
// along j) and plane = j for DIR = 2 (lines along k), so every neighbour
// line of a line has the other colour and the pass is deterministic.  A
// block of TX x P threads covers TX lines of one colour (every other i of
// a 2*TX-wide tile) and P segments of LS cells of each line.  Each thread
// eliminates its segment in registers, the P reduced equations of a line
// are solved by one thread (Thomas) from LDS, and every thread
// back-substitutes.  The tile's loads cover 2*TX consecutive i, half of
// them own cells and half the neighbour lines read by the stencil.
// ---------------------------------------------------------------------
template <typename T, int LS, int DIR>
__global__ void __launch_bounds__(1024) k_yzline(LevelArgs<T> Lv, const T *__restrict__ rhs, T *__restrict__ e,
                                                 int color, T omega) {
  extern __shared__ unsigned char smem_raw[];
  T *sh = reinterpret_cast<T *>(smem_raw);
  const int P = blockDim.y;
  const int tx = threadIdx.x, seg = threadIdx.y;
  const int plane = blockIdx.y;
  const int i = 2 * (blockIdx.x * blockDim.x + tx) + ((plane + color) & 1);
  int j, k;
  int n;
  long stride;
  if (DIR == 1) {
    k = plane;
    j = 0;
    n = Lv.cy;
    stride = Lv.sx;
  } else {
    j = plane;
    k = 0;
    n = Lv.cz;
    stride = Lv.sxy;
  }
  const bool plane_ok = (DIR == 1) ? (k < Lv.cz) : (j < Lv.cy);
  const bool ok = plane_ok && i < Lv.cx;
  T a[LS], b[LS], c[LS], d[LS];
  const int m0 = seg * LS;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    a[m] = T(0); b[m] = T(1); c[m] = T(0); d[m] = T(0);
    int q = m0 + m;
    if (ok && q < n) {
      int jj = (DIR == 1) ? q : j;
      int kk = (DIR == 1) ? k : q;
      long id = lidx(Lv.sx, Lv.sxy, i, jj, kk);
      uint8_t f = Lv.fl[id];
      T ev = e[id];
      if (f & FL_BLANK) {
        d[m] = (f & FL_GHOST) ? ev : T(0);
      } else {
        Row<T> w = make_row(Lv, f, i, jj, kk);
        T r;
        T am, ap;
        if (DIR == 1) {
          r = rhs[id] - e[id - 1] * w.amx - e[id + 1] * w.apx - e[id - Lv.sxy] * w.amz - e[id + Lv.sxy] * w.apz;
          am = w.amy; ap = w.apy;
        } else {
          r = rhs[id] - e[id - Lv.sx] * w.amy - e[id + Lv.sx] * w.apy - e[id - 1] * w.amx - e[id + 1] * w.apx;
          am = w.amz; ap = w.apz;
        }
        if (tabs(w.diag) < T(1.0e-12)) {
          d[m] = ev;
        } else {
          a[m] = am; b[m] = w.diag; c[m] = ap; d[m] = r;
        }
      }
    }
  }
  T D[LS], F[LS], G[LS], cl, dl, fl;
  seg_eliminate<T, LS>(a, b, c, d, D, F, G, cl, dl, fl);
  // LDS layout: [seg][field][tx]; fields: cl, dl, fl, D0, F0, G0, X
  auto S_ = [&](int s, int fld) -> T & { return sh[((long)s * 7 + fld) * blockDim.x + tx]; };
  S_(seg, 0) = cl; S_(seg, 1) = dl; S_(seg, 2) = fl;
  S_(seg, 3) = D[0]; S_(seg, 4) = F[0]; S_(seg, 5) = G[0];
  __syncthreads();
  if (seg == 0) {
    // Thomas on the P reduced equations of this line (column tx)
    // A_s X_{s-1} + B_s X_s + C_s X_{s+1} = R_s
    T cprev = T(0), dprev = T(0);
    for (int s = 0; s < P; s++) {
      T cls = S_(s, 0), dls = S_(s, 1), fls = S_(s, 2);
      T D0n = T(0), F0n = T(0), G0n = T(0);
      if (s + 1 < P) { D0n = S_(s + 1, 3); F0n = S_(s + 1, 4); G0n = S_(s + 1, 5); }
      T A = -fls, B = T(1) + cls * F0n, C = cls * G0n, R = dls - cls * D0n;
      T inv = T(1) / (B - A * cprev);
      cprev = C * inv;
      dprev = (R - A * dprev) * inv;
      S_(s, 6) = cprev;   // temporarily c'
      S_(s, 3) = dprev;   // d' (D0 of s no longer needed by s-1: already read)
    }
    T X = S_(P - 1, 3);
    S_(P - 1, 6) = X;
    for (int s = P - 2; s >= 0; s--) {
      X = S_(s, 3) - S_(s, 6) * X;
      S_(s, 6) = X;
    }
  }
  __syncthreads();
  T XS = S_(seg, 6);
  T XL = seg > 0 ? S_(seg - 1, 6) : T(0);
  if (!ok) return;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    int q = m0 + m;
    if (q < n) {
      T x = (m < LS - 1) ? (D[m] + F[m] * XL + G[m] * XS) : XS;
      int jj = (DIR == 1) ? q : j;
      int kk = (DIR == 1) ? k : q;
      long id = lidx(Lv.sx, Lv.sxy, i, jj, kk);
      e[id] = (T(1) - omega) * e[id] + omega * x;
    }
  }
  (void)stride;
}

// ---------------------------------------------------------------------
// coarsest level: e = G * rhs (dense pseudo-inverse of the coarsest
// operator with the fluid-volume mean constraint, computed in FP64 on the
// host).  One workgroup.
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_coarse_solve(LevelArgs<T> L, int n, const double *__restrict__ G,
                                                       const T *__restrict__ rhs, T *__restrict__ e) {
  extern __shared__ unsigned char smem_raw[];
  double *r = reinterpret_cast<double *>(smem_raw);
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    int i = q % L.cx, t = q / L.cx, j = t % L.cy, k = t / L.cy;
    r[q] = (double)rhs[lidx(L.sx, L.sxy, i, j, k)];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < n; p++) s += G[(long)p * n + q] * r[p];
    int i = q % L.cx, t = q / L.cx, j = t % L.cy, k = t / L.cy;
    e[lidx(L.sx, L.sxy, i, j, k)] = (T)s;
  }
}

// ---------------------------------------------------------------------
// Flexible GMRES on the reference layout (FP64 vectors u_q, zero on every
// non-fluid cell).  The basis is kept unnormalised (v_q = sigma_q u_q, the
// sigmas live on the host) so that no separate scaling pass is needed.
// ---------------------------------------------------------------------
constexpr int KRY_MAX = 16;
struct KryVecs {
  const double *u[KRY_MAX + 1];
  double c[KRY_MAX + 1];
  int n;
};

// z = e0 on fluid cells, 0 on blank cells (ghost rows are filled afterwards)
template <typename TE>
__global__ void __launch_bounds__(256) k_kry_z(RefArgs R, LevelArgs<double> L, const TE *__restrict__ e0,
                                               double *__restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  z[ridx(R, i, j, k)] = (L.fl[a] & FL_BLANK) ? 0.0 : (double)e0[a];
}

// w = A z on fluid cells (0 elsewhere) with z's ghost rows and outer layer
// already exact; per-block partial dots <w, u_q>, q < V.n, at part[blk*V.n+q]
template <int NT>
__global__ void __launch_bounds__(NT) k_kry_apply(RefArgs R, LevelArgs<double> L, const double *__restrict__ z,
                                                  double *__restrict__ w, KryVecs V, double *__restrict__ part) {
  __shared__ double s[NT / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  bool in = i < R.nxc && j < R.nyc;
  double wv = 0.0;
  long r = 0;
  if (in) {
    long a = lidx(L.sx, L.sxy, i, j, k);
    r = ridx(R, i, j, k);
    uint8_t f = L.fl[a];
    if (!(f & FL_BLANK)) {
      Row<double> rw = make_row(L, f, i, j, k);
      wv = z[r] * rw.diag + z[r - 1] * rw.amx + z[r + 1] * rw.apx + z[r - R.ex] * rw.amy + z[r + R.ex] * rw.apy +
           z[r - R.exy] * rw.amz + z[r + R.exy] * rw.apz;
    }
    w[r] = wv;
  }
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++) {
    if (q < V.n) {
      double d = block_sum<NT>(in ? wv * V.u[q][r] : 0.0, s);
      if (threadIdx.x == 0 && threadIdx.y == 0) part[blk * V.n + q] = d;
    }
  }
}

// u = w - sum_q c_q u_q on fluid cells (in place in w), level-0 rhs = u,
// e0 = 0, per-block partial |u|^2 at part[blk]
template <typename TE, int NT>
__global__ void __launch_bounds__(NT) k_kry_orth(RefArgs R, LevelArgs<double> L, double *__restrict__ w, KryVecs V,
                                                 TE *__restrict__ rhs0, TE *__restrict__ e0,
                                                 double *__restrict__ part) {
  __shared__ double s[NT / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double u = 0.0;
  if (i < R.nxc && j < R.nyc) {
    long a = lidx(L.sx, L.sxy, i, j, k);
    long r = ridx(R, i, j, k);
    if (!(L.fl[a] & FL_BLANK)) {
      u = w[r];
#pragma unroll
      for (int q = 0; q < KRY_MAX + 1; q++)
        if (q < V.n) u -= V.c[q] * V.u[q][r];
      w[r] = u;
    }
    rhs0[a] = (TE)u;
    e0[a] = TE(0);
  }
  double d = block_sum<NT>(u * u, s);
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blk] = d;
}

// p += sum_q c_q u_q on fluid cells
__global__ void __launch_bounds__(256) k_kry_update(RefArgs R, LevelArgs<double> L, double *__restrict__ p,
                                                    KryVecs V) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  if (L.fl[a] & FL_BLANK) return;
  long r = ridx(R, i, j, k);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++)
    if (q < V.n) acc += V.c[q] * V.u[q][r];
  p[r] = p[r] + acc;
}

// ---------------------------------------------------------------------
// layout conversions between the reference layout (FP64) and a level-0
// field (T), interior cells only
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_ref_to_level(RefArgs R, LevelArgs<T> L, const double *__restrict__ src,
                                                      T *__restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  dst[lidx(L.sx, L.sxy, i, j, k)] = (T)src[ridx(R, i, j, k)];
}
template <typename T>
__global__ void __launch_bounds__(256) k_level_to_ref(RefArgs R, LevelArgs<T> L, const T *__restrict__ src,
                                                      double *__restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  dst[ridx(R, i, j, k)] = (double)src[lidx(L.sx, L.sxy, i, j, k)];
}

// ---------------------------------------------------------------------
// batched tridiagonal solves through the line-solver cores (tdma parity)
// ---------------------------------------------------------------------
template <int S>
__global__ void __launch_bounds__(256) k_tdma_wave(int n, int nlines, const double *__restrict__ A,
                                                   const double *__restrict__ B, const double *__restrict__ C,
                                                   const double *__restrict__ Rr, double *__restrict__ U) {
  int lane = threadIdx.x & 63;
  int line = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (line >= nlines) return;
  long base = (long)line * n;
  double a[S], b[S], c[S], d[S], x[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    int q = lane * S + s;
    a[s] = 0; b[s] = 1; c[s] = 0; d[s] = 0;
    if (q < n) {
      a[s] = q > 0 ? A[base + q] : 0.0;
      b[s] = B[base + q];
      c[s] = q < n - 1 ? C[base + q] : 0.0;
      d[s] = Rr[base + q];
    }
  }
  wave_line_solve<double, S>(a, b, c, d, x);
#pragma unroll
  for (int s = 0; s < S; s++) {
    int q = lane * S + s;
    if (q < n) U[base + q] = x[s];
  }
}

template <int LS>
__global__ void __launch_bounds__(1024) k_tdma_seg(int n, int nlines, const double *__restrict__ A,
                                                   const double *__restrict__ B, const double *__restrict__ C,
                                                   const double *__restrict__ Rr, double *__restrict__ U) {
  extern __shared__ unsigned char smem_raw[];
  double *sh = reinterpret_cast<double *>(smem_raw);
  const int P = blockDim.y;
  const int tx = threadIdx.x, seg = threadIdx.y;
  const int line = blockIdx.x * blockDim.x + tx;
  const bool ok = line < nlines;
  double a[LS], b[LS], c[LS], d[LS];
  const int m0 = seg * LS;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    int q = m0 + m;
    a[m] = 0; b[m] = 1; c[m] = 0; d[m] = 0;
    if (ok && q < n) {
      // lines interleaved so that consecutive threads read consecutive
      // addresses: element q of line l at q*nlines + l
      long id = (long)q * nlines + line;
      a[m] = q > 0 ? A[id] : 0.0;
      b[m] = B[id];
      c[m] = q < n - 1 ? C[id] : 0.0;
      d[m] = Rr[id];
    }
  }
  double D[LS], F[LS], G[LS], cl, dl, fl;
  seg_eliminate<double, LS>(a, b, c, d, D, F, G, cl, dl, fl);
  auto S_ = [&](int s, int fld) -> double & { return sh[((long)s * 7 + fld) * blockDim.x + tx]; };
  S_(seg, 0) = cl; S_(seg, 1) = dl; S_(seg, 2) = fl;
  S_(seg, 3) = D[0]; S_(seg, 4) = F[0]; S_(seg, 5) = G[0];
  __syncthreads();
  if (seg == 0) {
    double cprev = 0, dprev = 0;
    for (int s = 0; s < P; s++) {
      double cls = S_(s, 0), dls = S_(s, 1), fls = S_(s, 2);
      double D0n = 0, F0n = 0, G0n = 0;
      if (s + 1 < P) { D0n = S_(s + 1, 3); F0n = S_(s + 1, 4); G0n = S_(s + 1, 5); }
      double Aa = -fls, Bb = 1.0 + cls * F0n, Cc = cls * G0n, Rv = dls - cls * D0n;
      double inv = 1.0 / (Bb - Aa * cprev);
      cprev = Cc * inv;
      dprev = (Rv - Aa * dprev) * inv;
      S_(s, 6) = cprev;
      S_(s, 3) = dprev;
    }
    double X = S_(P - 1, 3);
    S_(P - 1, 6) = X;
    for (int s = P - 2; s >= 0; s--) {
      X = S_(s, 3) - S_(s, 6) * X;
      S_(s, 6) = X;
    }
  }
  __syncthreads();
  double XS = S_(seg, 6);
  double XL = seg > 0 ? S_(seg - 1, 6) : 0.0;
  if (!ok) return;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    int q = m0 + m;
    if (q < n) U[(long)q * nlines + line] = (m < LS - 1) ? (D[m] + F[m] * XL + G[m] * XS) : XS;
  }
}

}  // namespace pp
