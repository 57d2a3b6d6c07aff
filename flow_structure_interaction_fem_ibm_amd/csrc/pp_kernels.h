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

#include <type_traits>

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

// chunk loads/stores of S consecutive elements (vectorised when the chunk
// is whole: 16-byte vectors when S is a multiple of their length and the
// chunk 16-byte aligned; FP32 chunks of an even length that is not get
// 8-byte pairs, the chunk then 8-byte aligned)
template <typename T, int S>
__device__ inline void load_chunk(const T *__restrict__ p, int nvalid, T (&v)[S]) {
  constexpr int VE = 16 / sizeof(T);
  if constexpr (sizeof(T) == 4 && S % VE != 0 && S % 2 == 0) {
    if (nvalid == S) {
#pragma unroll
      for (int q = 0; q < S / 2; q++) {
        const float2 w = reinterpret_cast<const float2 *>(p)[q];
        v[2 * q] = w.x; v[2 * q + 1] = w.y;
      }
      return;
    }
  }
  if (S % VE == 0 && nvalid == S) {
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
#pragma unroll
    for (int q = 0; q < S / VE; q++) {
      V w = reinterpret_cast<const V *>(p)[q];
      if constexpr (sizeof(T) == 4) {
        v[4 * q] = w.x; v[4 * q + 1] = w.y; v[4 * q + 2] = w.z; v[4 * q + 3] = w.w;
      } else {
        v[2 * q] = w.x; v[2 * q + 1] = w.y;
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < S; s++) v[s] = s < nvalid ? p[s] : T(0);
  }
}
template <typename T, int S>
__device__ inline void store_chunk(T *__restrict__ p, int nvalid, const T (&v)[S]) {
  constexpr int VE = 16 / sizeof(T);
  if constexpr (sizeof(T) == 4 && S % VE != 0 && S % 2 == 0) {
    if (nvalid == S) {
#pragma unroll
      for (int q = 0; q < S / 2; q++) reinterpret_cast<float2 *>(p)[q] = make_float2(v[2 * q], v[2 * q + 1]);
      return;
    }
  }
  if (S % VE == 0 && nvalid == S) {
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
#pragma unroll
    for (int q = 0; q < S / VE; q++) {
      V w;
      if constexpr (sizeof(T) == 4) {
        w.x = v[4 * q]; w.y = v[4 * q + 1]; w.z = v[4 * q + 2]; w.w = v[4 * q + 3];
      } else {
        w.x = v[2 * q]; w.y = v[2 * q + 1];
      }
      reinterpret_cast<V *>(p)[q] = w;
    }
  } else {
#pragma unroll
    for (int s = 0; s < S; s++)
      if (s < nvalid) p[s] = v[s];
  }
}
template <int S>
__device__ inline void load_flags(const uint8_t *__restrict__ p, int nvalid, uint8_t (&f)[S]) {
  if constexpr (S % 4 != 0 && S % 2 == 0) {
    // even chunks of other lengths: 2-byte pairs (the chunk 2-byte aligned)
    if (nvalid == S) {
#pragma unroll
      for (int q = 0; q < S / 2; q++) {
        const uint16_t w = reinterpret_cast<const uint16_t *>(p)[q];
        f[2 * q] = w & 0xff; f[2 * q + 1] = w >> 8;
      }
      return;
    }
  }
  if (S % 4 == 0 && nvalid == S) {
#pragma unroll
    for (int q = 0; q < S / 4; q++) {
      uint32_t w = reinterpret_cast<const uint32_t *>(p)[q];
      f[4 * q] = w & 0xff; f[4 * q + 1] = (w >> 8) & 0xff; f[4 * q + 2] = (w >> 16) & 0xff; f[4 * q + 3] = w >> 24;
    }
  } else {
#pragma unroll
    for (int s = 0; s < S; s++) f[s] = s < nvalid ? p[s] : uint8_t(FL_BLANK);
  }
}

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

// first stage of a two-stage reduction of partials: workgroup b combines
// partials [b*chunk, (b+1)*chunk) into out[3*b..3*b+2]
__global__ void k_reduce_summax_stage(const double *__restrict__ part, int nblk, int chunk, double *__restrict__ out) {
  SumMax v{0.0, -1.0, -1};
  int b0 = blockIdx.x * chunk, b1 = min(nblk, b0 + chunk);
  for (int b = b0 + threadIdx.x; b < b1; b += 256) {
    SumMax o{part[3 * b], part[3 * b + 1], (long)part[3 * b + 2]};
    v = summax_combine(v, o);
  }
  v = block_summax<256>(v);
  if (threadIdx.x == 0) { out[3 * blockIdx.x] = v.sum; out[3 * blockIdx.x + 1] = v.mx; out[3 * blockIdx.x + 2] = (double)v.arg; }
}

// first stage for nq independent sums: out[b*nq + q]
__global__ void k_reduce_sums_stage(const double *__restrict__ part, int nblk, int nq, int chunk,
                                    double *__restrict__ out) {
  __shared__ double s[256];
  int b0 = blockIdx.x * chunk, b1 = min(nblk, b0 + chunk);
  for (int q = 0; q < nq; q++) {
    double v = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += 256) v += part[(long)b * nq + q];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
      if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(long)blockIdx.x * nq + q] = s[0];
    __syncthreads();
  }
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

// Block sums of NV values per thread with one barrier: wavefront shuffles,
// the wavefront sums meet in LDS (s: NT/64 * NV doubles) and threads
// t < NV add them in a fixed order into part[blk * NV + t].
template <int NT, int NV>
__device__ inline void block_sums_n(double (&v)[NV], double *s, double *part, long blk) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x, w = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int t = 0; t < NV; t++) {
    double x = v[t];
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_down(x, d, 64);
    if (lane == 0) s[w * NV + t] = x;
  }
  __syncthreads();
  if (tid < NV) {
    double r = 0.0;
    for (int ww = 0; ww < NT / 64; ww++) r += s[ww * NV + tid];
    part[blk * NV + tid] = r;
  }
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
                                                       double *__restrict__ rout, TE *__restrict__ u0) {
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
    if (u0) u0[a] = (TE)res;
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

// rhs_poisson, non-cut-cell branch (CORE_POISSON.F90:281-288), reference
// layout, same operation order as the reference: nlu0 = dtinv*((u_{i+1}-u_i)
// dxinv + (v_{j+1}-v_j) dyinv + (w_{k+1}-w_k) dzinv)*(1-iblank); per-block
// partial sums of nlu0*dx*dy*dz at part[blk]
template <int NT>
__global__ void __launch_bounds__(NT) k_divergence(RefArgs R, const double *__restrict__ dx, const double *__restrict__ dy,
                                                   const double *__restrict__ dz, const double *__restrict__ dxinv,
                                                   const double *__restrict__ dyinv, const double *__restrict__ dzinv,
                                                   const int8_t *__restrict__ ib, const double *__restrict__ fu,
                                                   const double *__restrict__ fv, const double *__restrict__ fw,
                                                   double dtinv, double *__restrict__ out, double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[NT / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double vol = 0.0;
  if (i < R.nxc && j < R.nyc) {
    long id = ridx(R, i, j, k);
    int I = i + 1, J = j + 1, K = k + 1;
    double v = dtinv * ((fu[id + 1] - fu[id]) * dxinv[I] + (fv[id + R.ex] - fv[id]) * dyinv[J] +
                        (fw[id + R.exy] - fw[id]) * dzinv[K]);
    v = v * (double)(1 - ib[id]);
    out[id] = v;
    vol = v * dx[I] * dy[J] * dz[K];
  }
  double t = block_sum<NT>(vol, s);
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blk] = t;
}

// ---------------------------------------------------------------------
// Fractional step either side of the solve, FP64 reference layout.
// Per-direction metrics of the reference grid (0:n+1): fx (face
// interpolation weight, CORE_GRID.F90:172-182), dxinv, dxcinv.
// ---------------------------------------------------------------------
struct FracMetrics {
  const double *fx[3], *dinv[3], *dcinv[3];
};

// correct_vel (CORE_FRAC_STEP_ROUTINES.F90:270-444) and
// GCM_set_face_vel_body (GCM_GHOSTCELL_VEL_UPDATE_UNSTRUC.F90:387-423) for
// one interior cell: its three velocities, and its west/south/back faces
// (the last cell of a row also its east/north/front face), each face with
// the reference's overwrite and mask order:
//   f = f_old - dt ((p_c - p_w) dcinv)(1 - m_c);
//   f = f (1 - p_w) + bc_w p_w   (not for the first face of a row);
//   f = f (1 - m_c) + bc_c m_c;  0 if either cell is blank and not a ghost
// with m = ium/jum/kum and p = iup/jup/kup (the level-0 outer-face flags).
// One thread owns every face it writes, so the pass is race free.
template <int NT>
__global__ void __launch_bounds__(NT) k_correct_vel(RefArgs R, LevelArgs<double> L0, FracMetrics M, double dt,
                                                    const double *__restrict__ p, double *__restrict__ u,
                                                    double *__restrict__ v, double *__restrict__ w,
                                                    double *__restrict__ fu, double *__restrict__ fv,
                                                    double *__restrict__ fw, const double *__restrict__ bcu,
                                                    const double *__restrict__ bcv, const double *__restrict__ bcw) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  const long id = ridx(R, i, j, k);
  const long st[3] = {1, R.ex, R.exy};
  const int I[3] = {i + 1, j + 1, k + 1};
  const int nc[3] = {R.nxc, R.nyc, R.nzc};
  const long lst[3] = {1, (long)L0.sx, L0.sxy};
  const long la = lidx(L0.sx, L0.sxy, i, j, k);
  const uint8_t f = L0.fl[la];
  const uint8_t FM[3] = {FL_IM, FL_JM, FL_KM}, FP[3] = {FL_IP, FL_JP, FL_KP};
  const bool solid = (f & FL_BLANK) && !(f & FL_GHOST);
  double *vel[3] = {u, v, w}, *face[3] = {fu, fv, fw};
  const double *bc[3] = {bcu, bcv, bcw};
  const double pc = p[id];
  const double fl = (double)(1 - ((f & FL_BLANK) ? 1 : 0));
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int mflag = (f & FM[d]) ? 1 : 0, pflag = (f & FP[d]) ? 1 : 0;
    // cell velocity (:276-367)
    const double fe = M.fx[d][I[d] + 1], fw_ = M.fx[d][I[d]];
    const double pe = (fe * p[id + st[d]] + (1.0 - fe) * pc) * (double)(1 - pflag) + pc * (double)pflag;
    const double pw = (fw_ * pc + (1.0 - fw_) * p[id - st[d]]) * (double)(1 - mflag) + pc * (double)mflag;
    const double pg = (pe - pw) * M.dinv[d][I[d]];
    vel[d][id] = vel[d][id] - dt * pg * fl;
    // west face of the cell
    const double gw = (pc - p[id - st[d]]) * M.dcinv[d][I[d]];
    double fwv = face[d][id] - dt * gw * (double)(1 - mflag);
    bool zero = solid;
    if (I[d] > 1) {
      const uint8_t fo = L0.fl[la - lst[d]];
      const int po = (fo & FP[d]) ? 1 : 0;
      fwv = fwv * (double)(1 - po) + bc[d][id - st[d]] * (double)po;
      zero = zero || ((fo & FL_BLANK) && !(fo & FL_GHOST));
    }
    fwv = fwv * (double)(1 - mflag) + bc[d][id] * (double)mflag;
    face[d][id] = zero ? 0.0 : fwv;
    // east face of the last cell
    if (I[d] == nc[d]) {
      const double ge = (p[id + st[d]] - pc) * M.dcinv[d][I[d] + 1];
      double fev = face[d][id + st[d]] - dt * ge * (double)(1 - pflag);
      fev = fev * (double)(1 - pflag) + bc[d][id] * (double)pflag;
      face[d][id + st[d]] = solid ? 0.0 : fev;
    }
  }
}

// face_vel (CORE_FRAC_STEP_ROUTINES.F90:6-240) for one interior cell: its
// west/south/back faces (the last cell of a row also its east/north/front
// face), each the composition of the reference's sequential writes to it:
//   uT = u + dt grad_d p (1 - iblank) for VAN_KAN (:36-70), else u;
//   f  = (fx uT_c + (1 - fx) uT_w)(1 - m_c) + bc_c m_c          (:80-107)
//   VAN_KAN: f -= dt ((p_c - p_w) dcinv)(1 - m_c), then the outer-face
//   masks f = f (1 - p_w) + bc_w p_w; f = f (1 - m_c) + bc_c m_c (:111-175)
//   the same two masks again (:179-197)
//   VAN_KAN: f = (1 - g_w) f + g_w I; f = (1 - g_c) f + g_c I with I the
//   interpolation of u itself (ghost cells, :201-234)
//   0 if either cell is blank and not a ghost (:238)
// with m = ium/jum/kum and p = iup/jup/kup (the level-0 outer-face flags) and
// the west cell's terms only where it is an interior cell.  The cell
// gradient is the cell's own from fx-interpolated face pressures (the cell
// value on an outer face).  One thread owns every face it writes.
template <int NT>
__global__ void __launch_bounds__(NT) k_face_vel(RefArgs R, LevelArgs<double> L0, FracMetrics M, int van_kan,
                                                 double dt, const double *__restrict__ p,
                                                 const double *__restrict__ u, const double *__restrict__ v,
                                                 const double *__restrict__ w, const double *__restrict__ bcu,
                                                 const double *__restrict__ bcv, const double *__restrict__ bcw,
                                                 double *__restrict__ fu, double *__restrict__ fv,
                                                 double *__restrict__ fw) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  const long id = ridx(R, i, j, k);
  const long st[3] = {1, R.ex, R.exy};
  const int I[3] = {i + 1, j + 1, k + 1};
  const int nc[3] = {R.nxc, R.nyc, R.nzc};
  const long lst[3] = {1, (long)L0.sx, L0.sxy};
  const long la = lidx(L0.sx, L0.sxy, i, j, k);
  const uint8_t f = L0.fl[la];
  const uint8_t FM[3] = {FL_IM, FL_JM, FL_KM}, FP[3] = {FL_IP, FL_JP, FL_KP};
  const double *vel[3] = {u, v, w}, *bc[3] = {bcu, bcv, bcw};
  double *face[3] = {fu, fv, fw};
  // uTilde of the d-component at interior cell q (flags fq, index c along d)
  auto utilde = [&](int d, long q, uint8_t fq, int c) -> double {
    if (!van_kan) return vel[d][q];
    const int mq = (fq & FM[d]) ? 1 : 0, pq = (fq & FP[d]) ? 1 : 0;
    const double fe = M.fx[d][c + 1], fwt = M.fx[d][c], pc = p[q];
    const double pe = (fe * p[q + st[d]] + (1.0 - fe) * pc) * (double)(1 - pq) + pc * (double)pq;
    const double pw = (fwt * pc + (1.0 - fwt) * p[q - st[d]]) * (double)(1 - mq) + pc * (double)mq;
    const double pg = (pe - pw) * M.dinv[d][c];
    return vel[d][q] + dt * pg * (double)(1 - ((fq & FL_BLANK) ? 1 : 0));
  };
  const bool solid = (f & FL_BLANK) && !(f & FL_GHOST);
  const double gc = (f & FL_GHOST) ? 1.0 : 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int c = I[d];
    const int mc = (f & FM[d]) ? 1 : 0, pc = (f & FP[d]) ? 1 : 0;
    const double *U = vel[d];
    const double uTc = utilde(d, id, f, c);
    // west face
    {
      const bool has_w = c > 1;
      uint8_t fo = 0;
      if (has_w) fo = L0.fl[la - lst[d]];
      const double uTw = has_w ? utilde(d, id - st[d], fo, c - 1) : U[id - st[d]];
      const int pw = (fo & FP[d]) ? 1 : 0;
      const double fxc = M.fx[d][c];
      double x = (fxc * uTc + (1.0 - fxc) * uTw) * (double)(1 - mc) + bc[d][id] * (double)mc;
      auto masks = [&](double y) {
        if (has_w) y = y * (double)(1 - pw) + bc[d][id - st[d]] * (double)pw;
        return y * (double)(1 - mc) + bc[d][id] * (double)mc;
      };
      if (van_kan) {
        const double gw = (p[id] - p[id - st[d]]) * M.dcinv[d][c];
        x = masks(x - dt * gw * (double)(1 - mc));
      }
      x = masks(x);
      bool zero = solid;
      if (van_kan) {
        const double ip = fxc * U[id] + (1.0 - fxc) * U[id - st[d]];
        if (has_w) {
          const double gwf = (fo & FL_GHOST) ? 1.0 : 0.0;
          x = (1.0 - gwf) * x + gwf * ip;
        }
        x = (1.0 - gc) * x + gc * ip;
      }
      if (has_w) zero = zero || ((fo & FL_BLANK) && !(fo & FL_GHOST));
      face[d][id] = zero ? 0.0 : x;
    }
    // east face of the last cell of a row
    if (c == nc[d]) {
      const double fxe = M.fx[d][c + 1];
      double x = (fxe * U[id + st[d]] + (1.0 - fxe) * uTc) * (double)(1 - pc) + bc[d][id] * (double)pc;
      if (van_kan) {
        const double ge = (p[id + st[d]] - p[id]) * M.dcinv[d][c + 1];
        x = x - dt * ge * (double)(1 - pc);
        x = x * (double)(1 - pc) + bc[d][id] * (double)pc;
      }
      x = x * (double)(1 - pc) + bc[d][id] * (double)pc;
      if (van_kan) {
        const double ip = fxe * U[id + st[d]] + (1.0 - fxe) * U[id];
        x = (1.0 - gc) * x + gc * ip;
      }
      face[d][id + st[d]] = solid ? 0.0 : x;
    }
  }
}

// velocity ghost rows (GCM_vel_set_bc_internal): gather the ghost cells'
// current (u, v, w), Jacobi sweeps of
//   u_g = u_b bw - (sum_r cD_r x_r) iw
// (x_r: another ghost's iterate when kind >= 0, the field at off when
// kind == -1, the body-intercept value u_b when kind == -2), then scatter
template <typename T>
__global__ void k_vel_ghost_gather(int n, const long *__restrict__ gref, const T *__restrict__ u,
                                   const T *__restrict__ v, const T *__restrict__ w, T *__restrict__ g) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long r = gref[q];
  g[3 * q] = u[r]; g[3 * q + 1] = v[r]; g[3 * q + 2] = w[r];
}

__global__ void k_vel_ghost_sweep(int n, const double *__restrict__ cD, const int *__restrict__ kind,
                                  const long *__restrict__ off, double bw, double iw, const double *__restrict__ vbi,
                                  const double *__restrict__ u, const double *__restrict__ v,
                                  const double *__restrict__ w, const double *__restrict__ gin,
                                  double *__restrict__ gout) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double ub = vbi ? vbi[3 * q] : 0.0, vb = vbi ? vbi[3 * q + 1] : 0.0, wb = vbi ? vbi[3 * q + 2] : 0.0;
  double su = 0.0, sv = 0.0, sw = 0.0;
  for (int r = 0; r < 8; r++) {
    const long e = 8L * q + r;
    const double cf = cD[e];
    const int kd = kind[e];
    double xu, xv, xw;
    if (kd >= 0) {
      xu = gin[3L * kd]; xv = gin[3L * kd + 1]; xw = gin[3L * kd + 2];
    } else if (kd == -1) {
      const long o = off[e];
      xu = u[o]; xv = v[o]; xw = w[o];
    } else {
      xu = ub; xv = vb; xw = wb;
    }
    su += cf * xu; sv += cf * xv; sw += cf * xw;
  }
  gout[3L * q] = ub * bw - su * iw;
  gout[3L * q + 1] = vb * bw - sv * iw;
  gout[3L * q + 2] = wb * bw - sw * iw;
}

__global__ void k_vel_ghost_scatter(int n, const long *__restrict__ gref, const double *__restrict__ g,
                                    double *__restrict__ u, double *__restrict__ v, double *__restrict__ w) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long r = gref[q];
  u[r] = g[3 * q]; v[r] = g[3 * q + 1]; w[r] = g[3 * q + 2];
}

// update_pressure (CORE_FRAC_STEP_ROUTINES.F90:958-962) over the whole
// field: p = p + p' (van Kan) or p = p'
__global__ void k_update_pressure(long n, int van_kan, const double *__restrict__ pp, double *__restrict__ p) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  p[q] = van_kan ? p[q] + pp[q] : pp[q];
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
// q2 = sum(b*b), q3 = sum(V*p), q4 = sum(V).  CPT consecutive cells of a
// row per thread (the grid's x extent covers nxc / CPT), all loads issued
// before the flags are tested, so that a wavefront has CPT cells of every
// array in flight instead of one flag round trip ahead of one cell.
template <int NT, int CPT>
__global__ void __launch_bounds__(NT) k_fluid_sums(RefArgs R, LevelArgs<double> L, const double *__restrict__ w,
                                                   const double *__restrict__ b, const double *__restrict__ p,
                                                   double *__restrict__ part) {
  __shared__ double s[NT / 64 * 5];
  const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * CPT;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double q[5] = {0, 0, 0, 0, 0};
  if (i0 < R.nxc && j < R.nyc) {
    const long a0 = lidx(L.sx, L.sxy, i0, j, k);
    const long r0 = ridx(R, i0, j, k);
    uint8_t f[CPT];
    double wv[CPT], bv[CPT], pv[CPT], vol[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const bool in = i0 + c < R.nxc;
      f[c] = in ? L.fl[a0 + c] : (uint8_t)FL_BLANK;
      vol[c] = in ? L.wx[i0 + c] * L.wy[j] * L.wz[k] : 0.0;
      wv[c] = (in && w) ? w[r0 + c] : 0.0;
      bv[c] = (in && b) ? b[r0 + c] : 0.0;
      pv[c] = (in && p) ? p[r0 + c] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      if (f[c] & FL_BLANK) continue;
      const double wc = w ? wv[c] : vol[c];
      q[0] += wc * bv[c];
      q[1] += wc;
      q[2] += bv[c] * bv[c];
      q[3] += p ? pv[c] * vol[c] : 0.0;
      q[4] += vol[c];
    }
  }
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  block_sums_n<NT, 5>(q, s, part, blk);
}

// ---------------------------------------------------------------------
// Adjoint and left-null-vector set-up on the device (reference layout,
// fluid cells only; V = cell volume, mu = lambda / V):
//   k_fluid_sum_vol   partials {sum a, sum V} (a may be null)
//   k_adj_source      sv = (wobj - (V S) / vsum) / V, partial sum sv^2
//   k_mu_vol_sums     partials {sum mu V, sum w}
//   k_adj_lambda      lambda = mu V - zeta w
//   k_null_init       mu = mu_prev where nonzero (warm start), else 1
//   k_null_scale      w = (V mu) scale
// Blank cells are left untouched (the callers zero those fields first).
// ---------------------------------------------------------------------
template <int NT>
__global__ void __launch_bounds__(NT) k_fluid_sum_vol(RefArgs R, LevelArgs<double> L, const double *__restrict__ a,
                                                      double *__restrict__ part) {
  __shared__ double s[NT / 64 * 2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double q[2] = {0.0, 0.0};
  if (i < R.nxc && j < R.nyc && !(L.fl[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK)) {
    q[0] = a ? a[ridx(R, i, j, k)] : 0.0;
    q[1] = L.wx[i] * L.wy[j] * L.wz[k];
  }
  const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  block_sums_n<NT, 2>(q, s, part, blk);
}

template <int NT>
__global__ void __launch_bounds__(NT) k_adj_source(RefArgs R, LevelArgs<double> L, const double *__restrict__ wobj,
                                                   double S, double vsum, double *__restrict__ sv,
                                                   double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[NT / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double q[1] = {0.0};
  if (i < R.nxc && j < R.nyc && !(L.fl[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK)) {
    const long r = ridx(R, i, j, k);
    const double V = L.wx[i] * L.wy[j] * L.wz[k];
    const double v = (wobj[r] - V * S / vsum) / V;
    sv[r] = v;
    q[0] = v * v;
  }
  const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  block_sums_n<NT, 1>(q, s, part, blk);
}

template <int NT>
__global__ void __launch_bounds__(NT) k_mu_vol_sums(RefArgs R, LevelArgs<double> L, const double *__restrict__ mu,
                                                    const double *__restrict__ w, double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[NT / 64 * 2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double q[2] = {0.0, 0.0};
  if (i < R.nxc && j < R.nyc && !(L.fl[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK)) {
    const long r = ridx(R, i, j, k);
    q[0] = mu[r] * (L.wx[i] * L.wy[j] * L.wz[k]);
    q[1] = w ? w[r] : 0.0;
  }
  const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  block_sums_n<NT, 2>(q, s, part, blk);
}

__global__ void __launch_bounds__(256) k_adj_lambda(RefArgs R, LevelArgs<double> L, const double *__restrict__ mu,
                                                    const double *__restrict__ w, double zeta,
                                                    double *__restrict__ lambda) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc || (L.fl[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK)) return;
  const long r = ridx(R, i, j, k);
  lambda[r] = mu[r] * (L.wx[i] * L.wy[j] * L.wz[k]) - zeta * w[r];
}

__global__ void __launch_bounds__(256) k_null_init(RefArgs R, LevelArgs<double> L, const double *__restrict__ prev,
                                                   double *__restrict__ mu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc || (L.fl[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK)) return;
  const long r = ridx(R, i, j, k);
  mu[r] = (prev && prev[r] != 0.0) ? prev[r] : 1.0;
}

// shape of the adjoint's null space on level 0: s = w / V on fluid cells (w
// the left null vector of the operator, V the cell volume; s = 1 on the
// open domain far from bodies), 1 on blank cells, in the level layout
template <typename T>
__global__ void __launch_bounds__(256) k_null_shape(RefArgs R, LevelArgs<double> L, const double *__restrict__ w,
                                                    T *__restrict__ s, T *__restrict__ s_split, int sp_sx,
                                                    long sp_sxy, int hp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  const long a = lidx(L.sx, L.sxy, i, j, k);
  const T v = (L.fl[a] & FL_BLANK) ? T(1) : (T)(w[ridx(R, i, j, k)] / (L.wx[i] * L.wy[j] * L.wz[k]));
  s[a] = v;
  // the same values in the colour-split copy of level 0 (when given)
  if (s_split) s_split[sidx(sp_sx, sp_sxy, hp, i, j, k)] = v;
}

__global__ void __launch_bounds__(256) k_null_scale(RefArgs R, LevelArgs<double> L, const double *__restrict__ mu,
                                                    double scale, double *__restrict__ w) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc || (L.fl[lidx(L.sx, L.sxy, i, j, k)] & FL_BLANK)) return;
  const long r = ridx(R, i, j, k);
  w[r] = (L.wx[i] * L.wy[j] * L.wz[k] * mu[r]) * scale;
}

// dst[ref[q]] = vals[q] * scale
__global__ void __launch_bounds__(256) k_ghost_scatter_scaled(int n, const long *__restrict__ ref,
                                                              const double *__restrict__ vals, double scale,
                                                              double *__restrict__ dst) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[ref[q]] = vals[q] * scale;
}

// per ghost: out[q] = a * src[ref[q]] (+ b * add[q] when add is given)
__global__ void __launch_bounds__(256) k_ghost_combine(int n, const long *__restrict__ ref, const double *__restrict__ src,
                                                       double a, const double *__restrict__ add, double b,
                                                       const double *__restrict__ mul, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  double v = a * src[ref[q]];
  if (add) v = b * add[q] + v;
  out[q] = mul ? mul[q] * v : v;
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
  // no fluid non-fresh neighbour (a body that moved more than a cell): the
  // reference divides by zero; the cell keeps its value
  if (n > 0) p[r] = s / (double)n;
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
  target[t_idx[t]] = (base ? base[t] : T(0)) + vi * s;
}

// the solved transposed ghost coupling in one gather (transpose of the
// composite rows M, CSR by fluid cell):
//   target[ct_idx[t]] = base[t] + vinv[t] * sum_q M[g_q, f_t] rhsg[g_q]
// (base may be null: 0).  Fixed summation order: deterministic.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_composite(int nct, const long *__restrict__ ct_idx,
                                                       const int *__restrict__ ct_ptr,
                                                       const int *__restrict__ ct_ghost, const T *__restrict__ ct_w,
                                                       const T *__restrict__ vinv, const T *__restrict__ rhsg,
                                                       const T *__restrict__ base, T *__restrict__ target) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nct) return;
  T s = T(0);
  for (int q = ct_ptr[t]; q < ct_ptr[t + 1]; q++) s += ct_w[q] * rhsg[ct_ghost[q]];
  target[ct_idx[t]] = (base ? base[t] : T(0)) + vinv[t] * s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_gather_idx(int n, const long *__restrict__ idx, const T *__restrict__ src,
                                                    T *__restrict__ dst) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  dst[t] = src[idx[t]];
}

// ---------------------------------------------------------------------
// solved ghost rows in one gather: p[g] = sum_q w_q p[f_q] (+ beff[g]);
// reads fluid cells only, writes ghost cells only.  8 lanes per ghost
// split its row (fixed split and reduction order: deterministic).
template <typename T, typename TW>
__global__ void __launch_bounds__(256) k_ghost_composite(int n, const long *__restrict__ gidx,
                                                         const int *__restrict__ ptr, const long *__restrict__ fidx,
                                                         const TW *__restrict__ w, const double *__restrict__ beff,
                                                         T *__restrict__ p) {
  int t = blockIdx.x * 256 + threadIdx.x;
  int g = t >> 3, sub = t & 7;
  bool ok = g < n;
  TW s = TW(0);
  if (ok) {
    int q1 = ptr[g + 1];
    for (int q = ptr[g] + sub; q < q1; q += 8) s += w[q] * (TW)p[fidx[q]];
  }
  s += __shfl_down(s, 4, 8);
  s += __shfl_down(s, 2, 8);
  s += __shfl_down(s, 1, 8);
  if (ok && sub == 0) {
    if (beff) s += (TW)beff[g];
    p[gidx[g]] = (T)s;
  }
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

// The same residual, four consecutive cells per thread (vector loads and
// stores; rows whose interior starts 16-byte aligned and cx % 4 == 0).
template <typename T>
__global__ void __launch_bounds__(256) k_residual4(LevelArgs<T> L, const T *__restrict__ e, const T *__restrict__ rhs,
                                                   T *__restrict__ res) {
  const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i0 >= L.cx || j >= L.cy) return;
  const long a = lidx(L.sx, L.sxy, i0, j, k);
  T ec[4], ejm[4], ejp[4], ekm[4], ekp[4], rv[4], cmx[4], cpx[4], out[4];
  uint8_t fl[4];
  load_chunk<T, 4>(e + a, 4, ec);
  load_chunk<T, 4>(e + a - L.sx, 4, ejm);
  load_chunk<T, 4>(e + a + L.sx, 4, ejp);
  load_chunk<T, 4>(e + a - L.sxy, 4, ekm);
  load_chunk<T, 4>(e + a + L.sxy, 4, ekp);
  load_chunk<T, 4>(rhs + a, 4, rv);
  load_chunk<T, 4>(L.cmx + i0, 4, cmx);
  load_chunk<T, 4>(L.cpx + i0, 4, cpx);
  load_flags<4>(L.fl + a, 4, fl);
  const T el = e[a - 1], er = e[a + 4];
  const T my = L.cmy[j], py = L.cpy[j], mz = L.cmz[k], pz = L.cpz[k];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    uint8_t f = fl[s];
    T v = T(0);
    if (!(f & FL_BLANK)) {
      T amx = (f & FL_IM) ? T(0) : cmx[s], apx = (f & FL_IP) ? T(0) : cpx[s];
      T amy = (f & FL_JM) ? T(0) : my, apy = (f & FL_JP) ? T(0) : py;
      T amz = (f & FL_KM) ? T(0) : mz, apz = (f & FL_KP) ? T(0) : pz;
      T diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz);
      T left = s == 0 ? el : ec[s - 1], right = s == 3 ? er : ec[s + 1];
      v = rv[s] - ec[s] * diag - left * amx - right * apx - ejm[s] * amy - ejp[s] * apy - ekm[s] * amz - ekp[s] * apz;
    }
    out[s] = v;
  }
  store_chunk<T, 4>(res + a, 4, out);
}

// ---------------------------------------------------------------------
// weighted point Jacobi (MG_itsolv_Point_Jacobi, CORE_MG.F90:3716-3800),
// out of place: out = e + omega (s - e) with s = (rhs - sum_nb a_nb e_nb) /
// diag on fluid cells (the reference's operation order, contraction off)
// and s = ghostMark * e on blank cells.  A fluid row whose six faces are
// all closed keeps e (the diagonal guard of the line rows,
// CORE_MG.F90:1977-1982; the reference's point sweep would divide by zero).
// V consecutive cells per thread (V = 4 needs 16-byte aligned rows).
// ---------------------------------------------------------------------
template <typename T, int V>
__global__ void __launch_bounds__(256) k_jacobi(LevelArgs<T> L, const T *__restrict__ e, const T *__restrict__ rhs,
                                                T *__restrict__ out, T omega) {
#pragma clang fp contract(off)
  const int i0 = V * (blockIdx.x * blockDim.x + threadIdx.x);
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i0 >= L.cx || j >= L.cy) return;
  const int nv = min(V, L.cx - i0);
  const long a = lidx(L.sx, L.sxy, i0, j, k);
  T ec[V], ejm[V], ejp[V], ekm[V], ekp[V], rv[V], cmx[V], cpx[V], o[V];
  uint8_t fl[V];
  load_chunk<T, V>(e + a, nv, ec);
  load_chunk<T, V>(e + a - L.sx, nv, ejm);
  load_chunk<T, V>(e + a + L.sx, nv, ejp);
  load_chunk<T, V>(e + a - L.sxy, nv, ekm);
  load_chunk<T, V>(e + a + L.sxy, nv, ekp);
  load_chunk<T, V>(rhs + a, nv, rv);
  load_chunk<T, V>(L.cmx + i0, nv, cmx);
  load_chunk<T, V>(L.cpx + i0, nv, cpx);
  load_flags<V>(L.fl + a, nv, fl);
  const T el = e[a - 1], er = e[a + nv];
  const T my = L.cmy[j], py = L.cpy[j], mz = L.cmz[k], pz = L.cpz[k];
#pragma unroll
  for (int s = 0; s < V; s++) {
    const uint8_t f = fl[s];
    T sv;
    if (f & FL_BLANK) {
      sv = (f & FL_GHOST) ? ec[s] : T(0);
    } else {
      T amx = (f & FL_IM) ? T(0) : cmx[s], apx = (f & FL_IP) ? T(0) : cpx[s];
      T amy = (f & FL_JM) ? T(0) : my, apy = (f & FL_JP) ? T(0) : py;
      T amz = (f & FL_KM) ? T(0) : mz, apz = (f & FL_KP) ? T(0) : pz;
      T diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz);
      T left = s == 0 ? el : ec[s - 1], right = s + 1 < nv ? ec[s + 1] : er;
      T r = rv[s] - left * amx - right * apx - ejm[s] * amy - ejp[s] * apy - ekm[s] * amz - ekp[s] * apz;
      sv = tabs(diag) < T(1.0e-12) ? ec[s] : r / diag;
    }
    o[s] = ec[s] + omega * (sv - ec[s]);
  }
  store_chunk<T, V>(out + a, nv, o);
}

// ---------------------------------------------------------------------
// restriction: coarse rhs = fluid-volume-weighted mean of the children's
// residuals (the reference's masked averaging, MG_Restrict_X/Y/Z
// CORE_MG.F90:2511-2538, applied over all coarsened directions at once);
// 0 where no child is fluid.  Also zeroes the coarse correction.  The
// children of coarse cell I along a direction coarsened by f are f*I ..
// f*I+f-1, and the last coarse cell also takes the cell an odd fine count
// leaves over.
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(LevelArgs<T> F, LevelArgs<T> C, int fx, int fy, int fz,
                                                  const T *__restrict__ res, T *__restrict__ rhs_c,
                                                  T *__restrict__ e_c) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  int J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
  if (I >= C.cx || J >= C.cy) return;
  const int i0 = I * fx, i1 = I == C.cx - 1 ? F.cx : i0 + fx;
  const int j0 = J * fy, j1 = J == C.cy - 1 ? F.cy : j0 + fy;
  const int k0 = K * fz, k1 = K == C.cz - 1 ? F.cz : k0 + fz;
  T num = T(0), den = T(0);
  for (int k = k0; k < k1; k++) {
    for (int j = j0; j < j1; j++) {
      T wyz = F.wy[j] * F.wz[k];
      for (int i = i0; i < i1; i++) {
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

// The same restriction when every direction is coarsened by 2: the two
// x-children of a row are one vector load (same summation order).
template <typename T>
__global__ void __launch_bounds__(256) k_restrict2(LevelArgs<T> F, LevelArgs<T> C, const T *__restrict__ res,
                                                   T *__restrict__ rhs_c, T *__restrict__ e_c) {
  using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
  if (I >= C.cx || J >= C.cy) return;
  const V2 wx = *reinterpret_cast<const V2 *>(F.wx + 2 * I);
  T num = T(0), den = T(0);
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int k = 2 * K + c;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = 2 * J + b;
      const T wyz = F.wy[j] * F.wz[k];
      const long id = lidx(F.sx, F.sxy, 2 * I, j, k);
      const V2 r = *reinterpret_cast<const V2 *>(res + id);
      const uint16_t f2 = *reinterpret_cast<const uint16_t *>(F.fl + id);
      if (!(f2 & FL_BLANK)) {
        T v = wx.x * wyz;
        num += v * r.x;
        den += v;
      }
      if (!((f2 >> 8) & FL_BLANK)) {
        T v = wx.y * wyz;
        num += v * r.y;
        den += v;
      }
    }
  }
  const long cid = lidx(C.sx, C.sxy, I, J, K);
  rhs_c[cid] = den > T(0) ? num / den : T(0);
  e_c[cid] = T(0);
}

// The residual and the by-2 restriction in one pass: each coarse cell forms
// the residuals of its 8 children (the k_residual row, same operation
// order) and their fluid-volume-weighted mean directly, so the fine residual
// is never stored.  Also zeroes the coarse correction.
// SPL: the fine level is in the colour-split row layout (sidx, hp).
template <typename T, bool SPL = false>
__global__ void __launch_bounds__(256) k_resrestrict2(LevelArgs<T> F, LevelArgs<T> C, const T *__restrict__ e,
                                                      const T *__restrict__ rhs, T *__restrict__ rhs_c,
                                                      T *__restrict__ e_c, int hp) {
  using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
  if (I >= C.cx || J >= C.cy) return;
  const int i0 = 2 * I;
  const V2 wx = *reinterpret_cast<const V2 *>(F.wx + i0);
  const V2 cmx = *reinterpret_cast<const V2 *>(F.cmx + i0);
  const V2 cpx = *reinterpret_cast<const V2 *>(F.cpx + i0);
  T num = T(0), den = T(0);
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int k = 2 * K + c;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = 2 * J + b;
      const T wyz = F.wy[j] * F.wz[k];
      V2 ev, ejm, ejp, ekm, ekp, rv;
      T el, er;
      uint16_t f2;
      if constexpr (SPL) {
        // children 2I (even half) and 2I+1 (odd half, hp further)
        const long id = sidx(F.sx, F.sxy, hp, i0, j, k), io = id + hp;
        ev.x = e[id]; ev.y = e[io];
        el = e[io - 1]; er = e[id + 1];
        ejm.x = e[id - F.sx]; ejm.y = e[io - F.sx];
        ejp.x = e[id + F.sx]; ejp.y = e[io + F.sx];
        ekm.x = e[id - F.sxy]; ekm.y = e[io - F.sxy];
        ekp.x = e[id + F.sxy]; ekp.y = e[io + F.sxy];
        rv.x = rhs[id]; rv.y = rhs[io];
        f2 = (uint16_t)(F.fl[id] | (F.fl[io] << 8));
      } else {
        const long id = lidx(F.sx, F.sxy, i0, j, k);
        ev = *reinterpret_cast<const V2 *>(e + id);
        el = e[id - 1]; er = e[id + 2];
        ejm = *reinterpret_cast<const V2 *>(e + id - F.sx);
        ejp = *reinterpret_cast<const V2 *>(e + id + F.sx);
        ekm = *reinterpret_cast<const V2 *>(e + id - F.sxy);
        ekp = *reinterpret_cast<const V2 *>(e + id + F.sxy);
        rv = *reinterpret_cast<const V2 *>(rhs + id);
        f2 = *reinterpret_cast<const uint16_t *>(F.fl + id);
      }
      const T my = F.cmy[j], py = F.cpy[j], mz = F.cmz[k], pz = F.cpz[k];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const uint8_t f = (uint8_t)(f2 >> (8 * t));
        if (f & FL_BLANK) continue;
        const T amx = (f & FL_IM) ? T(0) : (t ? cmx.y : cmx.x), apx = (f & FL_IP) ? T(0) : (t ? cpx.y : cpx.x);
        const T amy = (f & FL_JM) ? T(0) : my, apy = (f & FL_JP) ? T(0) : py;
        const T amz = (f & FL_KM) ? T(0) : mz, apz = (f & FL_KP) ? T(0) : pz;
        const T diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz);
        const T ec = t ? ev.y : ev.x, left = t ? ev.x : el, right = t ? er : ev.y;
        const T r = (t ? rv.y : rv.x) - ec * diag - left * amx - right * apx - (t ? ejm.y : ejm.x) * amy -
                    (t ? ejp.y : ejp.x) * apy - (t ? ekm.y : ekm.x) * amz - (t ? ekp.y : ekp.x) * apz;
        const T v = (t ? wx.y : wx.x) * wyz;
        num += v * r;
        den += v;
      }
    }
  }
  const long cid = lidx(C.sx, C.sxy, I, J, K);
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
                                                 ProlongDir<T> pz, const T *__restrict__ e_c, T *__restrict__ e_f,
                                                 const T *__restrict__ scale) {
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
  if (den > T(0)) {
    const T corr = num / den;
    e_f[id] = e_f[id] + (scale ? scale[id] * corr : corr);
  }
}

// Prolongation + correction when every direction is coarsened by 2: one
// thread per coarse cell (I,J,K) updates its 2x2x2 children, reading the
// 3x3x3 coarse neighbourhood once (the children's interpolation partners
// are I-1 / I+1 etc.; a partner outside the level carries weight 0).
// Same weights, masks and renormalisation as k_prolong.  scale (fine level
// layout, may be NULL) multiplies each cell's correction.
// SPL: the fine level (and scale) in the colour-split row layout (sidx, hp).
template <typename T, bool SPL = false>
__global__ void __launch_bounds__(256) k_prolong2(LevelArgs<T> F, LevelArgs<T> C, const T *__restrict__ wx,
                                                  const T *__restrict__ wy, const T *__restrict__ wz,
                                                  const T *__restrict__ e_c, T *__restrict__ e_f,
                                                  const T *__restrict__ scale, int hp) {
  using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
  if (I >= C.cx || J >= C.cy) return;
  T ev[3][3][3];
  bool fl[3][3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int a = 0; a < 3; a++) {
        long q = lidx(C.sx, C.sxy, I + a - 1, J + b - 1, K + c - 1);
        ev[c][b][a] = e_c[q];
        fl[c][b][a] = !(C.fl[q] & FL_BLANK);
      }
  const T wxp[2] = {wx[2 * I], wx[2 * I + 1]};
  const T wyp[2] = {wy[2 * J], wy[2 * J + 1]};
  const T wzp[2] = {wz[2 * K], wz[2 * K + 1]};
#pragma unroll
  for (int kc = 0; kc < 2; kc++)
#pragma unroll
    for (int jc = 0; jc < 2; jc++) {
      const long id = SPL ? sidx(F.sx, F.sxy, hp, 2 * I, 2 * J + jc, 2 * K + kc)
                          : lidx(F.sx, F.sxy, 2 * I, 2 * J + jc, 2 * K + kc);
      V2 ef;
      uint16_t f2;
      if constexpr (SPL) {
        ef.x = e_f[id]; ef.y = e_f[id + hp];
        f2 = (uint16_t)(F.fl[id] | (F.fl[id + hp] << 8));
      } else {
        ef = *reinterpret_cast<const V2 *>(e_f + id);
        f2 = *reinterpret_cast<const uint16_t *>(F.fl + id);
      }
      T out[2] = {ef.x, ef.y};
#pragma unroll
      for (int ic = 0; ic < 2; ic++) {
        if ((f2 >> (8 * ic)) & FL_BLANK) continue;
        // partners: parent (index 1) and the neighbour on the child's side
        const int oa = ic ? 2 : 0, ob = jc ? 2 : 0, oc = kc ? 2 : 0;
        const T wi[2] = {wxp[ic], T(1) - wxp[ic]}, wj[2] = {wyp[jc], T(1) - wyp[jc]},
                wk[2] = {wzp[kc], T(1) - wzp[kc]};
        const int ai[2] = {1, oa}, bj[2] = {1, ob}, ck[2] = {1, oc};
        T num = T(0), den = T(0);
#pragma unroll
        for (int z = 0; z < 2; z++)
#pragma unroll
          for (int y = 0; y < 2; y++)
#pragma unroll
            for (int x = 0; x < 2; x++) {
              T w = wi[x] * wj[y] * wk[z];
              if (w == T(0) || !fl[ck[z]][bj[y]][ai[x]]) continue;
              num += w * ev[ck[z]][bj[y]][ai[x]];
              den += w;
            }
        if (den > T(0)) {
          const T corr = num / den;
          const long is = SPL ? (ic ? id + hp : id) : id + ic;
          out[ic] = out[ic] + (scale ? scale[is] * corr : corr);
        }
      }
      if constexpr (SPL) {
        e_f[id] = out[0];
        e_f[id + hp] = out[1];
      } else {
        ef.x = out[0];
        ef.y = out[1];
        *reinterpret_cast<V2 *>(e_f + id) = ef;
      }
    }
}

// ---------------------------------------------------------------------
// reciprocals: v_rcp_f32 (1 ulp) for FP32, v_rcp_f64 refined by two
// Newton steps for FP64 (the line solves are smoothers inside an FP64
// outer iteration; every division of their recurrences goes through here)
// ---------------------------------------------------------------------
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// ---------------------------------------------------------------------
// lane exchange inside a group of G consecutive lanes of a wavefront
// (G = 16: DPP row shifts, one VALU op; otherwise ds_bpermute).
//   from_lower<d>(v, edge): value of lane l-d of the group, `edge` for the
//   first d lanes; from_upper<d>: lane l+d, `edge` for the last d lanes.
// ---------------------------------------------------------------------
template <int CTRL>
__device__ inline float dpp_move(float v, float edge) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double dpp_move(double v, double edge) {
  long long vb = __double_as_longlong(v), eb = __double_as_longlong(edge);
  int lo = __builtin_amdgcn_update_dpp((int)(eb & 0xffffffff), (int)(vb & 0xffffffff), CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp((int)(eb >> 32), (int)(vb >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int G, int D, typename T>
__device__ inline T from_lower(T v, T edge) {
  if constexpr (G == 16) {
    return dpp_move<0x110 + D>(v, edge);  // row_shr:D
  } else {
    T r = __shfl_up(v, D, G);
    return ((threadIdx.x & (G - 1)) < D) ? edge : r;
  }
}
template <int G, int D, typename T>
__device__ inline T from_upper(T v, T edge) {
  if constexpr (G == 16) {
    return dpp_move<0x100 + D>(v, edge);  // row_shl:D
  } else {
    T r = __shfl_down(v, D, G);
    return ((threadIdx.x & (G - 1)) + D >= G) ? edge : r;
  }
}

// ---------------------------------------------------------------------
// Tridiagonal line solve held by a group of G lanes, S consecutive rows
// per lane (row q = lane_in_group*S + s):  a_q x_{q-1} + b_q x_q + c_q x_{q+1} = d_q.
// Each lane eliminates its segment while the rows arrive (push), which
// expresses x_m = D_m + F_m XL + G_m XS through the previous segment's last
// unknown XL and its own last unknown XS; the G reduced equations are
// solved by parallel cyclic reduction across the group, then every lane
// back-substitutes.  Rows past the end of the line are identity rows.
// ---------------------------------------------------------------------
template <typename T, int S>
struct SegElim {
  T cp[S], dp[S], fp[S];
  __device__ inline void push(int m, T a, T b, T c, T d) {
    if (m == 0) {
      T inv = fast_rcp(b);
      cp[0] = c * inv;
      dp[0] = d * inv;
      fp[0] = -a * inv;
    } else {
      T inv = fast_rcp(b - a * cp[m - 1]);
      cp[m] = c * inv;
      dp[m] = (d - a * dp[m - 1]) * inv;
      fp[m] = -a * fp[m - 1] * inv;
    }
  }
  // after the last push: back recurrences in place (dp -> D, fp -> F,
  // cp -> G with x_m = D_m + F_m XL + G_m XS); returns the last row's
  // forward coefficients
  __device__ inline void back(T &cl, T &dl, T &fl) {
    cl = cp[S - 1]; dl = dp[S - 1]; fl = fp[S - 1];
    T Dn = T(0), Fn = T(0), Gn = T(1);
#pragma unroll
    for (int m = S - 2; m >= 0; m--) {
      T c = cp[m];
      Dn = dp[m] - c * Dn;
      Fn = fp[m] - c * Fn;
      Gn = -c * Gn;
      dp[m] = Dn; fp[m] = Fn; cp[m] = Gn;
    }
    dp[S - 1] = T(0); fp[S - 1] = T(0); cp[S - 1] = T(1);
  }
  // x_m into dp[m] from the neighbouring reduced unknowns
  __device__ inline void combine(T XL, T XS) {
#pragma unroll
    for (int m = 0; m < S - 1; m++) dp[m] = dp[m] + fp[m] * XL + cp[m] * XS;
    dp[S - 1] = XS;
  }
  // the whole line inside one group of G lanes: on return dp[m] holds x_m
  template <int G>
  __device__ inline void solve() {
    T cl, dl, fl;
    back(cl, dl, fl);
    T D0n = from_upper<G, 1>(dp[0], T(0)), F0n = from_upper<G, 1>(fp[0], T(0)),
      G0n = from_upper<G, 1>(cp[0], T(0));
    T A = -fl, B = T(1) + cl * F0n, C = cl * G0n, R = dl - cl * D0n;
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
      T Am, Bm, Cm, Rm, Ap, Bp, Cp, Rp;
      switch (d) {
#define PP_PCR_STEP(DD)                                                                         \
  case DD:                                                                                      \
    Am = from_lower<G, DD>(A, T(0)); Bm = from_lower<G, DD>(B, T(1));                           \
    Cm = from_lower<G, DD>(C, T(0)); Rm = from_lower<G, DD>(R, T(0));                           \
    Ap = from_upper<G, DD>(A, T(0)); Bp = from_upper<G, DD>(B, T(1));                           \
    Cp = from_upper<G, DD>(C, T(0)); Rp = from_upper<G, DD>(R, T(0));                           \
    break;
        PP_PCR_STEP(1) PP_PCR_STEP(2) PP_PCR_STEP(4) PP_PCR_STEP(8) PP_PCR_STEP(16) PP_PCR_STEP(32)
#undef PP_PCR_STEP
        default: Am = Cm = Rm = Ap = Cp = Rp = T(0); Bm = Bp = T(1); break;
      }
      T k1 = A * fast_rcp(Bm), k2 = C * fast_rcp(Bp);
      T nA = -Am * k1, nC = -Cp * k2;
      T nB = B - Cm * k1 - Ap * k2, nR = R - Rm * k1 - Rp * k2;
      A = nA; B = nB; C = nC; R = nR;
    }
    T X = R * fast_rcp(B);
    combine(from_lower<G, 1>(X, T(0)), X);
  }
};

// Workgroup index remapped so that each XCD (workgroups are dispatched to
// the 8 XCDs round-robin) runs one contiguous range of the logical order:
// neighbouring lines then share an L2.
__device__ inline int xcd_block(int b, int nb) {
  int q = nb >> 3, rem = nb & 7, x = b & 7;
  return x * q + (x < rem ? x : rem) + (b >> 3);
}

// ---------------------------------------------------------------------
// x-line zebra pass (lines along the contiguous index): lines (j,k) with
// (j+k)%2 == color; a group of G lanes holds one line, S cells per lane.
// ---------------------------------------------------------------------
// SPL: the level is in the colour-split row layout (sidx, half-row offset
// hp; S even): a lane's S cells are S/2 consecutive even cells and the S/2
// odd cells beside them.
template <typename T, int G, int S, bool SPL = false>
__global__ void __launch_bounds__(256) k_xline(LevelArgs<T> L, const T *__restrict__ rhs, T *__restrict__ e,
                                               int color, T omega, int hp) {
  static_assert(!SPL || S % 2 == 0, "split rows need an even chunk");
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int line = (blk * 256 + (int)threadIdx.x) / G;
  const int sub = threadIdx.x & (G - 1);
  const int nyh = (L.cy + 1) >> 1;
  const int k = line / nyh;
  const int j = 2 * (line - k * nyh) + ((k + color) & 1);
  const bool valid = k < L.cz && j < L.cy;
  const int i0 = sub * S;
  const int nv = valid ? max(0, min(S, L.cx - i0)) : 0;
  const long base = valid ? (SPL ? sidx(L.sx, L.sxy, hp, i0, j, k) : lidx(L.sx, L.sxy, i0, j, k)) : 0;
  T eo[S], ejm[S], ejp[S], ekm[S], ekp[S], rv[S], cmx[S], cpx[S];
  uint8_t fl[S];
  auto ld = [&](const T *p, T(&v)[S]) {
    if constexpr (SPL) {
      T a[S / 2], b[S / 2];
      load_chunk<T, S / 2>(p, (nv + 1) >> 1, a);
      load_chunk<T, S / 2>(p + hp, nv >> 1, b);
#pragma unroll
      for (int q = 0; q < S / 2; q++) { v[2 * q] = a[q]; v[2 * q + 1] = b[q]; }
    } else {
      load_chunk<T, S>(p, nv, v);
    }
  };
  ld(e + base, eo);
  ld(e + base - L.sx, ejm);
  ld(e + base + L.sx, ejp);
  ld(e + base - L.sxy, ekm);
  ld(e + base + L.sxy, ekp);
  ld(rhs + base, rv);
  load_chunk<T, S>(L.cmx + i0, nv, cmx);
  load_chunk<T, S>(L.cpx + i0, nv, cpx);
  if constexpr (SPL) {
    uint8_t a[S / 2], b[S / 2];
    load_flags<S / 2>(L.fl + base, (nv + 1) >> 1, a);
    load_flags<S / 2>(L.fl + base + hp, nv >> 1, b);
#pragma unroll
    for (int q = 0; q < S / 2; q++) { fl[2 * q] = a[q]; fl[2 * q + 1] = b[q]; }
  } else {
    load_flags<S>(L.fl + base, nv, fl);
  }
  T my = valid ? L.cmy[j] : T(0), py = valid ? L.cpy[j] : T(0);
  T mz = valid ? L.cmz[k] : T(0), pz = valid ? L.cpz[k] : T(0);
  SegElim<T, S> se;
#pragma unroll
  for (int s = 0; s < S; s++) {
    uint8_t f = fl[s];
    T a = T(0), b = T(1), c = T(0), d = T(0);
    if (s < nv) {
      if (f & FL_BLANK) {
        d = (f & FL_GHOST) ? eo[s] : T(0);
      } else {
        T amx = (f & FL_IM) ? T(0) : cmx[s], apx = (f & FL_IP) ? T(0) : cpx[s];
        T amy = (f & FL_JM) ? T(0) : my, apy = (f & FL_JP) ? T(0) : py;
        T amz = (f & FL_KM) ? T(0) : mz, apz = (f & FL_KP) ? T(0) : pz;
        T diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz);
        T r = rv[s] - ejm[s] * amy - ejp[s] * apy - ekm[s] * amz - ekp[s] * apz;
        if (tabs(diag) < T(1.0e-12)) {
          d = eo[s];
        } else {
          a = amx; b = diag; c = apx; d = r;
        }
      }
    }
    se.push(s, a, b, c, d);
  }
  se.template solve<G>();
#pragma unroll
  for (int s = 0; s < S; s++) eo[s] = (T(1) - omega) * eo[s] + omega * se.dp[s];
  if constexpr (SPL) {
    T a[S / 2], b[S / 2];
#pragma unroll
    for (int q = 0; q < S / 2; q++) { a[q] = eo[2 * q]; b[q] = eo[2 * q + 1]; }
    store_chunk<T, S / 2>(e + base, (nv + 1) >> 1, a);
    store_chunk<T, S / 2>(e + base + hp, nv >> 1, b);
  } else {
    store_chunk<T, S>(e + base, nv, eo);
  }
}

// Reduced system of a line split into P segments held by the P threads
// (tx, 0..P-1) of a block column: back recurrences of each segment, then
// parallel cyclic reduction of the P equations (unknown = a segment's last
// cell) with one equation per thread, double-buffered in LDS
// [buf][seg][field][tx] (8*P*blockDim.x elements).  Every thread of the
// block must call it (it synchronises the block).  Returns the unknowns of
// the previous segment (XL, 0 for the first) and of this segment (XS).
template <typename T, int LS>
__device__ inline void seg_reduced_solve(T *sh, int P, int tx, int seg, SegElim<T, LS> &se, T &XL, T &XS) {
  T cl, dl, fl;
  se.back(cl, dl, fl);
  const int TXb = blockDim.x;
  T *bufA = sh, *bufB = sh + 4 * TXb * P;
  auto E_ = [&](T *buf, int s, int fld) -> T & { return buf[((long)s * 4 + fld) * TXb + tx]; };
  E_(bufB, seg, 0) = se.dp[0]; E_(bufB, seg, 1) = se.fp[0]; E_(bufB, seg, 2) = se.cp[0];
  __syncthreads();
  T D0n = T(0), F0n = T(0), G0n = T(0);
  if (seg + 1 < P) { D0n = E_(bufB, seg + 1, 0); F0n = E_(bufB, seg + 1, 1); G0n = E_(bufB, seg + 1, 2); }
  T A = -fl, B = T(1) + cl * F0n, C = cl * G0n, R = dl - cl * D0n;
  T *cur = bufA, *nxt = bufB;
  E_(cur, seg, 0) = A; E_(cur, seg, 1) = B; E_(cur, seg, 2) = C; E_(cur, seg, 3) = R;
  __syncthreads();
  for (int d = 1; d < P; d <<= 1) {
    T Am = T(0), Bm = T(1), Cm = T(0), Rm = T(0), Ap = T(0), Bp = T(1), Cp = T(0), Rp = T(0);
    if (seg - d >= 0) {
      Am = E_(cur, seg - d, 0); Bm = E_(cur, seg - d, 1); Cm = E_(cur, seg - d, 2); Rm = E_(cur, seg - d, 3);
    }
    if (seg + d < P) {
      Ap = E_(cur, seg + d, 0); Bp = E_(cur, seg + d, 1); Cp = E_(cur, seg + d, 2); Rp = E_(cur, seg + d, 3);
    }
    T k1 = A * fast_rcp(Bm), k2 = C * fast_rcp(Bp);
    T nA = -Am * k1, nC = -Cp * k2;
    T nB = B - Cm * k1 - Ap * k2, nR = R - Rm * k1 - Rp * k2;
    A = nA; B = nB; C = nC; R = nR;
    E_(nxt, seg, 0) = A; E_(nxt, seg, 1) = B; E_(nxt, seg, 2) = C; E_(nxt, seg, 3) = R;
    __syncthreads();
    T *t = cur; cur = nxt; nxt = t;
  }
  XS = R * fast_rcp(B);
  E_(nxt, seg, 0) = XS;
  __syncthreads();
  XL = seg > 0 ? E_(nxt, seg - 1, 0) : T(0);
}

// ---------------------------------------------------------------------
// y- and z-line zebra passes (lines across the contiguous index): lines
// (i, plane) with (i + plane) % 2 == colour, plane = k for DIR = 1 (lines
// along j) and plane = j for DIR = 2 (lines along k), so every neighbour
// line of a line has the other colour and the pass is deterministic.  A
// block of TX x P threads covers TX lines of one colour (every other i of
// a 2*TX-wide tile) and P segments of LS cells of each line.  Each thread
// eliminates its segment in registers, the P reduced equations of a line
// are solved by parallel cyclic reduction in LDS (one equation per thread),
// and every thread back-substitutes.  The tile's loads cover 2*TX consecutive i, half of
// them own cells and half the neighbour lines read by the stencil.
// ---------------------------------------------------------------------
template <typename T, int LS, int DIR, bool UPF = false>
__global__ void __launch_bounds__(1024) k_yzline(LevelArgs<T> Lv, const T *__restrict__ rhs, T *__restrict__ e,
                                                 int color, T omega) {
  extern __shared__ unsigned char smem_raw[];
  T *sh = reinterpret_cast<T *>(smem_raw);
  const int P = blockDim.y;
  const int tx = threadIdx.x, seg = threadIdx.y;
  const int lb = xcd_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int tile = lb % gridDim.x, plane = lb / gridDim.x;
  const int i = 2 * (tile * blockDim.x + tx) + ((plane + color) & 1);
  const int n = DIR == 1 ? Lv.cy : Lv.cz;
  const bool ok = ((DIR == 1) ? (plane < Lv.cz) : (plane < Lv.cy)) && i < Lv.cx;
  const int m0 = seg * LS;
  // coefficients constant along the line
  T mx = T(0), px = T(0), mo = T(0), po = T(0);
  if (ok) {
    mx = Lv.cmx[i]; px = Lv.cpx[i];
    mo = DIR == 1 ? Lv.cmz[plane] : Lv.cmy[plane];
    po = DIR == 1 ? Lv.cpz[plane] : Lv.cpy[plane];
  }
  const long ostride = DIR == 1 ? Lv.sxy : (long)Lv.sx;   // the other transverse direction
  const long lstride = DIR == 1 ? (long)Lv.sx : Lv.sxy;   // along the line
  const uint8_t FO_M = DIR == 1 ? FL_KM : FL_JM, FO_P = DIR == 1 ? FL_KP : FL_JP;
  const uint8_t FL_M = DIR == 1 ? FL_JM : FL_KM, FL_P = DIR == 1 ? FL_JP : FL_KP;
  const T *cml = DIR == 1 ? Lv.cmy : Lv.cmz;
  const T *cpl = DIR == 1 ? Lv.cpy : Lv.cpz;
  long id0 = 0;
  if (ok) id0 = DIR == 1 ? lidx(Lv.sx, Lv.sxy, i, m0, plane) : lidx(Lv.sx, Lv.sxy, i, plane, m0);
  // the line coefficients of this segment, loaded once ahead of the cells
  // (wavefront-uniform scalar loads when the tile is 64 lines wide)
  T cq_m[LS], cq_p[LS];
  const int m0u = blockDim.x >= 64 ? __builtin_amdgcn_readfirstlane(m0) : m0;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    const int q = m0u + m;
    cq_m[m] = q < n ? cml[q] : T(0);
    cq_p[m] = q < n ? cpl[q] : T(0);
  }
  // UPF: all loads of the segment are issued unconditionally (cells past the
  // line or the level read a valid cell instead and are then ignored), so
  // that they are all in flight before the first row is built (latency-bound
  // small levels; on large levels the extra registers cost occupancy)
  const long idsafe = lidx(Lv.sx, Lv.sxy, 0, 0, 0);
  constexpr int NU = UPF ? LS : 1;
  uint8_t fv[NU];
  T evv[NU], elv[NU], erv[NU], ezm[NU], ezp[NU], rvv[NU];
  auto load_cell = [&](int m, int slot) {
    const bool live = ok && m0 + m < n;
    const long id = live ? id0 + m * lstride : idsafe;
    fv[slot] = Lv.fl[id];
    evv[slot] = e[id];
    elv[slot] = e[id - 1];
    erv[slot] = e[id + 1];
    ezm[slot] = e[id - ostride];
    ezp[slot] = e[id + ostride];
    rvv[slot] = rhs[id];
  };
  if constexpr (UPF) {
#pragma unroll
    for (int m = 0; m < LS; m++) load_cell(m, m);
  }
  T eo[LS];
  SegElim<T, LS> se;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    T a = T(0), b = T(1), c = T(0), d = T(0);
    eo[m] = T(0);
    int q = m0 + m;
    const int sl = UPF ? m : 0;
    if constexpr (!UPF) {
      if (ok && q < n) load_cell(m, 0);
    }
    if (ok && q < n) {
      uint8_t f = fv[sl];
      T ev = evv[sl];
      eo[m] = ev;
      if (f & FL_BLANK) {
        d = (f & FL_GHOST) ? ev : T(0);
      } else {
        T amx = (f & FL_IM) ? T(0) : mx, apx = (f & FL_IP) ? T(0) : px;
        T amo = (f & FO_M) ? T(0) : mo, apo = (f & FO_P) ? T(0) : po;
        T am = (f & FL_M) ? T(0) : cq_m[m], ap = (f & FL_P) ? T(0) : cq_p[m];
        // (acx + acy + acz) in the reference's order
        T diag = DIR == 1 ? (-(amx + apx) + -(am + ap)) + -(amo + apo) : (-(amx + apx) + -(amo + apo)) + -(am + ap);
        T r = rvv[sl] - elv[sl] * amx - erv[sl] * apx - ezm[sl] * amo - ezp[sl] * apo;
        if (tabs(diag) < T(1.0e-12)) {
          d = ev;
        } else {
          a = am; b = diag; c = ap; d = r;
        }
      }
    }
    se.push(m, a, b, c, d);
  }
  T XL, XS;
  seg_reduced_solve<T, LS>(sh, P, tx, seg, se, XL, XS);
  if (!ok) return;
  se.combine(XL, XS);
#pragma unroll
  for (int m = 0; m < LS; m++) {
    if (m0 + m < n) e[id0 + m * lstride] = (T(1) - omega) * eo[m] + omega * se.dp[m];
  }
}

// k_yzline on the colour-split FP32 level 0 (sidx, half-row offset hp):
// a colour's cells are contiguous half rows, their x-neighbours the
// other half.  The cells of a segment are read PF ahead of their
// elimination (every load unconditional: cells past the line or the level
// read a valid cell and are ignored), and the register budget is held at
// six wavefronts per SIMD so that two 768-thread blocks share a CU.
template <int LS, int DIR>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_yzline_split(LevelArgs<float> Lv, int hp, const float *__restrict__ rhs, float *__restrict__ e, int color,
               float omega) {
  constexpr int PF = LS < 4 ? LS : 4;
  extern __shared__ unsigned char smem_raw[];
  float *sh = reinterpret_cast<float *>(smem_raw);
  const int P = blockDim.y;
  const int tx = threadIdx.x, seg = threadIdx.y;
  const int lb = xcd_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int tile = lb % gridDim.x, plane = lb / gridDim.x;
  const int par = (plane + color) & 1;
  const int i = 2 * (tile * blockDim.x + tx) + par;
  const int n = DIR == 1 ? Lv.cy : Lv.cz;
  const bool ok = ((DIR == 1) ? (plane < Lv.cz) : (plane < Lv.cy)) && i < Lv.cx;
  const int m0 = seg * LS;
  float mx = 0.f, px = 0.f, mo = 0.f, po = 0.f;
  if (ok) {
    mx = Lv.cmx[i]; px = Lv.cpx[i];
    mo = DIR == 1 ? Lv.cmz[plane] : Lv.cmy[plane];
    po = DIR == 1 ? Lv.cpz[plane] : Lv.cpy[plane];
  }
  const long ostride = DIR == 1 ? Lv.sxy : (long)Lv.sx;
  const long lstride = DIR == 1 ? (long)Lv.sx : Lv.sxy;
  const uint8_t FO_M = DIR == 1 ? FL_KM : FL_JM, FO_P = DIR == 1 ? FL_KP : FL_JP;
  const uint8_t FL_M = DIR == 1 ? FL_JM : FL_KM, FL_P = DIR == 1 ? FL_JP : FL_KP;
  const float *cml = DIR == 1 ? Lv.cmy : Lv.cmz;
  const float *cpl = DIR == 1 ? Lv.cpy : Lv.cpz;
  const long idsafe = sidx(Lv.sx, Lv.sxy, hp, 0, 0, 0);
  long id0 = idsafe;
  if (ok) id0 = DIR == 1 ? sidx(Lv.sx, Lv.sxy, hp, i, m0, plane) : sidx(Lv.sx, Lv.sxy, hp, i, plane, m0);
  // offsets of the x-neighbours i-1 and i+1 (the other half of the row)
  const long dlo = par ? -(long)hp : (long)hp - 1;
  const long dhi = par ? 1 - (long)hp : (long)hp;
  uint8_t fv[PF];
  float evv[PF], elv[PF], erv[PF], ezm[PF], ezp[PF], rvv[PF], cqm[PF], cqp[PF];
  auto load_cell = [&](int m, int slot) {
    const bool live = ok && m0 + m < n;
    const long id = live ? id0 + m * lstride : idsafe;
    const int q = live ? m0 + m : 0;
    fv[slot] = Lv.fl[id];
    evv[slot] = e[id];
    elv[slot] = e[id + dlo];
    erv[slot] = e[id + dhi];
    ezm[slot] = e[id - ostride];
    ezp[slot] = e[id + ostride];
    rvv[slot] = rhs[id];
    cqm[slot] = cml[q];
    cqp[slot] = cpl[q];
  };
#pragma unroll
  for (int m = 0; m < PF; m++) load_cell(m, m);
  float eo[LS];
  SegElim<float, LS> se;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    const int sl = m % PF;
    const uint8_t f = fv[sl];
    const float ev = evv[sl], el = elv[sl], er = erv[sl], zm = ezm[sl], zp = ezp[sl], rv = rvv[sl];
    const float cm = cqm[sl], cp = cqp[sl];
    if (m + PF < LS) load_cell(m + PF, sl);
    float a = 0.f, b = 1.f, c = 0.f, d = 0.f;
    const bool live = ok && m0 + m < n;
    eo[m] = live ? ev : 0.f;
    if (live) {
      if (f & FL_BLANK) {
        d = (f & FL_GHOST) ? ev : 0.f;
      } else {
        const float amx = (f & FL_IM) ? 0.f : mx, apx = (f & FL_IP) ? 0.f : px;
        const float amo = (f & FO_M) ? 0.f : mo, apo = (f & FO_P) ? 0.f : po;
        const float am = (f & FL_M) ? 0.f : cm, ap = (f & FL_P) ? 0.f : cp;
        // (acx + acy + acz) in the reference's order
        const float diag =
            DIR == 1 ? (-(amx + apx) + -(am + ap)) + -(amo + apo) : (-(amx + apx) + -(amo + apo)) + -(am + ap);
        const float r = rv - el * amx - er * apx - zm * amo - zp * apo;
        if (tabs(diag) < 1.0e-12f) {
          d = ev;
        } else {
          a = am; b = diag; c = ap; d = r;
        }
      }
    }
    se.push(m, a, b, c, d);
  }
  float XL, XS;
  seg_reduced_solve<float, LS>(sh, P, tx, seg, se, XL, XS);
  if (!ok) return;
  se.combine(XL, XS);
#pragma unroll
  for (int m = 0; m < LS; m++) {
    if (m0 + m < n) e[id0 + m * lstride] = (1.f - omega) * eo[m] + omega * se.dp[m];
  }
}

// The FP32 level-0 right-hand side into the colour-split copy, and the
// split correction zeroed (interior cells; two cells per thread).
__global__ void __launch_bounds__(256) k_split_in(LevelArgs<float> N, LevelArgs<float> S, int hp,
                                                  const float *__restrict__ rhs, float *__restrict__ rhs_s,
                                                  float *__restrict__ e_s) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  const int i = 2 * I;
  if (i >= N.cx || j >= N.cy) return;
  const long a = lidx(N.sx, N.sxy, i, j, k), se = sidx(S.sx, S.sxy, hp, i, j, k);
  rhs_s[se] = rhs[a];
  e_s[se] = 0.f;
  if (i + 1 < N.cx) {
    rhs_s[se + hp] = rhs[a + 1];
    e_s[se + hp] = 0.f;
  }
}

// the split correction back into the natural FP32 level-0 layout
__global__ void __launch_bounds__(256) k_split_out(LevelArgs<float> N, LevelArgs<float> S, int hp,
                                                   const float *__restrict__ e_s, float *__restrict__ e) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  const int i = 2 * I;
  if (i >= N.cx || j >= N.cy) return;
  const long a = lidx(N.sx, N.sxy, i, j, k), se = sidx(S.sx, S.sxy, hp, i, j, k);
  e[a] = e_s[se];
  if (i + 1 < N.cx) e[a + 1] = e_s[se + hp];
}

// ---------------------------------------------------------------------
// coarsest level (FP64 in both precisions): e = G r with the dense
// solution operator G of the coarsest level (bordered LU on the GPU at
// set-up: the coarse equations with the fluid-volume mean constraint per
// connected component), stored by output cell: GT[q*n + p] = response of
// cell q to a unit rhs at cell p (FP64 in both precisions, FP64
// accumulation); cell[q] = offset of coarse cell q (i fastest)
// in the level layout.  One wavefront per output cell streams
// its row of GT (coalesced) against the rhs staged in LDS, then reduces
// across the wavefront in a fixed order.  Rows go round-robin over the
// blocks (row q to block q % gridDim.x), so a grid of one block per CU
// gives every CU the same number of rows.
// ---------------------------------------------------------------------
constexpr int kCoarseBlocks = 256;   // blocks of the launch (one per CU)
constexpr int kCoarseMaxWaves = 16;  // wavefronts per block at most
template <typename T, typename TG>
__global__ void __launch_bounds__(64 * kCoarseMaxWaves) k_coarse_mv(int n, const int *__restrict__ cell,
                                                                    const TG *__restrict__ GT,
                                                                    const T *__restrict__ rhs, T *__restrict__ e) {
  extern __shared__ unsigned char smem_raw[];
  double *r = reinterpret_cast<double *>(smem_raw);
  for (int p = threadIdx.x; p < n; p += blockDim.x) r[p] = (double)rhs[cell[p]];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  // q is uniform per wavefront
  for (int q = wv * gridDim.x + blockIdx.x; q < n; q += nwv * gridDim.x) {
    const TG *row = GT + (long)q * n;
    // four independent row loads in flight per lane, four partial sums added
    // in a fixed order
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int p = lane;
    for (; p + 192 < n; p += 256) {
      const double a0 = (double)row[p], a1 = (double)row[p + 64], a2 = (double)row[p + 128], a3 = (double)row[p + 192];
      s0 += a0 * r[p];
      s1 += a1 * r[p + 64];
      s2 += a2 * r[p + 128];
      s3 += a3 * r[p + 192];
    }
    for (; p < n; p += 64) s0 += (double)row[p] * r[p];
    double s = (s0 + s1) + (s2 + s3);
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_down(s, d, 64);
    if (lane == 0) e[cell[q]] = (T)s;
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

// Block sums of n <= KRY_MAX + 1 values per thread in one barrier: each
// value is reduced across its wavefront by shuffles, the wavefront sums meet
// in LDS and threads q < n add them (fixed order) into part[blk * n + q].
template <int NT>
__device__ inline void block_sums_to(double (&v)[KRY_MAX + 1], int n, double *s, double *part, long blk) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x, w = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++) {
    if (q < n) {
      double x = v[q];
      for (int d = 32; d >= 1; d >>= 1) x += __shfl_down(x, d, 64);
      if (lane == 0) s[w * (KRY_MAX + 1) + q] = x;
    }
  }
  __syncthreads();
  if (tid < n) {
    double r = 0.0;
    for (int ww = 0; ww < NT / 64; ww++) r += s[ww * (KRY_MAX + 1) + tid];
    part[blk * n + tid] = r;
  }
}

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

// w = A z (- bsub) on fluid cells (0 elsewhere) with z's ghost rows and
// outer layer already exact; per-block partial dots <w, u_q>, q < V.n, at
// part[blk*V.n+q].  bsub (may be null): a field subtracted on fluid cells
// (the transposed ghost coupling of the adjoint operator).
template <int NT>
__global__ void __launch_bounds__(NT) k_kry_apply(RefArgs R, LevelArgs<double> L, const double *__restrict__ z,
                                                  double *__restrict__ w, KryVecs V, double *__restrict__ part,
                                                  const double *__restrict__ bsub) {
  __shared__ double s[(NT / 64) * (KRY_MAX + 1)];
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
      if (bsub) wv -= bsub[r];
    }
    w[r] = wv;
  }
  double acc[KRY_MAX + 1];
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++) acc[q] = (in && q < V.n) ? wv * V.u[q][r] : 0.0;
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  block_sums_to<NT>(acc, V.n, s, part, blk);
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
// Flexible GMRES for the mixed-precision path: basis and preconditioned
// vectors stored in FP32 in the level-0 layout (half the bytes of the FP64
// reference-layout vectors); products, dots and coefficients in FP64 on the
// stored values, so the Arnoldi relation holds for the vectors as stored.
// The outer FP64 residual at every restart keeps the solve exact to FP64.
// ---------------------------------------------------------------------
struct KryVecsF {
  const float *u[KRY_MAX + 1];
  double c[KRY_MAX + 1];
  int n;
};

// z = e0 (the V-cycle's correction, ghost rows solved), w = A z on fluid
// cells (FP64 coefficients), partial dots <w, u_q>
template <int NT>
__global__ void __launch_bounds__(NT) k_kryl_apply(LevelArgs<double> L, const float *__restrict__ e0,
                                                   float *__restrict__ z, float *__restrict__ w, KryVecsF V,
                                                   double *__restrict__ part) {
  __shared__ double s[NT / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  bool in = i < L.cx && j < L.cy;
  double wv = 0.0;
  long a = 0;
  if (in) {
    a = lidx(L.sx, L.sxy, i, j, k);
    float zc = e0[a];
    z[a] = zc;
    uint8_t f = L.fl[a];
    if (!(f & FL_BLANK)) {
      Row<double> rw = make_row(L, f, i, j, k);
      wv = (double)zc * rw.diag + (double)e0[a - 1] * rw.amx + (double)e0[a + 1] * rw.apx +
           (double)e0[a - L.sx] * rw.amy + (double)e0[a + L.sx] * rw.apy + (double)e0[a - L.sxy] * rw.amz +
           (double)e0[a + L.sxy] * rw.apz;
    }
    float wf = (float)wv;
    w[a] = wf;
    wv = (double)wf;
  }
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++) {
    if (q < V.n) {
      double d = block_sum<NT>(in ? wv * (double)V.u[q][a] : 0.0, s);
      if (threadIdx.x == 0 && threadIdx.y == 0) part[blk * V.n + q] = d;
    }
  }
}

// u = w - sum_q c_q u_q on fluid cells (in place in w), level-0 rhs = u,
// e0 = 0, partial |u|^2 of the stored values
template <int NT>
__global__ void __launch_bounds__(NT) k_kryl_orth(LevelArgs<double> L, float *__restrict__ w, KryVecsF V,
                                                  float *__restrict__ rhs0, float *__restrict__ e0,
                                                  double *__restrict__ part) {
  __shared__ double s[NT / 64];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double u2 = 0.0;
  if (i < L.cx && j < L.cy) {
    long a = lidx(L.sx, L.sxy, i, j, k);
    float uf = 0.0f;
    if (!(L.fl[a] & FL_BLANK)) {
      double u = (double)w[a];
#pragma unroll
      for (int q = 0; q < KRY_MAX + 1; q++)
        if (q < V.n) u -= V.c[q] * (double)V.u[q][a];
      uf = (float)u;
      w[a] = uf;
    }
    rhs0[a] = uf;
    e0[a] = 0.0f;
    u2 = (double)uf * (double)uf;
  }
  double d = block_sum<NT>(u2, s);
  long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blk] = d;
}

// k_kryl_apply for four consecutive cells per thread (rows whose interior
// starts 16-byte aligned, cx % 4 == 0): vector loads of every field and all
// dots reduced after one barrier.
// SPL: e0 is the colour-split copy (rows sp_sx, planes sp_sxy, half-row
// offset hp); z and w stay in the natural layout.
template <int NT, bool SPL = false>
__global__ void __launch_bounds__(NT) k_kryl_apply4(LevelArgs<double> L, const float *__restrict__ e0,
                                                    float *__restrict__ z, float *__restrict__ w, KryVecsF V,
                                                    double *__restrict__ part, int sp_sx, long sp_sxy, int hp) {
  __shared__ double s[(NT / 64) * (KRY_MAX + 1)];
  const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  const bool in = i0 < L.cx && j < L.cy;
  double acc[KRY_MAX + 1];
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++) acc[q] = 0.0;
  if (in) {
    const long a = lidx(L.sx, L.sxy, i0, j, k);
    float zc[4], ejm[4], ejp[4], ekm[4], ekp[4], wf[4];
    uint8_t fl[4];
    double cmx[4], cpx[4], wd[4];
    float el, er;
    if constexpr (SPL) {
      // cells i0, i0+2 at even half index i0/2, i0/2+1; i0+1, i0+3 in the odd half
      const long se = sidx(sp_sx, sp_sxy, hp, i0, j, k);
      auto ld = [&](long o, float(&v)[4]) {
        const float2 ev = *reinterpret_cast<const float2 *>(e0 + se + o);
        const float2 od = *reinterpret_cast<const float2 *>(e0 + se + hp + o);
        v[0] = ev.x; v[1] = od.x; v[2] = ev.y; v[3] = od.y;
      };
      ld(0, zc);
      ld(-sp_sx, ejm);
      ld(sp_sx, ejp);
      ld(-sp_sxy, ekm);
      ld(sp_sxy, ekp);
      el = e0[se + hp - 1];
      er = e0[se + 2];
    } else {
      load_chunk<float, 4>(e0 + a, 4, zc);
      load_chunk<float, 4>(e0 + a - L.sx, 4, ejm);
      load_chunk<float, 4>(e0 + a + L.sx, 4, ejp);
      load_chunk<float, 4>(e0 + a - L.sxy, 4, ekm);
      load_chunk<float, 4>(e0 + a + L.sxy, 4, ekp);
      el = e0[a - 1];
      er = e0[a + 4];
    }
    load_flags<4>(L.fl + a, 4, fl);
    load_chunk<double, 4>(L.cmx + i0, 4, cmx);
    load_chunk<double, 4>(L.cpx + i0, 4, cpx);
    const double my = L.cmy[j], py = L.cpy[j], mz = L.cmz[k], pz = L.cpz[k];
    store_chunk<float, 4>(z + a, 4, zc);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const uint8_t f = fl[t];
      double wv = 0.0;
      if (!(f & FL_BLANK)) {
        double amx = (f & FL_IM) ? 0.0 : cmx[t], apx = (f & FL_IP) ? 0.0 : cpx[t];
        double amy = (f & FL_JM) ? 0.0 : my, apy = (f & FL_JP) ? 0.0 : py;
        double amz = (f & FL_KM) ? 0.0 : mz, apz = (f & FL_KP) ? 0.0 : pz;
        double diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz);
        const float left = t == 0 ? el : zc[t - 1], right = t == 3 ? er : zc[t + 1];
        wv = (double)zc[t] * diag + (double)left * amx + (double)right * apx + (double)ejm[t] * amy +
             (double)ejp[t] * apy + (double)ekm[t] * amz + (double)ekp[t] * apz;
      }
      wf[t] = (float)wv;
      wd[t] = (double)wf[t];
    }
    store_chunk<float, 4>(w + a, 4, wf);
#pragma unroll
    for (int q = 0; q < KRY_MAX + 1; q++) {
      if (q < V.n) {
        float uq[4];
        load_chunk<float, 4>(V.u[q] + a, 4, uq);
        acc[q] = wd[0] * (double)uq[0] + wd[1] * (double)uq[1] + wd[2] * (double)uq[2] + wd[3] * (double)uq[3];
      }
    }
  }
  const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  block_sums_to<NT>(acc, V.n, s, part, blk);
}

// The adjoint operator's transposed ghost coupling on the FP32 level-0
// basis: after k_kryl_apply4 has written w = A z, each touched fluid cell
// gets w -= vinv * sum_q ct_w[q] rhsg[ct_ghost[q]] (the composite rows'
// transpose applied to the ghost sources rhsg of z), and the partial dots
// <w, u_q> of that launch are corrected by (w_new - w_old) u_q, written as
// blocks blk0, blk0 + 1, ... of part next to the apply launch's blocks.
template <int NT>
__global__ void __launch_bounds__(NT) k_kryl_adj_coupling(int nct, const long *__restrict__ ct_al,
                                                          const int *__restrict__ ct_ptr,
                                                          const int *__restrict__ ct_ghost,
                                                          const float *__restrict__ ct_w,
                                                          const float *__restrict__ vinv,
                                                          const float *__restrict__ rhsg, float *__restrict__ w,
                                                          KryVecsF V, double *__restrict__ part, long blk0) {
  __shared__ double s[(NT / 64) * (KRY_MAX + 1)];
  const int t = blockIdx.x * NT + threadIdx.x;
  double acc[KRY_MAX + 1];
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++) acc[q] = 0.0;
  if (t < nct) {
    float sum = 0.0f;
    for (int q = ct_ptr[t]; q < ct_ptr[t + 1]; q++) sum += ct_w[q] * rhsg[ct_ghost[q]];
    const long a = ct_al[t];
    const float wo = w[a];
    const float wn = wo - vinv[t] * sum;
    w[a] = wn;
    const double d = (double)wn - (double)wo;
#pragma unroll
    for (int q = 0; q < KRY_MAX + 1; q++)
      if (q < V.n) acc[q] = d * (double)V.u[q][a];
  }
  block_sums_to<NT>(acc, V.n, s, part, blk0 + blockIdx.x);
}

// k_kryl_orth for four consecutive cells per thread (cx % 4 == 0)
// SPL: rhs0 and e0 are the colour-split copies (rows sp_sx, planes sp_sxy,
// half-row offset hp); w stays in the natural layout.
template <int NT, bool SPL = false>
__global__ void __launch_bounds__(NT) k_kryl_orth4(LevelArgs<double> L, float *__restrict__ w, KryVecsF V,
                                                   float *__restrict__ rhs0, float *__restrict__ e0,
                                                   double *__restrict__ part, int sp_sx, long sp_sxy, int hp) {
  __shared__ double s[NT / 64];
  const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  double u2 = 0.0;
  if (i0 < L.cx && j < L.cy) {
    const long a = lidx(L.sx, L.sxy, i0, j, k);
    float wv[4], uf[4];
    uint8_t fl[4];
    load_chunk<float, 4>(w + a, 4, wv);
    load_flags<4>(L.fl + a, 4, fl);
    double u[4];
#pragma unroll
    for (int t = 0; t < 4; t++) u[t] = (double)wv[t];
#pragma unroll
    for (int q = 0; q < KRY_MAX + 1; q++) {
      if (q < V.n) {
        float uq[4];
        load_chunk<float, 4>(V.u[q] + a, 4, uq);
#pragma unroll
        for (int t = 0; t < 4; t++) u[t] -= V.c[q] * (double)uq[t];
      }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      uf[t] = (fl[t] & FL_BLANK) ? 0.0f : (float)u[t];
      u2 += (double)uf[t] * (double)uf[t];
    }
    store_chunk<float, 4>(w + a, 4, uf);
    if constexpr (SPL) {
      const long se = sidx(sp_sx, sp_sxy, hp, i0, j, k);
      *reinterpret_cast<float2 *>(rhs0 + se) = make_float2(uf[0], uf[2]);
      *reinterpret_cast<float2 *>(rhs0 + se + hp) = make_float2(uf[1], uf[3]);
      *reinterpret_cast<float2 *>(e0 + se) = make_float2(0.0f, 0.0f);
      *reinterpret_cast<float2 *>(e0 + se + hp) = make_float2(0.0f, 0.0f);
    } else {
      store_chunk<float, 4>(rhs0 + a, 4, uf);
      const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      store_chunk<float, 4>(e0 + a, 4, zero);
    }
  }
  const double d = block_sum<NT>(u2, s);
  const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blk] = d;
}

// p (FP64, reference layout) += sum_q c_q z_q (FP32, level layout) on fluid cells
__global__ void __launch_bounds__(256) k_kryl_update(RefArgs R, LevelArgs<double> L, double *__restrict__ p,
                                                     KryVecsF V) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  long a = lidx(L.sx, L.sxy, i, j, k);
  if (L.fl[a] & FL_BLANK) return;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < KRY_MAX + 1; q++)
    if (q < V.n) acc += V.c[q] * (double)V.u[q][a];
  long r = ridx(R, i, j, k);
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
template <int G, int S>
__global__ void __launch_bounds__(256) k_tdma_wave(int n, int nlines, const double *__restrict__ A,
                                                   const double *__restrict__ B, const double *__restrict__ C,
                                                   const double *__restrict__ Rr, double *__restrict__ U) {
  // one group of G lanes per line; lanes of a group past nlines solve
  // identity rows (every lane takes part in the group exchanges)
  const int line = (blockIdx.x * 256 + (int)threadIdx.x) / G;
  const int sub = threadIdx.x & (G - 1);
  const bool ok = line < nlines;
  const long base = ok ? (long)line * n : 0;
  SegElim<double, S> se;
#pragma unroll
  for (int s = 0; s < S; s++) {
    int q = sub * S + s;
    double a = 0, b = 1, c = 0, d = 0;
    if (ok && q < n) {
      a = q > 0 ? A[base + q] : 0.0;
      b = B[base + q];
      c = q < n - 1 ? C[base + q] : 0.0;
      d = Rr[base + q];
    }
    se.push(s, a, b, c, d);
  }
  se.template solve<G>();
#pragma unroll
  for (int s = 0; s < S; s++) {
    int q = sub * S + s;
    if (ok && q < n) U[base + q] = se.dp[s];
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
  const int m0 = seg * LS;
  SegElim<double, LS> se;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    int q = m0 + m;
    double a = 0, b = 1, c = 0, d = 0;
    if (ok && q < n) {
      // lines interleaved so that consecutive threads read consecutive
      // addresses: element q of line l at q*nlines + l
      long id = (long)q * nlines + line;
      a = q > 0 ? A[id] : 0.0;
      b = B[id];
      c = q < n - 1 ? C[id] : 0.0;
      d = Rr[id];
    }
    se.push(m, a, b, c, d);
  }
  double XL, XS;
  seg_reduced_solve<double, LS>(sh, P, tx, seg, se, XL, XS);
  if (!ok) return;
  se.combine(XL, XS);
#pragma unroll
  for (int m = 0; m < LS; m++) {
    int q = m0 + m;
    if (q < n) U[(long)q * nlines + line] = se.dp[m];
  }
}


// =====================================================================
// Implicit momentum (advection-diffusion) solve of the fractional step,
// FP64 on the reference layout: set_solve_ad, the line sweeps
// itsolv_ad_x/y/z and calc_residual_ad (CORE_SET_SOLVERS.F90:20-223,
// CORE_SOLVE_AD.F90:314-907) with the GCM fresh-cell rows, without periodic
// faces.
// The level-0 flags give iblank (FL_BLANK), ghostCellMark (FL_GHOST) and
// ium..kup (FL_IM..FL_KP).
// =====================================================================

// set_solve_ad: coef = amx_ad, apx_ad, amy_ad, apy_ad, amz_ad, apz_ad
// (nref apart) on the interior cells; the caller clears the outer layer.
template <int NT>
__global__ void __launch_bounds__(NT) k_set_solve_ad(RefArgs R, LevelArgs<double> L0, FracMetrics M, double dt,
                                                     const double *__restrict__ visc, const double *__restrict__ bxv,
                                                     const double *__restrict__ byv, const double *__restrict__ bzv,
                                                     double *__restrict__ coef, long nref) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  const long id = ridx(R, i, j, k);
  const uint8_t f = L0.fl[lidx(L0.sx, L0.sxy, i, j, k)];
  const long st[3] = {1, R.ex, R.exy};
  const int I[3] = {i + 1, j + 1, k + 1};
  const uint8_t FM[3] = {FL_IM, FL_JM, FL_KM}, FP[3] = {FL_IP, FL_JP, FL_KP};
  const double *bv[3] = {bxv, byv, bzv};
  const double hdt = 0.5 * dt, vc = visc[id];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int um = (f & FM[d]) ? 1 : 0, up = (f & FP[d]) ? 1 : 0;
    const int q = I[d];
    const double fp = M.fx[d][q + 1], fm = M.fx[d][q];
    const double nup = (fp * visc[id + st[d]] + (1.0 - fp) * vc) * (double)(1 - up) + bv[d][id] * (double)up;
    const double num = (fm * vc + (1.0 - fm) * visc[id - st[d]]) * (double)(1 - um) + bv[d][id] * (double)um;
    const double di = M.dinv[d][q];
    const double am = (M.dcinv[d][q] * (double)(1 - um) + di * (double)um * 2.0) * di;
    const double ap = (M.dcinv[d][q + 1] * (double)(1 - up) + di * (double)up * 2.0) * di;
    // amz_ad, apz_ad carry the factor ndim - 2 = 1 of a 3-D grid
    coef[(2 * d) * nref + id] = d == 2 ? -(hdt * num) * am * 1.0 : -(hdt * num) * am;
    coef[(2 * d + 1) * nref + id] = d == 2 ? -(hdt * nup) * ap * 1.0 : -(hdt * nup) * ap;
  }
}

// The operator of one cell as itsolv_ad_* and calc_residual_ad build it
// before masking (CORE_SOLVE_AD.F90:391-421): the viscous coefficients
// plus the Crank-Nicolson advection terms (half_dt = 0 for
// ADAMS_BASHFORTH2), and the diagonal sum acx + acy + acz.  Face
// velocities are fixed during one solve_ad call, so the rows are built
// once per call: row = amx, apx, amy, apy, amz, apz, sum (nref apart).
template <int NT>
__global__ void __launch_bounds__(NT) k_ad_rows(RefArgs R, LevelArgs<double> L0, FracMetrics M, double half_dt,
                                                const double *__restrict__ coef, const double *__restrict__ fu,
                                                const double *__restrict__ fv, const double *__restrict__ fw,
                                                const int8_t *__restrict__ fresh, double *__restrict__ row,
                                                long nref) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  if (i >= R.nxc || j >= R.nyc) return;
  const long id = ridx(R, i, j, k);
  const uint8_t f = L0.fl[lidx(L0.sx, L0.sxy, i, j, k)];
  const long st[3] = {1, R.ex, R.exy};
  const int I[3] = {i + 1, j + 1, k + 1};
  const uint8_t FM[3] = {FL_IM, FL_JM, FL_KM}, FP[3] = {FL_IP, FL_JP, FL_KP};
  const double *face[3] = {fu, fv, fw};
  double ac[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int um = (f & FM[d]) ? 1 : 0, up = (f & FP[d]) ? 1 : 0;
    const int q = I[d];
    double am = coef[(2 * d) * nref + id], ap = coef[(2 * d + 1) * nref + id];
    double a = -(am + ap);
    const double fm = M.fx[d][q], fp = M.fx[d][q + 1], di = M.dinv[d][q];
    const double vm = face[d][id], vp = face[d][id + st[d]];
    double t1 = 1.0 - fm, t2 = fp;
    am = am - half_dt * vm * t1 * di;
    ap = ap + half_dt * vp * t2 * di;
    t1 = fm * (double)(1 - um);
    t2 = (1.0 - fp) * (double)(1 - up);
    a = a + half_dt * (vp * t2 - vm * t1) * di;
    row[(2 * d) * nref + id] = am;
    row[(2 * d + 1) * nref + id] = ap;
    ac[d] = a;
  }
  // killForGCMFresh: a fresh cell's diagonal is 1 (CORE_SOLVE_AD.F90:425-427, 866)
  const double kill = fresh ? 1.0 - 1.0 * (double)fresh[id] : 1.0;
  row[6 * nref + id] = (ac[0] + ac[1] + ac[2]) * kill;
}

// Fresh-cell rows of set_solve_ad's GCM branch (CORE_SET_SOLVERS.F90:
// 175-218): amx_ad..apz_ad of each fresh cell replaced by -coeffGCMFreshD of
// its face neighbour in the stencil, 0 where the stencil has none (face: 6
// per cell, prepared with the tables).
__global__ void k_fresh_ad_coef(int n, const long *__restrict__ ids, const double *__restrict__ face,
                                double *__restrict__ coef, long nref) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long id = ids[q];
#pragma unroll
  for (int d = 0; d < 6; d++) coef[d * nref + id] = face[6 * q + d];
}

// The off-stencil terms of GCM_correct_rhs_ad (GCM_FRESHCELL_AD.F90:2-70)
// for a zebra line pass: rw = r + sum_m cc_m var(coff_m) at each fresh cell
// (the four stencil nodes that differ from the cell in two or more indices,
// in stencil order), from the fields as the pass starts, so every line of
// the pass reads the same values.  rw is the call's copy of r, equal to r
// off the fresh cells.  One thread per cell and field.
struct FreshCorr {
  const double *var[3];
  const double *r[3];
  double *rw[3];
};
__global__ void k_fresh_rhs(int n, int nv, const long *__restrict__ ids, const long *__restrict__ coff,
                            const double *__restrict__ cc, FreshCorr F) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nv) return;
  const int q = t % n, f = t / n;
  const double *var = F.var[f];
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < 4; m++) s = s + cc[4 * q + m] * var[coff[4 * q + m]];
  F.rw[f][ids[q]] = F.r[f][ids[q]] + s;
}

// index of reference offset id in the ascending fresh list, -1 if absent
__device__ inline int fresh_find(const long *__restrict__ ids, int n, long id) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const long v = ids[mid];
    if (v == id) return mid;
    if (v < id) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// SegElim for NR right-hand sides of one tridiagonal matrix: the
// elimination of the matrix (cp, fp) is done once, each right-hand side
// follows it with exactly the operations SegElim applies to its own.
template <typename T, int S, int NR>
struct SegElimN {
  T cp[S], fp[S], dp[NR][S];
  __device__ inline void push(int m, T a, T b, T c, const T (&d)[NR]) {
    if (m == 0) {
      T inv = fast_rcp(b);
      cp[0] = c * inv;
#pragma unroll
      for (int q = 0; q < NR; q++) dp[q][0] = d[q] * inv;
      fp[0] = -a * inv;
    } else {
      T inv = fast_rcp(b - a * cp[m - 1]);
      cp[m] = c * inv;
#pragma unroll
      for (int q = 0; q < NR; q++) dp[q][m] = (d[q] - a * dp[q][m - 1]) * inv;
      fp[m] = -a * fp[m - 1] * inv;
    }
  }
  __device__ inline void back(T &cl, T (&dl)[NR], T &fl) {
    cl = cp[S - 1]; fl = fp[S - 1];
    T Dn[NR], Fn = T(0), Gn = T(1);
#pragma unroll
    for (int q = 0; q < NR; q++) { dl[q] = dp[q][S - 1]; Dn[q] = T(0); }
#pragma unroll
    for (int m = S - 2; m >= 0; m--) {
      T c = cp[m];
#pragma unroll
      for (int q = 0; q < NR; q++) { Dn[q] = dp[q][m] - c * Dn[q]; dp[q][m] = Dn[q]; }
      Fn = fp[m] - c * Fn;
      Gn = -c * Gn;
      fp[m] = Fn; cp[m] = Gn;
    }
#pragma unroll
    for (int q = 0; q < NR; q++) dp[q][S - 1] = T(0);
    fp[S - 1] = T(0); cp[S - 1] = T(1);
  }
  __device__ inline void combine(int q, T XL, T XS) {
#pragma unroll
    for (int m = 0; m < S - 1; m++) dp[q][m] = dp[q][m] + fp[m] * XL + cp[m] * XS;
    dp[q][S - 1] = XS;
  }
};

// seg_reduced_solve for SegElimN: the reduced matrix (A, B, C) once, NR
// right-hand sides R; LDS [buf][seg][field][tx] with 3 + NR fields
// (2 * (3 + NR) * P * blockDim.x elements).  Every thread of the block must
// call it.
template <typename T, int LS, int NR>
__device__ inline void seg_reduced_solve_n(T *sh, int P, int tx, int seg, SegElimN<T, LS, NR> &se, T (&XL)[NR],
                                           T (&XS)[NR]) {
  constexpr int NF = 3 + NR;
  T cl, dl[NR], fl;
  se.back(cl, dl, fl);
  const int TXb = blockDim.x;
  T *bufA = sh, *bufB = sh + NF * TXb * P;
  auto E_ = [&](T *buf, int s, int fld) -> T & { return buf[((long)s * NF + fld) * TXb + tx]; };
#pragma unroll
  for (int q = 0; q < NR; q++) E_(bufB, seg, q) = se.dp[q][0];
  E_(bufB, seg, NR) = se.fp[0]; E_(bufB, seg, NR + 1) = se.cp[0];
  __syncthreads();
  T D0n[NR], F0n = T(0), G0n = T(0);
#pragma unroll
  for (int q = 0; q < NR; q++) D0n[q] = T(0);
  if (seg + 1 < P) {
#pragma unroll
    for (int q = 0; q < NR; q++) D0n[q] = E_(bufB, seg + 1, q);
    F0n = E_(bufB, seg + 1, NR); G0n = E_(bufB, seg + 1, NR + 1);
  }
  T A = -fl, B = T(1) + cl * F0n, C = cl * G0n, R[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) R[q] = dl[q] - cl * D0n[q];
  T *cur = bufA, *nxt = bufB;
  E_(cur, seg, 0) = A; E_(cur, seg, 1) = B; E_(cur, seg, 2) = C;
#pragma unroll
  for (int q = 0; q < NR; q++) E_(cur, seg, 3 + q) = R[q];
  __syncthreads();
  for (int d = 1; d < P; d <<= 1) {
    T Am = T(0), Bm = T(1), Cm = T(0), Ap = T(0), Bp = T(1), Cp = T(0), Rm[NR], Rp[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) { Rm[q] = T(0); Rp[q] = T(0); }
    if (seg - d >= 0) {
      Am = E_(cur, seg - d, 0); Bm = E_(cur, seg - d, 1); Cm = E_(cur, seg - d, 2);
#pragma unroll
      for (int q = 0; q < NR; q++) Rm[q] = E_(cur, seg - d, 3 + q);
    }
    if (seg + d < P) {
      Ap = E_(cur, seg + d, 0); Bp = E_(cur, seg + d, 1); Cp = E_(cur, seg + d, 2);
#pragma unroll
      for (int q = 0; q < NR; q++) Rp[q] = E_(cur, seg + d, 3 + q);
    }
    T k1 = A * fast_rcp(Bm), k2 = C * fast_rcp(Bp);
    T nA = -Am * k1, nC = -Cp * k2;
    T nB = B - Cm * k1 - Ap * k2;
#pragma unroll
    for (int q = 0; q < NR; q++) R[q] = R[q] - Rm[q] * k1 - Rp[q] * k2;
    A = nA; B = nB; C = nC;
    E_(nxt, seg, 0) = A; E_(nxt, seg, 1) = B; E_(nxt, seg, 2) = C;
#pragma unroll
    for (int q = 0; q < NR; q++) E_(nxt, seg, 3 + q) = R[q];
    __syncthreads();
    T *t = cur; cur = nxt; nxt = t;
  }
  const T rB = fast_rcp(B);
#pragma unroll
  for (int q = 0; q < NR; q++) {
    XS[q] = R[q] * rB;
    E_(nxt, seg, q) = XS[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NR; q++) XL[q] = seg > 0 ? E_(nxt, seg - 1, q) : T(0);
}

// One zebra pass of the momentum line sweeps along DIR (0 x, 1 y, 2 z)
// for u, v and w together: the three share the row coefficients, which are
// read once, and are independent until the ghost update, so solving them
// in one pass gives what three passes would.  Lines of colour `color` (the
// parity of the sum of their two transverse indices), so every off-line
// neighbour has the other colour and the pass is deterministic.  A block
// of TX lines x P segments of LS cells; each thread eliminates its segment
// of the three fields in registers (one elimination of the shared matrix)
// and the reduced system of a line is solved by cyclic reduction in LDS
// with three right-hand sides (seg_reduced_solve_n).  For
// x-lines the segments of one line are adjacent lanes (coalesced along the
// row); for y- and z-lines adjacent lanes are every other i.
// Row of a cell (itsolv_ad_x/y/z): off-line terms subtracted from r in the
// reference's grouping, masked line coefficients, diagonal
// 1 + sum * (1 - iblank), right-hand side r' (1 - iblank) + var * ghost.
struct AdLineFields {
  double *var[3];
  const double *r[3];
};
template <int DIR, int LS>
__global__ void __launch_bounds__(512) k_ad_line(RefArgs R, LevelArgs<double> L0, const double *__restrict__ row,
                                                 long nref, AdLineFields F, int color, int nlines) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem_raw[];
  double *sh = reinterpret_cast<double *>(smem_raw);
  const int TX = blockDim.x, P = blockDim.y;
  int tx, seg, i = 0, j = 0, k = 0;
  bool ok;
  if (DIR == 0) {
    const int flat = threadIdx.x + threadIdx.y * TX;
    seg = flat % P;
    tx = flat / P;
    const int line = blockIdx.x * TX + tx;
    const int nyh = (R.nyc + 1) >> 1;
    k = line / nyh;
    j = 2 * (line - k * nyh) + ((k + color) & 1);
    ok = line < nlines && k < R.nzc && j < R.nyc;
  } else {
    tx = threadIdx.x;
    seg = threadIdx.y;
    const int plane = blockIdx.y;
    i = 2 * (blockIdx.x * TX + tx) + ((plane + color) & 1);
    if (DIR == 1) k = plane; else j = plane;
    ok = i < R.nxc && plane < (DIR == 1 ? R.nzc : R.nyc);
  }
  const int n = DIR == 0 ? R.nxc : (DIR == 1 ? R.nyc : R.nzc);
  const long rst = DIR == 0 ? 1 : (DIR == 1 ? R.ex : R.exy);
  const long lst = DIR == 0 ? 1 : (DIR == 1 ? (long)L0.sx : L0.sxy);
  const int m0 = seg * LS;
  long id0 = 0, la0 = 0;
  if (ok) {
    const int i0 = DIR == 0 ? m0 : i, j0 = DIR == 1 ? m0 : j, k0 = DIR == 2 ? m0 : k;
    id0 = ridx(R, i0, j0, k0);
    la0 = lidx(L0.sx, L0.sxy, i0, j0, k0);
  }
  SegElimN<double, LS, 3> se;
#pragma unroll
  for (int m = 0; m < LS; m++) {
    double a = 0.0, b = 1.0, c = 0.0, d[3] = {0.0, 0.0, 0.0};
    if (ok && m0 + m < n) {
      const long id = id0 + m * rst;
      const uint8_t f = L0.fl[la0 + m * lst];
      const double rib = (f & FL_BLANK) ? 0.0 : 1.0;
      const double gh = (f & FL_GHOST) ? 1.0 : 0.0;
      const double amx = row[id], apx = row[nref + id], amy = row[2 * nref + id], apy = row[3 * nref + id];
      const double amz = row[4 * nref + id], apz = row[5 * nref + id], sum = row[6 * nref + id];
      const double oim = (f & FL_IM) ? 0.0 : 1.0, oip = (f & FL_IP) ? 0.0 : 1.0;
      const double ojm = (f & FL_JM) ? 0.0 : 1.0, ojp = (f & FL_JP) ? 0.0 : 1.0;
      const double okm = (f & FL_KM) ? 0.0 : 1.0, okp = (f & FL_KP) ? 0.0 : 1.0;
      if (DIR == 0) {
        a = amx * rib * oim;
        c = apx * rib * oip;
      } else if (DIR == 1) {
        a = amy * rib * ojm;
        c = apy * rib * ojp;
      } else {
        a = amz * rib * okm;
        c = apz * rib * okp;
      }
      b = 1.0 + sum * rib;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const double *var = F.var[q];
        double rhs;
        if (DIR == 0)
          rhs = F.r[q][id] - (var[id - R.ex] * amy * ojm + var[id + R.ex] * apy * ojp +
                              var[id - R.exy] * amz * okm + var[id + R.exy] * apz * okp);
        else if (DIR == 1)
          rhs = F.r[q][id] - (var[id - 1] * amx * oim + var[id + 1] * apx * oip + var[id - R.exy] * amz * okm +
                              var[id + R.exy] * apz * okp);
        else
          rhs = F.r[q][id] - var[id - R.ex] * amy * ojm - var[id + R.ex] * apy * ojp - var[id - 1] * amx * oim -
                var[id + 1] * apx * oip;
        d[q] = rhs * rib + var[id] * gh;
      }
    }
    se.push(m, a, b, c, d);
  }
  double XL[3], XS[3];
  seg_reduced_solve_n<double, LS, 3>(sh, P, tx, seg, se, XL, XS);
  if (!ok) return;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    se.combine(q, XL[q], XS[q]);
#pragma unroll
    for (int m = 0; m < LS; m++)
      if (m0 + m < n) F.var[q][id0 + m * rst] = se.dp[q][m];
  }
}

// calc_residual_ad for NV fields (u, v, w share the rows): per-block
// partials (0, max |res|, argmax) of each field at part + 3*nblk*q, the
// argmax the smallest reference-layout offset of the block's maximum.
struct AdFields {
  const double *var[3], *r[3];
};
// fresh cells of the residual: mask (reference layout, NULL = none), the
// ascending offsets and the four off-stencil nodes and weights of each
struct FreshRes {
  const int8_t *mask;
  const long *ids, *coff;
  const double *cc;
  int n;
};
template <int NT, int NV>
__global__ void __launch_bounds__(NT) k_ad_residual(RefArgs R, LevelArgs<double> L0, const double *__restrict__ row,
                                                    long nref, AdFields F, FreshRes FR, double *__restrict__ part,
                                                    long nblk, double *__restrict__ rout) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
  const bool in = i < R.nxc && j < R.nyc;
  long id = 0;
  uint8_t f = FL_BLANK;
  double amx = 0, apx = 0, amy = 0, apy = 0, amz = 0, apz = 0, sum = 0;
  if (in) {
    id = ridx(R, i, j, k);
    f = L0.fl[lidx(L0.sx, L0.sxy, i, j, k)];
    amx = row[id]; apx = row[nref + id]; amy = row[2 * nref + id]; apy = row[3 * nref + id];
    amz = row[4 * nref + id]; apz = row[5 * nref + id]; sum = row[6 * nref + id];
  }
  const double oim = (f & FL_IM) ? 0.0 : 1.0, oip = (f & FL_IP) ? 0.0 : 1.0;
  const double ojm = (f & FL_JM) ? 0.0 : 1.0, ojp = (f & FL_JP) ? 0.0 : 1.0;
  const double okm = (f & FL_KM) ? 0.0 : 1.0, okp = (f & FL_KP) ? 0.0 : 1.0;
  const double fl = (f & FL_BLANK) ? 0.0 : 1.0;
  // GCM_correct_res_ad (GCM_FRESHCELL_AD.F90:73-137): the off-stencil terms
  // of a fresh cell, added one by one in stencil order before the iblank mask
  const int fq = (in && FR.mask && FR.mask[id]) ? fresh_find(FR.ids, FR.n, id) : -1;
  const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
#pragma unroll
  for (int q = 0; q < NV; q++) {
    SumMax v{0.0, 0.0, -1};
    if (in) {
      const double *x = F.var[q];
      double s = F.r[q][id] - x[id] * (1.0 + sum * 1.0) - x[id - 1] * amx * oim - x[id + 1] * apx * oip -
                 x[id - R.ex] * amy * ojm - x[id + R.ex] * apy * ojp - x[id - R.exy] * amz * okm -
                 x[id + R.exy] * apz * okp;
      if (fq >= 0)
        for (int m = 0; m < 4; m++) s = s + FR.cc[4 * fq + m] * x[FR.coff[4 * fq + m]];
      s = s * fl;
      if (rout) rout[id] = s;
      v.mx = fabs(s);
      v.arg = id;
    }
    v = block_summax<NT>(v);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      double *pq = part + 3 * nblk * q;
      pq[3 * blk] = 0.0;
      pq[3 * blk + 1] = v.mx;
      pq[3 * blk + 2] = (double)v.arg;
    }
    __syncthreads();
  }
}

}  // namespace pp
