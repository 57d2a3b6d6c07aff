// C-ABI entry points (include/pp_hip.h) and the multigrid driver.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include <rocblas/rocblas.h>

#include "pp_kernels.h"
#include "pp_setup.h"

using namespace pp;

#define HIPCHK(x)                                                                 \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      return PP_HIP_ERROR;                                                        \
    }                                                                             \
  } while (0)

// ---------------------------------------------------------------------
// Scratch memory of the allocation-free entry point (pp_divergence_noalloc):
// one array per GPU for the whole process, holding the per-workgroup
// partial sums of the reduction and its two later stages.  Host code never
// uses the array's name as an address: pp_create looks up its device
// address once per device (hipGetSymbolAddress with the device current) and
// the entry point passes that address to the kernels.
// ---------------------------------------------------------------------
constexpr long kNoAllocPartials = 1L << 19;   // workgroups of the first stage
constexpr long kNoAllocDoubles = kNoAllocPartials + 256 + 64;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (once per device)
int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d: at most %d devices", dev, kMaxDevices);
    return PP_BAD_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return PP_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    return PP_HIP_ERROR;
  }
  g_scratch_ptr[dev] = static_cast<double *>(ptr);
  return PP_OK;
}
}  // namespace

namespace {

// ---------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------
// temporary device allocation of an entry point, freed on every return path
struct DevTemp {
  void *p = nullptr;
  DevTemp() = default;
  DevTemp(const DevTemp &) = delete;
  DevTemp &operator=(const DevTemp &) = delete;
  ~DevTemp() { if (p) hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
  double *d() const { return static_cast<double *>(p); }
};

struct CellLaunch {
  dim3 grid, block;
};
CellLaunch cell_launch(int cx, int cy, int cz) {
  // 256 threads per block always (the block reductions assume 4 waves)
  int bx = cx <= 64 ? 64 : (cx <= 128 ? 128 : 256);
  int by = 256 / bx;
  CellLaunch c;
  c.block = dim3(bx, by, 1);
  c.grid = dim3((cx + bx - 1) / bx, (cy + by - 1) / by, cz);
  return c;
}

// Byte accounting and opt-in timing of one launch: every launch site calls
// kbegin(...) before and kend(...) after the launch.
void kbegin(pp_ctx *c, int cls, int level, double bytes) {
  c->bytes_acc += bytes;
  c->kbytes[cls] += bytes;
  c->klaunch[cls]++;
  if (level == 0) { c->kbytes_l0[cls] += bytes; c->klaunch_l0[cls]++; }
  if (!c->prof_on) return;
  size_t q = 2 * c->prof_rec.size();
  while (c->prof_ev.size() < q + 2) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    c->prof_ev.push_back(e);
  }
  (void)hipEventRecord(c->prof_ev[q], c->stream);
  c->prof_rec.push_back({cls, level, bytes});
}
void kend(pp_ctx *c) {
  if (!c->prof_on) return;
  (void)hipEventRecord(c->prof_ev[2 * c->prof_rec.size() - 1], c->stream);
}
void kreset(pp_ctx *c, bool profile) {
  c->bytes_acc = 0.0;
  for (int q = 0; q < PP_K_CLASSES; q++) {
    c->kbytes[q] = c->kbytes_l0[q] = 0.0;
    c->klaunch[q] = c->klaunch_l0[q] = 0;
  }
  c->prof_on = profile;
  c->prof_rec.clear();
}
// after the stream has been synchronised
void kcollect(pp_ctx *c, pp_solve_stats &S) {
  for (int q = 0; q < PP_K_CLASSES; q++) {
    S.kern_bytes[q] = c->kbytes[q];
    S.kern_bytes_l0[q] = c->kbytes_l0[q];
    S.kern_launches[q] = c->klaunch[q];
    S.kern_launches_l0[q] = c->klaunch_l0[q];
    S.kern_ms[q] = S.kern_ms_l0[q] = 0.0;
  }
  if (c->prof_on) {
    for (size_t r = 0; r < c->prof_rec.size(); r++) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, c->prof_ev[2 * r], c->prof_ev[2 * r + 1]);
      S.kern_ms[c->prof_rec[r].cls] += ms;
      if (c->prof_rec[r].level == 0) S.kern_ms_l0[c->prof_rec[r].cls] += ms;
    }
  }
  c->prof_on = false;
  c->prof_rec.clear();
}

template <typename T>
LevelArgs<T> largs(const Level &L);
template <>
LevelArgs<float> largs<float>(const Level &L) {
  LevelArgs<float> a;
  a.cx = L.cx; a.cy = L.cy; a.cz = L.cz; a.sx = L.sx; a.sxy = L.sxy; a.fl = L.flags;
  a.cmx = L.cm32[0]; a.cpx = L.cp32[0]; a.cmy = L.cm32[1]; a.cpy = L.cp32[1]; a.cmz = L.cm32[2]; a.cpz = L.cp32[2];
  a.wx = L.w32[0]; a.wy = L.w32[1]; a.wz = L.w32[2];
  return a;
}
template <>
LevelArgs<double> largs<double>(const Level &L) {
  LevelArgs<double> a;
  a.cx = L.cx; a.cy = L.cy; a.cz = L.cz; a.sx = L.sx; a.sxy = L.sxy; a.fl = L.flags;
  a.cmx = L.cm64[0]; a.cpx = L.cp64[0]; a.cmy = L.cm64[1]; a.cpy = L.cp64[1]; a.cmz = L.cm64[2]; a.cpz = L.cp64[2];
  a.wx = L.w64[0]; a.wy = L.w64[1]; a.wz = L.w64[2];
  return a;
}
template <typename T>
T *lv_e(Level &L);
template <>
float *lv_e<float>(Level &L) { return L.e32; }
template <>
double *lv_e<double>(Level &L) { return L.e64; }
template <typename T>
T *lv_rhs(Level &L);
template <>
float *lv_rhs<float>(Level &L) { return L.rhs32; }
template <>
double *lv_rhs<double>(Level &L) { return L.rhs64; }
template <typename T>
T *lv_res(Level &L);
template <>
float *lv_res<float>(Level &L) { return L.res32; }
template <>
double *lv_res<double>(Level &L) { return L.res64; }
template <typename T>
ProlongDir<T> pdir(const Level &L, int d);
template <>
ProlongDir<float> pdir<float>(const Level &L, int d) { return {L.par_d[d], L.oth_d[d], L.wpar32[d]}; }
template <>
ProlongDir<double> pdir<double>(const Level &L, int d) { return {L.par_d[d], L.oth_d[d], L.wpar64[d]}; }
template <typename T>
const T *ghost_cw(const GhostTables &g);
template <>
const float *ghost_cw<float>(const GhostTables &g) { return g.c_w32; }
template <>
const double *ghost_cw<double>(const GhostTables &g) { return g.c_w64; }

// algorithmic bytes per cell (see pp_kernels.h)
double bytes_line(int s) { return 3.0 * s + 1.0; }

// ---------------------------------------------------------------------
// line passes
// ---------------------------------------------------------------------

template <typename Fn>
void with_line_shape(int G, int S, Fn &&fn) {
  using std::integral_constant;
#define PP_CASE_S(GG)                                                   \
  switch (S) {                                                          \
    case 1: fn(integral_constant<int, GG>{}, integral_constant<int, 1>{}); break;   \
    case 2: fn(integral_constant<int, GG>{}, integral_constant<int, 2>{}); break;   \
    case 3: fn(integral_constant<int, GG>{}, integral_constant<int, 3>{}); break;   \
    case 4: fn(integral_constant<int, GG>{}, integral_constant<int, 4>{}); break;   \
    case 6: fn(integral_constant<int, GG>{}, integral_constant<int, 6>{}); break;   \
    case 8: fn(integral_constant<int, GG>{}, integral_constant<int, 8>{}); break;   \
    case 12: fn(integral_constant<int, GG>{}, integral_constant<int, 12>{}); break; \
    default: fn(integral_constant<int, GG>{}, integral_constant<int, 16>{}); break; \
  }
  if (G == 16) {
    PP_CASE_S(16)
  } else if (G == 32) {
    PP_CASE_S(32)
  } else {
    PP_CASE_S(64)
  }
#undef PP_CASE_S
}

// the colour-split level 0 is in use (FP32 V-cycles, forward or transposed)
template <typename T>
bool split_on(const pp_ctx *c, int lvl) {
  return std::is_same<T, float>::value && lvl == 0 && c->sp.on;
}
// level arguments and arrays of the colour-split level 0
template <typename T>
LevelArgs<T> split_args(const pp_ctx *c, const Level &L) {
  LevelArgs<T> a = largs<T>(L);
  a.sx = c->sp.sx; a.sxy = c->sp.sxy; a.fl = c->sp.fl;
  return a;
}
template <typename T>
T *split_e(pp_ctx *c) { return std::is_same<T, float>::value ? reinterpret_cast<T *>(c->sp.e) : nullptr; }
template <typename T>
T *split_rhs(pp_ctx *c) { return std::is_same<T, float>::value ? reinterpret_cast<T *>(c->sp.rhs) : nullptr; }
template <typename T>
T *split_nshape(pp_ctx *c) { return std::is_same<T, float>::value ? reinterpret_cast<T *>(c->sp.nshape) : nullptr; }

template <typename T>
void launch_xline(pp_ctx *c, Level &L, int lvl, int color, T omega) {
  const bool spl = split_on<T>(c, lvl);
  LevelArgs<T> a = spl ? split_args<T>(c, L) : largs<T>(L);
  int G, S;
  line_shape(L.cx, G, S, (int)sizeof(T));
  long lines = (long)((L.cy + 1) / 2) * L.cz;
  dim3 grid((unsigned)((lines * G + 255) / 256)), block(256);
  T *e = spl ? split_e<T>(c) : lv_e<T>(L);
  const T *rhs = spl ? split_rhs<T>(c) : lv_rhs<T>(L);
  const int hp = c->sp.hp;
  hipStream_t st = c->stream;
  kbegin(c, PP_K_XLINE, lvl, 0.5 * (double)L.cx * L.cy * L.cz * bytes_line(sizeof(T)));
  with_line_shape(G, S, [&](auto g, auto sz) {
    constexpr int GG = decltype(g)::value, SS = decltype(sz)::value;
    // the split layout needs even chunks (Split0::ok guarantees them)
    if constexpr (SS % 2 == 0 && std::is_same<T, float>::value) {
      if (spl) {
        k_xline<T, GG, SS, true><<<grid, block, 0, st>>>(a, rhs, e, color, omega, hp);
        return;
      }
    }
    k_xline<T, GG, SS><<<grid, block, 0, st>>>(a, rhs, e, color, omega, 0);
  });
  kend(c);
}

// segment length LS and segments P for a line of n cells
// FP32: up to 16 segments of up to 16 cells (64 x P threads); FP64 keeps
// at most 12 cells per thread in registers and uses up to 32 segments
// (32 x P threads) for long lines.
void yz_shape(int n, int &LS, int &P, int &TX, int elem = 4) {
  // at most 8 cells per thread (FP64 register budget; FP32 measured no
  // faster with 12 or 16), at most 16 segments while that suffices
  static const int opts[] = {1, 2, 4, 8};
  int maxls = 8;
  LS = maxls;
  for (int o : opts) {
    if ((n + o - 1) / o <= 16) { LS = o; break; }
  }
  P = (n + LS - 1) / LS;
  // TX x P <= 1024 threads: lines of up to 1024 cells
  TX = P <= 16 ? 64 : (P <= 32 ? 32 : (P <= 64 ? 16 : 8));
  (void)elem;
}

template <typename T, int DIR>
void launch_yzline(pp_ctx *c, Level &L, int lvl, int color, T omega) {
  LevelArgs<T> a = largs<T>(L);
  int n = DIR == 1 ? L.cy : L.cz;
  int planes = DIR == 1 ? L.cz : L.cy;
  int LS, P, TX;
  yz_shape(n, LS, P, TX, (int)sizeof(T));
  int half = (L.cx + 1) / 2;
  // small levels are latency-bound: narrow tiles give more blocks (measured
  // on 128x96x96 ... 32x24x24: 20-25% faster than 64-line tiles)
  if ((long)((half + 7) / 8) * planes <= 2048) TX = 8;
  dim3 grid((half + TX - 1) / TX, planes), block(TX, P);
  size_t sh = sizeof(T) * 8 * TX * P;
  T *e = lv_e<T>(L);
  const T *rhs = lv_rhs<T>(L);
  hipStream_t st = c->stream;
  kbegin(c, DIR == 1 ? PP_K_YLINE : PP_K_ZLINE, lvl, 0.5 * (double)L.cx * L.cy * L.cz * bytes_line(sizeof(T)));
  if constexpr (std::is_same<T, float>::value) {
    if (split_on<T>(c, lvl)) {
      // at most 768 threads per block where tiles of 16 lines allow it: with
      // the split kernel's register budget (six wavefronts per SIMD) two
      // blocks then share a CU (narrower tiles cost more in coalescing)
      if (TX > 8) {
        TX = 64;
        while (TX > 16 && TX * P > 768) TX >>= 1;
        grid = dim3((half + TX - 1) / TX, planes);
        block = dim3(TX, P);
        sh = sizeof(T) * 8 * TX * P;
      }
      LevelArgs<float> as = split_args<float>(c, L);
      float *es = c->sp.e;
      const float *rs = c->sp.rhs;
      const int hp = c->sp.hp;
      switch (LS) {
        case 1: k_yzline_split<1, DIR><<<grid, block, sh, st>>>(as, hp, rs, es, color, omega); break;
        case 2: k_yzline_split<2, DIR><<<grid, block, sh, st>>>(as, hp, rs, es, color, omega); break;
        case 4: k_yzline_split<4, DIR><<<grid, block, sh, st>>>(as, hp, rs, es, color, omega); break;
        default: k_yzline_split<8, DIR><<<grid, block, sh, st>>>(as, hp, rs, es, color, omega); break;
      }
      kend(c);
      return;
    }
  }
  // loads issued up front on the narrow-tile (latency-bound) FP32 levels
  const bool upf = sizeof(T) == 4 && TX == 8;
  switch (LS) {
    case 1: k_yzline<T, 1, DIR><<<grid, block, sh, st>>>(a, rhs, e, color, omega); break;
    case 2: k_yzline<T, 2, DIR><<<grid, block, sh, st>>>(a, rhs, e, color, omega); break;
    case 4:
      if (upf) k_yzline<T, 4, DIR, true><<<grid, block, sh, st>>>(a, rhs, e, color, omega);
      else k_yzline<T, 4, DIR><<<grid, block, sh, st>>>(a, rhs, e, color, omega);
      break;
    default:
      if (upf) k_yzline<T, 8, DIR, true><<<grid, block, sh, st>>>(a, rhs, e, color, omega);
      else k_yzline<T, 8, DIR><<<grid, block, sh, st>>>(a, rhs, e, color, omega);
      break;
  }
  kend(c);
}

// solved homogeneous ghost rows on a level-0 correction (one gather)
// (spl: on the colour-split copy, whatever e)
template <typename T>
void ghost_sweep_level0(pp_ctx *c, T *e, bool spl) {
  GhostTables &g = c->gh;
  if (g.n == 0) return;
  dim3 grid((g.n + 31) / 32), block(256);
  double nnz = (double)g.h_c_ref.size();
  kbegin(c, PP_K_GHOST, 0, g.n * (12.0 + sizeof(T)) + nnz * (8.0 + 2.0 * sizeof(T)));
  k_ghost_composite<T, T><<<grid, block, 0, c->stream>>>(g.n, spl ? g.g_as : g.g_al, g.c_ptr, spl ? g.c_as : g.c_al,
                                                          ghost_cw<T>(g), nullptr, spl ? split_e<T>(c) : e);
  kend(c);
}

template <typename T>
struct AdjPtrs;
template <>
struct AdjPtrs<float> {
  static float *lam(GhostTables &g, int b) { return g.lam32[b]; }
  static float *rhsg(GhostTables &g) { return g.rhsg32; }
  static float *tbase(GhostTables &g) { return g.tbase32; }
  static const float *nbw(GhostTables &g) { return g.nb_w32; }
  static const float *tcoef(GhostTables &g) { return g.t_coef32; }
  static const float *tvinv(GhostTables &g) { return g.t_vinv32; }
  static const float *ctw(GhostTables &g) { return g.ct_w32; }
  static const float *ctvinv(GhostTables &g) { return g.ct_vinv32; }
  static float *ctbase(GhostTables &g) { return g.ctbase32; }
};
template <>
struct AdjPtrs<double> {
  static double *lam(GhostTables &g, int b) { return g.lam64[b]; }
  static double *rhsg(GhostTables &g) { return g.rhsg64; }
  static double *tbase(GhostTables &g) { return g.tbase64; }
  static const double *nbw(GhostTables &g) { return g.nb_w64; }
  static const double *tcoef(GhostTables &g) { return g.t_coef64; }
  static const double *tvinv(GhostTables &g) { return g.t_vinv64; }
  static const double *ctw(GhostTables &g) { return g.ct_w64; }
  static const double *ctvinv(GhostTables &g) { return g.ct_vinv64; }
  static double *ctbase(GhostTables &g) { return g.ctbase64; }
};

// Transposed ghost rows on the level-0 correction mu (homogeneous: the
// ghost sources belong to the outer residual): the ghost sources from the
// fluid neighbours of each ghost, then the fluid cells of the composite rows
// get rhs = base + (M^T sources)/V: the ghost-ghost coupling solved exactly,
// the transpose of the forward smoother's composite ghost rows.
template <typename T>
void adj_ghost_level0(pp_ctx *c, T *e, T *rhs) {
  GhostTables &g = c->gh;
  if (g.n == 0) return;
  hipStream_t st = c->stream;
  // on the colour-split copy while it is in use
  const bool spl = split_on<T>(c, 0);
  const long *nb = spl ? g.nb_as : g.nb_al, *ct = spl ? g.ct_as : g.ct_al;
  if (spl) {
    e = split_e<T>(c);
    rhs = split_rhs<T>(c);
  }
  dim3 grid((g.n + 255) / 256), block(256);
  const double nnz = (double)g.h_ct_ghost.size();
  kbegin(c, PP_K_GHOST, 0, g.n * 6.0 * (8 + sizeof(T)) + g.n * sizeof(T) +
                              (double)g.nct * (8 + 8 + 2 * sizeof(T)) + nnz * (4 + 2 * sizeof(T)));
  k_adj_ghost_rhs<T><<<grid, block, 0, st>>>(g.n, nb, AdjPtrs<T>::nbw(g), nullptr, e, AdjPtrs<T>::rhsg(g));
  if (g.nct > 0)
    k_adj_composite<T><<<dim3((g.nct + 255) / 256), block, 0, st>>>(g.nct, ct, g.ct_ptr, g.ct_ghost,
                                                                    AdjPtrs<T>::ctw(g), AdjPtrs<T>::ctvinv(g),
                                                                    AdjPtrs<T>::rhsg(g), AdjPtrs<T>::ctbase(g), rhs);
  kend(c);
}

// the rhs of the composite rows' fluid cells before a transposed V-cycle
// (the base the transposed ghost coupling adds to; spl: from the
// colour-split copy)
template <typename T>
void adj_ghost_base(pp_ctx *c, bool spl = false) {
  GhostTables &g = c->gh;
  if (g.nct > 0)
    k_gather_idx<T><<<dim3((g.nct + 255) / 256), 256, 0, c->stream>>>(
        g.nct, spl ? g.ct_as : g.ct_al, spl ? split_rhs<T>(c) : lv_rhs<T>(c->lv[0]), AdjPtrs<T>::ctbase(g));
}

// ghost rows of the level-0 correction before a smoothing step (forward:
// the solved ghost rows; adjoint: the transposed ghost coupling into rhs)
template <typename T>
void ghost_step_level0(pp_ctx *c, T *e, T *rhs) {
  if (c->adjoint_mode) adj_ghost_level0<T>(c, e, rhs);
  else ghost_sweep_level0<T>(c, e, split_on<T>(c, 0));
}

// one zebra line pass set: x, y, z lines, two colours each
template <typename T>
void line_passes(pp_ctx *c, int l, T omega) {
  Level &L = c->lv[l];
  for (int col = 0; col < 2; col++) launch_xline<T>(c, L, l, col, omega);
  for (int col = 0; col < 2; col++) launch_yzline<T, 1>(c, L, l, col, omega);
  for (int col = 0; col < 2; col++) launch_yzline<T, 2>(c, L, l, col, omega);
}

// one point-Jacobi sweep src -> dst on level l
template <typename T>
void jacobi_pass(pp_ctx *c, int l, const T *src, T *dst, T omega) {
  Level &L = c->lv[l];
  LevelArgs<T> a = largs<T>(L);
  kbegin(c, PP_K_JACOBI, l, (double)L.cx * L.cy * L.cz * (3.0 * sizeof(T) + 1.0));
  if (L.cx % 4 == 0) {
    CellLaunch q = cell_launch(L.cx / 4, L.cy, L.cz);
    k_jacobi<T, 4><<<q.grid, q.block, 0, c->stream>>>(a, src, lv_rhs<T>(L), dst, omega);
  } else {
    CellLaunch q = cell_launch(L.cx, L.cy, L.cz);
    k_jacobi<T, 1><<<q.grid, q.block, 0, c->stream>>>(a, src, lv_rhs<T>(L), dst, omega);
  }
  kend(c);
}

// `steps` smoothing steps on level l (the ghost rows first on level 0 when
// ghosts is set; zero_first: the correction is zero on entry, so the first
// step's ghost rows, homogeneous and linear in it, would change nothing).
// Point Jacobi alternates between e and the level's residual buffer (free
// while smoothing) and copies back after an odd count.
template <typename T>
void smooth_steps(pp_ctx *c, int l, const pp_solve_opts *o, int steps, bool ghosts = true, bool zero_first = false) {
  Level &L = c->lv[l];
  const T om = (T)o->omega;
  if (o->smoother != PP_SMOOTH_POINT_JACOBI) {
    for (int s = 0; s < steps; s++) {
      if (l == 0 && ghosts && !(s == 0 && zero_first)) ghost_step_level0<T>(c, lv_e<T>(L), lv_rhs<T>(L));
      line_passes<T>(c, l, om);
    }
    return;
  }
  T *buf[2] = {lv_e<T>(L), lv_res<T>(L)};
  int cur = 0;
  for (int s = 0; s < steps; s++) {
    if (l == 0 && ghosts && !(s == 0 && zero_first)) ghost_step_level0<T>(c, buf[cur], lv_rhs<T>(L));
    jacobi_pass<T>(c, l, buf[cur], buf[1 - cur], om);
    cur ^= 1;
  }
  if (cur) {
    kbegin(c, PP_K_JACOBI, l, 2.0 * sizeof(T) * L.total);
    hipMemcpyAsync(buf[0], buf[1], sizeof(T) * L.total, hipMemcpyDeviceToDevice, c->stream);
    kend(c);
  }
}

// level residual res = rhs - L e (four cells per thread when rows allow)
template <typename T>
void launch_residual(pp_ctx *c, Level &L) {
  LevelArgs<T> a = largs<T>(L);
  if (L.cx % 4 == 0) {
    CellLaunch q = cell_launch(L.cx / 4, L.cy, L.cz);
    k_residual4<T><<<q.grid, q.block, 0, c->stream>>>(a, lv_e<T>(L), lv_rhs<T>(L), lv_res<T>(L));
  } else {
    CellLaunch cl = cell_launch(L.cx, L.cy, L.cz);
    k_residual<T><<<cl.grid, cl.block, 0, c->stream>>>(a, lv_e<T>(L), lv_rhs<T>(L), lv_res<T>(L));
  }
}

// the level halves in all three directions (the fused by-2 kernels apply)
bool halves_by2(const Level &L, const Level &C) {
  return C.cf[0] == 2 && C.cf[1] == 2 && C.cf[2] == 2 && L.cx == 2 * C.cx && L.cy == 2 * C.cy && L.cz == 2 * C.cz;
}

// residual of level l restricted onto level l + 1 (its rhs; its correction
// zeroed): one fused pass where the level halves in all three directions,
// else the level residual and the general restriction
template <typename T>
void restrict_step(pp_ctx *c, int l) {
  Level &L = c->lv[l], &C = c->lv[l + 1];
  hipStream_t st = c->stream;
  const double cells = (double)L.cx * L.cy * L.cz;
  LevelArgs<T> a = largs<T>(L), ac = largs<T>(C);
  CellLaunch cc = cell_launch(C.cx, C.cy, C.cz);
  if (split_on<T>(c, l)) {
    // Split0::ok guarantees the by-2 halving
    kbegin(c, PP_K_RESTRICT, l, cells * (2.0 * sizeof(T) + 1.0) + 2.0 * sizeof(T) * C.cx * C.cy * C.cz);
    k_resrestrict2<T, true><<<cc.grid, cc.block, 0, st>>>(split_args<T>(c, L), ac, split_e<T>(c), split_rhs<T>(c),
                                                          lv_rhs<T>(C), lv_e<T>(C), c->sp.hp);
    kend(c);
  } else if (halves_by2(L, C)) {
    kbegin(c, PP_K_RESTRICT, l, cells * (2.0 * sizeof(T) + 1.0) + 2.0 * sizeof(T) * C.cx * C.cy * C.cz);
    k_resrestrict2<T><<<cc.grid, cc.block, 0, st>>>(a, ac, lv_e<T>(L), lv_rhs<T>(L), lv_rhs<T>(C), lv_e<T>(C), 0);
    kend(c);
  } else {
    kbegin(c, PP_K_RESIDUAL, l, cells * (3.0 * sizeof(T) + 1.0));
    launch_residual<T>(c, L);
    kend(c);
    kbegin(c, PP_K_RESTRICT, l, cells * (sizeof(T) + 1.0) + 2.0 * sizeof(T) * C.cx * C.cy * C.cz);
    k_restrict<T><<<cc.grid, cc.block, 0, st>>>(a, ac, C.cf[0], C.cf[1], C.cf[2], lv_res<T>(L), lv_rhs<T>(C),
                                                 lv_e<T>(C));
    kend(c);
  }
}

template <typename T>
const T *nshape(pp_ctx *c);
template <>
const float *nshape<float>(pp_ctx *c) { return c->d_nshape32; }
template <>
const double *nshape<double>(pp_ctx *c) { return c->d_nshape64; }

// the correction of level l + 1 prolongated onto level l and added to its
// correction on fluid cells
template <typename T>
void prolong_step(pp_ctx *c, int l) {
  Level &L = c->lv[l], &C = c->lv[l + 1];
  hipStream_t st = c->stream;
  const double cells = (double)L.cx * L.cy * L.cz;
  LevelArgs<T> a = largs<T>(L), ac = largs<T>(C);
  // adjoint: the level-0 correction takes the shape of the null space
  // (w/V), which the coarse levels, whose null space is the constants, miss
  const T *scale = (c->adjoint_mode && l == 0 && c->have_nshape) ? nshape<T>(c) : nullptr;
  kbegin(c, PP_K_PROLONG, l, cells * (2.0 * sizeof(T) + 1.0) + (sizeof(T) + 1.0) * C.cx * C.cy * C.cz);
  if (split_on<T>(c, l)) {
    CellLaunch cc = cell_launch(C.cx, C.cy, C.cz);
    const T *sscale = scale ? (const T *)split_nshape<T>(c) : nullptr;
    k_prolong2<T, true><<<cc.grid, cc.block, 0, st>>>(split_args<T>(c, L), ac, pdir<T>(L, 0).wpar,
                                                      pdir<T>(L, 1).wpar, pdir<T>(L, 2).wpar, lv_e<T>(C),
                                                      split_e<T>(c), sscale, c->sp.hp);
  } else if (halves_by2(L, C)) {
    CellLaunch cc = cell_launch(C.cx, C.cy, C.cz);
    k_prolong2<T><<<cc.grid, cc.block, 0, st>>>(a, ac, pdir<T>(L, 0).wpar, pdir<T>(L, 1).wpar, pdir<T>(L, 2).wpar,
                                                 lv_e<T>(C), lv_e<T>(L), scale, 0);
  } else {
    CellLaunch cl = cell_launch(L.cx, L.cy, L.cz);
    k_prolong<T><<<cl.grid, cl.block, 0, st>>>(a, ac, pdir<T>(L, 0), pdir<T>(L, 1), pdir<T>(L, 2), lv_e<T>(C),
                                                lv_e<T>(L), scale);
  }
  kend(c);
}

template <typename T>
void vcycle(pp_ctx *c, int l, const pp_solve_opts *o) {
  Level &L = c->lv[l];
  hipStream_t st = c->stream;
  if (l == (int)c->lv.size() - 1) {
    const CoarseSolve &cs = c->cs;
    int n = cs.n;
    // the dense FP64 solution operator in both precisions (the coarsest
    // solve is one of the FP64 parts of the mixed scheme)
    kbegin(c, PP_K_COARSE, l, 8.0 * n * n + 2.0 * sizeof(T) * n);
    // rows round-robin over kCoarseBlocks blocks of enough wavefronts for
    // one row each (up to kCoarseMaxWaves; beyond that each loops)
    const int nb = n < kCoarseBlocks ? n : kCoarseBlocks;
    int nw = (n + nb - 1) / nb;
    if (nw > kCoarseMaxWaves) nw = kCoarseMaxWaves;
    k_coarse_mv<T, double><<<nb, 64 * nw, sizeof(double) * n, st>>>(n, cs.cell, cs.G, lv_rhs<T>(L), lv_e<T>(L));
    kend(c);
    return;
  }
  const int npre = (l > 0 && o->nu_coarse > 0) ? o->nu_coarse : o->nu_pre;
  const int npost = (l > 0 && o->nu_coarse > 0) ? o->nu_coarse : o->nu_post;
  // the first cycle of a preconditioner application starts from a zero
  // level-0 correction (every caller zeroes it)
  const bool zero = l == 0 && c->e0_zero;
  if (l == 0) c->e0_zero = false;
  smooth_steps<T>(c, l, o, npre, true, zero);
  restrict_step<T>(c, l);
  vcycle<T>(c, l + 1, o);
  prolong_step<T>(c, l);
  smooth_steps<T>(c, l, o, npost);
}

// FP32 line-smoothed cycles run level 0 on its colour-split copy.
template <typename T>
bool split_cycles(const pp_ctx *c, const pp_solve_opts *o) {
  return std::is_same<T, float>::value && c->sp.ok && o->smoother == PP_SMOOTH_LINE;
}
void split_in(pp_ctx *c) {
  Level &L0 = c->lv[0];
  const CellLaunch q = cell_launch((L0.cx + 1) / 2, L0.cy, L0.cz);
  kbegin(c, PP_K_OTHER, 0, (double)L0.cx * L0.cy * L0.cz * 12.0);
  k_split_in<<<q.grid, q.block, 0, c->stream>>>(largs<float>(L0), split_args<float>(c, L0), c->sp.hp, L0.rhs32,
                                                c->sp.rhs, c->sp.e);
  kend(c);
}
void split_out(pp_ctx *c) {
  Level &L0 = c->lv[0];
  const CellLaunch q = cell_launch((L0.cx + 1) / 2, L0.cy, L0.cz);
  kbegin(c, PP_K_OTHER, 0, (double)L0.cx * L0.cy * L0.cz * 8.0);
  k_split_out<<<q.grid, q.block, 0, c->stream>>>(largs<float>(L0), split_args<float>(c, L0), c->sp.hp, c->sp.e,
                                                 L0.e32);
  kend(c);
}
// the preconditioner is captured into a hipGraph
bool graph_path(const pp_ctx *c, const pp_solve_opts *o) { return o->use_graph && !c->prof_on && !c->adjoint_mode; }

// The cycles of one preconditioner application.  On the colour-split level
// 0 the rhs goes in (and the split correction starts at zero) unless the
// caller has written both split copies already (in_split), and the
// correction comes back out into e0 unless the caller reads the split copy
// (out_split).
template <typename T>
void precondition_cycles(pp_ctx *c, const pp_solve_opts *o, bool in_split = false, bool out_split = false) {
  const bool spl = split_cycles<T>(c, o);
  if (spl) {
    if (!in_split) split_in(c);
    c->sp.on = true;
  }
  c->e0_zero = true;
  for (int cyc = 0; cyc < o->cycles_per_outer; cyc++) vcycle<T>(c, 0, o);
  c->e0_zero = false;
  if (spl) {
    c->sp.on = false;
    if (!out_split) split_out(c);
  }
}

// The transposed cycles of one adjoint preconditioner application, on the
// colour-split level 0 where it applies (in_split / out_split as for
// precondition_cycles).
template <typename T>
void adjoint_cycles(pp_ctx *c, const pp_solve_opts *o, pp_solve_stats &S, bool in_split = false,
                    bool out_split = false) {
  const bool spl = split_cycles<T>(c, o);
  if (spl && !in_split) split_in(c);
  adj_ghost_base<T>(c, spl);
  c->adjoint_mode = true;
  c->sp.on = spl;
  c->e0_zero = true;
  for (int cyc = 0; cyc < o->cycles_per_outer; cyc++) {
    vcycle<T>(c, 0, o);
    S.vcycles++;
  }
  c->e0_zero = false;
  c->sp.on = false;
  c->adjoint_mode = false;
  if (spl && !out_split) split_out(c);
}

// The preconditioner: cycles_per_outer V-cycles on the level-0 rhs (e0 = 0
// on entry).  With use_graph the launches of one application are captured
// once into a hipGraph (per precision and cycle shape; dropped when the
// geometry changes) and replayed.
// in_split / out_split as for precondition_cycles (eager launches only).
template <typename T>
int precondition(pp_ctx *c, const pp_solve_opts *o, pp_solve_stats &S, bool in_split = false,
                 bool out_split = false) {
  bool graph = graph_path(c, o);
  if (!graph) {
    precondition_cycles<T>(c, o, in_split, out_split);
    S.vcycles += o->cycles_per_outer;
    return PP_OK;
  }
  long key = (long)sizeof(T) | ((long)o->nu_pre << 4) | ((long)o->nu_post << 12) | ((long)std::max(o->nu_coarse, 0) << 8) |
             ((long)o->cycles_per_outer << 20) | ((long)std::lround(o->omega * 1e6) << 28) |
             ((long)(o->smoother & 3) << 56);
  if (!c->gexec || c->graph_key != key) {
    if (c->gexec) {
      HIPCHK(hipGraphExecDestroy(c->gexec));
      c->gexec = nullptr;
    }
    double b0 = c->bytes_acc;
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    precondition_cycles<T>(c, o);
    HIPCHK(hipStreamEndCapture(c->stream, &gr));
    hipError_t ie = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (ie != hipSuccess) {
      c->gexec = nullptr;
      set_error("hipGraphInstantiate failed: %s", hipGetErrorString(ie));
      return PP_HIP_ERROR;
    }
    c->graph_key = key;
    c->graph_bytes = c->bytes_acc - b0;
  } else {
    c->bytes_acc += c->graph_bytes;
  }
  HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  S.vcycles += o->cycles_per_outer;
  return PP_OK;
}

// ---------------------------------------------------------------------
// FP64 helpers on the reference layout
// ---------------------------------------------------------------------
CellLaunch ref_launch(const pp_ctx *c) { return cell_launch(c->nxc, c->nyc, c->nzc); }

// two-stage reductions of per-workgroup partials (kRedStage workgroups,
// then one): fixed order, so deterministic
constexpr int kRedStage = 256;
void launch_reduce_summax(pp_ctx *c, int nblk) {
  int chunk = (nblk + kRedStage - 1) / kRedStage;
  k_reduce_summax_stage<<<kRedStage, 256, 0, c->stream>>>(c->d_part, nblk, chunk, c->d_part2);
  k_reduce_summax<<<1, 256, 0, c->stream>>>(c->d_part2, kRedStage, c->d_res);
}
void launch_reduce_sums(pp_ctx *c, int nblk, int nq) {
  int chunk = (nblk + kRedStage - 1) / kRedStage;
  k_reduce_sums_stage<<<kRedStage, 256, 0, c->stream>>>(c->d_part, nblk, nq, chunk, c->d_part2);
  k_reduce_sums<<<1, 256, 0, c->stream>>>(c->d_part2, kRedStage, nq, c->d_res);
}

int reduce_summax(pp_ctx *c, int nblk) {
  launch_reduce_summax(c, nblk);
  HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

// exact ghost rows + outer ghost layer (unless outer is false) on an FP64
// reference-layout field; the composite rows read a stencil corner in the
// outer layer from the adjacent interior cell (its Neumann copy)
void ghost_exact_ref(pp_ctx *c, double *p, bool with_bnd, bool outer = true) {
  GhostTables &g = c->gh;
  hipStream_t st = c->stream;
  if (g.n > 0) {
    dim3 grid((g.n + 31) / 32), block(256);
    double nnz = (double)g.h_c_ref.size();
    kbegin(c, PP_K_GHOST, 0, g.n * (12.0 + 8.0 + (with_bnd ? 8.0 : 0.0)) + nnz * 24.0);
    k_ghost_composite<double, double><<<grid, block, 0, st>>>(g.n, g.g_ref, g.c_ptr, g.c_ref, g.c_w64,
                                                             with_bnd ? g.bnd_eff : nullptr, p);
    kend(c);
  }
  if (!outer) return;
  k_outer_ghost_x<<<dim3((c->nyc + 2 + 255) / 256, c->nzc + 2), 256, 0, st>>>(c->ref, p);
  k_outer_ghost_y<<<dim3((c->nxc + 2 + 255) / 256, c->nzc + 2), 256, 0, st>>>(c->ref, p);
  k_outer_ghost_z<<<dim3((c->nxc + 2 + 255) / 256, c->nyc + 2), 256, 0, st>>>(c->ref, p);
}

// Neumann data of the ghost rows from bc_src (host values): bnd = bself *
// bc_src for the compatibility projection and bnd_eff = (I - C_gg)^-1 bnd
// (kexact Jacobi sweeps, as the composite rows) for the solved ghost values
int set_ghost_bnd(pp_ctx *c, const std::vector<double> &bc) {
  GhostTables &g = c->gh;
  std::vector<double> hb(g.n), x(g.n), y(g.n);
  for (int q = 0; q < g.n; q++) hb[q] = g.h_bself[q] * bc[q];
  x = hb;
  for (int t = 1; t < g.kexact && !g.h_gg_idx.empty(); t++) {
    for (int q = 0; q < g.n; q++) {
      double s = hb[q];
      for (int e = g.h_gg_ptr[q]; e < g.h_gg_ptr[q + 1]; e++) s += g.h_gg_w[e] * x[g.h_gg_idx[e]];
      y[q] = s;
    }
    x.swap(y);
  }
  HIPCHK(hipMemcpyAsync(g.bnd, hb.data(), sizeof(double) * g.n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(g.bnd_eff, x.data(), sizeof(double) * g.n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

// fluid sums: out[0..4] = {sum w*b, sum w, sum b*b, sum V*p, sum V}
int fluid_sums(pp_ctx *c, const double *w, const double *b, const double *p) {
  constexpr int CPT = 4;
  CellLaunch cl = cell_launch((c->nxc + CPT - 1) / CPT, c->nyc, c->nzc);
  long nblk = (long)cl.grid.x * cl.grid.y * cl.grid.z;
  k_fluid_sums<256, CPT><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), w, b, p, c->d_part);
  launch_reduce_sums(c, (int)nblk, 5);
  return PP_OK;
}

// subtracting_average_pressure (CORE_POISSON.F90:419-490); with fresh =
// false the mean subtraction alone (the fresh-cell fill belongs to the
// initial guess only: the reference calls it once, before iterating, :535)
int mean_remove_ref(pp_ctx *c, double *p, double *mean_out, bool fresh = true) {
  CellLaunch cl = ref_launch(c);
  fluid_sums(c, nullptr, nullptr, p);
  k_mean_sub<<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), p, c->d_res);
  if (fresh && !c->h_fresh.empty()) {
    // fresh cells read the mean-removed snapshot
    HIPCHK(hipMemcpyAsync(c->d_tmp, p, sizeof(double) * c->nref, hipMemcpyDeviceToDevice, c->stream));
    k_fresh_fill<<<cl.grid, cl.block, 0, c->stream>>>(c->ref, c->d_iblank_ref, c->d_fresh, c->d_tmp, p);
  }
  if (mean_out) {
    HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *mean_out = c->h_res[3] / c->h_res[4];
  }
  return PP_OK;
}

bool opts_ok(const pp_solve_opts *o) {
  return o && (o->precision == PP_PREC_FP64 || o->precision == PP_PREC_MIXED) && o->max_outer >= 0 &&
         o->max_outer <= 10000 && o->cycles_per_outer >= 1 && o->nu_pre >= 0 && o->nu_post >= 0 &&
         o->omega > 0.0 && o->omega < 2.0 && o->krylov >= 0 && o->krylov <= KRY_MAX &&
         (o->smoother == PP_SMOOTH_LINE || o->smoother == PP_SMOOTH_POINT_JACOBI ||
          o->smoother == PP_SMOOTH_LINE_NATURAL);
}

// ---------------------------------------------------------------------
// flexible GMRES(m) on the fluid cells (FP64, reference layout)
// ---------------------------------------------------------------------
int ensure_krylov(pp_ctx *c, int m) {
  if (c->kry_m >= m) return PP_OK;
  for (double *q : c->d_kry) hipFree(q);
  c->d_kry.clear();
  c->kry_m = 0;
  size_t nb = sizeof(double) * c->nref;
  for (int q = 0; q < 2 * m + 1; q++) {
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, nb));
    c->d_kry.push_back(d);
    HIPCHK(hipMemset(d, 0, nb));
  }
  c->kry_m = m;
  return PP_OK;
}

int ensure_krylov_level(pp_ctx *c, int m) {
  if (c->kryl_m >= m) return PP_OK;
  for (float *q : c->d_kryl) hipFree(q);
  c->d_kryl.clear();
  c->kryl_m = 0;
  size_t nb = sizeof(float) * (size_t)c->lv[0].total;
  for (int q = 0; q < 2 * m + 1; q++) {
    float *d = nullptr;
    HIPCHK(hipMalloc(&d, nb));
    c->d_kryl.push_back(d);
    HIPCHK(hipMemset(d, 0, nb));
  }
  c->kryl_m = m;
  return PP_OK;
}

// sum of nq per-block partials -> c->h_res[0..nq)
int reduce_sums_host(pp_ctx *c, int nq, int nblk = -1) {
  launch_reduce_sums(c, nblk >= 0 ? nblk : c->nblk_red, nq);
  HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, sizeof(double) * nq, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

// One restart cycle.  On entry d_kry[0] holds r = b - A p (fluid cells) and
// the level-0 rhs holds r in T with e0 = 0; beta = |r|.  Each step applies
// the V-cycle(s) to the current basis vector (right preconditioning), keeps
// the preconditioned vector z_j (flexible variant: the V-cycle in FP32 is
// not exactly linear), orthogonalises A z_j by classical Gram-Schmidt and
// updates the Givens QR of the Hessenberg matrix.  Stops at m steps, at the
// iteration limit or when the least-squares residual meets the tolerance;
// then p += Z y and the ghost rows are made exact.
// ---------------------------------------------------------------------
// Flexible GMRES restart cycle, generic over the operator/storage policy
// Ops (forward mixed, forward FP64, adjoint):
//   Ops::V / Ops::E       vector-list type and element type of the vectors
//   u(q), z(q)            basis vector q (m+1) and preconditioned vector q (m)
//   precondition()        V-cycle(s) on the level-0 rhs (holds basis j)
//   apply(j, V)           z_j from the V-cycle's correction, u_{j+1} = Op z_j,
//                         partial dots <u_{j+1}, u_q> (q <= j)
//   orth(j, V)            u_{j+1} -= sum c_q u_q, its norm, level-0 rhs = u_{j+1}
//   update(V)             solution += sum c_q z_q (+ what follows an update)
// On entry u(0) holds the residual r (fluid cells), the level-0 rhs holds r
// and e0 = 0; beta = |r|.  The basis is kept unnormalised (v_q = sigma_q u_q)
// and orthogonalised by classical Gram-Schmidt; the Hessenberg matrix is
// reduced by Givens rotations on the host.  Stops at m steps, at the
// iteration limit or when the least-squares residual meets the tolerance.
// ---------------------------------------------------------------------
template <typename Ops>
int gmres_cycle(pp_ctx *c, const pp_solve_opts *o, Ops &ops, double beta, double bden, int &outer,
                pp_solve_stats &S) {
  const int m = o->krylov;
  std::vector<double> H((m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0), sig(m + 1, 0.0);
  auto h = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
  sig[0] = 1.0 / beta;
  gv[0] = beta;
  int k = 0;
  for (int j = 0; j < m && outer < o->max_outer; j++) {
    int rc = ops.precondition();
    if (rc) return rc;
    typename Ops::V V{};
    V.n = j + 1;
    for (int q = 0; q <= j; q++) V.u[q] = ops.u(q);
    ops.apply(j, V);
    rc = reduce_sums_host(c, j + 1, ops.blocks());
    if (rc) return rc;
    for (int q = 0; q <= j; q++) {
      double d = c->h_res[q];
      h(q, j) = sig[j] * sig[q] * d;
      V.c[q] = sig[q] * sig[q] * d;
    }
    ops.orth(j, V);
    rc = reduce_sums_host(c, 1, ops.blocks());
    if (rc) return rc;
    double hn = std::sqrt(c->h_res[0]);
    h(j + 1, j) = sig[j] * hn;
    sig[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
    for (int i = 0; i < j; i++) {
      double t = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
      h(i + 1, j) = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
      h(i, j) = t;
    }
    double rho = std::hypot(h(j, j), h(j + 1, j));
    if (!std::isfinite(rho)) {
      set_error("non-finite Krylov coefficient after %d iterations", outer);
      return PP_NONFINITE;
    }
    if (rho == 0.0) break;  // Op z_j adds nothing: keep the first j columns
    cs[j] = h(j, j) / rho;
    sn[j] = h(j + 1, j) / rho;
    h(j, j) = rho;
    h(j + 1, j) = 0.0;
    gv[j + 1] = -sn[j] * gv[j];
    gv[j] = cs[j] * gv[j];
    k = j + 1;
    outer++;
    double est = std::fabs(gv[j + 1]);
    if (outer <= 64) S.history[outer - 1] = est / bden;
    bool done = (o->rel_tol > 0.0 && est <= o->rel_tol * bden) || (o->abs_linf_tol > 0.0 && est <= o->abs_linf_tol);
    if (done || hn == 0.0) break;
  }
  if (k == 0) {
    set_error("Krylov breakdown: the preconditioned residual has no component in the range");
    return PP_NOT_CONVERGED;
  }
  std::vector<double> y(k);
  for (int i = k - 1; i >= 0; i--) {
    double t = gv[i];
    for (int q = i + 1; q < k; q++) t -= h(i, q) * y[q];
    y[i] = t / h(i, i);
  }
  typename Ops::V V{};
  V.n = k;
  for (int q = 0; q < k; q++) {
    V.u[q] = ops.z(q);
    V.c[q] = y[q] * sig[q];
  }
  ops.update(V);
  return PP_OK;
}

// forward operator, FP64 vectors on the reference layout (FP64 V-cycles)
struct FwdOps64 {
  using V = KryVecs;
  pp_ctx *c;
  const pp_solve_opts *o;
  pp_solve_stats *S;
  double *p;
  bool have_bnd;
  double cells;
  double *u(int q) { return c->d_kry[q]; }
  double *z(int q) { return c->d_kry[c->kry_m + 1 + q]; }
  int blocks() const { return c->nblk_red; }
  int precondition() { return ::precondition<double>(c, o, *S); }
  void apply(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = ref_launch(c);
    LevelArgs<double> a0 = largs<double>(L0);
    kbegin(c, PP_K_KRYLOV, 0, cells * (8.0 + 1.0 + 8.0));
    k_kry_z<double><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a0, lv_e<double>(L0), z(j));
    kend(c);
    ghost_exact_ref(c, z(j), false);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 8.0 * (j + 1)));
    k_kry_apply<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a0, z(j), u(j + 1), vv, c->d_part, nullptr);
    kend(c);
  }
  void orth(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = ref_launch(c);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 8.0 * (j + 1) + 16.0));
    k_kry_orth<double, 256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(L0), u(j + 1), vv,
                                                                  lv_rhs<double>(L0), lv_e<double>(L0), c->d_part);
    kend(c);
  }
  void update(const V &vv) {
    CellLaunch cl = ref_launch(c);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 8.0 * vv.n));
    k_kry_update<<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), p, vv);
    kend(c);
    ghost_exact_ref(c, p, have_bnd);
  }
};

// forward operator, FP32 vectors on the level-0 layout (FP32 V-cycles):
// products, dots and coefficients in FP64 on the stored values
struct FwdOps32 {
  using V = KryVecsF;
  pp_ctx *c;
  const pp_solve_opts *o;
  pp_solve_stats *S;
  double *p;
  bool have_bnd;
  double cells;
  bool rhs_split = false;  // the last orth wrote the split rhs (and zeroed the split correction)
  float *u(int q) { return c->d_kryl[q]; }
  float *z(int q) { return c->d_kryl[c->kryl_m + 1 + q]; }
  // four cells per thread where the level-0 rows allow it
  bool by4() const { return c->lv[0].cx % 4 == 0; }
  // the preconditioner's input and output stay in the colour-split copy:
  // orth writes it, apply reads it (eager cycles, four-cell kernels)
  bool sio() const { return by4() && split_cycles<float>(c, o) && !graph_path(c, o); }
  int precondition() { return ::precondition<float>(c, o, *S, sio() && rhs_split, sio()); }
  CellLaunch launch() const {
    const Level &L0 = c->lv[0];
    return by4() ? cell_launch(L0.cx / 4, L0.cy, L0.cz) : ref_launch(c);
  }
  int blocks() const {
    CellLaunch cl = launch();
    return (int)((long)cl.grid.x * cl.grid.y * cl.grid.z);
  }
  void apply(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = launch();
    const bool spl = sio();
    ghost_sweep_level0<float>(c, lv_e<float>(L0), spl);
    kbegin(c, PP_K_KRYLOV, 0, cells * (4.0 + 1.0 + 4.0 + 4.0 + 4.0 * (j + 1)));
    if (spl)
      k_kryl_apply4<256, true><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), c->sp.e, z(j), u(j + 1), vv,
                                                                    c->d_part, c->sp.sx, c->sp.sxy, c->sp.hp);
    else if (by4())
      k_kryl_apply4<256><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), lv_e<float>(L0), z(j), u(j + 1),
                                                              vv, c->d_part, 0, 0, 0);
    else
      k_kryl_apply<256><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), lv_e<float>(L0), z(j), u(j + 1), vv,
                                                             c->d_part);
    kend(c);
  }
  void orth(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = launch();
    kbegin(c, PP_K_KRYLOV, 0, cells * (1.0 + 8.0 + 4.0 * (j + 1) + 8.0));
    rhs_split = sio();
    if (rhs_split)
      k_kryl_orth4<256, true><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), u(j + 1), vv, c->sp.rhs,
                                                                   c->sp.e, c->d_part, c->sp.sx, c->sp.sxy,
                                                                   c->sp.hp);
    else if (by4())
      k_kryl_orth4<256><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), u(j + 1), vv, lv_rhs<float>(L0),
                                                             lv_e<float>(L0), c->d_part, 0, 0, 0);
    else
      k_kryl_orth<256><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), u(j + 1), vv, lv_rhs<float>(L0),
                                                            lv_e<float>(L0), c->d_part);
    kend(c);
  }
  void update(const V &vv) {
    CellLaunch cl = ref_launch(c);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 4.0 * vv.n));
    k_kryl_update<<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), p, vv);
    kend(c);
    ghost_exact_ref(c, p, have_bnd);
  }
};

template <typename T>
struct FwdOpsFor;
template <>
struct FwdOpsFor<double> { using type = FwdOps64; };
template <>
struct FwdOpsFor<float> { using type = FwdOps32; };

template <typename T>
int krylov_cycle(pp_ctx *c, const pp_solve_opts *o, double *p, bool have_bnd, double beta, double bden,
                 int &outer, pp_solve_stats &S) {
  typename FwdOpsFor<T>::type ops{c, o, &S, p, have_bnd, (double)c->nxc * c->nyc * c->nzc};
  return gmres_cycle(c, o, ops, beta, bden, outer, S);
}

// ---------------------------------------------------------------------
// the solve (FP64 iterative refinement or flexible GMRES around
// T-precision V-cycles)
// ---------------------------------------------------------------------
template <typename T>
int solve_impl(pp_ctx *c, double *p, const double *rhs, const double *bc_src, const pp_solve_opts *o,
               pp_solve_stats *st) {
  hipStream_t s = c->stream;
  Level &L0 = c->lv[0];
  LevelArgs<double> a0 = largs<double>(L0);
  CellLaunch cl = ref_launch(c);
  long nblk = (long)cl.grid.x * cl.grid.y * cl.grid.z;
  pp_solve_stats S;
  memset(&S, 0, sizeof(S));
  kreset(c, o->profile != 0);
  hipEvent_t ev0, ev1;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventRecord(ev0, s));
  double cells = (double)c->nxc * c->nyc * c->nzc;
  const int m = o->krylov;
  const bool kry64 = m > 0 && sizeof(T) == 8, kry32 = m > 0 && sizeof(T) == 4;
  if (kry64) {
    int rc = ensure_krylov(c, m);
    if (rc) return rc;
  }
  if (kry32) {
    int rc = ensure_krylov_level(c, m);
    if (rc) return rc;
  }
  // ghost Neumann data: bnd = (c_self + probeLength) * uNormal
  GhostTables &g = c->gh;
  bool have_bnd = bc_src != nullptr && g.n > 0;
  if (g.n > 0) {
    if (have_bnd) {
      std::vector<double> hs(g.n);
      HIPCHK(hipMemcpy(hs.data(), bc_src, sizeof(double) * g.n, hipMemcpyDeviceToHost));
      int rc = set_ghost_bnd(c, hs);
      if (rc) return rc;
    }
  }
  // b = rhs on fluid cells
  k_mask_rhs<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, rhs, c->d_b);
  // compatibility projection: b_f -= (w.b + wg.bnd) / sum_f w
  if (o->project_rhs) {
    fluid_sums(c, c->d_w, c->d_b, nullptr);
    HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double num = c->h_res[0], den = c->h_res[1];
    if (have_bnd) {
      // ghost-row part of w^T b
      std::vector<double> hw(g.n), hb(g.n);
      k_gather_idx<double><<<dim3((g.n + 255) / 256), 256, 0, s>>>(g.n, g.g_ref, c->d_w, c->d_tmp);
      HIPCHK(hipStreamSynchronize(s));
      HIPCHK(hipMemcpy(hw.data(), c->d_tmp, sizeof(double) * g.n, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hb.data(), g.bnd, sizeof(double) * g.n, hipMemcpyDeviceToHost));
      for (int q = 0; q < g.n; q++) num += hw[q] * hb[q];
    }
    double shift = num / den;
    S.rhs_shift = shift;
    c->h_res[0] = shift;
    c->h_res[1] = 1.0;
    HIPCHK(hipMemcpyAsync(c->d_res + 8, c->h_res, 2 * sizeof(double), hipMemcpyHostToDevice, s));
    k_fluid_shift<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_b, c->d_res + 8, c->d_res + 9, -1.0);
  }
  fluid_sums(c, nullptr, c->d_b, nullptr);
  HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double bnorm = std::sqrt(c->h_res[2]);
  S.b_l2 = bnorm;
  if (!std::isfinite(bnorm)) {
    set_error("non-finite right-hand side");
    S.status = PP_NONFINITE;
    if (st) *st = S;
    return PP_NONFINITE;
  }
  // initial guess: subtracting_average_pressure (CORE_POISSON.F90:535), then
  // ghost rows and outer layer (GCM_correct_pressure semantics)
  mean_remove_ref(c, p, nullptr);
  ghost_exact_ref(c, p, have_bnd);
  double bden = bnorm > 0.0 ? bnorm : 1.0;
  int status = PP_NOT_CONVERGED;
  int outer = 0;
  for (;;) {
    // r = b - A p -> level-0 rhs (T), e0 = 0, norms
    kbegin(c, PP_K_OUTER_RESIDUAL, 0, cells * (17.0 + 2.0 * sizeof(T)));
    k_outer_residual<T, 256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, p, c->d_b, lv_rhs<T>(L0), lv_e<T>(L0),
                                                          c->d_part, kry64 ? c->d_kry[0] : nullptr,
                                                          kry32 ? (T *)c->d_kryl[0] : (T *)nullptr);
    kend(c);
    int rc = reduce_summax(c, (int)nblk);
    if (rc) return rc;
    double rl2 = std::sqrt(c->h_res[0]);
    double rinf = c->h_res[1];
    long argr = (long)c->h_res[2];
    if (!std::isfinite(rl2)) {
      set_error("non-finite residual after %d outer iterations", outer);
      status = PP_NONFINITE;
      break;
    }
    S.rel_l2 = rl2 / bden;
    S.abs_linf = rinf;
    if (rinf > 0.0 && argr >= 0) {
      long q = argr;
      S.loc[0] = (int)(q % c->ref.ex);
      S.loc[1] = (int)((q / c->ref.ex) % (c->nyc + 3));
      S.loc[2] = (int)(q / c->ref.exy);
    } else {
      S.loc[0] = S.loc[1] = S.loc[2] = 0;
    }
    if (outer > 0 && outer <= 64) S.history[outer - 1] = S.rel_l2;
    bool conv = (o->rel_tol > 0.0 && rl2 <= o->rel_tol * bden) || (o->abs_linf_tol > 0.0 && rinf <= o->abs_linf_tol) ||
                bnorm == 0.0;
    if (conv) { status = PP_OK; break; }
    if (outer >= o->max_outer) break;
    if (m > 0) {
      int krc = krylov_cycle<T>(c, o, p, have_bnd, rl2, bden, outer, S);
      if (krc == PP_NOT_CONVERGED || krc == PP_NONFINITE) { status = krc; break; }
      if (krc) return krc;
      continue;
    }
    int prc = precondition<T>(c, o, S);
    if (prc) return prc;
    // p += e on fluid cells, ghost rows exact
    kbegin(c, PP_K_UPDATE, 0, cells * (17.0 + sizeof(T)));
    k_update<T><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, p, lv_e<T>(L0));
    kend(c);
    ghost_exact_ref(c, p, have_bnd);
    outer++;
  }
  if (o->mean_remove) {
    // zero fluid-volume mean of the solution; fresh cells keep their solved
    // values (the reference fills them on the initial guess only)
    double mean = 0.0;
    mean_remove_ref(c, p, &mean, false);
    S.mean_removed = mean;
    ghost_exact_ref(c, p, have_bnd);
  }
  HIPCHK(hipEventRecord(ev1, s));
  HIPCHK(hipEventSynchronize(ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
  hipEventDestroy(ev0);
  hipEventDestroy(ev1);
  HIPCHK(hipGetLastError());
  S.ms = ms;
  S.bytes = c->bytes_acc;
  S.outer_iterations = outer;
  S.status = status;
  kcollect(c, S);
  if (st) *st = S;
  return status;
}

// ---------------------------------------------------------------------
// the transposed solve A^T lambda = s in mu = lambda/V on fluid cells,
// lambda_g on ghosts: FP64 refinement around T-precision V-cycles.
// d_mu (reference layout) holds the initial guess and the result (ghost
// and solid entries stay 0); d_sv = s_f/V_f; sg = s_g (device, may be
// null).  ref_norm < 0: converge relative to the first residual.
// On return the exact lambda_g of the final mu is in gh.lam64[*lam_idx].
// ---------------------------------------------------------------------
// transposed ghost coupling of a correction z (fluid cells): lambda_g from
// the fluid neighbours of each ghost (solved ghost-ghost rows), then
// out = sum_g c lambda_g / V on the touched fluid cells, 0 elsewhere
void adjoint_ghost_coupling(pp_ctx *c, const double *z, double *out) {
  GhostTables &g = c->gh;
  hipStream_t s = c->stream;
  hipMemsetAsync(out, 0, sizeof(double) * c->nref, s);
  if (g.n == 0) return;
  dim3 gg((g.n + 255) / 256), blk(256);
  k_adj_ghost_rhs<double><<<gg, blk, 0, s>>>(g.n, g.nb_ref, g.nb_w64, nullptr, z, g.orhsg64);
  if (g.nct > 0)
    k_adj_composite<double><<<dim3((g.nct + 255) / 256), blk, 0, s>>>(g.nct, g.ct_ref, g.ct_ptr, g.ct_ghost,
                                                                      g.ct_w64, g.ct_vinv64, g.orhsg64, nullptr,
                                                                      out);
}

// adjoint operator B mu = A mu - (transposed ghost coupling of mu) on the
// fluid cells (mu = lambda/V), FP64 vectors on the reference layout,
// preconditioned by the transposed V-cycle in precision T
template <typename T>
struct AdjOps {
  using V = KryVecs;
  pp_ctx *c;
  const pp_solve_opts *o;
  pp_solve_stats *S;
  double cells;
  double *u(int q) { return c->d_kry[q]; }
  double *z(int q) { return c->d_kry[c->kry_m + 1 + q]; }
  int blocks() const { return c->nblk_red; }
  int precondition() {
    adjoint_cycles<T>(c, o, *S);
    return PP_OK;
  }
  void apply(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = ref_launch(c);
    LevelArgs<double> a0 = largs<double>(L0);
    kbegin(c, PP_K_KRYLOV, 0, cells * (sizeof(T) + 1.0 + 8.0));
    k_kry_z<T><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a0, lv_e<T>(L0), z(j));
    kend(c);
    adjoint_ghost_coupling(c, z(j), c->d_tmp);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 8.0 * (j + 1)));
    k_kry_apply<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a0, z(j), u(j + 1), vv, c->d_part, c->d_tmp);
    kend(c);
  }
  void orth(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = ref_launch(c);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 8.0 * (j + 1) + 2.0 * sizeof(T)));
    k_kry_orth<T, 256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(L0), u(j + 1), vv,
                                                             lv_rhs<T>(L0), lv_e<T>(L0), c->d_part);
    kend(c);
  }
  void update(const V &vv) {
    CellLaunch cl = ref_launch(c);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 8.0 * vv.n));
    k_kry_update<<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), c->d_mu, vv);
    kend(c);
  }
};

// The same operator on the FP32 level-0 basis of the mixed forward solve
// (four cells per thread): w = A z by k_kryl_apply4 (z = e, which is 0 on
// blank cells in adjoint mode), then the transposed ghost coupling of z on
// its touched cells with the matching dot corrections.  For FP32
// V-cycles on levels whose rows split into fours.
struct AdjOps32 {
  using V = KryVecsF;
  pp_ctx *c;
  const pp_solve_opts *o;
  pp_solve_stats *S;
  double cells;
  float *u(int q) { return c->d_kryl[q]; }
  float *z(int q) { return c->d_kryl[c->kryl_m + 1 + q]; }
  CellLaunch launch() const {
    const Level &L0 = c->lv[0];
    return cell_launch(L0.cx / 4, L0.cy, L0.cz);
  }
  long apply_blocks() const {
    CellLaunch cl = launch();
    return (long)cl.grid.x * cl.grid.y * cl.grid.z;
  }
  long coupling_blocks() const { return (c->gh.nct + 255) / 256; }
  int blocks() const { return (int)(apply_blocks() + coupling_blocks()); }
  // the dots of both launches fit the partial-sum buffer
  static bool fits(pp_ctx *c) {
    if (c->lv[0].cx % 4 != 0) return false;
    AdjOps32 t{c, nullptr, nullptr, 0.0};
    return t.blocks() <= c->nblk_red;
  }
  // the preconditioner's input and output stay in the colour-split copy:
  // orth writes it, apply reads it
  bool rhs_split = false;
  bool sio() const { return split_cycles<float>(c, o); }
  int precondition() {
    adjoint_cycles<float>(c, o, *S, sio() && rhs_split, sio());
    return PP_OK;
  }
  void apply(int j, const V &vv) {
    Level &L0 = c->lv[0];
    GhostTables &g = c->gh;
    CellLaunch cl = launch();
    const bool spl = sio();
    kbegin(c, PP_K_KRYLOV, 0, cells * (4.0 + 1.0 + 4.0 + 4.0 + 4.0 * (j + 1)));
    if (spl)
      k_kryl_apply4<256, true><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), c->sp.e, z(j), u(j + 1), vv,
                                                                    c->d_part, c->sp.sx, c->sp.sxy, c->sp.hp);
    else
      k_kryl_apply4<256><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), lv_e<float>(L0), z(j), u(j + 1),
                                                              vv, c->d_part, 0, 0, 0);
    kend(c);
    // with no touched cells the coupling launch still writes its (zero) partial block
    kbegin(c, PP_K_GHOST, 0, (double)g.n * 6.0 * 12.0 + (double)g.nct * (8.0 + 8.0 + 4.0 * (j + 2)));
    if (g.n > 0)
      k_adj_ghost_rhs<float><<<dim3((g.n + 255) / 256), 256, 0, c->stream>>>(
          g.n, spl ? g.nb_as : g.nb_al, g.nb_w32, nullptr, spl ? c->sp.e : lv_e<float>(L0), g.rhsg32);
    k_kryl_adj_coupling<256><<<dim3((unsigned)std::max<long>(coupling_blocks(), 1)), 256, 0, c->stream>>>(
        g.nct, g.ct_al, g.ct_ptr, g.ct_ghost, g.ct_w32, g.ct_vinv32, g.rhsg32, u(j + 1), vv, c->d_part,
        apply_blocks());
    kend(c);
  }
  void orth(int j, const V &vv) {
    Level &L0 = c->lv[0];
    CellLaunch cl = launch();
    kbegin(c, PP_K_KRYLOV, 0, cells * (1.0 + 8.0 + 4.0 * (j + 1) + 8.0));
    rhs_split = sio();
    if (rhs_split)
      k_kryl_orth4<256, true><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), u(j + 1), vv, c->sp.rhs,
                                                                   c->sp.e, c->d_part, c->sp.sx, c->sp.sxy,
                                                                   c->sp.hp);
    else
      k_kryl_orth4<256><<<cl.grid, cl.block, 0, c->stream>>>(largs<double>(L0), u(j + 1), vv, lv_rhs<float>(L0),
                                                             lv_e<float>(L0), c->d_part, 0, 0, 0);
    kend(c);
    // the orthogonalisation's norm is reduced over the apply launch's blocks
    // only: zero the coupling blocks' slots it does not write
    const long nb = coupling_blocks();
    if (nb > 0) k_fill<double><<<dim3((unsigned)((nb + 255) / 256)), 256, 0, c->stream>>>(
        c->d_part + apply_blocks(), (int)nb, 0.0);
  }
  void update(const V &vv) {
    CellLaunch cl = ref_launch(c);
    kbegin(c, PP_K_KRYLOV, 0, cells * (17.0 + 4.0 * vv.n));
    k_kryl_update<<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), c->d_mu, vv);
    kend(c);
  }
};

template <typename T>
int adjoint_impl(pp_ctx *c, const double *sg, const pp_solve_opts *o, double ref_norm, pp_solve_stats *st,
                 int *lam_idx) {
  hipStream_t s = c->stream;
  Level &L0 = c->lv[0];
  LevelArgs<double> a0 = largs<double>(L0);
  CellLaunch cl = ref_launch(c);
  long nblk = (long)cl.grid.x * cl.grid.y * cl.grid.z;
  GhostTables &g = c->gh;
  pp_solve_stats S;
  memset(&S, 0, sizeof(S));
  kreset(c, o->profile != 0);
  double cells = (double)c->nxc * c->nyc * c->nzc;
  hipEvent_t ev0, ev1;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventRecord(ev0, s));
  const int m = o->krylov;
  // FP32 V-cycles use the FP32 level-0 basis where the rows split into fours
  const bool kry32 = m > 0 && sizeof(T) == 4 && AdjOps32::fits(c);
  if (m > 0) {
    int rc = kry32 ? ensure_krylov_level(c, m) : ensure_krylov(c, m);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpyAsync(c->d_bwork, c->d_sv, sizeof(double) * c->nref, hipMemcpyDeviceToDevice, s));
  dim3 gg((g.n + 255) / 256), gt((g.nt + 255) / 256), blk(256);
  if (g.nt > 0) k_gather_idx<double><<<gt, blk, 0, s>>>(g.nt, g.t_ref, c->d_sv, g.otbase64);
  int cur = 0;
  int status = PP_NOT_CONVERGED;
  int outer = 0;
  for (;;) {
    if (g.n > 0) {
      k_adj_ghost_rhs<double><<<gg, blk, 0, s>>>(g.n, g.nb_ref, g.nb_w64, sg, c->d_mu, g.orhsg64);
      k_fill<double><<<gg, blk, 0, s>>>(g.olam64[0], g.n, 0.0);
      cur = 0;
      for (int q = 0; q < g.kexact; q++) {
        k_adj_ghost_jacobi<double><<<gg, blk, 0, s>>>(g.n, g.g_tidx, g.t_ptr, g.t_ghost, g.t_coef64, g.orhsg64,
                                                     g.olam64[cur], g.olam64[1 - cur]);
        cur ^= 1;
      }
      if (g.nt > 0)
        k_adj_touch<double><<<gt, blk, 0, s>>>(g.nt, g.t_ref, g.t_ptr, g.t_ghost, g.t_coef64, g.t_vinv64,
                                               g.olam64[cur], g.otbase64, c->d_bwork);
    }
    kbegin(c, PP_K_OUTER_RESIDUAL, 0, cells * (17.0 + 2.0 * sizeof(T)));
    k_outer_residual<T, 256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, c->d_bwork, lv_rhs<T>(L0),
                                                          lv_e<T>(L0), c->d_part,
                                                          (m > 0 && !kry32) ? c->d_kry[0] : nullptr,
                                                          kry32 ? (T *)c->d_kryl[0] : (T *)nullptr);
    kend(c);
    int rc = reduce_summax(c, (int)nblk);
    if (rc) return rc;
    double rl2 = std::sqrt(c->h_res[0]);
    if (!std::isfinite(rl2)) {
      set_error("non-finite adjoint residual after %d outer iterations", outer);
      status = PP_NONFINITE;
      break;
    }
    if (ref_norm < 0.0) ref_norm = rl2 > 0.0 ? rl2 : 1.0;
    double rel = rl2 / (ref_norm > 0.0 ? ref_norm : 1.0);
    S.rel_l2 = rel;
    S.abs_linf = c->h_res[1];
    if (outer > 0 && outer <= 64) S.history[outer - 1] = rel;
    if ((o->rel_tol > 0.0 && rel <= o->rel_tol) || rl2 == 0.0) { status = PP_OK; break; }
    if (outer >= o->max_outer) break;
    if (m > 0) {
      int krc;
      if (kry32) {
        AdjOps32 ops{c, o, &S, cells};
        krc = gmres_cycle(c, o, ops, rl2, ref_norm > 0.0 ? ref_norm : 1.0, outer, S);
      } else {
        AdjOps<T> ops{c, o, &S, cells};
        krc = gmres_cycle(c, o, ops, rl2, ref_norm > 0.0 ? ref_norm : 1.0, outer, S);
      }
      if (krc == PP_NOT_CONVERGED || krc == PP_NONFINITE) { status = krc; break; }
      if (krc) return krc;
      continue;
    }
    adjoint_cycles<T>(c, o, S);
    kbegin(c, PP_K_UPDATE, 0, cells * (17.0 + sizeof(T)));
    k_update<T><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, lv_e<T>(L0));
    kend(c);
    outer++;
  }
  HIPCHK(hipEventRecord(ev1, s));
  HIPCHK(hipEventSynchronize(ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  HIPCHK(hipGetLastError());
  S.ms = ms;
  S.bytes = c->bytes_acc;
  S.outer_iterations = outer;
  S.status = status;
  S.b_l2 = ref_norm;
  kcollect(c, S);
  if (st) *st = S;
  if (lam_idx) *lam_idx = cur;
  return status;
}

// Left null vector of the ghost-cell operator: the transposed iteration
// with s = 0 started from mu = 1 (the fluid volumes, exact without ghost
// rows) converges to w; normalised so sum_f w_f = sum_f V_f.
// the level-0 shape arrays w/V from c->d_w
int null_shape_arrays(pp_ctx *c) {
  hipStream_t s = c->stream;
  Level &L0 = c->lv[0];
  CellLaunch cl = ref_launch(c);
  LevelArgs<double> a0 = largs<double>(L0);
  if (!c->d_nshape32) {
    HIPCHK(hipMalloc(&c->d_nshape32, sizeof(float) * L0.total));
    HIPCHK(hipMalloc(&c->d_nshape64, sizeof(double) * L0.total));
    HIPCHK(hipMemsetAsync(c->d_nshape32, 0, sizeof(float) * L0.total, s));
    HIPCHK(hipMemsetAsync(c->d_nshape64, 0, sizeof(double) * L0.total, s));
  }
  k_null_shape<float><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_w, c->d_nshape32, c->sp.ok ? c->sp.nshape : nullptr,
                                                   c->sp.sx, c->sp.sxy, c->sp.hp);
  k_null_shape<double><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_w, c->d_nshape64, nullptr, 0, 0, 0);
  HIPCHK(hipGetLastError());
  return PP_OK;
}

// the shape arrays of the current null-vector iterate mu (w = V mu scaled
// to sum_f w = vsum, written to c->d_w)
int null_shape_from_mu(pp_ctx *c, double vsum) {
  hipStream_t s = c->stream;
  LevelArgs<double> a0 = largs<double>(c->lv[0]);
  CellLaunch cl = ref_launch(c);
  k_mu_vol_sums<256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, nullptr, c->d_part);
  int rc = reduce_sums_host(c, 2);
  if (rc) return rc;
  if (!(c->h_res[0] > 0.0)) return PP_OK;   // no usable iterate: keep the unit shape
  k_null_scale<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, vsum / c->h_res[0], c->d_w);
  return null_shape_arrays(c);
}

int compute_null_vector(pp_ctx *c) {
  hipStream_t s = c->stream;
  LevelArgs<double> a0 = largs<double>(c->lv[0]);
  CellLaunch cl = ref_launch(c);
  // vsum = sum of fluid volumes
  k_fluid_sum_vol<256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, nullptr, c->d_part);
  int rc = reduce_sums_host(c, 2);
  if (rc) return rc;
  const double vsum = c->h_res[1];
  HIPCHK(hipMemsetAsync(c->d_w, 0, sizeof(double) * c->nref, s));
  if (c->gh.n > 0) {
    // mu: the fluid indicator; s = 0.  Every geometry starts cold: a moving
    // body's previous null vector as the start took 22-23 transposed
    // iterations on the 384x256x256 moving sphere, a cold start 19, and the
    // result then depends on this geometry alone
    HIPCHK(hipMemsetAsync(c->d_mu, 0, sizeof(double) * c->nref, s));
    HIPCHK(hipMemsetAsync(c->d_sv, 0, sizeof(double) * c->nref, s));
    k_null_init<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, nullptr, c->d_mu);
    pp_solve_opts o;
    pp_default_opts(&o);
    o.precision = PP_PREC_MIXED;
    o.rel_tol = 1e-11;

    // When every ghost stencil reads fluid or ghost cells, the composite
    // operator maps constants to zero and A^T has an exact null vector,
    // solved to 1e-11.  A stencil corner on a solid non-ghost cell (which
    // holds 0) breaks that by the size of its coefficient: A is then
    // nonsingular, the iteration (A^T mu = 0 from mu = 1) levels off within
    // a few cycles, and what remains of mu is A^T's near-null mode, the
    // vector the compatibility projection needs; it is taken from a bounded
    // number of cycles and a loose floor.
    const bool exact = c->gh.solid_in_stencil == 0;
    o.max_outer = exact ? 200 : 40;
    const double accept = exact ? 1e-10 : 1e-4;
    pp_solve_stats st;
    int lam = 0;
    // The transposed cycles' coarse correction takes the null vector's
    // shape: a first few cycles run without it, the shape of that iterate
    // then shapes the correction of the rest.
    c->have_nshape = false;
    pp_solve_opts o1 = o;
    o1.max_outer = 3;
    rc = adjoint_impl<float>(c, nullptr, &o1, -1.0, &st, &lam);
    if (rc != PP_OK && rc != PP_NOT_CONVERGED) return rc;
    const double ref = st.b_l2;
    const int stage1 = st.outer_iterations;
    rc = null_shape_from_mu(c, vsum);
    if (rc) return rc;
    c->have_nshape = c->d_nshape32 != nullptr;
    o.max_outer -= stage1;  // the bound counts both stages
    // the second stage converges relative to the first one's initial residual
    rc = adjoint_impl<float>(c, nullptr, &o, ref, &st, &lam);
    if (rc != PP_OK && rc != PP_NOT_CONVERGED) return rc;
    st.outer_iterations += stage1;
    st.b_l2 = ref;
    if (!(st.rel_l2 <= accept)) {
      set_error("left null vector did not converge (relative residual %g after %d iterations, %d ghost "
                "stencil entries on solid cells)", st.rel_l2, st.outer_iterations, c->gh.solid_in_stencil);
      return PP_BAD_GEOMETRY;
    }
    // w = V mu normalised to sum_f w = sum_f V; ghost entries lambda_g likewise
    k_mu_vol_sums<256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, nullptr, c->d_part);
    rc = reduce_sums_host(c, 2);
    if (rc) return rc;
    const double scale = vsum / c->h_res[0];
    k_null_scale<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, scale, c->d_w);
    k_ghost_scatter_scaled<<<(c->gh.n + 255) / 256, 256, 0, s>>>(c->gh.n, c->gh.g_ref, c->gh.olam64[lam], scale,
                                                                 c->d_w);
    c->null_iters = st.outer_iterations;
    c->null_res = st.rel_l2;
    // the null space shape w/V for the adjoint V-cycle's coarse correction
    rc = null_shape_arrays(c);
    if (rc) return rc;
    c->have_nshape = true;
  } else {
    // no ghost rows: the operator is volume-symmetric, w = V
    HIPCHK(hipMemsetAsync(c->d_mu, 0, sizeof(double) * c->nref, s));
    k_null_init<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, nullptr, c->d_mu);
    k_null_scale<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, 1.0, c->d_w);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return PP_OK;
}

void free_fresh_rows(pp_ctx *c) {
  hipFree(c->fr.ids); hipFree(c->fr.face); hipFree(c->fr.coff); hipFree(c->fr.cc);
  c->fr = FreshRows{};
}

}  // namespace

// ======================================================================
// C-ABI
// ======================================================================
extern "C" {

void pp_default_opts(pp_solve_opts *o) {
  if (!o) return;
  o->precision = PP_PREC_MIXED;
  o->rel_tol = 1e-8;
  o->abs_linf_tol = 0.0;
  o->max_outer = 100;
  o->cycles_per_outer = 1;
  o->nu_pre = 2;
  o->nu_post = 2;
  // over-relaxed zebra lines: 9-15% fewer V-cycle sweeps to the same
  // residual than omega = 1 on stretched and moving-body grids
  o->omega = 1.3;
  o->mean_remove = 1;
  o->project_rhs = 1;
  o->use_graph = 0;
  o->profile = 0;
  o->krylov = 8;
  o->nu_coarse = 1;
  o->smoother = PP_SMOOTH_LINE;
}

const char *pp_last_error(void) { return pp::last_error(); }

int32_t pp_create(pp_ctx **out, const pp_grid_desc *g) {
  if (!out || !g) { set_error("pp_create: null argument"); return PP_BAD_ARG; }
  *out = nullptr;
  if (g->nxc < 4 || g->nyc < 4 || g->nzc < 4 || !g->x || !g->y || !g->z) {
    set_error("pp_create: grid must have at least 4 cells per direction and node arrays");
    return PP_BAD_ARG;
  }
  if (g->nxc > kMaxLine || g->nyc > kMaxLine || g->nzc > kMaxLine) {
    set_error("pp_create: %dx%dx%d cells; the line solvers hold at most %d cells per line", g->nxc, g->nyc,
              g->nzc, kMaxLine);
    return PP_BAD_ARG;
  }
  for (int f = 0; f < 6; f++)
    if (g->bc[f] != PP_BC_NEUMANN) {
      set_error("pp_create: face %d has pressure BC %d; only PP_BC_NEUMANN (2) is implemented", f, g->bc[f]);
      return PP_BAD_ARG;
    }
  {
    // every grid coarsens to at most 4096 cells; the check guards the
    // dense coarsest-level operator (n^2 entries) against a future rule
    const long nc = coarsest_cells(g->nxc, g->nyc, g->nzc);
    if (nc > kCoarseDenseMax) {
      set_error("pp_create: %dx%dx%d cells coarsen to a %ld-cell coarsest level; at most %ld cells are solved "
                "directly", g->nxc, g->nyc, g->nzc, nc, kCoarseDenseMax);
      return PP_BAD_ARG;
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("pp_create: no HIP device available");
    return PP_HIP_ERROR;
  }
  if (g->device < 0 || g->device >= ndev) { set_error("pp_create: bad device %d", g->device); return PP_BAD_ARG; }
  HIPCHK(hipSetDevice(g->device));
  {
    int rs = scratch_setup(g->device);
    if (rs) return rs;
  }
  pp_ctx *c = new pp_ctx();
  c->device = g->device;
  c->nxc = g->nxc; c->nyc = g->nyc; c->nzc = g->nzc;
  c->ref.nxc = g->nxc; c->ref.nyc = g->nyc; c->ref.nzc = g->nzc;
  c->ref.ex = g->nxc + 3;
  c->ref.exy = (long)(g->nxc + 3) * (g->nyc + 3);
  c->nref = c->ref.exy * (g->nzc + 3);
  const double *xs[3] = {g->x, g->y, g->z};
  int ns[3] = {g->nxc, g->nyc, g->nzc};
  for (int d = 0; d < 3; d++) {
    for (int q = 1; q <= ns[d]; q++)
      if (!(xs[d][q] > xs[d][q - 1])) {
        set_error("pp_create: node coordinates of direction %d are not increasing at %d", d, q);
        delete c;
        return PP_BAD_ARG;
      }
    make_grid_dir(ns[d] + 1, xs[d], c->hx[d], c->rxc[d], c->rdx[d], c->rdxinv[d], c->rdxc[d], c->rdxcinv[d]);
  }
  // a blocking stream: ordered with the legacy default stream, on which the
  // synchronous hipMemcpy/hipMemset calls of the entry points and a caller's
  // default-stream work (the producer of pp_solve_device's inputs) run
  HIPCHK(hipStreamCreate(&c->stream));
  for (int d = 0; d < 3; d++) {
    size_t nd = sizeof(double) * c->rdx[d].size();
    HIPCHK(hipMalloc(&c->d_rdx[d], nd));
    HIPCHK(hipMalloc(&c->d_rdxinv[d], nd));
    HIPCHK(hipMemcpy(c->d_rdx[d], c->rdx[d].data(), nd, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rdxinv[d], c->rdxinv[d].data(), nd, hipMemcpyHostToDevice));
    // fx(i) = (x(i) - xc(i-1)) / (xc(i) - xc(i-1)), i = 1..n (CORE_GRID.F90:172-182)
    std::vector<double> fx(c->rdx[d].size(), 0.0);
    const int nn = ns[d] + 1;
    for (int i = 1; i <= nn; i++) fx[i] = (c->hx[d][i] - c->rxc[d][i - 1]) / (c->rxc[d][i] - c->rxc[d][i - 1]);
    HIPCHK(hipMalloc(&c->d_rfx[d], nd));
    HIPCHK(hipMalloc(&c->d_rdxcinv[d], nd));
    HIPCHK(hipMemcpy(c->d_rfx[d], fx.data(), nd, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rdxcinv[d], c->rdxcinv[d].data(), nd, hipMemcpyHostToDevice));
  }
  size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMalloc(&c->d_p, nb));
  HIPCHK(hipMalloc(&c->d_b, nb));
  HIPCHK(hipMalloc(&c->d_w, nb));
  HIPCHK(hipMalloc(&c->d_bwork, nb));
  HIPCHK(hipMalloc(&c->d_tmp, nb));
  HIPCHK(hipMalloc(&c->d_mu, nb));
  HIPCHK(hipMalloc(&c->d_sv, nb));
  HIPCHK(hipMemset(c->d_b, 0, nb));
  HIPCHK(hipMemset(c->d_w, 0, nb));
  CellLaunch cl = cell_launch(c->nxc, c->nyc, c->nzc);
  c->nblk_red = (int)((long)cl.grid.x * cl.grid.y * cl.grid.z);
  // partial sums: up to KRY_MAX+1 per block (Krylov dots), 5 (fluid sums)
  HIPCHK(hipMalloc(&c->d_part, sizeof(double) * (KRY_MAX + 1) * (size_t)c->nblk_red + 64));
  HIPCHK(hipMalloc(&c->d_part2, sizeof(double) * (KRY_MAX + 1) * kRedStage));
  HIPCHK(hipMalloc(&c->d_res, sizeof(double) * 32));
  HIPCHK(hipHostMalloc(&c->h_res, sizeof(double) * 32));
  *out = c;
  return PP_OK;
}

int32_t pp_destroy(pp_ctx *c) {
  if (!c) return PP_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_levels(c);
  free_ghosts(c);
  hipFree(c->d_p); hipFree(c->d_b); hipFree(c->d_w); hipFree(c->d_bwork); hipFree(c->d_tmp);
  hipFree(c->d_fresh); hipFree(c->d_iblank_ref); hipFree(c->d_count); hipFree(c->d_mu); hipFree(c->d_sv);
  hipFree(c->d_ad_coef);
  hipFree(c->d_nshape32); hipFree(c->d_nshape64);
  free_fresh_rows(c);
  hipFree(c->d_part); hipFree(c->d_part2); hipFree(c->d_res);
  for (double *q : c->d_kry) hipFree(q);
  for (float *q : c->d_kryl) hipFree(q);
  for (int d = 0; d < 3; d++) {
    hipFree(c->d_rdx[d]); hipFree(c->d_rdxinv[d]); hipFree(c->d_rfx[d]); hipFree(c->d_rdxcinv[d]);
  }
  if (c->h_res) hipHostFree(c->h_res);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->blas) rocblas_destroy_handle((rocblas_handle)c->blas);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PP_OK;
}

int32_t pp_set_geometry(pp_ctx *c, const pp_geometry *geo) {
  if (!c || !geo || !geo->iblank) { set_error("pp_set_geometry: null argument"); return PP_BAD_ARG; }
  if (geo->n_ghost < 0 || (geo->n_ghost > 0 && (!geo->i_ghost || !geo->j_ghost || !geo->k_ghost || !geo->i_cell ||
                                                 !geo->j_cell || !geo->k_cell || !geo->coeff || !geo->probe_length))) {
    set_error("pp_set_geometry: ghost arrays missing");
    return PP_BAD_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  c->have_geom = false;
  c->have_ad = false;  // set_solve_ad follows every set_boundary
  c->have_nshape = false;
  free_fresh_rows(c);
  if (c->gexec) {  // captured launches point at the old hierarchy
    HIPCHK(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
    c->graph_key = -1;
  }
  c->h_iblank.assign(geo->iblank, geo->iblank + c->nref);
  {
    std::atomic<long> bad(-1);
    host_parallel(c->nref, [&](long a, long b) {
      for (long r = a; r < b; r++)
        if (c->h_iblank[r] != 0 && c->h_iblank[r] != 1) { bad.store(r); break; }
    });
    if (bad.load() >= 0) {
      long r = bad.load();
      set_error("pp_set_geometry: iblank must be 0 or 1 (entry %ld = %d)", r, (int)c->h_iblank[r]);
      return PP_BAD_GEOMETRY;
    }
  }
  // the outer layer carries no geometry
  host_parallel(c->nzc + 3, [&](long k0, long k1) {
    for (int k = (int)k0; k < (int)k1; k++)
      for (int j = 0; j < c->nyc + 3; j++)
        for (int i = 0; i < c->nxc + 3; i++)
          if (i == 0 || j == 0 || k == 0 || i > c->nxc || j > c->nyc || k > c->nzc)
            c->h_iblank[(long)i + c->ref.ex * j + c->ref.exy * k] = 0;
  });
  free_ghosts(c);
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  auto t0 = now();
  int rc = PP_OK;
  if (c->lv.empty()) {  // the hierarchy depends on the grid only
    rc = build_levels(c);
    if (rc) return rc;
    rc = upload_levels(c);
    if (rc) return rc;
  }
  if (!c->d_iblank_ref) HIPCHK(hipMalloc(&c->d_iblank_ref, c->nref));
  HIPCHK(hipMemcpyAsync(c->d_iblank_ref, c->h_iblank.data(), c->nref, hipMemcpyHostToDevice, c->stream));
  long nmark = 0;
  rc = build_flags(c, c->d_iblank_ref, &nmark);
  if (rc) return rc;
  if (nmark != geo->n_ghost) {
    set_error("pp_set_geometry: iblank has %ld ghost cells, the ghost list has %d", nmark, geo->n_ghost);
    return PP_BAD_GEOMETRY;
  }
  auto t1 = now();
  rc = build_ghosts(c, geo);
  if (rc) return rc;
  auto t2 = now();
  rc = build_coarse(c);
  if (rc) return rc;
  auto t3 = now();
  // fresh cells
  c->h_fresh.clear();
  hipFree(c->d_fresh);
  c->d_fresh = nullptr;
  if (geo->fresh_cell) {
    std::atomic<bool> any(false);
    host_parallel(c->nref, [&](long a, long b) {
      for (long r = a; r < b; r++)
        if (geo->fresh_cell[r] == 1) { any.store(true); break; }
    });
    if (any.load()) {
      c->h_fresh.assign(geo->fresh_cell, geo->fresh_cell + c->nref);
      HIPCHK(hipMalloc(&c->d_fresh, c->nref));
      HIPCHK(hipMemcpy(c->d_fresh, c->h_fresh.data(), c->nref, hipMemcpyHostToDevice));
    }
  }
  c->have_geom = true;
  auto t4 = now();
  rc = compute_null_vector(c);
  if (rc) { c->have_geom = false; return rc; }
  HIPCHK(hipDeviceSynchronize());
  c->have_geom = true;
  auto t5 = now();
  c->setup_ms[0] = ms(t0, t1);
  c->setup_ms[1] = ms(t1, t2);
  c->setup_ms[2] = ms(t2, t3);
  c->setup_ms[3] = ms(t4, t5);
  return PP_OK;
}

int32_t pp_setup_info(pp_ctx *c, double *phase_ms, int32_t *null_iterations, double *null_residual) {
  if (!c) { set_error("pp_setup_info: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_setup_info: no geometry"); return PP_NO_GEOMETRY; }
  if (phase_ms)
    for (int q = 0; q < 4; q++) phase_ms[q] = c->setup_ms[q];
  if (null_iterations) *null_iterations = c->null_iters;
  if (null_residual) *null_residual = c->null_res;
  return PP_OK;
}

int32_t pp_solve_device(pp_ctx *c, double *d_p, const double *d_rhs, const double *d_bc, const pp_solve_opts *o,
                        pp_solve_stats *st) {
  if (!c || !d_p || !d_rhs) { set_error("pp_solve_device: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_solve_device: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  pp_solve_opts def;
  if (!o) { pp_default_opts(&def); o = &def; }
  if (!opts_ok(o)) { set_error("pp_solve_device: invalid options"); return PP_BAD_ARG; }
  HIPCHK(hipSetDevice(c->device));
  if (o->precision == PP_PREC_MIXED) return solve_impl<float>(c, d_p, d_rhs, d_bc, o, st);
  return solve_impl<double>(c, d_p, d_rhs, d_bc, o, st);
}

int32_t pp_solve(pp_ctx *c, double *p, const double *rhs, const double *bc, const pp_solve_opts *o,
                 pp_solve_stats *st) {
  if (!c || !p || !rhs) { set_error("pp_solve: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_solve: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMemcpy(c->d_p, p, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_bwork, rhs, nb, hipMemcpyHostToDevice));
  DevTemp dbc;
  if (bc && c->gh.n > 0) {
    HIPCHK(dbc.alloc(sizeof(double) * c->gh.n));
    HIPCHK(hipMemcpy(dbc.d(), bc, sizeof(double) * c->gh.n, hipMemcpyHostToDevice));
  }
  int rc = pp_solve_device(c, c->d_p, c->d_bwork, dbc.d(), o, st);
  if (rc == PP_OK || rc == PP_NOT_CONVERGED) HIPCHK(hipMemcpy(p, c->d_p, nb, hipMemcpyDeviceToHost));
  return rc;
}

int32_t pp_calc_residual(pp_ctx *c, const double *var, const double *r, double *res, double *resm, int32_t *loc,
                         double *l2) {
  if (!c || !var || !r) { set_error("pp_calc_residual: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_calc_residual: no geometry"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMemcpy(c->d_p, var, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_bwork, r, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_tmp, 0, nb));
  CellLaunch cl = ref_launch(c);
  long nblk = (long)cl.grid.x * cl.grid.y * cl.grid.z;
  k_ref_residual<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), c->d_p, c->d_bwork,
                                                           c->d_tmp, c->d_part);
  int rc = reduce_summax(c, (int)nblk);
  if (rc) return rc;
  if (res) HIPCHK(hipMemcpy(res, c->d_tmp, nb, hipMemcpyDeviceToHost));
  double mx = c->h_res[1];
  if (resm) *resm = mx;
  if (l2) *l2 = std::sqrt(c->h_res[0]);
  if (loc) {
    long q = (long)c->h_res[2];
    if (mx > 0.0 && q >= 0) {
      loc[0] = (int)(q % c->ref.ex);
      loc[1] = (int)((q / c->ref.ex) % (c->nyc + 3));
      loc[2] = (int)(q / c->ref.exy);
    } else {
      loc[0] = loc[1] = loc[2] = 0;
    }
  }
  return PP_OK;
}

int32_t pp_ghost_rows(pp_ctx *c, double *p, const double *bc) {
  if (!c || !p) { set_error("pp_ghost_rows: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_ghost_rows: no geometry"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMemcpy(c->d_p, p, nb, hipMemcpyHostToDevice));
  GhostTables &g = c->gh;
  bool have = bc && g.n > 0;
  if (have) {
    int rc = set_ghost_bnd(c, std::vector<double>(bc, bc + g.n));
    if (rc) return rc;
  }
  ghost_exact_ref(c, c->d_p, have);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(p, c->d_p, nb, hipMemcpyDeviceToHost));
  return PP_OK;
}

int32_t pp_set_vel_ghosts(pp_ctx *c, const double *coeffD, double bi_weight, double ip_weight) {
  if (!c) { set_error("pp_set_vel_ghosts: null context"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_set_vel_ghosts: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  GhostTables &g = c->gh;
  if (g.n > 0 && !coeffD) { set_error("pp_set_vel_ghosts: null coefficients"); return PP_BAD_ARG; }
  if (!std::isfinite(bi_weight) || !std::isfinite(ip_weight)) {
    set_error("pp_set_vel_ghosts: weights must be finite");
    return PP_BAD_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  hipFree(g.v_cD); hipFree(g.v_buf[0]); hipFree(g.v_buf[1]); hipFree(g.v_kind); hipFree(g.v_off);
  g.v_cD = g.v_buf[0] = g.v_buf[1] = nullptr; g.v_kind = nullptr; g.v_off = nullptr;
  g.have_vel = false;
  const int n = g.n;
  std::vector<int> kind(8L * n);
  std::vector<long> off(8L * n);
  std::vector<double> cd(coeffD ? coeffD : nullptr, coeffD ? coeffD + 8L * n : nullptr);
  std::unordered_map<long, int> gidx;
  gidx.reserve(2L * n + 1);
  for (int q = 0; q < n; q++) gidx[g.h_g_ref[q]] = q;
  for (int q = 0; q < n; q++)
    for (int r = 0; r < 8; r++) {
      const long e = 8L * q + r;
      if (!std::isfinite(cd[e])) { set_error("pp_set_vel_ghosts: ghost %d coefficient %d is not finite", q, r); return PP_BAD_ARG; }
      const long o = g.h_s_raw[e];
      off[e] = o;
      if (o == g.h_g_ref[q]) kind[e] = -2;
      else {
        auto it = gidx.find(o);
        kind[e] = it != gidx.end() ? it->second : -1;
      }
    }
  // sweeps that bring the ghost-ghost coupling to round-off: the same
  // Jacobi iteration on the host, field values and body-intercept values
  // fixed pseudo-random numbers
  int sweeps = 1;
  if (n > 0) {
    auto hv = [](long a) {
      uint64_t h = (uint64_t)a * 0x9E3779B97F4A7C15ull + 777u;
      h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 29;
      return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    };
    std::vector<double> x(n, 0.0), y(n);
    double dmax = 1.0;
    sweeps = 0;
    while (sweeps < 1000 && dmax > 1e-15) {
      dmax = 0.0;
      for (int q = 0; q < n; q++) {
        const double b = hv(-1 - (long)q);
        double sum = 0.0;
        for (int r = 0; r < 8; r++) {
          const long e = 8L * q + r;
          const double xv = kind[e] >= 0 ? x[kind[e]] : (kind[e] == -1 ? hv(off[e]) : b);
          sum += cd[e] * xv;
        }
        y[q] = b * bi_weight - sum * ip_weight;
      }
      for (int q = 0; q < n; q++) {
        dmax = std::max(dmax, std::fabs(y[q] - x[q]) / std::max(1.0, std::fabs(y[q])));
        x[q] = y[q];
      }
      sweeps++;
    }
    if (dmax > 1e-15) {
      set_error("pp_set_vel_ghosts: the velocity ghost rows do not converge (change %g after %d sweeps)", dmax, sweeps);
      return PP_BAD_GEOMETRY;
    }
    HIPCHK(hipMalloc(&g.v_cD, sizeof(double) * 8 * n));
    HIPCHK(hipMalloc(&g.v_kind, sizeof(int) * 8 * n));
    HIPCHK(hipMalloc(&g.v_off, sizeof(long) * 8 * n));
    HIPCHK(hipMalloc(&g.v_buf[0], sizeof(double) * 3 * n));
    HIPCHK(hipMalloc(&g.v_buf[1], sizeof(double) * 3 * n));
    HIPCHK(hipMemcpy(g.v_cD, cd.data(), sizeof(double) * 8 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.v_kind, kind.data(), sizeof(int) * 8 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.v_off, off.data(), sizeof(long) * 8 * n, hipMemcpyHostToDevice));
  }
  g.v_sweeps = sweeps;
  g.v_bw = bi_weight;
  g.v_iw = ip_weight;
  g.have_vel = true;
  return PP_OK;
}

}  // extern "C"

namespace {
// GCM_vel_set_bc_internal on device fields: gather the ghost cells'
// (u, v, w), the Jacobi sweeps fixed by pp_set_vel_ghosts, scatter
void vel_ghosts(pp_ctx *c, double *d_u, double *d_v, double *d_w, const double *d_vbi) {
  GhostTables &g = c->gh;
  if (g.n == 0) return;
  hipStream_t s = c->stream;
  dim3 gg((g.n + 255) / 256), blk(256);
  k_vel_ghost_gather<double><<<gg, blk, 0, s>>>(g.n, g.g_ref, d_u, d_v, d_w, g.v_buf[0]);
  int cur = 0;
  for (int q = 0; q < g.v_sweeps; q++) {
    k_vel_ghost_sweep<<<gg, blk, 0, s>>>(g.n, g.v_cD, g.v_kind, g.v_off, g.v_bw, g.v_iw, d_vbi, d_u, d_v, d_w,
                                         g.v_buf[cur], g.v_buf[1 - cur]);
    cur ^= 1;
  }
  k_vel_ghost_scatter<<<gg, blk, 0, s>>>(g.n, g.g_ref, g.v_buf[cur], d_u, d_v, d_w);
}
}  // namespace

extern "C" {

int32_t pp_correct_vel_device(pp_ctx *c, double dt, const double *d_p, double *d_u, double *d_v, double *d_w,
                              double *d_fu, double *d_fv, double *d_fw, const double *d_bcxu, const double *d_bcyv,
                              const double *d_bczw, const double *d_vbi) {
  if (!c || !d_p || !d_u || !d_v || !d_w || !d_fu || !d_fv || !d_fw || !d_bcxu || !d_bcyv || !d_bczw) {
    set_error("pp_correct_vel: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_correct_vel: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  GhostTables &g = c->gh;
  if (g.n > 0 && !g.have_vel) {
    set_error("pp_correct_vel: no velocity ghost coefficients (call pp_set_vel_ghosts after pp_set_geometry)");
    return PP_NO_GEOMETRY;
  }
  if (!std::isfinite(dt)) { set_error("pp_correct_vel: dt must be finite"); return PP_BAD_ARG; }
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  FracMetrics M;
  for (int d = 0; d < 3; d++) { M.fx[d] = c->d_rfx[d]; M.dinv[d] = c->d_rdxinv[d]; M.dcinv[d] = c->d_rdxcinv[d]; }
  CellLaunch cl = ref_launch(c);
  k_correct_vel<256><<<cl.grid, cl.block, 0, s>>>(c->ref, largs<double>(c->lv[0]), M, dt, d_p, d_u, d_v, d_w, d_fu,
                                                   d_fv, d_fw, d_bcxu, d_bcyv, d_bczw);
  vel_ghosts(c, d_u, d_v, d_w, d_vbi);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return PP_OK;
}

int32_t pp_correct_vel(pp_ctx *c, double dt, const double *p, double *u, double *v, double *w, double *fu, double *fv,
                       double *fw, const double *bcxu, const double *bcyv, const double *bczw, const double *vbi) {
  if (!c || !p || !u || !v || !w || !fu || !fv || !fw || !bcxu || !bcyv || !bczw) {
    set_error("pp_correct_vel: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_correct_vel: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp d, dvbi;
  HIPCHK(d.alloc(10 * nb));
  double *dv[10];
  for (int q = 0; q < 10; q++) dv[q] = d.d() + q * c->nref;
  const double *hin[10] = {p, u, v, w, fu, fv, fw, bcxu, bcyv, bczw};
  for (int q = 0; q < 10; q++) HIPCHK(hipMemcpy(dv[q], hin[q], nb, hipMemcpyHostToDevice));
  if (vbi && c->gh.n > 0) {
    HIPCHK(dvbi.alloc(sizeof(double) * 3 * c->gh.n));
    HIPCHK(hipMemcpy(dvbi.d(), vbi, sizeof(double) * 3 * c->gh.n, hipMemcpyHostToDevice));
  }
  int rc = pp_correct_vel_device(c, dt, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7], dv[8], dv[9],
                                 dvbi.d());
  if (rc == PP_OK) {
    double *hout[6] = {u, v, w, fu, fv, fw};
    for (int q = 0; q < 6; q++) HIPCHK(hipMemcpy(hout[q], dv[1 + q], nb, hipMemcpyDeviceToHost));
  }
  return rc;
}

int32_t pp_update_pressure_device(pp_ctx *c, int32_t van_kan, const double *d_pprime, double *d_p,
                                  const double *d_bc) {
  if (!c || !d_pprime || !d_p) { set_error("pp_update_pressure: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_update_pressure: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  GhostTables &g = c->gh;
  const bool have = d_bc && g.n > 0;
  if (have) {
    std::vector<double> hs(g.n);
    HIPCHK(hipMemcpy(hs.data(), d_bc, sizeof(double) * g.n, hipMemcpyDeviceToHost));
    int rc = set_ghost_bnd(c, hs);
    if (rc) return rc;
  }
  const long n = c->nref;
  k_update_pressure<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(n, van_kan ? 1 : 0, d_pprime, d_p);
  // GCM_p_set_bc_internal only: the outer layer keeps p + p' (both hold
  // the Neumann copy after a solve, so their sum does too)
  ghost_exact_ref(c, d_p, have, false);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int32_t pp_update_pressure(pp_ctx *c, int32_t van_kan, const double *pprime, double *p, const double *bc) {
  if (!c || !pprime || !p) { set_error("pp_update_pressure: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_update_pressure: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMemcpy(c->d_p, p, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_bwork, pprime, nb, hipMemcpyHostToDevice));
  DevTemp dbc;
  if (bc && c->gh.n > 0) {
    HIPCHK(dbc.alloc(sizeof(double) * c->gh.n));
    HIPCHK(hipMemcpy(dbc.d(), bc, sizeof(double) * c->gh.n, hipMemcpyHostToDevice));
  }
  int rc = pp_update_pressure_device(c, van_kan, c->d_bwork, c->d_p, dbc.d());
  if (rc == PP_OK) HIPCHK(hipMemcpy(p, c->d_p, nb, hipMemcpyDeviceToHost));
  return rc;
}

int32_t pp_mean_remove(pp_ctx *c, double *p, double *mean) {
  if (!c || !p) { set_error("pp_mean_remove: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_mean_remove: no geometry"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMemcpy(c->d_p, p, nb, hipMemcpyHostToDevice));
  double m = 0.0;
  int rc = mean_remove_ref(c, c->d_p, &m);
  if (rc) return rc;
  HIPCHK(hipMemcpy(p, c->d_p, nb, hipMemcpyDeviceToHost));
  if (mean) *mean = m;
  return PP_OK;
}

}  // extern "C"

namespace {
template <typename T>
int time_kernel(pp_ctx *c, int cls, int level, int reps, double *avg_us, double *bytes) {
  Level &L = c->lv[level];
  hipStream_t st = c->stream;
  kreset(c, false);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  // the line passes of the FP32 level 0 run on its colour-split copy, as
  // in the mixed solve
  c->sp.on = std::is_same<T, float>::value && level == 0 && c->sp.ok && cls != PP_K_RESIDUAL;
  for (int r = 0; r < reps; r++) {
    switch (cls) {
      case PP_K_XLINE: launch_xline<T>(c, L, level, r & 1, T(1)); break;
      case PP_K_YLINE: launch_yzline<T, 1>(c, L, level, r & 1, T(1)); break;
      case PP_K_ZLINE: launch_yzline<T, 2>(c, L, level, r & 1, T(1)); break;
      default:
        kbegin(c, PP_K_RESIDUAL, level, (double)L.cx * L.cy * L.cz * (3.0 * sizeof(T) + 1.0));
        launch_residual<T>(c, L);
        kend(c);
        break;
    }
  }
  c->sp.on = false;
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipGetLastError());
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *avg_us = 1e3 * (double)ms / reps;
  *bytes = c->bytes_acc / reps;
  return PP_OK;
}
}  // namespace

extern "C" {

int32_t pp_time_kernel(pp_ctx *c, int32_t cls, int32_t level, int32_t precision, int32_t reps, double *avg_us,
                       double *bytes) {
  if (!c || !avg_us || !bytes) { set_error("pp_time_kernel: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_time_kernel: no geometry"); return PP_NO_GEOMETRY; }
  if (cls != PP_K_XLINE && cls != PP_K_YLINE && cls != PP_K_ZLINE && cls != PP_K_RESIDUAL) {
    set_error("pp_time_kernel: kernel class %d is not timed separately", cls);
    return PP_BAD_ARG;
  }
  if (level < 0 || level >= (int)c->lv.size() - 1 || reps < 1 || reps > 100000 ||
      (precision != PP_PREC_FP64 && precision != PP_PREC_MIXED)) {
    set_error("pp_time_kernel: bad level %d, reps %d or precision %d", level, reps, precision);
    return PP_BAD_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  return precision == PP_PREC_MIXED ? time_kernel<float>(c, cls, level, reps, avg_us, bytes)
                                    : time_kernel<double>(c, cls, level, reps, avg_us, bytes);
}

int32_t pp_level_data(pp_ctx *c, int32_t level, uint8_t *flags, double *w, double *xc, double *cm, double *cp) {
  if (!c) { set_error("pp_level_data: null context"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_level_data: no geometry"); return PP_NO_GEOMETRY; }
  if (level < 0 || level >= (int)c->lv.size()) { set_error("pp_level_data: bad level %d", level); return PP_BAD_ARG; }
  const Level &L = c->lv[level];
  if (flags) {
    HIPCHK(hipSetDevice(c->device));
    std::vector<uint8_t> hf(L.total);
    HIPCHK(hipMemcpyAsync(hf.data(), L.flags, L.total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < L.cz; k++)
      for (int j = 0; j < L.cy; j++)
        for (int i = 0; i < L.cx; i++)
          flags[i + (long)L.cx * (j + (long)L.cy * k)] = hf[lidx(L.sx, L.sxy, i, j, k)];
  }
  double *outs[4] = {w, xc, cm, cp};
  const std::vector<double> *src[4] = {L.w, L.xc, L.cm, L.cp};
  for (int q = 0; q < 4; q++) {
    if (!outs[q]) continue;
    long off = 0;
    for (int d = 0; d < 3; d++) {
      std::copy(src[q][d].begin(), src[q][d].end(), outs[q] + off);
      off += (long)src[q][d].size();
    }
  }
  return PP_OK;
}

}  // extern "C"

namespace {
// dense interior (i fastest) <-> level layout, host side, converted to T
template <typename T>
int level_upload(pp_ctx *c, const Level &L, const double *src, T *dst) {
  std::vector<T> h((size_t)L.total, T(0));
  for (int k = 0; k < L.cz; k++)
    for (int j = 0; j < L.cy; j++)
      for (int i = 0; i < L.cx; i++)
        h[lidx(L.sx, L.sxy, i, j, k)] = (T)src[i + (long)L.cx * (j + (long)L.cy * k)];
  HIPCHK(hipMemcpy(dst, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return PP_OK;
}
template <typename T>
int level_download(pp_ctx *c, const Level &L, const T *src, double *dst) {
  std::vector<T> h((size_t)L.total);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(h.data(), src, sizeof(T) * h.size(), hipMemcpyDeviceToHost));
  for (int k = 0; k < L.cz; k++)
    for (int j = 0; j < L.cy; j++)
      for (int i = 0; i < L.cx; i++) dst[i + (long)L.cx * (j + (long)L.cy * k)] = (double)h[lidx(L.sx, L.sxy, i, j, k)];
  return PP_OK;
}

template <typename T>
int level_apply(pp_ctx *c, int op, int l, double omega, const double *a, const double *b, double *out) {
  Level &L = c->lv[l];
  kreset(c, false);
  int rc = PP_OK;
  switch (op) {
    case PP_LOP_RESIDUAL:
      if ((rc = level_upload<T>(c, L, a, lv_e<T>(L))) || (rc = level_upload<T>(c, L, b, lv_rhs<T>(L)))) return rc;
      launch_residual<T>(c, L);
      return level_download<T>(c, L, lv_res<T>(L), out);
    case PP_LOP_RESTRICT: {
      Level &C = c->lv[l + 1];
      if ((rc = level_upload<T>(c, L, a, lv_res<T>(L)))) return rc;
      CellLaunch cc = cell_launch(C.cx, C.cy, C.cz);
      k_restrict<T><<<cc.grid, cc.block, 0, c->stream>>>(largs<T>(L), largs<T>(C), C.cf[0], C.cf[1], C.cf[2],
                                                          lv_res<T>(L), lv_rhs<T>(C), lv_e<T>(C));
      return level_download<T>(c, C, lv_rhs<T>(C), out);
    }
    case PP_LOP_RESRESTRICT:
      if ((rc = level_upload<T>(c, L, a, lv_e<T>(L))) || (rc = level_upload<T>(c, L, b, lv_rhs<T>(L)))) return rc;
      restrict_step<T>(c, l);
      return level_download<T>(c, c->lv[l + 1], lv_rhs<T>(c->lv[l + 1]), out);
    case PP_LOP_PROLONG:
      if ((rc = level_upload<T>(c, c->lv[l + 1], a, lv_e<T>(c->lv[l + 1]))) ||
          (rc = level_upload<T>(c, L, b, lv_e<T>(L))))
        return rc;
      prolong_step<T>(c, l);
      return level_download<T>(c, L, lv_e<T>(L), out);
    default:  // PP_LOP_LINE_PASS
      if ((rc = level_upload<T>(c, L, a, lv_e<T>(L))) || (rc = level_upload<T>(c, L, b, lv_rhs<T>(L)))) return rc;
      line_passes<T>(c, l, (T)omega);
      return level_download<T>(c, L, lv_e<T>(L), out);
  }
}
}  // namespace

extern "C" {

int32_t pp_level_apply(pp_ctx *c, int32_t op, int32_t level, int32_t precision, double omega, const double *a,
                       const double *b, double *out) {
  if (!c || !a || !out) { set_error("pp_level_apply: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_level_apply: no geometry"); return PP_NO_GEOMETRY; }
  const int nl = (int)c->lv.size();
  const bool two_levels = op == PP_LOP_RESTRICT || op == PP_LOP_RESRESTRICT || op == PP_LOP_PROLONG;
  if (op < PP_LOP_RESIDUAL || op > PP_LOP_LINE_PASS || level < 0 || level >= nl - (two_levels ? 1 : 0) ||
      (precision != PP_PREC_FP64 && precision != PP_PREC_MIXED) || (op != PP_LOP_RESTRICT && !b) ||
      (op == PP_LOP_LINE_PASS && !(omega > 0.0 && omega < 2.0))) {
    set_error("pp_level_apply: bad op %d, level %d (of %d), precision %d, omega %g or missing b", op, level, nl,
              precision, omega);
    return PP_BAD_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  int rc = precision == PP_PREC_MIXED ? level_apply<float>(c, op, level, omega, a, b, out)
                                      : level_apply<double>(c, op, level, omega, a, b, out);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return PP_OK;
}

int32_t pp_level_info(pp_ctx *c, int32_t *n_levels, int32_t *dims) {
  if (!c || !n_levels) { set_error("pp_level_info: null argument"); return PP_BAD_ARG; }
  *n_levels = (int32_t)c->lv.size();
  if (dims)
    for (size_t l = 0; l < c->lv.size(); l++) {
      dims[3 * l] = c->lv[l].cx; dims[3 * l + 1] = c->lv[l].cy; dims[3 * l + 2] = c->lv[l].cz;
    }
  return PP_OK;
}

int32_t pp_null_vector(pp_ctx *c, double *w) {
  if (!c || !w) { set_error("pp_null_vector: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_null_vector: no geometry"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(w, c->d_w, sizeof(double) * c->nref, hipMemcpyDeviceToHost));
  return PP_OK;
}

int32_t pp_smooth(pp_ctx *c, int32_t smoother, double omega, int32_t sweeps, double *var, const double *r) {
  if (!c || !var || !r) { set_error("pp_smooth: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_smooth: no geometry"); return PP_NO_GEOMETRY; }
  if ((smoother != PP_SMOOTH_LINE && smoother != PP_SMOOTH_POINT_JACOBI) || sweeps < 0 || sweeps > 10000 ||
      !(omega > 0.0 && omega < 2.0)) {
    set_error("pp_smooth: smoother must be PP_SMOOTH_LINE or PP_SMOOTH_POINT_JACOBI, 0 <= sweeps <= 10000, "
              "0 < omega < 2");
    return PP_BAD_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * c->nref;
  HIPCHK(hipMemcpy(c->d_p, var, nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_bwork, r, nb, hipMemcpyHostToDevice));
  Level &L = c->lv[0];
  LevelArgs<double> a = largs<double>(L);
  CellLaunch cl = ref_launch(c);
  k_ref_to_level<double><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a, c->d_p, L.e64);
  k_ref_to_level<double><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a, c->d_bwork, L.rhs64);
  pp_solve_opts o;
  pp_default_opts(&o);
  o.smoother = smoother;
  o.omega = omega;
  const bool adj = c->adjoint_mode;
  c->adjoint_mode = false;
  smooth_steps<double>(c, 0, &o, sweeps, false);
  c->adjoint_mode = adj;
  k_level_to_ref<double><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, a, L.e64, c->d_p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(var, c->d_p, nb, hipMemcpyDeviceToHost));
  return PP_OK;
}

int32_t pp_tdma_batched(pp_ctx *c, const double *a, const double *b, const double *cc, const double *r, double *u,
                        int32_t n, int32_t nl, int32_t method) {
  if (!c || !a || !b || !cc || !r || !u || n < 1 || nl < 1) { set_error("pp_tdma_batched: bad argument"); return PP_BAD_ARG; }
  if (n > kMaxLine) { set_error("pp_tdma_batched: the line solvers hold at most %d rows", kMaxLine); return PP_BAD_ARG; }
  if (method != 0 && method != 1) { set_error("pp_tdma_batched: method must be 0 or 1"); return PP_BAD_ARG; }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * (size_t)n * nl;
  DevTemp dt[5];
  double *d[5];
  for (int q = 0; q < 5; q++) {
    HIPCHK(dt[q].alloc(nb));
    d[q] = dt[q].d();
  }
  std::vector<double> tmp;
  const double *src[4] = {a, b, cc, r};
  for (int q = 0; q < 4; q++) {
    if (method == 1) {
      // interleave: element m of line l at m*nl + l
      tmp.resize((size_t)n * nl);
      for (int l = 0; l < nl; l++)
        for (int m = 0; m < n; m++) tmp[(size_t)m * nl + l] = src[q][(size_t)l * n + m];
      HIPCHK(hipMemcpy(d[q], tmp.data(), nb, hipMemcpyHostToDevice));
    } else {
      HIPCHK(hipMemcpy(d[q], src[q], nb, hipMemcpyHostToDevice));
    }
  }
  hipStream_t s = c->stream;
  if (method == 0) {
    int G, S;
    line_shape(n, G, S, 8);
    dim3 grid((unsigned)(((long)nl * G + 255) / 256)), block(256);
    with_line_shape(G, S, [&](auto g, auto sz) {
      k_tdma_wave<decltype(g)::value, decltype(sz)::value><<<grid, block, 0, s>>>(n, nl, d[0], d[1], d[2], d[3], d[4]);
    });
  } else {
    int LS, P, TX;
    yz_shape(n, LS, P, TX, 8);
    dim3 grid((nl + TX - 1) / TX), block(TX, P);
    size_t sh = sizeof(double) * 8 * TX * P;
    switch (LS) {
      case 1: k_tdma_seg<1><<<grid, block, sh, s>>>(n, nl, d[0], d[1], d[2], d[3], d[4]); break;
      case 2: k_tdma_seg<2><<<grid, block, sh, s>>>(n, nl, d[0], d[1], d[2], d[3], d[4]); break;
      case 4: k_tdma_seg<4><<<grid, block, sh, s>>>(n, nl, d[0], d[1], d[2], d[3], d[4]); break;
      default: k_tdma_seg<8><<<grid, block, sh, s>>>(n, nl, d[0], d[1], d[2], d[3], d[4]); break;
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  if (method == 1) {
    tmp.resize((size_t)n * nl);
    HIPCHK(hipMemcpy(tmp.data(), d[4], nb, hipMemcpyDeviceToHost));
    for (int l = 0; l < nl; l++)
      for (int m = 0; m < n; m++) u[(size_t)l * n + m] = tmp[(size_t)m * nl + l];
  } else {
    HIPCHK(hipMemcpy(u, d[4], nb, hipMemcpyDeviceToHost));
  }
  return PP_OK;
}

int32_t pp_divergence_noalloc(pp_ctx *c, const double *d_fu, const double *d_fv, const double *d_fw, double dtinv,
                              double *d_nlu0, double *sum) {
  if (!c || !d_fu || !d_fv || !d_fw || !d_nlu0) { set_error("pp_divergence_noalloc: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_divergence_noalloc: no geometry (iblank)"); return PP_NO_GEOMETRY; }
  if (c->nblk_red > kNoAllocPartials) {
    set_error("pp_divergence_noalloc: %dx%dx%d cells need %d partial sums; the scratch array holds %ld",
              c->nxc, c->nyc, c->nzc, c->nblk_red, kNoAllocPartials);
    return PP_BAD_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  double *ptr = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!ptr) { set_error("pp_divergence_noalloc: no scratch address for device %d", c->device); return PP_HIP_ERROR; }
  double *part = ptr, *part2 = ptr + kNoAllocPartials, *res = ptr + kNoAllocPartials + 256;
  hipStream_t s = c->stream;
  // the output's outer layer and blank cells are 0, as pp_divergence leaves
  // them (a fill kernel, not a memset, so the call can be captured)
  k_fill<double><<<(unsigned)((c->nref + 255) / 256), 256, 0, s>>>(d_nlu0, c->nref, 0.0);
  CellLaunch cl = ref_launch(c);
  k_divergence<256><<<cl.grid, cl.block, 0, s>>>(c->ref, c->d_rdx[0], c->d_rdx[1], c->d_rdx[2], c->d_rdxinv[0],
                                                 c->d_rdxinv[1], c->d_rdxinv[2], c->d_iblank_ref, d_fu, d_fv, d_fw,
                                                 dtinv, d_nlu0, part);
  const int nblk = c->nblk_red, chunk = (nblk + kRedStage - 1) / kRedStage;
  k_reduce_sums_stage<<<kRedStage, 256, 0, s>>>(part, nblk, 1, chunk, part2);
  k_reduce_sums<<<1, 256, 0, s>>>(part2, kRedStage, 1, res);
  HIPCHK(hipGetLastError());
  if (sum) {
    HIPCHK(hipMemcpyAsync(c->h_res, res, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *sum = c->h_res[0];
  }
  return PP_OK;
}

int32_t pp_divergence(pp_ctx *c, const double *fu, const double *fv, const double *fw, double dtinv, double *nlu0,
                      double *sum) {
  if (!c || !fu || !fv || !fw || !nlu0) { set_error("pp_divergence: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_divergence: no geometry (iblank)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  size_t nb = sizeof(double) * c->nref;
  double *d[4] = {nullptr, nullptr, nullptr, nullptr};
  const double *src[3] = {fu, fv, fw};
  int rc = PP_OK;
  for (int q = 0; q < 4 && rc == PP_OK; q++)
    if (hipMalloc(&d[q], nb) != hipSuccess) { set_error("pp_divergence: hipMalloc failed"); rc = PP_HIP_ERROR; }
  if (rc == PP_OK) {
    for (int q = 0; q < 3; q++) hipMemcpyAsync(d[q], src[q], nb, hipMemcpyHostToDevice, c->stream);
    hipMemsetAsync(d[3], 0, nb, c->stream);
    CellLaunch cl = ref_launch(c);
    k_divergence<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, c->d_rdx[0], c->d_rdx[1], c->d_rdx[2],
                                                           c->d_rdxinv[0], c->d_rdxinv[1], c->d_rdxinv[2],
                                                           c->d_iblank_ref, d[0], d[1], d[2], dtinv, d[3],
                                                           c->d_part);
    launch_reduce_sums(c, c->nblk_red, 1);
    hipMemcpyAsync(c->h_res, c->d_res, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(nlu0, d[3], nb, hipMemcpyDeviceToHost, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) { set_error("pp_divergence: %s", hipGetErrorString(e)); rc = PP_HIP_ERROR; }
    else if (sum) *sum = c->h_res[0];
  }
  for (int q = 0; q < 4; q++) hipFree(d[q]);
  return rc;
}

int32_t pp_adjoint_device(pp_ctx *c, const double *d_wobj, double *d_lambda, double *d_dJ_dg,
                          const pp_solve_opts *o, pp_solve_stats *st) {
  if (!c || !d_wobj || !d_lambda) { set_error("pp_adjoint: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_adjoint: no geometry"); return PP_NO_GEOMETRY; }
  pp_solve_opts def;
  if (!o) { pp_default_opts(&def); o = &def; }
  if (!opts_ok(o)) { set_error("pp_adjoint: invalid options"); return PP_BAD_ARG; }
  HIPCHK(hipSetDevice(c->device));
  GhostTables &g = c->gh;
  hipStream_t s = c->stream;
  LevelArgs<double> a0 = largs<double>(c->lv[0]);
  CellLaunch cl = ref_launch(c);
  // J = wobj . M p with M the fluid-volume mean removal: s = M^T wobj,
  // S = sum of wobj over fluid and ghost cells, vsum = fluid volume
  k_fluid_sum_vol<256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, d_wobj, c->d_part);
  int rc = reduce_sums_host(c, 2);
  if (rc) return rc;
  double S = c->h_res[0];
  const double vsum = c->h_res[1];
  DevTemp dsg;
  if (g.n > 0) {
    HIPCHK(dsg.alloc(sizeof(double) * g.n));
    k_gather_idx<double><<<(g.n + 255) / 256, 256, 0, s>>>(g.n, g.g_ref, d_wobj, dsg.d());
    std::vector<double> hs(g.n);
    HIPCHK(hipMemcpyAsync(hs.data(), dsg.d(), sizeof(double) * g.n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int q = 0; q < g.n; q++) S += hs[q];
  }
  HIPCHK(hipMemsetAsync(c->d_sv, 0, sizeof(double) * c->nref, s));
  HIPCHK(hipMemsetAsync(c->d_mu, 0, sizeof(double) * c->nref, s));
  k_adj_source<256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, d_wobj, S, vsum, c->d_sv, c->d_part);
  rc = reduce_sums_host(c, 1);
  if (rc) return rc;
  const double ref = std::sqrt(c->h_res[0]);
  int lam = 0;
  rc = o->precision == PP_PREC_MIXED ? adjoint_impl<float>(c, dsg.d(), o, ref > 0 ? ref : 1.0, st, &lam)
                                     : adjoint_impl<double>(c, dsg.d(), o, ref > 0 ? ref : 1.0, st, &lam);
  if (rc != PP_OK && rc != PP_NOT_CONVERGED) return rc;
  // dJ/db = Pi^T lambda with Pi the compatibility projection:
  // lambda_f = mu V - zeta w, zeta = sum_f mu V / sum_f w
  k_mu_vol_sums<256><<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, c->d_w, c->d_part);
  int rc2 = reduce_sums_host(c, 2);
  if (rc2) return rc2;
  const double zeta = c->h_res[0] / c->h_res[1];
  HIPCHK(hipMemsetAsync(d_lambda, 0, sizeof(double) * c->nref, s));
  k_adj_lambda<<<cl.grid, cl.block, 0, s>>>(c->ref, a0, c->d_mu, c->d_w, zeta, d_lambda);
  if (d_dJ_dg && g.n > 0)
    k_ghost_combine<<<(g.n + 255) / 256, 256, 0, s>>>(g.n, g.g_ref, c->d_w, -zeta, g.olam64[lam], 1.0, g.bself,
                                                      d_dJ_dg);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return rc;
}

int32_t pp_adjoint(pp_ctx *c, const double *wobj, double *lambda, double *dJ_dg, const pp_solve_opts *o,
                   pp_solve_stats *st) {
  if (!c || !wobj || !lambda) { set_error("pp_adjoint: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_adjoint: no geometry"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp dw, dl, dg;
  HIPCHK(dw.alloc(nb));
  HIPCHK(dl.alloc(nb));
  HIPCHK(hipMemcpy(dw.d(), wobj, nb, hipMemcpyHostToDevice));
  const int ng = c->gh.n;
  if (dJ_dg && ng > 0) HIPCHK(dg.alloc(sizeof(double) * ng));
  int rc = pp_adjoint_device(c, dw.d(), dl.d(), dg.d(), o, st);
  if (rc != PP_OK && rc != PP_NOT_CONVERGED) return rc;
  HIPCHK(hipMemcpy(lambda, dl.d(), nb, hipMemcpyDeviceToHost));
  if (dJ_dg && ng > 0) HIPCHK(hipMemcpy(dJ_dg, dg.d(), sizeof(double) * ng, hipMemcpyDeviceToHost));
  return rc;
}

}  // extern "C"

// =====================================================================
// implicit momentum solve (set_solve_ad, solve_ad, calc_residual_ad)
// =====================================================================
namespace {
// cells per thread of a line segment (x-lines: 2 and 8 measured slower,
// 31 and 39 ms against 27.5 ms for the 256x192x192 momentum solve)
constexpr int kAdLS = 4, kAdLSx = 4;

FracMetrics ref_metrics(const pp_ctx *c) {
  FracMetrics M;
  for (int d = 0; d < 3; d++) { M.fx[d] = c->d_rfx[d]; M.dinv[d] = c->d_rdxinv[d]; M.dcinv[d] = c->d_rdxcinv[d]; }
  return M;
}

struct AdLineLaunch {
  dim3 grid, block;
  size_t lds;
  int nlines;
};
// TX lines x P segments per block, TX a power of two up to 32 with at most
// 512 threads (one segment per thread: P = ceil(n / LS) <= 256)
AdLineLaunch ad_line_launch(const pp_ctx *c, int dir) {
  const int n = dir == 0 ? c->nxc : (dir == 1 ? c->nyc : c->nzc);
  const int ls = dir == 0 ? kAdLSx : kAdLS;
  const int P = (n + ls - 1) / ls;
  int tx = 1;
  while (tx < 32 && 2 * tx * P <= 512) tx *= 2;
  AdLineLaunch L;
  L.block = dim3(tx, P);
  L.lds = sizeof(double) * 12 * (size_t)P * tx;  // seg_reduced_solve_n with 3 right-hand sides
  if (dir == 0) {
    L.nlines = ((c->nyc + 1) / 2) * c->nzc;
    L.grid = dim3((L.nlines + tx - 1) / tx);
  } else {
    L.nlines = 0;
    L.grid = dim3((c->nxc + 2 * tx - 1) / (2 * tx), dir == 1 ? c->nzc : c->nyc);
  }
  return L;
}

// Fresh cells (fluid now, blank at the previous geometry) take the
// GCM_set_freshcell rows (CORE_SET_SOLVERS.F90:175-218): a geometry with a
// fresh_cell mask needs the tables of pp_set_fresh_cells first.
int ad_fresh_ready(const pp_ctx *c, const char *who) {
  if (c->h_fresh.empty() || c->fr.n > 0) return PP_OK;
  set_error("%s: the geometry has fresh cells (pp_geometry.fresh_cell): call pp_set_fresh_cells with their "
            "tables (pp_geom_fresh) after pp_set_geometry", who);
  return PP_NO_GEOMETRY;
}

FreshRes fresh_res(const pp_ctx *c) {
  FreshRes r{};
  if (c->fr.n > 0) {
    r.mask = c->d_fresh;
    r.ids = c->fr.ids;
    r.coff = c->fr.coff;
    r.cc = c->fr.cc;
    r.n = c->fr.n;
  }
  return r;
}

int ad_half_dt(int32_t scheme, double dt, double *half_dt, const char *who) {
  if (scheme == PP_ADAMS_BASHFORTH2) { *half_dt = 0.0; return PP_OK; }
  if (scheme == PP_CRANK_NICOLSON1 || scheme == PP_CRANK_NICOLSON2) { *half_dt = 0.5 * dt; return PP_OK; }
  set_error("%s: unknown advec_scheme %d", who, (int)scheme);
  return PP_BAD_ARG;
}

// the rows of k_ad_rows (7 fields) for the current coefficients
void ad_build_rows(pp_ctx *c, double half_dt, const double *fu, const double *fv, const double *fw, double *row) {
  CellLaunch cl = ref_launch(c);
  k_ad_rows<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), ref_metrics(c), half_dt,
                                                      c->d_ad_coef, fu, fv, fw, c->fr.n > 0 ? c->d_fresh : nullptr,
                                                      row, c->nref);
}

// itsolv_ad of u, v and w: x, y and z line sweeps, two zebra passes each
// (the three fields in each pass); with fresh cells F.r are the call's
// copies of the right-hand sides that k_fresh_rhs refreshes before each pass
void ad_sweeps(pp_ctx *c, const double *row, AdLineFields F, const FreshCorr *fc) {
  LevelArgs<double> L0 = largs<double>(c->lv[0]);
  for (int dir = 0; dir < 3; dir++) {
    AdLineLaunch L = ad_line_launch(c, dir);
    for (int color = 0; color < 2; color++) {
      if (fc)
        k_fresh_rhs<<<(3 * c->fr.n + 255) / 256, 256, 0, c->stream>>>(c->fr.n, 3, c->fr.ids, c->fr.coff, c->fr.cc,
                                                                      *fc);
      if (dir == 0)
        k_ad_line<0, kAdLSx><<<L.grid, L.block, L.lds, c->stream>>>(c->ref, L0, row, c->nref, F, color, L.nlines);
      else if (dir == 1)
        k_ad_line<1, kAdLS><<<L.grid, L.block, L.lds, c->stream>>>(c->ref, L0, row, c->nref, F, color, L.nlines);
      else
        k_ad_line<2, kAdLS><<<L.grid, L.block, L.lds, c->stream>>>(c->ref, L0, row, c->nref, F, color, L.nlines);
    }
  }
}

// calc_residual_ad of nv fields: h_res[3q + 1] = max |res|, h_res[3q + 2]
// = its reference-layout offset (first in k, j, i order), rout the field
// of the first when not NULL
int ad_residual(pp_ctx *c, const double *row, int nv, AdFields F, double *rout) {
  CellLaunch cl = ref_launch(c);
  const long nblk = c->nblk_red;
  LevelArgs<double> L0 = largs<double>(c->lv[0]);
  hipStream_t s = c->stream;
  const FreshRes FR = fresh_res(c);
  if (nv == 3)
    k_ad_residual<256, 3><<<cl.grid, cl.block, 0, s>>>(c->ref, L0, row, c->nref, F, FR, c->d_part, nblk, rout);
  else
    k_ad_residual<256, 1><<<cl.grid, cl.block, 0, s>>>(c->ref, L0, row, c->nref, F, FR, c->d_part, nblk, rout);
  const int chunk = (int)((nblk + kRedStage - 1) / kRedStage);
  for (int q = 0; q < nv; q++) {
    k_reduce_summax_stage<<<kRedStage, 256, 0, s>>>(c->d_part + 3 * nblk * q, (int)nblk, chunk,
                                                   c->d_part2 + 3 * kRedStage * q);
    k_reduce_summax<<<1, 256, 0, s>>>(c->d_part2 + 3 * kRedStage * q, kRedStage, c->d_res + 3 * q);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_res, c->d_res, 3 * nv * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return PP_OK;
}

void ad_loc(const pp_ctx *c, double mx, double arg, int32_t loc[3]) {
  if (mx > 0.0 && arg >= 0.0) {
    const long q = (long)arg;
    loc[0] = (int32_t)(q % c->ref.ex);
    loc[1] = (int32_t)((q / c->ref.ex) % (c->nyc + 3));
    loc[2] = (int32_t)(q / c->ref.exy);
  } else {
    loc[0] = loc[1] = loc[2] = 0;
  }
}
}  // namespace

extern "C" {

int32_t pp_set_fresh_cells(pp_ctx *c, int32_t n, const int32_t *ifr, const int32_t *jfr, const int32_t *kfr,
                           const int32_t *ic, const int32_t *jc, const int32_t *kc, const double *coeffD) {
  if (!c || n < 0 || (n > 0 && (!ifr || !jfr || !kfr || !ic || !jc || !kc || !coeffD))) {
    set_error("pp_set_fresh_cells: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_set_fresh_cells: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  c->have_ad = false;  // the coefficients take the new rows in pp_set_solve_ad
  free_fresh_rows(c);
  long nmask = 0;
  for (long r = 0; r < (long)c->h_fresh.size(); r++) nmask += c->h_fresh[r] == 1;
  if (nmask != n) {
    set_error("pp_set_fresh_cells: the geometry has %ld fresh cells, %d given", nmask, n);
    return PP_BAD_ARG;
  }
  if (n == 0) return PP_OK;
  // stencil order of coeffGCMFreshD (incI, incJ, incK of the ghost stencils)
  static const int inc[3][8] = {{0, 0, 1, 1, 0, 0, 1, 1}, {0, 1, 1, 0, 0, 1, 1, 0}, {0, 0, 0, 0, 1, 1, 1, 1}};
  const int nc[3] = {c->nxc, c->nyc, c->nzc};
  std::vector<int> order(n);
  std::vector<long> off(n);
  for (int q = 0; q < n; q++) {
    const int cell[3] = {ifr[q], jfr[q], kfr[q]}, lo[3] = {ic[q], jc[q], kc[q]};
    for (int d = 0; d < 3; d++)
      if (cell[d] < 1 || cell[d] > nc[d] || lo[d] < cell[d] - 1 || lo[d] > cell[d] || lo[d] < 1) {
        set_error("pp_set_fresh_cells: fresh cell %d (%d,%d,%d) with stencil corner (%d,%d,%d): the corner must "
                  "be the cell or the one below it", q, cell[0], cell[1], cell[2], lo[0], lo[1],
                  lo[2]);
        return PP_BAD_ARG;
      }
    off[q] = (long)cell[0] + c->ref.ex * cell[1] + c->ref.exy * cell[2];
    if (c->h_fresh[off[q]] != 1) {
      set_error("pp_set_fresh_cells: cell (%d,%d,%d) is not a fresh cell of the geometry", cell[0], cell[1], cell[2]);
      return PP_BAD_ARG;
    }
    for (int m = 0; m < 8; m++)
      if (!std::isfinite(coeffD[8 * q + m])) {
        set_error("pp_set_fresh_cells: coefficient %d of fresh cell %d is not finite", m, q);
        return PP_BAD_ARG;
      }
    order[q] = q;
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return off[a] < off[b]; });
  for (int q = 1; q < n; q++)
    if (off[order[q]] == off[order[q - 1]]) {
      set_error("pp_set_fresh_cells: cell (%d,%d,%d) given twice", ifr[order[q]], jfr[order[q]], kfr[order[q]]);
      return PP_BAD_ARG;
    }
  std::vector<long> h_ids(n), h_coff(4 * (size_t)n);
  std::vector<double> h_face(6 * (size_t)n, 0.0), h_cc(4 * (size_t)n);
  for (int t = 0; t < n; t++) {
    const int q = order[t];
    const int cell[3] = {ifr[q], jfr[q], kfr[q]}, lo[3] = {ic[q], jc[q], kc[q]};
    h_ids[t] = off[q];
    int ncorner = 0;
    for (int m = 0; m < 8; m++) {
      int node[3], ndiff = 0, dd = -1;
      for (int d = 0; d < 3; d++) {
        node[d] = lo[d] + inc[d][m];
        if (node[d] != cell[d]) { ndiff++; dd = d; }
      }
      const double cD = coeffD[8 * q + m];
      if (ndiff == 1)  // face neighbour: amx_ad/apx_ad/... = -coeffGCMFreshD
        h_face[6 * t + 2 * dd + (node[dd] > cell[dd] ? 1 : 0)] = -cD;
      else if (ndiff >= 2) {
        h_coff[4 * t + ncorner] = (long)node[0] + c->ref.ex * node[1] + c->ref.exy * node[2];
        h_cc[4 * t + ncorner] = cD;
        ncorner++;
      }
    }
  }
  FreshRows &F = c->fr;
  HIPCHK(hipMalloc(&F.ids, sizeof(long) * n));
  HIPCHK(hipMalloc(&F.face, sizeof(double) * 6 * n));
  HIPCHK(hipMalloc(&F.coff, sizeof(long) * 4 * n));
  HIPCHK(hipMalloc(&F.cc, sizeof(double) * 4 * n));
  HIPCHK(hipMemcpy(F.ids, h_ids.data(), sizeof(long) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(F.face, h_face.data(), sizeof(double) * 6 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(F.coff, h_coff.data(), sizeof(long) * 4 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(F.cc, h_cc.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice));
  F.n = n;
  return PP_OK;
}

int32_t pp_set_solve_ad_device(pp_ctx *c, double dt, const double *d_visc, const double *d_bxv, const double *d_byv,
                               const double *d_bzv) {
  if (!c || !d_visc || !d_bxv || !d_byv || !d_bzv) { set_error("pp_set_solve_ad: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_set_solve_ad: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  if (!std::isfinite(dt)) { set_error("pp_set_solve_ad: dt must be finite"); return PP_BAD_ARG; }
  if (int rf = ad_fresh_ready(c, "pp_set_solve_ad")) return rf;
  HIPCHK(hipSetDevice(c->device));
  c->have_ad = false;
  const size_t nb = sizeof(double) * 6 * (size_t)c->nref;
  if (!c->d_ad_coef) HIPCHK(hipMalloc(&c->d_ad_coef, nb));
  HIPCHK(hipMemsetAsync(c->d_ad_coef, 0, nb, c->stream));
  CellLaunch cl = ref_launch(c);
  k_set_solve_ad<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), ref_metrics(c), dt,
                                                           d_visc, d_bxv, d_byv, d_bzv, c->d_ad_coef, c->nref);
  if (c->fr.n > 0)
    k_fresh_ad_coef<<<(c->fr.n + 255) / 256, 256, 0, c->stream>>>(c->fr.n, c->fr.ids, c->fr.face, c->d_ad_coef,
                                                                 c->nref);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ad_dt = dt;
  c->have_ad = true;
  return PP_OK;
}

int32_t pp_set_solve_ad(pp_ctx *c, double dt, const double *visc, const double *bxv, const double *byv,
                        const double *bzv) {
  if (!c || !visc || !bxv || !byv || !bzv) { set_error("pp_set_solve_ad: null argument"); return PP_BAD_ARG; }
  if (!c->have_geom) { set_error("pp_set_solve_ad: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  if (int rf = ad_fresh_ready(c, "pp_set_solve_ad")) return rf;
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp d;
  HIPCHK(d.alloc(4 * nb));
  const double *hin[4] = {visc, bxv, byv, bzv};
  for (int q = 0; q < 4; q++) HIPCHK(hipMemcpy(d.d() + q * c->nref, hin[q], nb, hipMemcpyHostToDevice));
  return pp_set_solve_ad_device(c, dt, d.d(), d.d() + c->nref, d.d() + 2 * c->nref, d.d() + 3 * c->nref);
}

}  // extern "C"

namespace {
// the arguments of face_vel for a CRANK_NICOLSON2 solve (device pointers)
struct FaceVelArgs {
  int van_kan;
  double dt;
  const double *p, *bcxu, *bcyv, *bczw;
};

void face_vel_launch(pp_ctx *c, const FaceVelArgs &a, const double *u, const double *v, const double *w, double *fu,
                     double *fv, double *fw) {
  FracMetrics M;
  for (int d = 0; d < 3; d++) { M.fx[d] = c->d_rfx[d]; M.dinv[d] = c->d_rdxinv[d]; M.dcinv[d] = c->d_rdxcinv[d]; }
  CellLaunch cl = ref_launch(c);
  k_face_vel<256><<<cl.grid, cl.block, 0, c->stream>>>(c->ref, largs<double>(c->lv[0]), M, a.van_kan, a.dt, a.p, u,
                                                       v, w, a.bcxu, a.bcyv, a.bczw, fu, fv, fw);
}

// The GCM loop of solve_ad (CORE_SOLVE_AD.F90:218-247) on device fields;
// with cn2 the face velocities are recomputed by face_vel after each ghost
// update (CRANK_NICOLSON2) and written to d_fu, d_fv, d_fw.
int32_t solve_ad_run(pp_ctx *c, int32_t advec_scheme, double *d_u, double *d_v, double *d_w, const double *d_nlu,
                     const double *d_nlv, const double *d_nlw, double *d_fu, double *d_fv, double *d_fw,
                     const FaceVelArgs *cn2, const double *d_vbi, int32_t itermax, double restol, pp_ad_stats *st) {
  if (!c || !d_u || !d_v || !d_w || !d_nlu || !d_nlv || !d_nlw || !d_fu || !d_fv || !d_fw) {
    set_error("pp_solve_ad: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_solve_ad: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  if (!c->have_ad) {
    set_error("pp_solve_ad: no momentum coefficients (call pp_set_solve_ad after pp_set_geometry)");
    return PP_NO_GEOMETRY;
  }
  if (c->gh.n > 0 && !c->gh.have_vel) {
    set_error("pp_solve_ad: no velocity ghost coefficients (call pp_set_vel_ghosts after pp_set_geometry)");
    return PP_NO_GEOMETRY;
  }
  if (itermax < 0 || !std::isfinite(restol)) { set_error("pp_solve_ad: bad itermax_ad or restol_ad"); return PP_BAD_ARG; }
  if (int rf = ad_fresh_ready(c, "pp_solve_ad")) return rf;
  if (!cn2 && advec_scheme == PP_CRANK_NICOLSON2) {
    set_error("pp_solve_ad: CRANK_NICOLSON2 recomputes the face velocities after every ghost update (face_vel): "
              "call pp_solve_ad_cn2 with the pressure and the outer-face velocities");
    return PP_BAD_ARG;
  }
  double half_dt = 0.0;
  int rc = ad_half_dt(advec_scheme, c->ad_dt, &half_dt, "pp_solve_ad");
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  DevTemp rows;
  HIPCHK(rows.alloc(sizeof(double) * 7 * (size_t)c->nref));
  // the call's timing events, destroyed on every return path
  struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() { if (a) hipEventDestroy(a); if (b) hipEventDestroy(b); }
  } ev;
  HIPCHK(hipEventCreate(&ev.a));
  HIPCHK(hipEventCreate(&ev.b));
  hipEvent_t e0 = ev.a, e1 = ev.b;
  HIPCHK(hipEventRecord(e0, s));
  ad_build_rows(c, half_dt, d_fu, d_fv, d_fw, rows.d());
  double *vars[3] = {d_u, d_v, d_w};
  const double *rhs[3] = {d_nlu, d_nlv, d_nlw};
  AdFields F;
  for (int q = 0; q < 3; q++) { F.var[q] = vars[q]; F.r[q] = rhs[q]; }
  // fresh cells: the line passes read copies of nlu, nlv, nlw whose fresh
  // entries carry the off-stencil terms (k_fresh_rhs)
  DevTemp rw;
  FreshCorr fc{};
  const bool fresh = c->fr.n > 0;
  if (fresh) {
    const size_t nb = sizeof(double) * (size_t)c->nref;
    HIPCHK(rw.alloc(3 * nb));
    for (int q = 0; q < 3; q++) {
      fc.var[q] = vars[q];
      fc.r[q] = rhs[q];
      fc.rw[q] = rw.d() + q * c->nref;
      HIPCHK(hipMemcpyAsync(fc.rw[q], rhs[q], nb, hipMemcpyDeviceToDevice, s));
    }
  }
  int it = 0;
  double maxres = 1.0e10, ruvw[3] = {0.0, 0.0, 0.0};
  int32_t loc[3] = {0, 0, 0};
  while (it < itermax && maxres > restol) {
    AdLineFields LF;
    for (int q = 0; q < 3; q++) { LF.var[q] = vars[q]; LF.r[q] = fresh ? fc.rw[q] : rhs[q]; }
    ad_sweeps(c, rows.d(), LF, fresh ? &fc : nullptr);
    vel_ghosts(c, d_u, d_v, d_w, d_vbi);
    if (cn2) {  // face_vel, then the rows of the new face velocities
      face_vel_launch(c, *cn2, d_u, d_v, d_w, d_fu, d_fv, d_fw);
      ad_build_rows(c, half_dt, d_fu, d_fv, d_fw, rows.d());
    }
    rc = ad_residual(c, rows.d(), 3, F, nullptr);
    if (rc) break;
    int qm = 0;
    for (int q = 0; q < 3; q++) {
      ruvw[q] = c->h_res[3 * q + 1];
      if (ruvw[q] > ruvw[qm]) qm = q;
    }
    maxres = ruvw[qm];
    ad_loc(c, c->h_res[3 * qm + 1], c->h_res[3 * qm + 2], loc);
    it++;
  }
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  if (rc) return rc;
  const int32_t status = maxres > restol ? PP_NOT_CONVERGED : PP_OK;
  if (st) {
    st->status = status;
    st->iterations = it;
    st->max_residual = maxres;
    for (int q = 0; q < 3; q++) { st->residual_uvw[q] = ruvw[q]; st->loc[q] = loc[q]; }
    st->ms = ms;
  }
  if (status != PP_OK) set_error("pp_solve_ad: residual %g after %d iterations", maxres, it);
  return status;
}

}  // namespace

extern "C" {

int32_t pp_solve_ad_device(pp_ctx *c, int32_t advec_scheme, double *d_u, double *d_v, double *d_w,
                           const double *d_nlu, const double *d_nlv, const double *d_nlw, const double *d_fu,
                           const double *d_fv, const double *d_fw, const double *d_vbi, int32_t itermax,
                           double restol, pp_ad_stats *st) {
  // the face velocities are read only (no CRANK_NICOLSON2 here)
  return solve_ad_run(c, advec_scheme, d_u, d_v, d_w, d_nlu, d_nlv, d_nlw, const_cast<double *>(d_fu),
                      const_cast<double *>(d_fv), const_cast<double *>(d_fw), nullptr, d_vbi, itermax, restol, st);
}

int32_t pp_face_vel_device(pp_ctx *c, int32_t van_kan, double dt, const double *d_p, const double *d_u,
                           const double *d_v, const double *d_w, const double *d_bcxu, const double *d_bcyv,
                           const double *d_bczw, double *d_fu, double *d_fv, double *d_fw) {
  if (!c || (van_kan && !d_p) || !d_u || !d_v || !d_w || !d_bcxu || !d_bcyv || !d_bczw || !d_fu || !d_fv || !d_fw) {
    set_error("pp_face_vel: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_face_vel: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  if (!std::isfinite(dt)) { set_error("pp_face_vel: dt must be finite"); return PP_BAD_ARG; }
  HIPCHK(hipSetDevice(c->device));
  FaceVelArgs a{van_kan ? 1 : 0, dt, d_p ? d_p : d_u, d_bcxu, d_bcyv, d_bczw};
  face_vel_launch(c, a, d_u, d_v, d_w, d_fu, d_fv, d_fw);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return PP_OK;
}

int32_t pp_face_vel(pp_ctx *c, int32_t van_kan, double dt, const double *p, const double *u, const double *v,
                    const double *w, const double *bcxu, const double *bcyv, const double *bczw, double *fu,
                    double *fv, double *fw) {
  if (!c || (van_kan && !p) || !u || !v || !w || !bcxu || !bcyv || !bczw || !fu || !fv || !fw) {
    set_error("pp_face_vel: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_face_vel: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp d;
  HIPCHK(d.alloc(10 * nb));
  double *dv[10];
  for (int q = 0; q < 10; q++) dv[q] = d.d() + q * c->nref;
  const double *hin[10] = {p ? p : u, u, v, w, bcxu, bcyv, bczw, fu, fv, fw};
  for (int q = 0; q < 10; q++) HIPCHK(hipMemcpy(dv[q], hin[q], nb, hipMemcpyHostToDevice));
  int rc = pp_face_vel_device(c, van_kan, dt, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7], dv[8], dv[9]);
  if (rc == PP_OK) {
    double *hout[3] = {fu, fv, fw};
    for (int q = 0; q < 3; q++) HIPCHK(hipMemcpy(hout[q], dv[7 + q], nb, hipMemcpyDeviceToHost));
  }
  return rc;
}

int32_t pp_solve_ad_cn2_device(pp_ctx *c, double *d_u, double *d_v, double *d_w, const double *d_nlu,
                               const double *d_nlv, const double *d_nlw, double *d_fu, double *d_fv, double *d_fw,
                               int32_t van_kan, const double *d_p, const double *d_bcxu, const double *d_bcyv,
                               const double *d_bczw, const double *d_vbi, int32_t itermax, double restol,
                               pp_ad_stats *st) {
  if (!c || (van_kan && !d_p) || !d_bcxu || !d_bcyv || !d_bczw) {
    set_error("pp_solve_ad_cn2: null argument");
    return PP_BAD_ARG;
  }
  FaceVelArgs a{van_kan ? 1 : 0, c->ad_dt, d_p ? d_p : d_bcxu, d_bcxu, d_bcyv, d_bczw};
  return solve_ad_run(c, PP_CRANK_NICOLSON2, d_u, d_v, d_w, d_nlu, d_nlv, d_nlw, d_fu, d_fv, d_fw, &a, d_vbi,
                      itermax, restol, st);
}

int32_t pp_solve_ad_cn2(pp_ctx *c, double *u, double *v, double *w, const double *nlu, const double *nlv,
                        const double *nlw, double *fu, double *fv, double *fw, int32_t van_kan, const double *p,
                        const double *bcxu, const double *bcyv, const double *bczw, const double *vbi,
                        int32_t itermax, double restol, pp_ad_stats *st) {
  if (!c || !u || !v || !w || !nlu || !nlv || !nlw || !fu || !fv || !fw || (van_kan && !p) || !bcxu || !bcyv ||
      !bczw) {
    set_error("pp_solve_ad_cn2: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_solve_ad: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp d, dvbi;
  HIPCHK(d.alloc(13 * nb));
  double *dv[13];
  for (int q = 0; q < 13; q++) dv[q] = d.d() + q * c->nref;
  const double *hin[13] = {u, v, w, nlu, nlv, nlw, fu, fv, fw, p ? p : u, bcxu, bcyv, bczw};
  for (int q = 0; q < 13; q++) HIPCHK(hipMemcpy(dv[q], hin[q], nb, hipMemcpyHostToDevice));
  if (vbi && c->gh.n > 0) {
    HIPCHK(dvbi.alloc(sizeof(double) * 3 * c->gh.n));
    HIPCHK(hipMemcpy(dvbi.d(), vbi, sizeof(double) * 3 * c->gh.n, hipMemcpyHostToDevice));
  }
  int rc = pp_solve_ad_cn2_device(c, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7], dv[8], van_kan, dv[9],
                                  dv[10], dv[11], dv[12], dvbi.d(), itermax, restol, st);
  if (rc == PP_OK || rc == PP_NOT_CONVERGED) {
    double *hout[6] = {u, v, w, fu, fv, fw};
    for (int q = 0; q < 6; q++) HIPCHK(hipMemcpy(hout[q], dv[q < 3 ? q : q + 3], nb, hipMemcpyDeviceToHost));
  }
  return rc;
}

int32_t pp_solve_ad(pp_ctx *c, int32_t advec_scheme, double *u, double *v, double *w, const double *nlu,
                    const double *nlv, const double *nlw, const double *fu, const double *fv, const double *fw,
                    const double *vbi, int32_t itermax, double restol, pp_ad_stats *st) {
  if (!c || !u || !v || !w || !nlu || !nlv || !nlw || !fu || !fv || !fw) {
    set_error("pp_solve_ad: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_solve_ad: no geometry (call pp_set_geometry first)"); return PP_NO_GEOMETRY; }
  if (int rf = ad_fresh_ready(c, "pp_solve_ad")) return rf;
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp d, dvbi;
  HIPCHK(d.alloc(9 * nb));
  double *dv[9];
  for (int q = 0; q < 9; q++) dv[q] = d.d() + q * c->nref;
  const double *hin[9] = {u, v, w, nlu, nlv, nlw, fu, fv, fw};
  for (int q = 0; q < 9; q++) HIPCHK(hipMemcpy(dv[q], hin[q], nb, hipMemcpyHostToDevice));
  if (vbi && c->gh.n > 0) {
    HIPCHK(dvbi.alloc(sizeof(double) * 3 * c->gh.n));
    HIPCHK(hipMemcpy(dvbi.d(), vbi, sizeof(double) * 3 * c->gh.n, hipMemcpyHostToDevice));
  }
  int rc = pp_solve_ad_device(c, advec_scheme, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7], dv[8],
                              dvbi.d(), itermax, restol, st);
  if (rc == PP_OK || rc == PP_NOT_CONVERGED) {
    double *hout[3] = {u, v, w};
    for (int q = 0; q < 3; q++) HIPCHK(hipMemcpy(hout[q], dv[q], nb, hipMemcpyDeviceToHost));
  }
  return rc;
}

int32_t pp_calc_residual_ad(pp_ctx *c, int32_t advec_scheme, const double *var, const double *r, const double *fu,
                            const double *fv, const double *fw, double *res, double *resm, int32_t loc[3]) {
  if (!c || !var || !r || !fu || !fv || !fw || !resm || !loc) {
    set_error("pp_calc_residual_ad: null argument");
    return PP_BAD_ARG;
  }
  if (!c->have_geom) { set_error("pp_calc_residual_ad: no geometry"); return PP_NO_GEOMETRY; }
  if (!c->have_ad) { set_error("pp_calc_residual_ad: no momentum coefficients (call pp_set_solve_ad)"); return PP_NO_GEOMETRY; }
  if (int rf = ad_fresh_ready(c, "pp_calc_residual_ad")) return rf;
  double half_dt = 0.0;
  int rc = ad_half_dt(advec_scheme, c->ad_dt, &half_dt, "pp_calc_residual_ad");
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  const size_t nb = sizeof(double) * c->nref;
  DevTemp d, rows;
  HIPCHK(d.alloc(6 * nb));
  HIPCHK(rows.alloc(7 * nb));
  double *dv[6];
  for (int q = 0; q < 6; q++) dv[q] = d.d() + q * c->nref;
  const double *hin[5] = {var, r, fu, fv, fw};
  for (int q = 0; q < 5; q++) HIPCHK(hipMemcpy(dv[q], hin[q], nb, hipMemcpyHostToDevice));
  HIPCHK(hipMemsetAsync(dv[5], 0, nb, c->stream));
  ad_build_rows(c, half_dt, dv[2], dv[3], dv[4], rows.d());
  AdFields F{};
  F.var[0] = dv[0];
  F.r[0] = dv[1];
  rc = ad_residual(c, rows.d(), 1, F, dv[5]);
  if (rc) return rc;
  *resm = c->h_res[1];
  ad_loc(c, c->h_res[1], c->h_res[2], loc);
  if (res) HIPCHK(hipMemcpy(res, dv[5], nb, hipMemcpyDeviceToHost));
  return PP_OK;
}

}  // extern "C"
