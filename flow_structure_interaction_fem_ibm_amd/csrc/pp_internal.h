// Internal declarations of the MI355X pressure-Poisson library.
//
// Device data layout (one level of the hierarchy, FP32 or FP64):
//   cell (i,j,k), 0-based, i in [-1,cx], j in [-1,cy], k in [-1,cz]
//   (one halo cell each side) lives at
//       OX + i + sx*(j+1) + sxy*(k+1)
//   with OX = 4 and the row pitch sx a multiple of 16 elements, so the
//   interior of every row starts 16-byte aligned and every row starts
//   64-byte aligned.  One flag byte per cell uses the same offsets:
//       bit0 blank (iblank=1), bit1 ghost (ghostCellMark=1),
//       bit2..7 face closed (Neumann) on the i-,i+,j-,j+,k-,k+ face.
// The FP64 solution and RHS stay in the caller's reference layout
// ((0:nx+1,0:ny+1,0:nz+1), i fastest) so no layout copy is needed at the
// boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pp_hip.h"

namespace pp {

constexpr int OX = 4;
constexpr uint8_t FL_BLANK = 1, FL_GHOST = 2, FL_IM = 4, FL_IP = 8,
                  FL_JM = 16, FL_JP = 32, FL_KM = 64, FL_KP = 128;

__host__ __device__ inline long lidx(int sx, long sxy, int i, int j, int k) {
  return (long)OX + i + (long)sx * (j + 1) + sxy * (k + 1);
}

// Colour-split rows of the finest level's FP32 smoothing arrays: row (j,k)
// holds its even cells contiguously from SPLIT_E0 + i/2 and its odd cells
// from SPLIT_E0 + hp + (i-1)/2 (the halo cell i = -1 sits just below the
// odd half).  A zebra colour of y- or z-lines then owns one half of every
// row: whole cache lines read and written instead of every other element.
constexpr int SPLIT_E0 = 32;
__host__ __device__ inline long sidx(int sx, long sxy, int hp, int i, int j, int k) {
  return (long)sx * (j + 1) + sxy * (k + 1) + SPLIT_E0 + ((i & 1) ? hp + ((i - 1) >> 1) : (i >> 1));
}

// Per-level device view passed by value to kernels.
template <typename T>
struct LevelArgs {
  int cx, cy, cz;
  int sx;
  long sxy;
  const uint8_t *fl;
  const T *cmx, *cpx, *cmy, *cpy, *cmz, *cpz;  // face coefficients per cell
  const T *wx, *wy, *wz;                        // cell widths
};

// Prolongation map of one direction of a fine level onto its coarse parent.
template <typename T>
struct ProlongDir {
  const int *par, *oth;   // parent and second coarse neighbour per fine index
  const T *wpar;          // weight of the parent
};

// Reference-layout view of the finest level (FP64 fields of the caller).
struct RefArgs {
  int nxc, nyc, nzc;
  long ex, exy;           // extents nxc+3, (nxc+3)*(nyc+3)
};
__host__ __device__ inline long ridx(const RefArgs &r, int i, int j, int k) {
  // 0-based interior cell (i,j,k) = reference (i+1,j+1,k+1)
  return (long)(i + 1) + r.ex * (j + 1) + r.exy * (k + 1);
}

struct Level {
  int cx = 0, cy = 0, cz = 0, sx = 0, sy = 0;
  long sxy = 0, total = 0;
  int cf[3] = {1, 1, 1};  // coarsening factor of this level w.r.t. the previous
  std::vector<double> node[3], xc[3], w[3], cm[3], cp[3];
  // maps from this (fine) level onto level+1
  std::vector<int> par[3], oth[3];
  std::vector<double> wpar[3];
  std::vector<uint8_t> hflags;
  // device
  uint8_t *flags = nullptr;
  double *cm64[3] = {}, *cp64[3] = {}, *w64[3] = {}, *wpar64[3] = {};
  float *cm32[3] = {}, *cp32[3] = {}, *w32[3] = {}, *wpar32[3] = {};
  int *par_d[3] = {}, *oth_d[3] = {};
  float *e32 = nullptr, *rhs32 = nullptr, *res32 = nullptr;
  double *e64 = nullptr, *rhs64 = nullptr, *res64 = nullptr;
};

struct GhostTables {
  int n = 0;
  int kexact = 1;                 // Jacobi sweeps that solve the ghost rows
  // forward gather (device)
  long *g_ref = nullptr, *g_al = nullptr;       // ghost cell offsets
  long *g_as = nullptr, *c_as = nullptr;        // ghost and fluid offsets in the split level 0
  long *nb_as = nullptr, *ct_as = nullptr;      // the adjoint tables' offsets in the split level 0
  double *bself = nullptr;                      // c_self + probeLength
  double *bnd = nullptr;                        // bself * uNormal (per solve)
  // composite ghost rows: the solved ghost value as a combination of fluid
  // values, g = sum_q w_q p[f_q] + bnd_eff (ghost-ghost coupling folded in)
  int *c_ptr = nullptr;                         // CSR by ghost (n+1)
  long *c_ref = nullptr, *c_al = nullptr;       // fluid cell offsets
  double *c_w64 = nullptr;
  float *c_w32 = nullptr;
  double *bnd_eff = nullptr;                    // (I - C_gg)^-1 bnd (per solve)
  std::vector<int> h_c_ptr, h_gg_ptr, h_gg_idx;  // h_gg: ghost-ghost coupling CSR
  std::vector<long> h_c_ref, h_c_al;
  std::vector<double> h_c_w, h_gg_w;
  // adjoint: fluid neighbours of each ghost and their face conductances
  // (coefficient of the ghost in the neighbour's row times its volume)
  long *nb_ref = nullptr, *nb_al = nullptr;     // 6 per ghost (-1: none)
  double *nb_w64 = nullptr;
  float *nb_w32 = nullptr;
  // transpose of the stencil: touched cells (fluid or ghost) and the
  // ghosts whose stencils read them, CSR by touched cell
  int nt = 0;
  long *t_ref = nullptr, *t_al = nullptr;
  int *t_ptr = nullptr, *t_ghost = nullptr;
  double *t_coef64 = nullptr, *t_vinv64 = nullptr;
  float *t_coef32 = nullptr, *t_vinv32 = nullptr;  // vinv = 1/V (fluid), 0 (ghost)
  int *g_tidx = nullptr;                        // touched index of a ghost cell or -1
  // transpose of the composite rows, CSR by fluid cell: the solved
  // transposed ghost coupling of the fluid rows in one gather,
  // out_f = base_f + vinv_f sum_g M[g,f] r_g with r_g the ghost sources
  int nct = 0;
  long *ct_ref = nullptr, *ct_al = nullptr;
  int *ct_ptr = nullptr, *ct_ghost = nullptr;
  double *ct_w64 = nullptr, *ct_vinv64 = nullptr, *ctbase64 = nullptr;
  float *ct_w32 = nullptr, *ct_vinv32 = nullptr, *ctbase32 = nullptr;
  std::vector<long> h_ct_ref, h_ct_al;
  std::vector<int> h_ct_ptr, h_ct_ghost;
  std::vector<double> h_ct_w, h_ct_vinv;
  // adjoint work buffers (inner V-cycle, per precision) and the FP64
  // outer loop's own copies
  double *lam64[2] = {nullptr, nullptr}, *rhsg64 = nullptr, *tbase64 = nullptr;
  double *olam64[2] = {nullptr, nullptr}, *orhsg64 = nullptr, *otbase64 = nullptr;
  float *lam32[2] = {nullptr, nullptr}, *rhsg32 = nullptr, *tbase32 = nullptr;
  // host copies
  std::vector<long> h_g_ref, h_g_al, h_s_ref, h_s_al, h_nb_ref, h_nb_al, h_t_ref, h_t_al;
  std::vector<double> h_coef, h_bself, h_nb_w, h_t_coef, h_t_vinv;
  std::vector<int> h_t_ptr, h_t_ghost, h_g_tidx;
  std::vector<int> h_ig, h_jg, h_kg;            // 1-based reference indices
  int solid_in_stencil = 0;                     // stencil entries on solid non-ghost cells
  // velocity ghost rows (GCM_vel_set_bc_internal), set by pp_set_vel_ghosts:
  // Dirichlet coefficients, per stencil entry the ghost index it reads
  // (>= 0), -1 for the field at v_off, -2 for the body-intercept value
  std::vector<long> h_s_raw;                    // stencil offsets, outer layer not clamped
  double *v_cD = nullptr, *v_buf[2] = {nullptr, nullptr};
  int *v_kind = nullptr;
  long *v_off = nullptr;
  int v_sweeps = 0;                             // Jacobi sweeps that reach round-off
  double v_bw = 0.0, v_iw = 0.0;
  bool have_vel = false;
};

// fresh cells of the momentum solve (pp_set_fresh_cells)
struct FreshRows {
  int n = 0;
  long *ids = nullptr;      // reference offsets, ascending
  double *face = nullptr;   // 6 per cell: the fresh row's amx_ad..apz_ad
  long *coff = nullptr;     // 4 per cell: offsets of the off-stencil nodes, stencil order
  double *cc = nullptr;     // 4 per cell: their coeffGCMFreshD
};

// the colour-split copy of level 0 used by the FP32 V-cycles, forward and
// transposed (flags, correction, right-hand side, null-space shape; the
// level-0 FP32 arrays stay in the natural layout for everything else)
struct Split0 {
  bool ok = false;    // level 0 qualifies (halves by 2, even x-line chunks)
  bool on = false;    // set while FP32 cycles run on the copy
  int hp = 0, sx = 0;
  long sxy = 0, total = 0;
  float *e = nullptr, *rhs = nullptr;
  float *nshape = nullptr;  // w/V of the null vector (the adjoint cycles' level-0 correction scale)
  uint8_t *fl = nullptr;
};

struct CoarseSolve {
  int n = 0;                      // cells of the coarsest level (at most 4096)
  double *G = nullptr;            // n*n by output cell: G[q*n+p] = d e_q / d r_p (FP64 in both precisions)
  int *cell = nullptr;            // level-layout offset of coarse cell q (i fastest)
};

}  // namespace pp

struct pp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nxc = 0, nyc = 0, nzc = 0;
  pp::RefArgs ref{};
  long nref = 0;                       // reference field size
  std::vector<double> hx[3];           // reference node coords x(0..n+1)
  std::vector<double> rxc[3], rdx[3], rdxinv[3], rdxc[3], rdxcinv[3];
  std::vector<pp::Level> lv;
  bool have_geom = false;
  std::vector<int8_t> h_iblank, h_fresh;  // reference layout
  pp::GhostTables gh;
  pp::CoarseSolve cs;
  pp::Split0 sp;
  void *blas = nullptr;                // rocblas_handle of the coarse factorisation
  double *d_rdx[3] = {}, *d_rdxinv[3] = {};  // make_grid dx, dxinv (0:n+1) per direction
  double *d_rdxcinv[3] = {}, *d_rfx[3] = {};  // dxcinv and fx (0:n+1) per direction
  // FP64 work arrays in reference layout
  double *d_p = nullptr, *d_b = nullptr, *d_w = nullptr, *d_bwork = nullptr;
  double *d_tmp = nullptr;
  int8_t *d_fresh = nullptr;
  int8_t *d_iblank_ref = nullptr;
  unsigned long long *d_count = nullptr;  // set-up counter (ghost cells of the iblank)
  // reductions
  double *d_part = nullptr;            // partial sums
  double *d_part2 = nullptr;           // first-stage results of the two-stage reductions
  double *d_res = nullptr;             // final reduction results
  double *h_res = nullptr;             // pinned host mirror
  int nblk_red = 0;
  bool adjoint_mode = false;
  bool e0_zero = false;                // the level-0 correction is zero (start of a preconditioner application)
  double setup_ms[4] = {};             // last pp_set_geometry: flags, ghost tables, coarse operator, null vector
  // w/V of the null vector in the level-0 layout (FP32, FP64): the adjoint
  // V-cycle scales its level-0 coarse-grid correction by it
  float *d_nshape32 = nullptr;
  double *d_nshape64 = nullptr;
  bool have_nshape = false;
  // implicit momentum solve: amx_ad..apz_ad (6 fields, reference layout)
  // of the last pp_set_solve_ad, and its dt
  double *d_ad_coef = nullptr;
  double ad_dt = 0.0;
  bool have_ad = false;
  pp::FreshRows fr;                    // fresh-cell rows of the current geometry
  int null_iters = 0;
  double null_res = 0.0;           // level-0 smoother uses the transposed ghost rows
  std::vector<double *> d_kry;         // flexible GMRES (FP64 path): m+1 basis + m preconditioned vectors
  int kry_m = 0;                       // restart length d_kry is sized for
  std::vector<float *> d_kryl;         // the same for the mixed path: FP32, level-0 layout
  int kryl_m = 0;
  double *d_mu = nullptr;              // adjoint unknown lambda/V (reference layout)
  double *d_sv = nullptr;              // adjoint rhs s/V (reference layout)
  // graph cache
  hipGraphExec_t gexec = nullptr;        // captured preconditioner (use_graph)
  long graph_key = -1;
  double graph_bytes = 0.0;
  double bytes_acc = 0;
  // opt-in per-launch timing (pp_solve_opts.profile)
  bool prof_on = false;
  std::vector<hipEvent_t> prof_ev;       // pool, 2 per launch
  struct ProfRec {
    int cls, level;
    double bytes;
  };
  std::vector<ProfRec> prof_rec;
  double kbytes[PP_K_CLASSES] = {}, kbytes_l0[PP_K_CLASSES] = {};
  int klaunch[PP_K_CLASSES] = {}, klaunch_l0[PP_K_CLASSES] = {};
};

namespace pp {
void set_error(const char *fmt, ...);
}
