// Host-side set-up routines shared by the API translation unit.
#pragma once

#include <functional>

#include "pp_internal.h"

namespace pp {
const char *last_error();
void host_parallel(long n, const std::function<void(long, long)> &fn);
void make_grid_dir(int n, const double *xn, std::vector<double> &x, std::vector<double> &xc,
                   std::vector<double> &dx, std::vector<double> &dxinv, std::vector<double> &dxc,
                   std::vector<double> &dxcinv);
int build_levels(pp_ctx *c);
// Cells of the coarsest level build_levels reaches for a grid, and the most
// cells the coarsest level may have: its dense FP64 solution operator
// (n^2 entries) and the bordered LU behind it grow as n^2.
long coarsest_cells(int nxc, int nyc, int nzc);
constexpr long kCoarseDenseMax = 16384;
// longest line the x-line (wavefront) and y/z-line (partitioned) solvers hold
constexpr int kMaxLine = 1024;
int build_flags(pp_ctx *c, const int8_t *d_iblank, long *n_ghost);
// lane-group width G and rows per lane S of the x-line pass (elem: bytes)
void line_shape(int n, int &G, int &S, int elem);
int upload_levels(pp_ctx *c);
void free_levels(pp_ctx *c);
void free_ghosts(pp_ctx *c);
// whether reference offset r is a ghost cell of c->h_iblank
bool ref_is_ghost(const pp_ctx *c, long r);
int build_ghosts(pp_ctx *c, const pp_geometry *g);
int build_coarse(pp_ctx *c);
}  // namespace pp
