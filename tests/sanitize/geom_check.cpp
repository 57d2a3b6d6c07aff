// Runs the host side of the C-ABI that needs no GPU -- pp_geom_build in both
// modes with a previous iblank (fresh cells), pp_geom_build_surfaces on a
// canonical body's own mesh, and every pp_geom getter -- so that an
// AddressSanitizer / UndefinedBehaviorSanitizer build of pp_geom.cpp and
// pp_geom_tri.cpp reports any invalid access or undefined operation.
// Prints "ok" and exits 0 when everything ran.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pp_hip.h"

// the library's error sink lives with the device code; this build has none
namespace pp {
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  fputc('\n', stderr);
  va_end(ap);
}
}  // namespace pp

static int check(int rc, const char *what) {
  if (rc != PP_OK) { std::printf("%s failed: %d\n", what, rc); std::exit(1); }
  return rc;
}

static void read_all(pp_geom *g, long cells) {
  int32_t ng = 0, nb = 0, nf = 0;
  check(pp_geom_counts(g, &ng, &nb, &nf), "counts");
  std::vector<int8_t> ib(cells), fr(cells);
  const size_t m = ng > 0 ? ng : 1, mf = nf > 0 ? nf : 1;
  std::vector<int32_t> idx(6 * m), be(2 * m), fi(8 * mf);
  std::vector<double> coeff(8 * m), probe(m), v(9 * m), cd(8 * m), fc(8 * mf), fb(3 * mf);
  check(pp_geom_get(g, ib.data(), fr.data(), &idx[0], &idx[m], &idx[2 * m], &idx[3 * m], &idx[4 * m], &idx[5 * m],
                    coeff.data(), probe.data(), &v[0], &v[3 * m], &v[6 * m]), "get");
  check(pp_geom_ghost_elements(g, &be[0], &be[m]), "elements");
  check(pp_geom_dirichlet(g, cd.data()), "dirichlet");
  pp_geom_fresh(g, &fi[0], &fi[mf], &fi[2 * mf], &fi[3 * mf], &fi[4 * mf], &fi[5 * mf], fc.data(), fb.data(),
                &fi[6 * mf], &fi[7 * mf]);
  std::printf("ghosts %d blank %d fresh %d\n", ng, nb, nf);
}

int main() {
  const int nx = 28, ny = 24, nz = 22;
  std::vector<double> xn(nx + 1), yn(ny + 1), zn(nz + 1);
  auto nodes = [](std::vector<double> &x, double beta) {
    const int n = (int)x.size() - 1;
    for (int i = 0; i <= n; i++)
      x[i] = 0.5 * (1.0 + std::atanh((2.0 * i / n - 1.0) * std::tanh(beta)) / beta);
  };
  nodes(xn, 1.2); nodes(yn, 0.9); nodes(zn, 1.4);
  pp_grid_desc gd{};
  gd.nxc = nx; gd.nyc = ny; gd.nzc = nz;
  gd.x = xn.data(); gd.y = yn.data(); gd.z = zn.data();
  for (int f = 0; f < 6; f++) gd.bc[f] = PP_BC_NEUMANN;
  const long cells = (long)(nx + 3) * (ny + 3) * (nz + 3);
  pp_body_desc b0{PP_BODY_ELLIPSOID, {0.46, 0.5, 0.52}, {0.22, 0.2, 0.24}, 25.0, 24};
  pp_body_desc b1{PP_BODY_ELLIPSOID, {0.49, 0.5, 0.51}, {0.22, 0.2, 0.24}, 25.0, 24};
  for (int mode : {PP_GEOM_TRIANGULATED, PP_GEOM_ANALYTIC}) {
    pp_geom *g0 = nullptr, *g1 = nullptr;
    check(pp_geom_build(&g0, &gd, 1, &b0, mode, 2.0, nullptr), "build");
    std::vector<int8_t> prev(cells);
    check(pp_geom_get(g0, prev.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr, nullptr, nullptr, nullptr), "get iblank");
    check(pp_geom_build(&g1, &gd, 1, &b1, mode, 2.0, prev.data()), "build with iblank_prev");
    read_all(g1, cells);
    pp_geom_destroy(g0);
    pp_geom_destroy(g1);
  }
  int32_t nm = 0, ne = 0;
  check(pp_geom_canonical_surface(&b1, &nm, &ne, nullptr, nullptr), "surface counts");
  std::vector<double> xyz(3 * nm);
  std::vector<int32_t> tri(3 * ne);
  check(pp_geom_canonical_surface(&b1, &nm, &ne, xyz.data(), tri.data()), "surface");
  pp_surface_desc sd{nm, ne, xyz.data(), tri.data(), {0.9, 0.9, 0.9}};
  pp_geom *gs = nullptr;
  check(pp_geom_build_surfaces(&gs, &gd, 1, &sd, 2.0, nullptr), "surfaces");
  read_all(gs, cells);
  pp_geom_destroy(gs);
  std::printf("ok\n");
  return 0;
}
