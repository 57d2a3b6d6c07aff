// Triangulated-surface geometry (filled in with the reference restatement).
#include "pp_geom.h"

namespace ppg {
int tri_build(const Grid &, const std::vector<Body> &, TriBodies &) { return 99; }
int tri_iblank(const Grid &, const TriBodies &, std::vector<int8_t> &, std::vector<int> &) { return 99; }
void tri_intercept(const Grid &, const TriBodies &, int, int, int, int, double *, double *, int &) {}
}  // namespace ppg
