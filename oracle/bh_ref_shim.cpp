// C entry point around the reference's own Barnes-Hut tree
// (neon_old_codebase/src/cpp/BarnesHutTree.{h,cpp}), compiled from the
// reference sources where they lie. TEST INFRASTRUCTURE ONLY: the library it
// builds (oracle/_ref/libbh_ref.so) checks the oracle's literal Barnes-Hut
// restatement; nothing in the product links it.
#include <cstdint>

#include "BarnesHutTree.h"

extern "C" __attribute__((visibility("default"))) void ref_bh_forces(
    int32_t n, const double* x, const double* y, const double* q, double theta,
    double k, double min_dist, double* fx, double* fy) {
  fluxion::BarnesHutTree tree(theta);
  tree.build(x, y, q, n);
  for (int32_t i = 0; i < n; ++i)
    tree.computeForcesForParticle(i, x[i], y[i], q[i], k, min_dist, fx[i], fy[i]);
}
