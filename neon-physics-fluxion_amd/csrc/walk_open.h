#pragma once
// Opening test of the exact Barnes-Hut walk (BarnesHutTree.cpp:151-157) as a
// threshold on the squared distance.
//
// For a cell of size `size` whose centre is at (dx, dy) from the particle,
// the walk computes d2 = dx*dx + dy*dy, dist = max(sqrt(d2), min_dist) with
// an IEEE square root, and accepts the cell iff fl(size / dist) < theta. For
// positive operands fl(a/b) < theta iff a < theta_m * b, or a == theta_m * b
// when that tie rounds down (tie_le), with theta_m the midpoint of theta and
// its FP32 predecessor (the product is exact in FP64). That predicate grows
// with dist, dist grows with d2, so the decision is d2 >= S for one FP32
// value S per cell size: the walk compares d2 against S and takes the square
// root only for accepted cells.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace fx {

struct OpenRule {
  double theta_m;  // midpoint of theta and its FP32 predecessor
  int tie_le;      // a == theta_m * b rounds to a value below theta
  float min_dist;  // distance clamp, > 0
};

// The FP32 value `steps` units in the last place from v (v >= 0).
__host__ __device__ inline float f32_step(float v, int steps) {
  uint32_t u;
  __builtin_memcpy(&u, &v, 4);
  u = (uint32_t)((int32_t)u + steps);
  __builtin_memcpy(&v, &u, 4);
  return v;
}

// size / dist < theta, evaluated exactly (dist > 0).
__host__ __device__ inline bool open_accepts(float size, float dist, const OpenRule& r) {
  const double th = r.theta_m * (double)dist;
  const double a = (double)size;
  return a < th || (r.tie_le && a == th);
}

// The smallest FP32 d2 >= 0 whose cell of size `size` the walk accepts: 0
// when every distance accepts, +inf when none does. *ok is cleared if the
// search did not settle (it settles in a few steps, the FP64 quotient being
// within an ulp of the boundary; the bound only keeps the loop finite).
// Requires min_dist > 0 and size >= 0 finite.
__host__ __device__ inline float open_threshold_sq(float size, const OpenRule& r, int* ok) {
  const float md = r.min_dist;
  if (!(md > 0.0f) || !(size >= 0.0f) || !(size <= FLT_MAX)) {
    *ok = 0;
    return INFINITY;
  }
  if (open_accepts(size, md, r)) return 0.0f;
  if (!(r.theta_m > 0.0) || !open_accepts(size, FLT_MAX, r)) return INFINITY;
  // T: the smallest FP32 distance that accepts; T > md as md does not.
  const double c = (double)size / r.theta_m;
  float t = c < (double)FLT_MAX ? (float)c : FLT_MAX;
  if (!(t > md)) t = f32_step(md, 1);
  int steps = 0;
  while (!open_accepts(size, t, r) && steps < 64) {
    t = f32_step(t, 1);
    ++steps;
  }
  while (steps < 64) {
    const float p = f32_step(t, -1);
    if (p <= md || !open_accepts(size, p, r)) break;
    t = p;
    ++steps;
  }
  if (steps >= 64) *ok = 0;
  // sqrt(d2) rounds to T or above iff sqrt(d2) exceeds the midpoint m of T
  // and its predecessor, i.e. d2 > m^2 (m has 25 significant bits, so m^2 is
  // exact in FP64 and never an FP32 value): S is m^2 rounded up to FP32.
  const double m = ((double)f32_step(t, -1) + (double)t) * 0.5;
  const double m2 = m * m;
  float s = (float)m2;
  if ((double)s <= m2) s = f32_step(s, 1);
  return s;
}

}  // namespace fx
