// CPU check of the exact walk's opening thresholds (csrc/walk_open.h) against
// the opening test in its FP32 form, BarnesHutTree.cpp:151-157:
//   dist = sqrt(dx*dx + dy*dy); if (dist < min_dist) dist = min_dist;
//   accept iff size / dist < theta
// For each random (size, theta, min_dist) the threshold S must satisfy
// (d2 >= S) == accept(d2) on d2 within 3 ulps of S, on both sides of
// (size/theta)^2 and min_dist^2, and on random d2. x86-64 SSE division and
// square root are IEEE, so the FP32 form is evaluated exactly as written.
// Prints the counts; exits 1 on any mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "walk_open.h"

using fx::OpenRule;

static bool fp32_accept(float size, float d2, float theta, float md) {
  volatile float dist = sqrtf(d2);
  if (dist < md) dist = md;
  volatile float q = size / dist;
  return q < theta;
}

// theta's rounding boundary and tie rule, as the library derives them.
static OpenRule rule_for(float th, float md) {
  OpenRule r;
  const float lo = nextafterf(th, -INFINITY);
  const double hi = (std::isinf(th) && th > 0) ? ldexp(1.0, 128) : (double)th;
  r.theta_m = ((double)lo + hi) / 2;
  uint32_t b;
  memcpy(&b, &th, 4);
  r.tie_le = (std::isfinite(th) && (b & 1u)) ? 1 : 0;
  r.min_dist = md;
  return r;
}

int main(int argc, char** argv) {
  const int trials = argc > 1 ? atoi(argv[1]) : 200000;
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> U(0, 1);
  long long checks = 0, bad = 0, unsettled = 0;
  const float thetas[] = {0.5f, 0.3f, 0.7f, 1.0f, 0.25f, 0.9f, 1.5f, 0.123456f};
  for (int trial = 0; trial < trials; ++trial) {
    const float theta = trial < trials * 2 / 5 ? thetas[trial % 8]
                                               : (float)ldexp(1.0 + U(g), (int)(U(g) * 20) - 10);
    const float md = (trial % 3 == 0) ? 0.1f : (float)ldexp(1.0 + U(g), (int)(U(g) * 40) - 20);
    const float root = (float)ldexp(1.0 + U(g), (int)(U(g) * 40) - 18);
    const float size = ldexpf(root, -(int)(U(g) * 32));
    const OpenRule r = rule_for(theta, md);
    int ok = 1;
    const float S = fx::open_threshold_sq(size, r, &ok);
    if (!ok) {
      ++unsettled;
      continue;
    }
    auto chk = [&](float d2) {
      if (!(d2 >= 0) || !(d2 < 1e30f)) return;
      ++checks;
      const bool a = d2 >= S, b = fp32_accept(size, d2, theta, md);
      if (a != b) {
        if (bad < 10)
          printf("mismatch size=%a theta=%a min_dist=%a S=%a d2=%a got %d want %d\n", size, theta,
                 md, S, d2, a, b);
        ++bad;
      }
    };
    if (S > 0 && S < INFINITY)
      for (int k = -3; k <= 3; ++k) chk(fx::f32_step(S, k));
    for (int k = 0; k < 8; ++k) chk((float)(S * ldexp(1.0, (int)(U(g) * 8) - 4) * U(g)));
    chk(0.0f);
    chk(md * md);
    chk(fx::f32_step(md * md, 1));
    chk(fx::f32_step(md * md, -1));
    const float e = (float)((double)size / theta), e2 = e * e;
    for (int k = -4; k <= 4; ++k) chk(fx::f32_step(e2, k));
  }
  printf("checks %lld mismatches %lld unsettled %lld\n", checks, bad, unsettled);
  return (bad || unsettled) ? 1 : 0;
}
