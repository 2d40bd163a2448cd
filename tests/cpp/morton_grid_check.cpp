// Grid-path quadrant keys (csrc/morton_grid.h) against the level-by-level
// descent (BarnesHutTree.cpp:100-134) on roots built like the tree's (float
// bounding box, margin 1.0, squared about the centre; BarnesHutTree.cpp:24-46)
// and on random FP64 roots, at points inside, outside, on the cell bounds of
// every depth and one ulp either side of them, NaN and infinities.
// Usage: morton_grid_check <roots>. Prints "... mismatches 0 ..." on success.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "morton_grid.h"

using fx::KeyGrid;

static void make_root(double mnx, double mxx, double mny, double mxy, double* root, int* D) {
  const double min_x = mnx - 1.0, max_x = mxx + 1.0, min_y = mny - 1.0, max_y = mxy + 1.0;
  const double size = std::fmax(max_x - min_x, max_y - min_y);
  const double cx = (min_x + max_x) / 2.0, cy = (min_y + max_y) / 2.0;
  root[0] = cx - size / 2;
  root[1] = cy - size / 2;
  root[2] = cx + size / 2;
  root[3] = cy + size / 2;
  double xl = root[0], yl = root[1], xh = root[2], yh = root[3];
  int d = 0;
  while (d < 32 && std::fmax(xh - xl, yh - yl) >= 1e-8) {
    xh = (xl + xh) / 2;
    yh = (yl + yh) / 2;
    ++d;
  }
  *D = d;
}

int main(int argc, char** argv) {
  const int roots = argc > 1 ? atoi(argv[1]) : 2000;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  long long points = 0, mismatches = 0, exact_roots = 0, tried = 0, bound_points = 0, ksum = 0;
  for (int r = 0; r < roots; ++r) {
    double root[4];
    int D = 0;
    const int kind = r % 4;
    if (kind < 3) {
      // float boxes of several magnitudes (the tree's roots)
      const double scale[3] = {1000.0, 3.0e6, 1.0e-3};
      const double s = scale[kind] * (0.5 + u01(rng));
      const double off = (r & 4) ? -s * u01(rng) : 0.0;
      float a = (float)(off + s * u01(rng)), b = (float)(off + s * u01(rng));
      float c = (float)(off + s * u01(rng)), d = (float)(off + s * u01(rng));
      make_root(std::fmin(a, b), std::fmax(a, b), std::fmin(c, d), std::fmax(c, d), root, &D);
    } else {
      // arbitrary FP64 squares and depths
      const double lo = (u01(rng) - 0.5) * std::ldexp(1.0, (int)(u01(rng) * 40) - 20);
      const double side = std::ldexp(1.0 + u01(rng), (int)(u01(rng) * 30) - 10);
      root[0] = lo;
      root[1] = lo * 0.75;
      root[2] = lo + side;
      root[3] = lo * 0.75 + side;
      D = 1 + (int)(u01(rng) * 32);
    }
    if (D < 1) continue;
    ++tried;
    KeyGrid kg;
    fx::key_grid_setup(root, D, &kg);
    if (kg.K > 0) ++exact_roots;
    ksum += kg.K;
    if (r < 8) printf("root kind %d D %d grid levels %d\n", kind, D, kg.K);
    std::vector<double> xs, ys;
    for (int k = 0; k < 200; ++k) {
      xs.push_back((double)(float)(root[0] + (root[2] - root[0]) * u01(rng)));
      ys.push_back((double)(float)(root[1] + (root[3] - root[1]) * u01(rng)));
    }
    // bounds of every depth along random paths, and their neighbours
    for (int k = 0; k < 40; ++k) {
      double xl = root[0], xh = root[2], yl = root[1], yh = root[3];
      const int stop = 1 + (int)(u01(rng) * D);
      for (int d = 0; d < stop; ++d) {
        const double mx = (xl + xh) / 2, my = (yl + yh) / 2;
        for (double v : {mx, std::nextafter(mx, -INFINITY), std::nextafter(mx, INFINITY)}) {
          xs.push_back(v);
          ys.push_back((double)(float)(yl + (yh - yl) * u01(rng)));
        }
        for (double v : {my, std::nextafter(my, -INFINITY), std::nextafter(my, INFINITY)}) {
          ys.push_back(v);
          xs.push_back((double)(float)(xl + (xh - xl) * u01(rng)));
        }
        bound_points += 6;
        if (u01(rng) < 0.5) xh = mx; else xl = mx;
        if (u01(rng) < 0.5) yh = my; else yl = my;
      }
    }
    const double special[] = {root[0], root[2], root[1], root[3], root[0] - 1.0, root[2] + 1.0,
                              NAN, INFINITY, -INFINITY, 0.0, -0.0, DBL_MAX, -DBL_MAX};
    for (double a : special)
      for (double b : special) {
        xs.push_back(a);
        ys.push_back(b);
      }
    for (size_t k = 0; k < xs.size(); ++k) {
      const uint64_t want = fx::key_descent(xs[k], ys[k], root, D);
      const uint64_t got = fx::key_grid(xs[k], ys[k], root, D, kg);
      ++points;
      if (want != got) {
        if (mismatches < 10)
          printf("MISMATCH root %.17g %.17g %.17g %.17g D %d p %.17g %.17g: %016llx vs %016llx\n",
                 root[0], root[1], root[2], root[3], D, xs[k], ys[k], (unsigned long long)want,
                 (unsigned long long)got);
        ++mismatches;
      }
    }
  }
  printf("roots %lld with_grid %lld mean_grid_levels %.2f points %lld bound_points %lld "
         "mismatches %lld \n", tried, exact_roots, tried ? (double)ksum / tried : 0.0, points,
         bound_points, mismatches);
  return mismatches == 0 && exact_roots > tried / 2 ? 0 : 1;
}
