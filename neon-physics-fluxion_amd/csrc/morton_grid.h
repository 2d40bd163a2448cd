#pragma once
// Quadrant paths of the Barnes-Hut tree (insertIntoChild,
// BarnesHutTree.cpp:100-134): the first levels read off a grid, the rest
// descended level by level.
//
// The descent starts from the root square [x0, x2] x [y0, y2]; at each of D
// levels mx = (xl + xh) / 2 in FP64, west iff px <= mx, south iff py <= my,
// and the quadrant's digit (NW 0, NE 1, SW 2, SE 3) enters the key from the
// top (key_descent). While every sum and midpoint of the descent is exact in
// FP64 — the first K levels (grid_levels) — the level-d bounds are exactly
// x0 + k (x2 - x0) 2^-d, so the first K west/east decisions spell, most
// significant first, the index j of the depth-K cell (x0 + j c, x0 + (j+1) c]
// holding px, c = (x2 - x0) 2^-K (j = 0 at or below x0 + c, 2^K - 1 above
// x2 - c). The key's top 2K bits interleave the two indices, and the descent
// continues from that cell's bounds (exact FP64 values), rounding as the
// reference's does. An index is an FP64 estimate corrected and then checked
// by exact comparisons with the cell bounds (fma(j, c, x0) is exact: the
// bound is an FP64 value); a point the check does not settle (NaN) takes the
// whole descent.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace fx {

// Levels d0..D-1 of the descent from the cell [xl, xh] x [yl, yh] at depth
// d0, added into key.
__host__ __device__ inline uint64_t descend(double px, double py, double xl, double xh, double yl,
                                            double yh, int d0, int D, uint64_t key) {
  for (int d = d0; d < D; ++d) {
    const double mx = (xl + xh) / 2, my = (yl + yh) / 2;
    const bool west = px <= mx, south = py <= my;
    xh = west ? mx : xh;
    xl = west ? xl : mx;
    yh = south ? my : yh;
    yl = south ? yl : my;
    const uint64_t quad = (west ? 0u : 1u) | (south ? 2u : 0u);
    key |= quad << (62 - 2 * d);
  }
  return key;
}

// The whole descent, as the reference inserts a particle: root = x0, y0, x2, y2.
__host__ __device__ inline uint64_t key_descent(double px, double py, const double* root, int D) {
  return descend(px, py, root[0], root[2], root[1], root[3], 0, D, 0);
}

__host__ __device__ inline uint64_t f64_bits(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  return b;
}

// For finite non-zero v: the exponent of its lowest set bit (v = m 2^e, m
// odd) and an E with |v| < 2^E.
__host__ __device__ inline int f64_lsb_exp(double v) {
  const uint64_t b = f64_bits(v);
  const int ef = (int)((b >> 52) & 0x7FF);
  uint64_t m = b & ((1ull << 52) - 1);
  int e = -1074;
  if (ef) {
    m |= 1ull << 52;
    e = ef - 1075;
  }
  return e + __builtin_ctzll(m);
}
__host__ __device__ inline int f64_mag_exp(double v) {
  const int ef = (int)((f64_bits(v) >> 52) & 0x7FF);
  return ef ? ef - 1022 : -1022;
}

// How many of the first D (<= 32) levels of the descent over [lo, hi] take
// exact steps only: lo, hi are multiples of 2^b with magnitudes below 2^E, so
// every level-k bound is a multiple of 2^(b - k) below 2^E and every sum of
// two below 2^(E + 1): at most E + 1 - b + k significant bits, kept within 52.
__host__ __device__ inline int grid_levels(double lo, double hi, int D) {
  if (!(hi > lo) || !(fabs(lo) <= 1.0e300) || !(fabs(hi) <= 1.0e300) || D < 1 || D > 32) return 0;
  int b = 4096, E = -4096;
  if (lo != 0.0) {
    b = f64_lsb_exp(lo);
    E = f64_mag_exp(lo);
  }
  if (hi != 0.0) {
    const int bh = f64_lsb_exp(hi), eh = f64_mag_exp(hi);
    b = bh < b ? bh : b;
    E = eh > E ? eh : E;
  }
  const int k = 52 - (E + 1 - b);
  return k < 0 ? 0 : (k < D ? k : D);
}

// The depth-K cell index of p along an axis with bounds lo + j c (c = cell
// side, inv_c ~ 1 / c, jmax = 2^K - 1); false when p is not settled (NaN).
__host__ __device__ inline bool grid_index(double p, double lo, double c, double inv_c,
                                           int64_t jmax, uint32_t* out) {
  double q = (p - lo) * inv_c;
  q = q > 0.0 ? q : 0.0;
  q = q < (double)jmax ? q : (double)jmax;
  int64_t j = (int64_t)q;
  if (j > 0 && p <= fma((double)j, c, lo)) --j;
  else if (j < jmax && p > fma((double)(j + 1), c, lo)) ++j;
  *out = (uint32_t)j;
  return (j == 0 || p > fma((double)j, c, lo)) && (j == jmax || p <= fma((double)(j + 1), c, lo));
}

// Bits of v at the even positions 0, 2, ..., 62.
__host__ __device__ inline uint64_t spread_bits(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// Constants of the grid levels: K of them (0: the whole descent), cell sides
// and their reciprocals at depth K.
struct KeyGrid {
  int K;
  double cx, cy, inv_cx, inv_cy;
  int64_t jmax;
};

__host__ __device__ inline void key_grid_setup(const double* root, int D, KeyGrid* kg) {
  const int kx = grid_levels(root[0], root[2], D), ky = grid_levels(root[1], root[3], D);
  kg->K = kx < ky ? kx : ky;
  kg->cx = ldexp(root[2] - root[0], -kg->K);
  kg->cy = ldexp(root[3] - root[1], -kg->K);
  kg->inv_cx = 1.0 / kg->cx;
  kg->inv_cy = 1.0 / kg->cy;
  kg->jmax = (int64_t)((1ull << kg->K) - 1);
}

// The key of (px, py): grid levels, then the descent (D >= 1).
__host__ __device__ inline uint64_t key_grid(double px, double py, const double* root, int D,
                                             const KeyGrid& kg) {
  uint32_t jx, jy;
  if (kg.K == 0 || !grid_index(px, root[0], kg.cx, kg.inv_cx, kg.jmax, &jx) ||
      !grid_index(py, root[1], kg.cy, kg.inv_cy, kg.jmax, &jy))
    return key_descent(px, py, root, D);
  // x digit: east = index bit 1; y digit: south = index bit 0.
  const uint32_t sy = (uint32_t)kg.jmax ^ jy;
  const uint64_t top = (spread_bits(jx) | (spread_bits(sy) << 1)) << (64 - 2 * kg.K);
  return descend(px, py, fma((double)jx, kg.cx, root[0]), fma((double)jx + 1.0, kg.cx, root[0]),
                 fma((double)jy, kg.cy, root[1]), fma((double)jy + 1.0, kg.cy, root[1]), kg.K, D,
                 top);
}

}  // namespace fx
