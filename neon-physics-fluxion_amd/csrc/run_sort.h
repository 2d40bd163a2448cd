#pragma once
// Ordering of one run of the Barnes-Hut key sort's fix-up: the sort orders
// the keys by their top 32 bits; each run of equal top halves is then ordered
// by (full key, particle index) by one thread. Short runs (the usual case: a
// few particles closer than 2^-16 of the root) take an insertion sort; long
// ones (many particles packed into one such cell) a bottom-up merge sort
// through scratch arrays of the run's length, so the cost stays
// O(len log len) rather than O(len^2).

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fx {

__host__ __device__ inline bool run_less(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb) {
  return ka < kb || (ka == kb && va < vb);
}

__host__ __device__ inline void run_insertion(uint64_t* k, uint32_t* v, int len) {
  for (int a = 1; a < len; ++a) {
    const uint64_t key = k[a];
    const uint32_t val = v[a];
    int b = a - 1;
    while (b >= 0 && run_less(key, val, k[b], v[b])) {
      k[b + 1] = k[b];
      v[b + 1] = v[b];
      --b;
    }
    k[b + 1] = key;
    v[b + 1] = val;
  }
}

constexpr int kRunInsertionMax = 32;  // longer runs are merge-sorted

// Orders k/v[0, len) by (key, value); ks/vs: scratch of len entries.
__host__ __device__ inline void sort_run(uint64_t* k, uint32_t* v, uint64_t* ks, uint32_t* vs,
                                         int len) {
  if (len <= kRunInsertionMax) {
    run_insertion(k, v, len);
    return;
  }
  constexpr int kChunk = 16;
  for (int lo = 0; lo < len; lo += kChunk)
    run_insertion(k + lo, v + lo, len - lo < kChunk ? len - lo : kChunk);
  uint64_t *sk = k, *dk = ks;
  uint32_t *sv = v, *dv = vs;
  for (int w = kChunk; w < len; w *= 2) {
    for (int lo = 0; lo < len; lo += 2 * w) {
      const int mid = lo + w < len ? lo + w : len;
      const int hi = lo + 2 * w < len ? lo + 2 * w : len;
      int a = lo, b = mid, o = lo;
      while (a < mid && b < hi) {
        if (run_less(sk[b], sv[b], sk[a], sv[a])) {
          dk[o] = sk[b];
          dv[o++] = sv[b++];
        } else {
          dk[o] = sk[a];
          dv[o++] = sv[a++];
        }
      }
      while (a < mid) {
        dk[o] = sk[a];
        dv[o++] = sv[a++];
      }
      while (b < hi) {
        dk[o] = sk[b];
        dv[o++] = sv[b++];
      }
    }
    uint64_t* tk = sk;
    sk = dk;
    dk = tk;
    uint32_t* tv = sv;
    sv = dv;
    dv = tv;
  }
  if (sk != k) {
    for (int a = 0; a < len; ++a) {
      k[a] = sk[a];
      v[a] = sv[a];
    }
  }
}

}  // namespace fx
