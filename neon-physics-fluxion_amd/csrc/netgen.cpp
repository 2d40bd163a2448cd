// Seeded synthetic netlist generators for the placement workloads.
//
// The reference ships generators of its own (neon_old_codebase/examples/
// run_million_gates.py:25-79, benchmarks/run_benchmarks.py:15-44) whose degree
// and locality distributions differ from the workloads this library is
// measured on, so the distributions below are stated here and generated
// deterministically:
//   flat netlists : net degree 2 (p=0.6), U{3..5} (p=0.3), U{6..20} (p=0.1);
//                   each net picks an anchor block of 64 consecutive gate ids
//                   and draws every pin from that block with p=0.8, else from
//                   all gates; duplicate pins inside a net are redrawn.
//                   Initial positions are uniform on [0,die_w) x [0,die_h).
//   leaf batches  : leaf sizes log-uniform on [min_n, max_n] (rounded),
//                   round(1.1 n) nets per leaf, degree U{2..8} (capped at n),
//                   pins uniform inside the leaf, one square region per leaf.
// Random numbers: SplitMix64 seeds a PCG32 (XSH-RR) stream.

#include <cmath>
#include <cstdint>
#include <cstring>

#include "fluxion_hip.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

struct Pcg32 {
  uint64_t state, inc;
  explicit Pcg32(uint64_t seed) {
    SplitMix64 sm{seed};
    state = sm.next();
    inc = sm.next() | 1ull;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
  // Integer in [0, n) by multiply-shift.
  uint32_t below(uint32_t n) { return (uint32_t)(((uint64_t)next() * n) >> 32); }
  // Float in [0, 1) with 24 random bits.
  float unit() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

const uint64_t kPositionStream = 0xA5A5A5A5A5A5A5A5ull;

bool contains(const int32_t* a, int n, int32_t v) {
  for (int i = 0; i < n; ++i)
    if (a[i] == v) return true;
  return false;
}

}  // namespace

extern "C" int64_t fx_netgen_flat(int32_t num_gates, int32_t num_nets, uint64_t seed,
                                  float die_w, float die_h, int32_t* net_ptr,
                                  int32_t* net_pins, int64_t pin_capacity, float* x,
                                  float* y) {
  if (num_gates < 2 || num_nets < 0) return -1;
  Pcg32 rng(seed);
  const uint32_t blocks = (uint32_t)((num_gates + 63) / 64);
  int64_t p = 0;
  int32_t pins[20];
  if (net_ptr) net_ptr[0] = 0;
  for (int32_t e = 0; e < num_nets; ++e) {
    float u = rng.unit();
    int deg;
    if (u < 0.6f) deg = 2;
    else if (u < 0.9f) deg = 3 + (int)rng.below(3);
    else deg = 6 + (int)rng.below(15);
    if (deg > num_gates) deg = num_gates;
    uint32_t b = rng.below(blocks);
    int32_t base = (int32_t)(b * 64u);
    uint32_t bsize = (uint32_t)(num_gates - base < 64 ? num_gates - base : 64);
    for (int k = 0; k < deg; ++k) {
      int32_t g;
      do {
        if (rng.unit() < 0.8f) g = base + (int32_t)rng.below(bsize);
        else g = (int32_t)rng.below((uint32_t)num_gates);
      } while (contains(pins, k, g));
      pins[k] = g;
    }
    if (net_pins) {
      if (p + deg > pin_capacity) return -2;
      std::memcpy(net_pins + p, pins, sizeof(int32_t) * deg);
    }
    p += deg;
    if (net_ptr) net_ptr[e + 1] = (int32_t)p;
  }
  if (x && y) {
    Pcg32 pos(seed ^ kPositionStream);
    for (int32_t i = 0; i < num_gates; ++i) {
      x[i] = pos.unit() * die_w;
      y[i] = pos.unit() * die_h;
    }
  }
  return p;
}

extern "C" int64_t fx_netgen_leaves(int32_t num_leaves, uint64_t seed, int32_t min_n,
                                    int32_t max_n, float region, int32_t* leaf_ptr,
                                    float* leaf_region, int32_t* net_ptr,
                                    int32_t* net_pins, int64_t* counts) {
  // counts[0] = gates, counts[1] = nets, counts[2] = pins. Returns pins or <0.
  if (num_leaves < 1 || min_n < 2 || max_n < min_n) return -1;
  Pcg32 rng(seed);
  const int grid = (int)std::ceil(std::sqrt((double)num_leaves));
  const float lmin = std::log((float)min_n), lmax = std::log((float)max_n);
  int64_t gates = 0, nets = 0, pins_total = 0;
  int32_t pins[8];
  if (leaf_ptr) leaf_ptr[0] = 0;
  if (net_ptr) net_ptr[0] = 0;
  for (int32_t c = 0; c < num_leaves; ++c) {
    int n = (int)std::lround(std::exp(lmin + rng.unit() * (lmax - lmin)));
    if (n < min_n) n = min_n;
    if (n > max_n) n = max_n;
    int e = (int)std::lround(1.1f * (float)n);
    for (int j = 0; j < e; ++j) {
      int deg = 2 + (int)rng.below(7);
      if (deg > n) deg = n;
      for (int k = 0; k < deg; ++k) {
        int32_t g;
        do {
          g = (int32_t)gates + (int32_t)rng.below((uint32_t)n);
        } while (contains(pins, k, g));
        pins[k] = g;
      }
      if (net_pins) std::memcpy(net_pins + pins_total, pins, sizeof(int32_t) * deg);
      pins_total += deg;
      ++nets;
      if (net_ptr) net_ptr[nets] = (int32_t)pins_total;
    }
    gates += n;
    if (leaf_ptr) leaf_ptr[c + 1] = (int32_t)gates;
    if (leaf_region) {
      leaf_region[4 * c + 0] = ((float)(c % grid) + 0.5f) * region;
      leaf_region[4 * c + 1] = ((float)(c / grid) + 0.5f) * region;
      leaf_region[4 * c + 2] = region;
      leaf_region[4 * c + 3] = region;
    }
  }
  if (counts) {
    counts[0] = gates;
    counts[1] = nets;
    counts[2] = pins_total;
  }
  return pins_total;
}

// Circuit::build_adjacency (HAPR_fluxion_v4/src/circuit.rs:154-171): for every
// net and every pin pair (i<j), append b to adj[a] and a to adj[b] unless
// already present, keeping first-insertion order. Output is CSR; when adj is
// null only the entry count is returned.
extern "C" int64_t fx_build_adjacency(int32_t num_gates, int32_t num_nets,
                                      const int32_t* net_ptr, const int32_t* net_pins,
                                      int32_t* adj_ptr, int32_t* adj, int64_t capacity) {
  if (num_gates < 0 || num_nets < 0 || !net_ptr || !net_pins) return -1;
  // Per-gate growable lists in one arena; a gate's list only grows while nets
  // are processed, so lists are chained blocks.
  struct List { int32_t* v; int32_t n, cap; };
  List* lists = new List[num_gates]();
  auto push_unique = [&](int32_t a, int32_t b) {
    List& l = lists[a];
    for (int32_t i = 0; i < l.n; ++i)
      if (l.v[i] == b) return;
    if (l.n == l.cap) {
      int32_t nc = l.cap ? l.cap * 2 : 4;
      int32_t* nv = new int32_t[nc];
      if (l.n) std::memcpy(nv, l.v, sizeof(int32_t) * l.n);
      delete[] l.v;
      l.v = nv;
      l.cap = nc;
    }
    l.v[l.n++] = b;
  };
  for (int32_t e = 0; e < num_nets; ++e) {
    const int32_t s = net_ptr[e], t = net_ptr[e + 1];
    for (int32_t i = s; i < t; ++i)
      for (int32_t j = i + 1; j < t; ++j) {
        const int32_t a = net_pins[i], b = net_pins[j];
        if (a < 0 || a >= num_gates || b < 0 || b >= num_gates) continue;
        push_unique(a, b);
        push_unique(b, a);
      }
  }
  int64_t total = 0;
  for (int32_t g = 0; g < num_gates; ++g) total += lists[g].n;
  int64_t rc = total;
  if (adj_ptr) {
    if (adj && total > capacity) {
      rc = -2;
    } else {
      int64_t o = 0;
      adj_ptr[0] = 0;
      for (int32_t g = 0; g < num_gates; ++g) {
        if (adj) std::memcpy(adj + o, lists[g].v, sizeof(int32_t) * lists[g].n);
        o += lists[g].n;
        adj_ptr[g + 1] = (int32_t)o;
      }
    }
  }
  for (int32_t g = 0; g < num_gates; ++g) delete[] lists[g].v;
  delete[] lists;
  return rc;
}

// Initial vector of compute_fiedler_vector (partitioner.rs:145-147):
// v_i = sin(i as f32 * 2.71828) in FP32 with the host C library's sinf, the
// function Rust's f32::sin calls on this platform.
extern "C" int fx_fiedler_init(int32_t n, float* v) {
  if (n < 0 || (n && !v)) return FX_E_INVALID;
  for (int32_t i = 0; i < n; ++i) v[i] = std::sin((float)i * 2.71828f);
  return FX_OK;
}
