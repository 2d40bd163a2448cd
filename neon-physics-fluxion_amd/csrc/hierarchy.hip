// Recursive spectral bisection: build_hierarchy -> recursive_bisect ->
// compute_fiedler_vector -> split_by_fiedler
// (HAPR_fluxion_v4/src/algorithm/partitioner.rs:36-58, 64-100, 106-185,
// 189-213), and the region layout of hapr_placement (placer.rs:34-74,
// 111-141).
//
// The shape of the hierarchy depends only on the gate count and the
// configuration: a node of n gates is a leaf when n <= max_leaf_size or
// n <= min_cluster_size, otherwise it splits at
//   split = min(max(n/2, ceil(n*balance_factor)), n - ceil(n*balance_factor))
// (partitioner.rs:205-207) and becomes a leaf when a side is empty (85-87).
// The host therefore plans every node (pre-order ids, as 65-66 and 91-92
// assign them) before any GPU work; the GPU decides which gates go where.
//
// Gates live in one permutation `perm` in which every node is the
// contiguous slice [start, start+count), in the node's local order
// (partitioner.rs:116-121 numbers a node's gates in that order). All nodes
// of one depth are bisected together:
//   1. label: pos[g] = position of gate g, cl[g] = label of its node (-1 for
//      gates in leaves);
//   2. the level's induced subgraphs as one CSR over positions: each row
//      keeps the neighbours of adjacency order whose label matches
//      (partitioner.rs:128-137), so the SpMV subtracts in the reference's
//      adj_pairs order (159-165); degree = row length;
//   3. power iteration per node (145-182): FP32 SpMV, then the mean and the
//      L2 norm as the mode says (below);
//   4. stable sort of each node's slice by Fiedler value (196-202, -0.0 ==
//      +0.0); the left child is the first `split` entries.
// Nodes of more than kWgMax gates use grid-wide kernels (one block per 256
// positions) and a radix sort of (slice index, value) keys; smaller nodes
// run the whole of steps 3-4 in one workgroup with the vectors in LDS.
//
// Mean and norm, two modes, each identical on both paths:
//  * FX_HIER_EXACT: the reference's arithmetic (partitioner.rs:149-155,
//    168-179): FP32 sums accumulated one element after another in local
//    order, mean = sum/n, lv -= mean, norm = sqrt(sum of lv*lv),
//    v = lv/norm when norm > 1e-10f. One lane per node runs each sum (LDS
//    double-buffered by its workgroup); everything else is parallel.
//  * FX_HIER_FP64: FP64 sums in a blocked order. A node's local order is
//    cut into 256-element chunks; a chunk is reduced by a stride-halving
//    pairwise tree (slot k += slot k+h for h = 128, 64, ..., 1, missing
//    slots +0.0); the chunk results P_0..P_{m-1} fold into 256 slots
//    Q_k = P_k + P_{k+256} + ... (in index order) and the same tree reduces
//    Q. mean = (float)(sum/n); v = (float)((double)lv/norm) when
//    norm > 1e-10.

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace fx {
namespace {

constexpr int kB = 256;      // threads per block = reduction chunk length
constexpr int kWgMax = 2048; // nodes up to this size are bisected by one workgroup
constexpr int kRows = kWgMax / kB;
constexpr int kInlineChunks = 1024;  // above this a node's totals get a kernel of their own

struct Seg {
  int32_t start, n, c0, m;  // slice, first chunk, chunk count
};
struct Tile {
  int32_t start, n, label, pad;  // label >= 0: a grid-path node being sorted
};

__device__ __forceinline__ uint32_t order_bits(float f, int* nan_flag) {
  if (f != f) {
    atomicOr(nan_flag, 1);
    f = 0.0f;
  }
  if (f == 0.0f) f = 0.0f;  // -0.0 == +0.0 under partial_cmp
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Stride-halving tree over one value per thread; the result is returned to
// every thread.
__device__ double tree256(double x, double* sh, double* out) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
  if (t < 128) sh[t] = sh[t] + sh[t + 128];
  __syncthreads();
  if (t < 64) {
    double v = sh[t] + sh[t + 64];
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) v += __shfl_down(v, h, 64);
    if (t == 0) *out = v;
  }
  __syncthreads();
  const double r = *out;
  __syncthreads();
  return r;
}

// Total of a node from its m chunk results (part[0..m)).
__device__ double fold_total(const double* part, int m, double* sh, double* out) {
  const int t = threadIdx.x;
  double q = 0.0;
  if (t < m) {
    q = part[t];
    for (int i = t + kB; i < m; i += kB) q += part[i];
  }
  return tree256(q, sh, out);
}

constexpr int kNbChunk = 8;  // neighbours whose loads are issued together

__global__ __launch_bounds__(kB) void label_kernel(const Seg* __restrict__ segs,
                                                   const int* __restrict__ perm,
                                                   int* __restrict__ pos, int* __restrict__ cl) {
  const Seg s = segs[blockIdx.x];
  for (int i = threadIdx.x; i < s.n; i += kB) {
    const int p = s.start + i;
    const int g = perm[p];
    pos[g] = p;
    cl[g] = blockIdx.x;
  }
}

__global__ __launch_bounds__(kB) void count_kernel(int n, const int* __restrict__ perm,
                                                   const int* __restrict__ cl,
                                                   const int* __restrict__ aptr,
                                                   const int* __restrict__ adj,
                                                   int* __restrict__ cnt) {
  const int p = blockIdx.x * kB + threadIdx.x;
  if (p >= n) return;
  const int g = perm[p];
  const int L = cl[g];
  int c = 0;
  if (L >= 0) {
    // Neighbours in chunks: their indices, then their labels, are loaded
    // together (independent loads in flight instead of two dependent loads
    // per neighbour).
    const int k1 = aptr[g + 1];
    for (int k0 = aptr[g]; k0 < k1; k0 += kNbChunk) {
      int nb[kNbChunk];
#pragma unroll
      for (int q = 0; q < kNbChunk; ++q) nb[q] = k0 + q < k1 ? adj[k0 + q] : -1;
#pragma unroll
      for (int q = 0; q < kNbChunk; ++q) c += nb[q] >= 0 && cl[nb[q]] == L;
    }
  }
  cnt[p] = c;
}

__global__ __launch_bounds__(kB) void fill_col_kernel(int n, const int* __restrict__ perm,
                                                      const int* __restrict__ cl,
                                                      const int* __restrict__ pos,
                                                      const int* __restrict__ aptr,
                                                      const int* __restrict__ adj,
                                                      const int* __restrict__ rp,
                                                      int* __restrict__ col) {
  const int p = blockIdx.x * kB + threadIdx.x;
  if (p >= n) return;
  const int g = perm[p];
  const int L = cl[g];
  if (L < 0) return;
  int o = rp[p];
  const int k1 = aptr[g + 1];
  for (int k0 = aptr[g]; k0 < k1; k0 += kNbChunk) {
    int nb[kNbChunk], lb[kNbChunk], ps[kNbChunk];
#pragma unroll
    for (int q = 0; q < kNbChunk; ++q) nb[q] = k0 + q < k1 ? adj[k0 + q] : -1;
#pragma unroll
    for (int q = 0; q < kNbChunk; ++q) {
      lb[q] = nb[q] >= 0 ? cl[nb[q]] : -1;
      ps[q] = nb[q] >= 0 ? pos[nb[q]] : 0;
    }
#pragma unroll
    for (int q = 0; q < kNbChunk; ++q)  // adjacency order
      if (nb[q] >= 0 && lb[q] == L) col[o++] = ps[q];
  }
}

// ---- grid path: one block per 256-position chunk of a node ----
__device__ __forceinline__ void chunk_of(const Seg* segs, const int* ch_seg, int& si, Seg& s,
                                         int& i) {
  si = ch_seg[blockIdx.x];
  s = segs[si];
  i = (blockIdx.x - s.c0) * kB + threadIdx.x;
}

template <bool kFp64>
__global__ __launch_bounds__(kB) void grid_init(const Seg* __restrict__ segs,
                                                const int* __restrict__ ch_seg,
                                                const float* __restrict__ sinit,
                                                float* __restrict__ lv, double* __restrict__ psq) {
  __shared__ double sh[kB];
  __shared__ double out;
  int si, i;
  Seg s;
  chunk_of(segs, ch_seg, si, s, i);
  double q = 0.0;
  if (i < s.n) {
    const float a = sinit[i];
    lv[s.start + i] = a;
    q = (double)a * (double)a;
  }
  if (!kFp64) return;
  const double r = tree256(q, sh, &out);
  if (threadIdx.x == 0) psq[blockIdx.x] = r;
}

// Node totals when a node has too many chunks for every block to fold them.
__global__ __launch_bounds__(kB) void grid_totals(const Seg* __restrict__ segs,
                                                  const double* __restrict__ part,
                                                  double* __restrict__ tot) {
  __shared__ double sh[kB];
  __shared__ double out;
  const Seg s = segs[blockIdx.x];
  const double r = fold_total(part + s.c0, s.m, sh, &out);
  if (threadIdx.x == 0) tot[blockIdx.x] = r;
}

__device__ __forceinline__ double node_total(const Seg& s, int si, const double* part,
                                             const double* tot, double* sh, double* out) {
  if (s.m > kInlineChunks) return tot[si];
  return fold_total(part + s.c0, s.m, sh, out);
}

__global__ __launch_bounds__(kB) void grid_normalise(const Seg* __restrict__ segs,
                                                     const int* __restrict__ ch_seg,
                                                     const float* __restrict__ lv,
                                                     const double* __restrict__ psq,
                                                     const double* __restrict__ tot,
                                                     float* __restrict__ v) {
  __shared__ double sh[kB];
  __shared__ double out;
  int si, i;
  Seg s;
  chunk_of(segs, ch_seg, si, s, i);
  const double nrm = sqrt(node_total(s, si, psq, tot, sh, &out));
  if (i < s.n) {
    const float a = lv[s.start + i];
    v[s.start + i] = nrm > 1e-10 ? (float)((double)a / nrm) : a;
  }
}

template <bool kFp64>
__global__ __launch_bounds__(kB) void grid_spmv(const Seg* __restrict__ segs,
                                                const int* __restrict__ ch_seg,
                                                const int* __restrict__ rp,
                                                const int* __restrict__ col,
                                                const float* __restrict__ v,
                                                float* __restrict__ lv,
                                                double* __restrict__ psum) {
  __shared__ double sh[kB];
  __shared__ double out;
  int si, i;
  Seg s;
  chunk_of(segs, ch_seg, si, s, i);
  double q = 0.0;
  if (i < s.n) {
    const int p = s.start + i;
    int k = rp[p];
    const int e = rp[p + 1];
    float a = (float)(e - k) * v[p];
    for (; k + 8 <= e; k += 8) {
      int c[8];
      float g[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) c[u] = col[k + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) g[u] = v[c[u]];
#pragma unroll
      for (int u = 0; u < 8; ++u) a -= g[u];
    }
    for (; k < e; ++k) a -= v[col[k]];
    lv[p] = a;
    q = (double)a;
  }
  if (!kFp64) return;
  const double r = tree256(q, sh, &out);
  if (threadIdx.x == 0) psum[blockIdx.x] = r;
}

__global__ __launch_bounds__(kB) void grid_center(const Seg* __restrict__ segs,
                                                  const int* __restrict__ ch_seg,
                                                  float* __restrict__ lv,
                                                  const double* __restrict__ psum,
                                                  const double* __restrict__ tot,
                                                  double* __restrict__ psq) {
  __shared__ double sh[kB];
  __shared__ double out;
  int si, i;
  Seg s;
  chunk_of(segs, ch_seg, si, s, i);
  const float mean = (float)(node_total(s, si, psum, tot, sh, &out) / (double)s.n);
  double q = 0.0;
  if (i < s.n) {
    const float t = lv[s.start + i] - mean;
    lv[s.start + i] = t;
    q = (double)t * (double)t;
  }
  const double r = tree256(q, sh, &out);
  if (threadIdx.x == 0) psq[blockIdx.x] = r;
}

// ---- exact mode: FP32 sums in local order, one lane per node ----
constexpr int kStage = 2048;
constexpr int kPad = 32;  // floats readable past the end of a serial_sum input

// s += b_i for i = 0..len-1, in order, over values in LDS (b must hold
// len + kPad floats). One lane runs
// it: each add waits for the previous one, so the loop body is nothing but
// the adds and the LDS reads that feed them.
__device__ __forceinline__ void add32(float& s, const float4 (&q)[8]) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    s += q[u].x;
    s += q[u].y;
    s += q[u].z;
    s += q[u].w;
  }
}

__device__ __forceinline__ float serial_sum(const float* b, int len, float s) {
  const float4* b4 = reinterpret_cast<const float4*>(b);
  int i = 0;
  // Two register sets alternate: while one set's 32 values are added, the
  // other set's LDS reads are in flight (no register copies in the loop).
  float4 qa[8], qb[8];
  if (len >= 64) {
#pragma unroll
    for (int u = 0; u < 8; ++u) qa[u] = b4[u];
  }
  for (; i + 64 <= len; i += 64) {
#pragma unroll
    for (int u = 0; u < 8; ++u) qb[u] = b4[((i + 32) >> 2) + u];
    add32(s, qa);
    // Unconditional: b holds kPad floats past its last value, and values
    // read past len are never added.
#pragma unroll
    for (int u = 0; u < 8; ++u) qa[u] = b4[((i + 64) >> 2) + u];
    add32(s, qb);
  }
  for (; i < len; ++i) s += b[i];
  return s;
}

// (x - mean) * (x - mean) in FP32, the reference's square of a centred
// entry (partitioner.rs:169-174); x itself when !kSquare.
template <bool kSquare>
__device__ __forceinline__ float term(float x, float mean) {
  if (!kSquare) return x;
  const float t = x - mean;
  return t * t;
}

// The same over the terms of n values in global memory: the workgroup
// computes and stages kStage terms at a time in LDS (double-buffered)
// while lane 0 accumulates. Result valid in thread 0.
template <bool kSquare>
__device__ float chain_sum(const float* __restrict__ src, int n, float mean,
                           float (*buf)[kStage + kPad]) {
  const int t = threadIdx.x;
  const int nch = (n + kStage - 1) / kStage;
  constexpr int kPer = kStage / kB;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = k * kB + t;
    if (i < n) buf[0][i] = term<kSquare>(src[i], mean);
  }
  __syncthreads();
  float s = 0.0f;
  for (int c = 0; c < nch; ++c) {
    const int next = (c + 1) * kStage;
    float r[kPer];
    if (c + 1 < nch) {
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = next + k * kB + t;
        r[k] = i < n ? src[i] : 0.0f;
      }
    }
    if (t == 0) s = serial_sum(buf[c & 1], min(kStage, n - c * kStage), s);
    if (c + 1 < nch) {
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = k * kB + t;
        if (next + i < n) buf[(c + 1) & 1][i] = term<kSquare>(r[k], mean);
      }
    }
    __syncthreads();
  }
  return s;
}

// Per grid-path node: mean (center != 0; else 0) and norm of lv.
__global__ __launch_bounds__(kB) void chain_kernel(const Seg* __restrict__ segs,
                                                   const float* __restrict__ lv, int center,
                                                   float2* __restrict__ stats) {
  __shared__ __attribute__((aligned(16))) float buf[2][kStage + kPad];
  __shared__ float sh_mean;
  const Seg s = segs[blockIdx.x];
  const float* src = lv + s.start;
  float mean = 0.0f;
  if (center) {
    const float sum = chain_sum<false>(src, s.n, 0.0f, buf);
    if (threadIdx.x == 0) sh_mean = sum / (float)s.n;
    __syncthreads();
    mean = sh_mean;
  }
  const float sq = chain_sum<true>(src, s.n, mean, buf);
  if (threadIdx.x == 0) stats[blockIdx.x] = make_float2(mean, sqrtf(sq));
}

__global__ __launch_bounds__(kB) void grid_normalise_exact(const Seg* __restrict__ segs,
                                                           const int* __restrict__ ch_seg,
                                                           const float* __restrict__ lv,
                                                           const float2* __restrict__ stats,
                                                           float* __restrict__ v) {
  int si, i;
  Seg s;
  chunk_of(segs, ch_seg, si, s, i);
  if (i >= s.n) return;
  const float2 st = stats[si];
  const float a = lv[s.start + i] - st.x;
  v[s.start + i] = st.y > 1e-10f ? a / st.y : a;
}

// Sort keys: (tile index, value) for grid-path nodes, (tile index, 0) for
// every other tile, which the stable sort leaves in place.
__global__ __launch_bounds__(kB) void tile_keys(const Tile* __restrict__ tiles,
                                                const float* __restrict__ v,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals, int* nan_flag) {
  const Tile t = tiles[blockIdx.x];
  const uint64_t hi = (uint64_t)blockIdx.x << 32;
  for (int i = threadIdx.x; i < t.n; i += kB) {
    const int p = t.start + i;
    keys[p] = hi | (t.label >= 0 ? order_bits(v[p], nan_flag) : 0u);
    vals[p] = (uint32_t)p;
  }
}

__global__ __launch_bounds__(kB) void gather_perm(int n, const uint32_t* __restrict__ src,
                                                  const int* __restrict__ perm,
                                                  int* __restrict__ out) {
  const int q = blockIdx.x * kB + threadIdx.x;
  if (q < n) out[q] = perm[src[q]];
}

// ---- workgroup path: one workgroup bisects one node of <= kWgMax gates ----
// Reduces kRows chunks held one value per thread each (chunks >= m are
// absent); returns the node total to every thread.
__device__ double wg_total(const double (&x)[kRows], int m, double (*red)[kB], double* out) {
  const int t = threadIdx.x;
  for (int c = 0; c < m; ++c) red[c][t] = x[c];
  __syncthreads();
  if (t < 128)
    for (int c = 0; c < m; ++c) red[c][t] = red[c][t] + red[c][t + 128];
  __syncthreads();
  if (t < 64) {
    double P[kRows];
#pragma unroll
    for (int c = 0; c < kRows; ++c) {
      double v = 0.0;
      if (c < m) {
        v = red[c][t] + red[c][t + 64];
        for (int h = 32; h > 0; h >>= 1) v += __shfl_down(v, h, 64);
      }
      P[c] = v;
    }
    if (t == 0) {
      // The fold's tree over Q = (P_0..P_7, +0.0, ...): levels h = 128..8
      // only add +0.0, then h = 4, 2, 1.
#pragma unroll
      for (int c = 0; c < kRows; ++c) P[c] = P[c] + 0.0;
      double s0 = P[0] + P[4], s1 = P[1] + P[5], s2 = P[2] + P[6], s3 = P[3] + P[7];
      s0 = s0 + s2;
      s1 = s1 + s3;
      *out = s0 + s1;
    }
  }
  __syncthreads();
  const double r = *out;
  __syncthreads();
  return r;
}

template <bool kFp64>
__global__ __launch_bounds__(kB) void wg_bisect(const Seg* __restrict__ segs,
                                                const int* __restrict__ rp,
                                                const int* __restrict__ col,
                                                const float* __restrict__ sinit, int iters,
                                                int* __restrict__ perm, int* nan_flag) {
  __shared__ __attribute__((aligned(16))) float v[kWgMax + kPad];
  __shared__ __attribute__((aligned(16))) float lv[kWgMax + kPad];
  __shared__ uint64_t kbuf[kWgMax];  // reduction slots, then sort keys
  __shared__ double out;
  __shared__ float outf;
  double(*red)[kB] = reinterpret_cast<double(*)[kB]>(kbuf);
  const Seg s = segs[blockIdx.x];
  const int t = threadIdx.x;
  const int n = s.n, m = (n + kB - 1) / kB;
  double x[kRows];
  // Mean (center) and norm of lv; v = (lv - mean) / norm.
  auto normalise = [&](bool center) {
    if (kFp64) {
      float mean = 0.0f;
      if (center) {
#pragma unroll
        for (int c = 0; c < kRows; ++c) {
          const int i = c * kB + t;
          x[c] = i < n ? (double)lv[i] : 0.0;
        }
        mean = (float)(wg_total(x, m, red, &out) / (double)n);
      }
#pragma unroll
      for (int c = 0; c < kRows; ++c) {
        const int i = c * kB + t;
        x[c] = 0.0;
        if (i < n) {
          const float u = lv[i] - mean;
          lv[i] = u;
          x[c] = (double)u * (double)u;
        }
      }
      const double nrm = sqrt(wg_total(x, m, red, &out));
#pragma unroll
      for (int c = 0; c < kRows; ++c) {
        const int i = c * kB + t;
        if (i < n) v[i] = nrm > 1e-10 ? (float)((double)lv[i] / nrm) : lv[i];
      }
    } else {
      float mean = 0.0f;
      if (center) {
        if (t == 0) outf = serial_sum(lv, n, 0.0f) / (float)n;
        __syncthreads();
        mean = outf;
      }
      // Squares in parallel (v is free until the normalisation below).
      for (int i = t; i < n; i += kB) v[i] = term<true>(lv[i], mean);
      __syncthreads();
      if (t == 0) outf = sqrtf(serial_sum(v, n, 0.0f));
      __syncthreads();
      const float nrm = outf;
#pragma unroll
      for (int c = 0; c < kRows; ++c) {
        const int i = c * kB + t;
        if (i < n) {
          const float u = lv[i] - mean;
          v[i] = nrm > 1e-10f ? u / nrm : u;
        }
      }
    }
    __syncthreads();
  };

  // Initial vector and its normalisation (partitioner.rs:145-155).
  for (int i = t; i < n; i += kB) lv[i] = sinit[i];
  __syncthreads();
  normalise(false);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < kRows; ++c) {
      const int i = c * kB + t;
      if (i < n) {
        const int p = s.start + i;
        const int e = rp[p + 1];
        int k = rp[p];
        float a = (float)(e - k) * v[i];
        for (; k < e; ++k) a -= v[col[k] - s.start];
        lv[i] = a;
      }
    }
    __syncthreads();
    normalise(true);
  }

  // Stable sort by value: (value bits, local index) keys are distinct.
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (int i = t; i < np2; i += kB)
    kbuf[i] = i < n ? ((uint64_t)order_bits(v[i], nan_flag) << 32) | (uint32_t)i : ~0ull;
  int* old = reinterpret_cast<int*>(lv);
  for (int i = t; i < n; i += kB) old[i] = perm[s.start + i];
  __syncthreads();
  for (int k = 2; k <= np2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < np2; i += kB) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = kbuf[i], b = kbuf[ixj];
          if ((a > b) == ((i & k) == 0)) {
            kbuf[i] = b;
            kbuf[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int r = t; r < n; r += kB) perm[s.start + r] = old[(uint32_t)kbuf[r]];
}

// ---- host planning ----
struct PNode {
  int32_t start, count, depth, left, right, split;
};

bool config_ok(const fx_partition_config* c) {
  return c && c->max_leaf_size >= 0 && c->min_cluster_size >= 0 && c->power_iterations >= 0 &&
         c->balance_factor >= 0.0f && c->balance_factor <= 1.0f;
}

int32_t split_of(int32_t n, float bf) {
  const int32_t min_size = (int32_t)std::ceil((float)n * bf);
  return std::min(std::max(n / 2, min_size), n - min_size);
}

// recursive_bisect's node sequence (partitioner.rs:64-100) from sizes alone.
std::vector<PNode> plan(int32_t num_gates, const fx_partition_config& c) {
  std::vector<PNode> nodes;
  struct Item { int32_t start, n, depth, parent, side; };
  std::vector<Item> stack{{0, num_gates, 0, -1, 0}};
  while (!stack.empty()) {
    const Item it = stack.back();
    stack.pop_back();
    const int32_t id = (int32_t)nodes.size();
    nodes.push_back({it.start, it.n, it.depth, -1, -1, 0});
    if (it.parent >= 0) (it.side ? nodes[it.parent].right : nodes[it.parent].left) = id;
    if (it.n <= c.max_leaf_size || it.n <= c.min_cluster_size) continue;
    const int32_t split = split_of(it.n, c.balance_factor);
    if (split <= 0 || split >= it.n) continue;
    nodes[id].split = split;
    stack.push_back({it.start + split, it.n - split, it.depth + 1, id, 1});
    stack.push_back({it.start, split, it.depth + 1, id, 0});
  }
  return nodes;
}

}  // namespace
}  // namespace fx

extern "C" void fx_partition_defaults(fx_partition_config* c) {
  if (!c) return;
  c->max_leaf_size = 64;
  c->min_cluster_size = 8;
  c->balance_factor = 0.4f;
  c->power_iterations = 30;
}

extern "C" int fx_hierarchy_size(int32_t num_gates, const fx_partition_config* cfg,
                                 int32_t* num_nodes, int32_t* max_depth) {
  fx::clear_error();
  FX_CHECK(num_gates >= 0 && fx::config_ok(cfg) && num_nodes, FX_E_INVALID,
           "fx_hierarchy_size: bad arguments");
  const std::vector<fx::PNode> nodes = fx::plan(num_gates, *cfg);
  *num_nodes = (int32_t)nodes.size();
  if (max_depth) {
    int32_t d = 0;
    for (const auto& nd : nodes) d = std::max(d, nd.depth);
    *max_depth = d;
  }
  return FX_OK;
}

extern "C" int fx_build_hierarchy(fx_ctx* ctx, int32_t num_gates, const int32_t* adj_ptr,
                                  const int32_t* adj, const fx_partition_config* cfg,
                                  int32_t mode, int32_t* order, int32_t* node_start,
                                  int32_t* node_count,
                                  int32_t* node_depth, int32_t* node_left, int32_t* node_right,
                                  int32_t capacity, int32_t* num_nodes) {
  using namespace fx;
  clear_error();
  FX_CHECK(ctx && num_gates >= 0 && (num_gates == 0 || (adj_ptr && order)) && config_ok(cfg) &&
               node_start && node_count && node_depth && node_left && node_right && num_nodes,
           FX_E_INVALID, "fx_build_hierarchy: bad arguments");
  FX_CHECK(mode == FX_HIER_EXACT || mode == FX_HIER_FP64, FX_E_INVALID,
           "fx_build_hierarchy: unknown mode %d", mode);
  const bool fp64 = mode == FX_HIER_FP64;
  const std::vector<PNode> nodes = plan(num_gates, *cfg);
  FX_CHECK((int64_t)nodes.size() <= capacity, FX_E_INVALID,
           "fx_build_hierarchy: %zu nodes exceed capacity %d", nodes.size(), capacity);
  const int64_t nnz = num_gates ? adj_ptr[num_gates] : 0;
  FX_CHECK(adj_ptr == nullptr || adj_ptr[0] == 0, FX_E_INVALID, "adj_ptr[0] must be 0");
  FX_CHECK(nnz >= 0 && nnz < INT32_MAX && (nnz == 0 || adj), FX_E_INVALID,
           "fx_build_hierarchy: bad adjacency");
  for (int32_t g = 0; g < num_gates; ++g)
    FX_CHECK(adj_ptr[g + 1] >= adj_ptr[g], FX_E_INVALID, "adj_ptr not monotone at %d", g);
  for (int64_t k = 0; k < nnz; ++k)
    FX_CHECK(adj[k] >= 0 && adj[k] < num_gates, FX_E_INVALID, "adjacency entry %lld out of range",
             (long long)k);
  *num_nodes = (int32_t)nodes.size();
  for (size_t i = 0; i < nodes.size(); ++i) {
    node_start[i] = nodes[i].start;
    node_count[i] = nodes[i].count;
    node_depth[i] = nodes[i].depth;
    node_left[i] = nodes[i].left;
    node_right[i] = nodes[i].right;
  }
  ctx->hierarchy_ms = 0.0;
  int32_t max_depth = 0;
  int32_t max_split_n = 0;
  for (const auto& nd : nodes) {
    max_depth = std::max(max_depth, nd.depth);
    if (nd.split) max_split_n = std::max(max_split_n, nd.count);
  }
  if (max_split_n == 0) {  // the root is a leaf
    for (int32_t g = 0; g < num_gates; ++g) order[g] = g;
    return FX_OK;
  }

  // Per-level tables, all uploaded once before the first kernel.
  struct Level {
    int64_t seg_off = 0, ch_off = 0, tile_off = 0;
    int32_t nbig = 0, nsmall = 0, nchunks = 0, ntiles = 0, max_m = 0;
  };
  std::vector<Level> levels(max_depth + 1);
  std::vector<int32_t> tab;  // Seg, chunk map and Tile records, int32 words
  std::vector<std::vector<int32_t>> by_depth(max_depth + 1);
  for (int32_t id = 0; id < (int32_t)nodes.size(); ++id) by_depth[nodes[id].depth].push_back(id);
  std::vector<int32_t> tiling;  // nodes that tile [0, N) at the current depth, by start
  for (int32_t d = 0; d <= max_depth; ++d) {
    Level& L = levels[d];
    // tiling at depth d: leaves above d plus every node at depth d
    std::vector<int32_t> t2;
    for (int32_t id : tiling)
      if (nodes[id].split == 0) t2.push_back(id);
    for (int32_t id : by_depth[d]) t2.push_back(id);
    std::sort(t2.begin(), t2.end(),
              [&](int32_t a, int32_t b) { return nodes[a].start < nodes[b].start; });
    tiling.swap(t2);
    std::vector<int32_t> big, small;
    for (int32_t id : by_depth[d])
      if (nodes[id].split) (nodes[id].count > kWgMax ? big : small).push_back(id);
    L.nbig = (int32_t)big.size();
    L.nsmall = (int32_t)small.size();
    L.seg_off = (int64_t)tab.size();
    int32_t c0 = 0;
    for (int32_t id : big) {
      const int32_t m = (int32_t)div_up(nodes[id].count, kB);
      tab.insert(tab.end(), {nodes[id].start, nodes[id].count, c0, m});
      c0 += m;
      L.max_m = std::max(L.max_m, m);
    }
    for (int32_t id : small)
      tab.insert(tab.end(), {nodes[id].start, nodes[id].count, 0, (int32_t)div_up(nodes[id].count, kB)});
    L.nchunks = c0;
    L.ch_off = (int64_t)tab.size();
    for (int32_t b = 0; b < L.nbig; ++b)
      for (int32_t c = 0; c < tab[L.seg_off + 4 * b + 3]; ++c) tab.push_back(b);
    L.tile_off = (int64_t)tab.size();
    if (L.nbig) {
      std::vector<int32_t> label_of(nodes.size(), -1);
      for (int32_t b = 0; b < L.nbig; ++b) label_of[big[b]] = b;
      for (int32_t id : tiling) tab.insert(tab.end(), {nodes[id].start, nodes[id].count, label_of[id], 0});
      L.ntiles = (int32_t)tiling.size();
    }
  }
  int32_t max_chunks = 1, max_segs = 1;
  int64_t max_tiles = 1;
  for (const Level& L : levels) {
    max_chunks = std::max(max_chunks, L.nchunks);
    max_segs = std::max(max_segs, L.nbig + L.nsmall);
    max_tiles = std::max<int64_t>(max_tiles, L.ntiles);
  }
  FX_CHECK(max_tiles < (1ll << 31), FX_E_INVALID, "too many tiles");
  // Initial vector sin(i * 2.71828) in FP32 (partitioner.rs:145-147),
  // computed with the host's libm.
  std::vector<float> sinit(max_split_n);
  fx_fiedler_init(max_split_n, sinit.data());

  std::lock_guard<std::mutex> lock(ctx->mu);
  FX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t N = num_gates;
  const int64_t scan_tmp = scan_tmp_elems(N);
  DevBuf d_aptr, d_adj, d_perm, d_perm2, d_pos, d_cl, d_cnt, d_rp, d_col, d_v, d_lv, d_psum, d_psq,
      d_tot, d_stats, d_keys, d_vals, d_sinit, d_tab, d_tmp, d_flag;
  SortWorkspace ws;
  int rc;
  if ((rc = d_aptr.alloc(ctx, 4 * (N + 1))) || (rc = d_adj.alloc(ctx, 4 * (nnz + 1))) ||
      (rc = d_perm.alloc(ctx, 4 * N)) || (rc = d_perm2.alloc(ctx, 4 * N)) ||
      (rc = d_pos.alloc(ctx, 4 * N)) || (rc = d_cl.alloc(ctx, 4 * N)) ||
      (rc = d_cnt.alloc(ctx, 4 * N)) || (rc = d_rp.alloc(ctx, 4 * (N + 1))) ||
      (rc = d_col.alloc(ctx, 4 * (nnz + 1))) || (rc = d_v.alloc(ctx, 4 * N)) ||
      (rc = d_lv.alloc(ctx, 4 * N)) || (rc = d_psum.alloc(ctx, 8 * (size_t)max_chunks)) ||
      (rc = d_psq.alloc(ctx, 8 * (size_t)max_chunks)) || (rc = d_tot.alloc(ctx, 8 * (size_t)max_segs)) ||
      (rc = d_stats.alloc(ctx, sizeof(float2) * (size_t)max_segs)) ||
      (rc = d_keys.alloc(ctx, 8 * N)) || (rc = d_vals.alloc(ctx, 4 * N)) ||
      (rc = d_sinit.alloc(ctx, 4 * (size_t)max_split_n)) ||
      (rc = d_tab.alloc(ctx, 4 * (tab.size() + 4))) || (rc = d_tmp.alloc(ctx, 4 * scan_tmp)) ||
      (rc = d_flag.alloc(ctx, 4)))
    return rc;
  std::vector<int32_t> iota(N);
  for (int32_t g = 0; g < num_gates; ++g) iota[g] = g;
  FX_HIP(hipMemcpyAsync(d_aptr.p, adj_ptr, 4 * (N + 1), hipMemcpyHostToDevice, st));
  if (nnz) FX_HIP(hipMemcpyAsync(d_adj.p, adj, 4 * nnz, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(d_perm.p, iota.data(), 4 * N, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(d_sinit.p, sinit.data(), 4 * (size_t)max_split_n, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(d_tab.p, tab.data(), 4 * tab.size(), hipMemcpyHostToDevice, st));
  if ((rc = fill_i32(st, d_flag.as<int32_t>(), 1, 0))) return rc;
  FX_HIP(hipEventRecord(ctx->ev0, st));
  const dim3 blk(kB);
  const unsigned nblk = div_up(N, kB);
  const int iters = cfg->power_iterations;
  int* perm = d_perm.as<int>();
  int* perm2 = d_perm2.as<int>();
  for (int32_t d = 0; d <= max_depth; ++d) {
    const Level& L = levels[d];
    const int nseg = L.nbig + L.nsmall;
    if (!nseg) continue;
    const Seg* segs = reinterpret_cast<const Seg*>(d_tab.as<int32_t>() + L.seg_off);
    const int* ch_seg = d_tab.as<int32_t>() + L.ch_off;
    if ((rc = fill_i32(st, d_cl.as<int32_t>(), N, -1))) return rc;
    hipLaunchKernelGGL(label_kernel, dim3(nseg), blk, 0, st, segs, perm, d_pos.as<int>(),
                       d_cl.as<int>());
    hipLaunchKernelGGL(count_kernel, dim3(nblk), blk, 0, st, (int)N, perm, d_cl.as<int>(),
                       d_aptr.as<int>(), d_adj.as<int>(), d_cnt.as<int>());
    if ((rc = scan_exclusive_i32(st, d_cnt.as<int32_t>(), d_rp.as<int32_t>(), N, d_tmp.as<int32_t>(),
                                 scan_tmp)))
      return rc;
    hipLaunchKernelGGL(fill_col_kernel, dim3(nblk), blk, 0, st, (int)N, perm, d_cl.as<int>(),
                       d_pos.as<int>(), d_aptr.as<int>(), d_adj.as<int>(), d_rp.as<int>(),
                       d_col.as<int>());
    FX_HIP(hipGetLastError());
    if (L.nbig) {
      const dim3 cg(L.nchunks);
      const bool need_tot = L.max_m > kInlineChunks;
      double* tot = d_tot.as<double>();
      double* psum = d_psum.as<double>();
      double* psq = d_psq.as<double>();
      float2* stats = d_stats.as<float2>();
      float* v = d_v.as<float>();
      float* lv = d_lv.as<float>();
      const int* rpp = d_rp.as<int>();
      const int* colp = d_col.as<int>();
      if (fp64) {
        hipLaunchKernelGGL(grid_init<true>, cg, blk, 0, st, segs, ch_seg, d_sinit.as<float>(), lv, psq);
        if (need_tot) hipLaunchKernelGGL(grid_totals, dim3(L.nbig), blk, 0, st, segs, psq, tot);
        hipLaunchKernelGGL(grid_normalise, cg, blk, 0, st, segs, ch_seg, lv, psq, tot, v);
        for (int it = 0; it < iters; ++it) {
          hipLaunchKernelGGL(grid_spmv<true>, cg, blk, 0, st, segs, ch_seg, rpp, colp, v, lv, psum);
          if (need_tot) hipLaunchKernelGGL(grid_totals, dim3(L.nbig), blk, 0, st, segs, psum, tot);
          hipLaunchKernelGGL(grid_center, cg, blk, 0, st, segs, ch_seg, lv, psum, tot, psq);
          if (need_tot) hipLaunchKernelGGL(grid_totals, dim3(L.nbig), blk, 0, st, segs, psq, tot);
          hipLaunchKernelGGL(grid_normalise, cg, blk, 0, st, segs, ch_seg, lv, psq, tot, v);
        }
      } else {
        hipLaunchKernelGGL(grid_init<false>, cg, blk, 0, st, segs, ch_seg, d_sinit.as<float>(), lv, psq);
        hipLaunchKernelGGL(chain_kernel, dim3(L.nbig), blk, 0, st, segs, lv, 0, stats);
        hipLaunchKernelGGL(grid_normalise_exact, cg, blk, 0, st, segs, ch_seg, lv, stats, v);
        for (int it = 0; it < iters; ++it) {
          hipLaunchKernelGGL(grid_spmv<false>, cg, blk, 0, st, segs, ch_seg, rpp, colp, v, lv, psum);
          hipLaunchKernelGGL(chain_kernel, dim3(L.nbig), blk, 0, st, segs, lv, 1, stats);
          hipLaunchKernelGGL(grid_normalise_exact, cg, blk, 0, st, segs, ch_seg, lv, stats, v);
        }
      }
      const Tile* tiles = reinterpret_cast<const Tile*>(d_tab.as<int32_t>() + L.tile_off);
      hipLaunchKernelGGL(tile_keys, dim3(L.ntiles), blk, 0, st, tiles, v, d_keys.as<uint64_t>(),
                         d_vals.as<uint32_t>(), d_flag.as<int>());
      int tile_bits = 1;
      while ((1ll << tile_bits) < L.ntiles) ++tile_bits;
      if ((rc = radix_sort_pairs(ctx, st, ws, d_keys.as<uint64_t>(), d_vals.as<uint32_t>(), N, 0,
                                 32 + tile_bits)))
        return rc;
      hipLaunchKernelGGL(gather_perm, dim3(nblk), blk, 0, st, (int)N, d_vals.as<uint32_t>(), perm,
                         perm2);
      std::swap(perm, perm2);
      FX_HIP(hipGetLastError());
    }
    if (L.nsmall) {
      if (fp64)
        hipLaunchKernelGGL(wg_bisect<true>, dim3(L.nsmall), blk, 0, st, segs + L.nbig, d_rp.as<int>(),
                           d_col.as<int>(), d_sinit.as<float>(), iters, perm, d_flag.as<int>());
      else
        hipLaunchKernelGGL(wg_bisect<false>, dim3(L.nsmall), blk, 0, st, segs + L.nbig, d_rp.as<int>(),
                           d_col.as<int>(), d_sinit.as<float>(), iters, perm, d_flag.as<int>());
      FX_HIP(hipGetLastError());
    }
  }
  FX_HIP(hipEventRecord(ctx->ev1, st));
  int nan_flag = 0;
  FX_HIP(hipMemcpyAsync(&nan_flag, d_flag.p, 4, hipMemcpyDeviceToHost, st));
  FX_HIP(hipMemcpyAsync(order, perm, 4 * N, hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  bool faulted = false;
  if ((rc = radix_fault(ws, st, &faulted))) return rc;
  FX_CHECK(!faulted, FX_E_DEVICE, "radix sort look-back timed out");
  FX_CHECK(!nan_flag, FX_E_INVALID,
           "fx_build_hierarchy: NaN Fiedler value (partial_cmp has no order for NaN)");
  float ms = 0.f;
  FX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->hierarchy_ms = ms;
  return FX_OK;
}

extern "C" double fx_hierarchy_last_ms(const fx_ctx* ctx) { return ctx ? ctx->hierarchy_ms : 0.0; }

// Regions of hapr_placement: the root gets the die (placer.rs:43-48), every
// level's children a grid inside their parent (place_level_recursive,
// 111-141), in FP32 as the reference computes them. The spectral branch for
// depth 0 (52-58) never runs: attention has no level for the single root
// cluster (attention.rs:40-43), so attn_level is None.
extern "C" int fx_hierarchy_regions(int32_t num_nodes, const int32_t* node_left,
                                    const int32_t* node_right, float chip_width,
                                    float chip_height, float* region) {
  fx::clear_error();
  FX_CHECK(num_nodes >= 0 && (num_nodes == 0 || (node_left && node_right && region)),
           FX_E_INVALID, "fx_hierarchy_regions: bad arguments");
  if (!num_nodes) return FX_OK;
  region[0] = chip_width / 2.0f;
  region[1] = chip_height / 2.0f;
  region[2] = chip_width;
  region[3] = chip_height;
  // Pre-order ids: a parent precedes its children.
  for (int32_t id = 0; id < num_nodes; ++id) {
    const int32_t kids[2] = {node_left[id], node_right[id]};
    if (kids[0] < 0 && kids[1] < 0) continue;
    FX_CHECK(kids[0] > id && kids[1] > id && kids[0] < num_nodes && kids[1] < num_nodes,
             FX_E_INVALID, "node %d: children must follow it in pre-order", id);
    const int n = 2;
    const float* P = region + 4 * id;
    const float half_w = P[2] / 2.0f, half_h = P[3] / 2.0f;
    const float origin_x = P[0] - half_w, origin_y = P[1] - half_h;
    const int cols = (int)std::ceil(std::sqrt((float)n));
    const int rows = (n + cols - 1) / cols;
    for (int i = 0; i < n; ++i) {
      const int col = i % cols, row = i / cols;
      const float cw = P[2] / (float)cols;
      const float ch = P[3] / (float)rows;
      float* C = region + 4 * kids[i];
      C[0] = origin_x + ((float)col + 0.5f) * cw;
      C[1] = origin_y + ((float)row + 0.5f) * ch;
      C[2] = cw;
      C[3] = ch;
    }
  }
  return FX_OK;
}
