// Device-wide primitives: stable LSD radix sort of (u64 key, u32 value)
// pairs and exclusive prefix scans (int32, int64 pairs).
//
// Radix sort: 8-bit digits. Per pass, an upsweep kernel counts each tile's
// digits into a digit-major histogram, one workgroup scans it, and a
// downsweep kernel ranks the tile's items stably — items are processed in
// rounds of 256 in input order; inside a 64-lane wave, eight ballots match
// lanes with equal digits (rank = popcount of lower matching lanes), and
// per-wave digit counts combine through LDS — then scatters them.

#include "internal.h"

namespace fx {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(kSortThreads) void radix_upsweep(const uint64_t* __restrict__ keys,
                                                              int64_t n, int shift,
                                                              uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  hist[(int64_t)blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

// Exclusive prefix over the tiles of every digit, in place, then the
// exclusive prefix of the digit totals into dbase. 1024 threads: digit d is
// scanned by four threads, each over a quarter of the tiles (loads coalesced
// across digits), joined through LDS.
__global__ __launch_bounds__(1024) void radix_scan(uint32_t* __restrict__ hist, int nblocks,
                                                   uint32_t* __restrict__ dbase) {
  __shared__ uint32_t seg[4][256];
  const int d = threadIdx.x & 255, q = threadIdx.x >> 8;
  const int per = (nblocks + 3) / 4;
  const int b0 = q * per, b1 = min(nblocks, b0 + per);
  uint32_t s = 0;
  for (int b = b0; b < b1; ++b) s += hist[(int64_t)b * 256 + d];
  seg[q][d] = s;
  __syncthreads();
  uint32_t run = 0;
  for (int k = 0; k < q; ++k) run += seg[k][d];
  for (int b = b0; b < b1; ++b) {
    const uint32_t v = hist[(int64_t)b * 256 + d];
    hist[(int64_t)b * 256 + d] = run;
    run += v;
  }
  // Digit totals -> exclusive prefix: threads 0..255 (q == 0, waves 0-3) scan
  // their wave with shuffles; wave sums combine through LDS. Every thread
  // reaches both barriers.
  __shared__ uint32_t wsum[4];
  const uint32_t mine = seg[0][d] + seg[1][d] + seg[2][d] + seg[3][d];
  uint32_t incl = mine;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (q == 0) {
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
  }
  __syncthreads();
  if (q == 0) {
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += wsum[w];
    dbase[d] = add + incl - mine;
  }
}

__global__ __launch_bounds__(kSortThreads) void radix_downsweep(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ dbase,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcount[4][256];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  base[t] = hist[(int64_t)blockIdx.x * 256 + t] + dbase[t];
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int64_t tile = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = tile + r * kSortThreads + t;
    const bool valid = i < n;
    uint64_t k = 0;
    uint32_t v = 0;
    if (valid) {
      k = kin[i];
      v = vin[i];
    }
    const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
    wcount[0][t] = 0;
    wcount[1][t] = 0;
    wcount[2][t] = 0;
    wcount[3][t] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t off = 0;
    if (valid)
      for (int w = 0; w < wave; ++w) off += wcount[w][d];
    const uint32_t pos = base[d] + off + rank;
    __syncthreads();
    // Advance per-digit bases by this round's totals (thread t owns digit t).
    base[t] += wcount[0][t] + wcount[1][t] + wcount[2][t] + wcount[3][t];
    if (valid) {
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
  }
}

int SortWorkspace::ensure(fx_ctx* ctx, int64_t n) {
  if (n <= capacity) return FX_OK;
  const int64_t nb = (n + kSortTile - 1) / kSortTile;
  int rc;
  if ((rc = keys_alt.alloc(ctx, sizeof(uint64_t) * n)) || (rc = vals_alt.alloc(ctx, sizeof(uint32_t) * n)) ||
      (rc = hist.alloc(ctx, sizeof(uint32_t) * 256 * (nb + 1))))
    return rc;
  capacity = n;
  blocks = nb;
  return FX_OK;
}

int radix_sort_pairs(fx_ctx* ctx, hipStream_t st, SortWorkspace& ws, uint64_t* keys,
                     uint32_t* vals, int64_t n, int lo_bit, int hi_bit) {
  if (n <= 1) return FX_OK;
  int rc = ws.ensure(ctx, n);
  if (rc) return rc;
  const int nb = (int)((n + kSortTile - 1) / kSortTile);
  uint64_t* ka = keys;
  uint32_t* va = vals;
  uint64_t* kb = ws.keys_alt.as<uint64_t>();
  uint32_t* vb = ws.vals_alt.as<uint32_t>();
  uint32_t* hist = ws.hist.as<uint32_t>();
  uint32_t* dbase = hist + (size_t)256 * nb;
  int passes = 0;
  for (int shift = lo_bit; shift < hi_bit; shift += 8, ++passes) {
    hipLaunchKernelGGL(radix_upsweep, dim3(nb), dim3(kSortThreads), 0, st, ka, n, shift, hist);
    hipLaunchKernelGGL(radix_scan, dim3(1), dim3(1024), 0, st, hist, nb, dbase);
    hipLaunchKernelGGL(radix_downsweep, dim3(nb), dim3(kSortThreads), 0, st, ka, va, n, shift,
                       (const uint32_t*)hist, (const uint32_t*)dbase, kb, vb);
    FX_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (passes & 1) {
    FX_HIP(hipMemcpyAsync(keys, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
    FX_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  }
  return FX_OK;
}

// ---- exclusive scans ----
// Reduce-then-scan over tiles of 4096 elements: per-tile sums, a
// single-workgroup scan of the sums, then per-tile local scans.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T> struct ScanOps;
template <> struct ScanOps<int32_t> {
  __device__ static int32_t zero() { return 0; }
  __device__ static int32_t add(int32_t a, int32_t b) { return a + b; }
};
struct I64x2 {
  long long a, b;
};
template <> struct ScanOps<I64x2> {
  __device__ static I64x2 zero() { return {0, 0}; }
  __device__ static I64x2 add(I64x2 x, I64x2 y) { return {x.a + y.a, x.b + y.b}; }
};

template <typename T>
__device__ T block_exclusive_scan(T v, T* sh, T* total) {
  using O = ScanOps<T>;
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    T u = t >= off ? sh[t - off] : O::zero();
    __syncthreads();
    sh[t] = O::add(sh[t], u);
    __syncthreads();
  }
  T incl = sh[t];
  *total = sh[kScanThreads - 1];
  __syncthreads();
  T ex = t ? sh[t - 1] : O::zero();
  (void)incl;
  return ex;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(const T* __restrict__ in, int64_t n,
                                                               T* __restrict__ sums) {
  using O = ScanOps<T>;
  __shared__ T sh[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T s = O::zero();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s = O::add(s, in[base + k]);
  T total;
  (void)block_exclusive_scan(s, sh, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void scan_sums(T* __restrict__ sums, int64_t m,
                                                          T* __restrict__ grand) {
  using O = ScanOps<T>;
  __shared__ T sh[kScanThreads];
  const int64_t per = (m + kScanThreads - 1) / kScanThreads;
  const int64_t s = threadIdx.x * per, e = min(m, s + per);
  T acc = O::zero();
  for (int64_t i = s; i < e; ++i) acc = O::add(acc, sums[i]);
  T total;
  T run = block_exclusive_scan(acc, sh, &total);
  for (int64_t i = s; i < e; ++i) {
    const T v = sums[i];
    sums[i] = run;
    run = O::add(run, v);
  }
  if (threadIdx.x == 0 && grand) *grand = total;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void scan_tiles(const T* __restrict__ in, int64_t n,
                                                           const T* __restrict__ sums,
                                                           T* __restrict__ out) {
  using O = ScanOps<T>;
  __shared__ T sh[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = O::zero();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : O::zero();
    s = O::add(s, v[k]);
  }
  T total;
  T run = O::add(sums[blockIdx.x], block_exclusive_scan(s, sh, &total));
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) {
      out[base + k] = run;
      run = O::add(run, v[k]);
    }
  // The thread holding element n-1 writes the grand total to out[n].
  if (base <= n - 1 && n - 1 < base + kScanItems) out[n] = run;
}

template <typename T>
static int scan_exclusive_impl(hipStream_t st, const T* in, T* out, int64_t n, T* tmp,
                               int64_t tmp_elems) {
  if (n <= 0) return FX_OK;
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb > tmp_elems) return fail(FX_E_INVALID, "scan workspace too small");
  hipLaunchKernelGGL(scan_tile_sums<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, tmp);
  hipLaunchKernelGGL(scan_sums<T>, dim3(1), dim3(kScanThreads), 0, st, tmp, nb, (T*)nullptr);
  hipLaunchKernelGGL(scan_tiles<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n,
                     (const T*)tmp, out);
  FX_HIP(hipGetLastError());
  return FX_OK;
}

// out[0..n] = exclusive scan of in[0..n) followed by the total at out[n].
int scan_exclusive_i32(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, int32_t* tmp,
                       int64_t tmp_elems) {
  return scan_exclusive_impl<int32_t>(st, in, out, n, tmp, tmp_elems);
}
int scan_exclusive_i64x2(hipStream_t st, const void* in, void* out, int64_t n, void* tmp,
                         int64_t tmp_elems) {
  return scan_exclusive_impl<I64x2>(st, (const I64x2*)in, (I64x2*)out, n, (I64x2*)tmp, tmp_elems);
}
int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

}  // namespace fx

// ---- public sort entry points ----
namespace {

__global__ void iota_u32(uint32_t* v, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

// Order-preserving map of f32 to u32 with -0.0 == +0.0 (partial_cmp).
__global__ void f32_keys(const float* __restrict__ in, uint64_t* __restrict__ out, int64_t n,
                         int* nan_flag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f = in[i];
  if (f != f) {
    atomicOr(nan_flag, 1);
    f = 0.0f;
  }
  if (f == 0.0f) f = 0.0f;  // canonicalise -0.0
  uint32_t u = __float_as_uint(f);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  out[i] = (uint64_t)u;
}

int sort_common(fx_ctx* ctx, int64_t n, const uint64_t* hkeys, const float* hf, uint32_t* perm) {
  fx::clear_error();
  FX_CHECK(ctx && perm && (hkeys || hf) && n >= 0, FX_E_INVALID, "fx_sort: bad arguments");
  FX_CHECK(n < (int64_t)UINT32_MAX, FX_E_INVALID, "fx_sort: n too large");
  if (n == 0) return FX_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  FX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  fx::DevBuf k, v, fin, flag;
  fx::SortWorkspace ws;
  int rc;
  if ((rc = k.alloc(ctx, 8 * n)) || (rc = v.alloc(ctx, 4 * n)) || (rc = flag.alloc(ctx, 4)))
    return rc;
  int hi = 64;
  if (hkeys) {
    FX_HIP(hipMemcpyAsync(k.p, hkeys, 8 * n, hipMemcpyHostToDevice, st));
  } else {
    if ((rc = fin.alloc(ctx, 4 * n))) return rc;
    FX_HIP(hipMemcpyAsync(fin.p, hf, 4 * n, hipMemcpyHostToDevice, st));
    FX_HIP(hipMemsetAsync(flag.p, 0, 4, st));
    hipLaunchKernelGGL(f32_keys, dim3(fx::div_up(n, 256)), dim3(256), 0, st, fin.as<float>(),
                       k.as<uint64_t>(), n, flag.as<int>());
    hi = 32;
  }
  hipLaunchKernelGGL(iota_u32, dim3(fx::div_up(n, 256)), dim3(256), 0, st, v.as<uint32_t>(), n);
  FX_HIP(hipGetLastError());
  FX_HIP(hipEventRecord(ctx->ev0, st));
  if ((rc = fx::radix_sort_pairs(ctx, st, ws, k.as<uint64_t>(), v.as<uint32_t>(), n, 0, hi)))
    return rc;
  FX_HIP(hipEventRecord(ctx->ev1, st));
  int nanf = 0;
  if (!hkeys) FX_HIP(hipMemcpyAsync(&nanf, flag.p, 4, hipMemcpyDeviceToHost, st));
  FX_HIP(hipMemcpyAsync(perm, v.p, 4 * n, hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  FX_CHECK(!nanf, FX_E_INVALID, "fx_sort_f32: NaN key (partial_cmp has no order for NaN)");
  return FX_OK;
}

}  // namespace

extern "C" int fx_sort_u64(fx_ctx* ctx, int64_t n, const uint64_t* keys, uint32_t* perm) {
  return sort_common(ctx, n, keys, nullptr, perm);
}

extern "C" int fx_sort_f32(fx_ctx* ctx, int64_t n, const float* keys, uint32_t* perm) {
  return sort_common(ctx, n, nullptr, keys, perm);
}

// split_by_fiedler (partitioner.rs:189-213).
extern "C" int fx_split_by_fiedler(fx_ctx* ctx, int32_t n, const float* fiedler,
                                   const int32_t* gates, float balance_factor, int32_t* out,
                                   int32_t* left) {
  fx::clear_error();
  FX_CHECK(ctx && fiedler && gates && out && left && n >= 0, FX_E_INVALID,
           "fx_split_by_fiedler: bad arguments");
  if (n == 0) {
    *left = 0;
    return FX_OK;
  }
  std::vector<uint32_t> perm(n);
  int rc = fx_sort_f32(ctx, n, fiedler, perm.data());
  if (rc) return rc;
  const int min_size = (int)std::ceil((float)n * balance_factor);
  int split = n / 2;
  split = std::min(std::max(split, min_size), n - min_size);
  for (int i = 0; i < n; ++i) out[i] = gates[perm[i]];
  *left = split;
  return FX_OK;
}
