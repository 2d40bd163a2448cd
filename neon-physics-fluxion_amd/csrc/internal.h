// Internal declarations shared by the HIP sources of libfluxion_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "fluxion_hip.h"

struct fx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  char name[256] = {0};
  int num_cus = 0;
  std::atomic<int64_t> bytes{0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double leaves_ms = 0.0;
  fx_leaf_tile leaf_tile[2] = {{1, 64, 256}, {1, 64, 256}};  // per leaf mode
  double fiedler_ms = 0.0;
  double hierarchy_ms = 0.0;
  double attention_ms = 0.0;
  std::mutex mu;  // serialises calls that use the context's own stream
};

namespace fx {

// Error reporting: record a thread-local message and return the code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define FX_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t fx_e_ = (expr);                                                          \
    if (fx_e_ != hipSuccess)                                                            \
      return ::fx::fail(FX_E_DEVICE, "%s failed: %s (%s:%d)", #expr,                    \
                        hipGetErrorString(fx_e_), __FILE__, __LINE__);                  \
  } while (0)

#define FX_CHECK(cond, code, ...)                                                       \
  do {                                                                                  \
    if (!(cond)) return ::fx::fail(code, __VA_ARGS__);                                  \
  } while (0)

// Device buffer owned by a context; counts bytes against the context.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  fx_ctx* ctx = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int alloc(fx_ctx* c, size_t n);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ---- radix sort (sort.hip) ----
// Workspace for sorting n (u64 key, u32 value) pairs.
struct SortWorkspace {
  DevBuf keys_alt, vals_alt, hist;
  int64_t capacity = 0;
  int64_t blocks = 0;
  uint64_t spin_ticks = 0;  // look-back time limit in wall-clock ticks
  // Allocates for n items and clears the look-back state on `st`.
  int ensure(fx_ctx* ctx, hipStream_t st, int64_t n);
};
// Stable LSD radix sort over key bits [lo_bit, hi_bit). keys/vals are sorted
// in place (ping-pong through the workspace). Enqueued on `stream`.
int radix_sort_pairs(fx_ctx* ctx, hipStream_t stream, SortWorkspace& ws, uint64_t* keys,
                     uint32_t* vals, int64_t n, int lo_bit, int hi_bit);
// Whether a look-back wait of any sort on this workspace timed out (host sync).
int radix_fault(SortWorkspace& ws, hipStream_t stream, bool* faulted);

// ---- scans (prims.hip) ----
// out[0..n] = exclusive scan of in[0..n), total at out[n]. tmp holds
// scan_tmp_elems(n) elements of the same type.
int scan_exclusive_i32(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, int32_t* tmp,
                       int64_t tmp_elems);
// Same for pairs of int64 (16-byte elements).
int scan_exclusive_i64x2(hipStream_t st, const void* in, void* out, int64_t n, void* tmp,
                         int64_t tmp_elems);
int64_t scan_tmp_elems(int64_t n);

// Fill kernel (used instead of hipMemsetAsync inside graph-captured work).
int fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v);

// ---- deterministic reductions ----
constexpr int kReduceBlocks = 256;

}  // namespace fx
