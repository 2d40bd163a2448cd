// DCT Poisson density force (neon_old_codebase/src/python/fluxion/
// force_electrostatic.py:47-157) on MFMA.
//
//   rho[r][c]  = count(r,c) / (bin_w*bin_h)           point-in-bin (47-60)
//   a          = rho - target                          (122-123)
//   rho_hat    = C a C^T                                dctn type 2 ortho (69)
//   phi_hat    = -rho_hat / lambda, phi_hat[0][0] = 0   (72-86)
//   phi        = C^T phi_hat C                          idctn (89)
//   E          = centred differences of phi, one-sided at the edges (91-103)
// with C[k][m] = s_k cos(pi k (2m+1) / 2G), s_0 = sqrt(1/G), s_k = sqrt(2/G),
// and lambda = (wx cos(pi kx/G) - wx)^2 + (wy cos(pi ky/G) - wy)^2, lambda_00 = 1
// (the reference's eigenvalue expression, kept as written). The difference
// operators are linear, so they are folded into the inverse transform:
//   Ex = (C^T phi_hat) (C Dx),   Ey = (Dy C^T) (phi_hat C),
// where Dx, Dy are the reference's difference stencils as matrices; C Dx and
// Dy C^T are formed once in FP64. This is the same E as differencing phi,
// but each product is accurate relative to E itself: the reference's
// eigenvalue grows as k^4, phi is dominated by the lowest modes, and its
// differences are ~G/pi times smaller than phi, so differencing an FP32 phi
// loses two digits at G=512. The per-gate gather of E is fused into the
// force-combine kernel.
//
// Each product is a GxGxG GEMM on v_mfma_f32_32x32x8_bf16 with FP32 operands
// split into bf16 parts, accumulated in FP32:
//   forward (C a C^T): x = hi + lo, products hi*hi + hi*lo + lo*hi (3 passes;
//     4 with lo*lo), or with dct_passes = 6 the three-part split below;
//   inverse (C^T phi_hat C): x = hi + mid + lo, the six products whose
//     orders sum to at most 2 (hi*hi, hi*mid, mid*hi, hi*lo, mid*mid, lo*hi).
// The forward input is taken about the mean bin density, a = (count - N/G^2)
// * w: this differs from rho - target by a constant, which only changes the
// DC coefficient, and the DC coefficient is discarded (phi_hat[0][0] = 0), so
// phi is unchanged; removing the mean keeps the split precision relative to
// the density fluctuations rather than to the background.
// A workgroup owns a 32x32 output tile, each of its waves an equal share of
// the K range (summed through LDS in fixed order); operands for the next K
// step are loaded while the current step's MFMAs issue (eight waves for
// G >= 256). The two independent
// products of each inverse stage (U and V, then Ex and Ey) share one launch,
// so an iteration issues four GEMM launches of 256-512 workgroups. Bin counts are
// deposited with int32 atomics (exact, order-free).

#include <cmath>
#include <vector>

#include "job.h"

namespace fx {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t bf16_rne(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t u = __float_as_uint(f);
#else
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
#endif
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf16_to_f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(u);
#else
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
#endif
}

template <int SPLIT>
__device__ __forceinline__ void split4(const float v[4], short4_t& p0, short4_t& p1,
                                       short4_t& p2) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint16_t h = bf16_rne(v[k]);
    const float r = v[k] - bf16_to_f(h);
    const uint16_t m = bf16_rne(r);
    p0[k] = (short)h;
    p1[k] = (short)m;
    if (SPLIT == 3) p2[k] = (short)bf16_rne(r - bf16_to_f(m));
  }
}

enum OpKind { kConst = 0, kData = 1, kCounts = 2 };
enum Epilogue { kStore = 0, kEigen = 1 };

struct GemmArgs {
  int G;
  // left operand A[M][K]
  const uint16_t* a_p[3];  // kConst: bf16 parts, row-major [row][k]
  const float* a_f;        // kData: fp32 row-major [row][k]
  // right operand B[K][N]
  const uint16_t* bt_p[3];  // kConst: bf16 parts stored transposed, bt[col][k]
  const float* b_f;         // kData: fp32 row-major [k][col]
  const int* b_cnt;         // kCounts: int32 row-major [k][col]
  float cnt_w, mean;
  const float* eig;  // kEigen: lambda [row][col]
  float* out;
  int passes;  // SPLIT == 2: 3 or 4
};

template <int KIND, int SPLIT, bool LEFT>
__device__ __forceinline__ void load_operand(const GemmArgs& g, int idx_row, int k,
                                             short4_t& p0, short4_t& p1, short4_t& p2) {
  const int G = g.G;
  if (KIND == kConst) {
    const uint16_t* const* parts = LEFT ? g.a_p : g.bt_p;
    const size_t o = (size_t)idx_row * G + k;
    p0 = __builtin_bit_cast(short4_t, *reinterpret_cast<const uint2*>(parts[0] + o));
    p1 = __builtin_bit_cast(short4_t, *reinterpret_cast<const uint2*>(parts[1] + o));
    if (SPLIT == 3) p2 = __builtin_bit_cast(short4_t, *reinterpret_cast<const uint2*>(parts[2] + o));
  } else if (LEFT) {
    const float4 v4 = *reinterpret_cast<const float4*>(g.a_f + (size_t)idx_row * G + k);
    const float v[4] = {v4.x, v4.y, v4.z, v4.w};
    split4<SPLIT>(v, p0, p1, p2);
  } else {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const size_t idx = (size_t)(k + j) * G + idx_row;
      if (KIND == kCounts) v[j] = (float)g.b_cnt[idx] * g.cnt_w - g.mean;
      else v[j] = g.b_f[idx];
    }
    split4<SPLIT>(v, p0, p1, p2);
  }
}

#define MFMA(a, b) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0)

// One workgroup (NW waves: 8 for G >= 256, else 4) per 32x32 output tile;
// wave w reduces the K range [w*G/NW, (w+1)*G/NW) and the partial tiles are
// summed through LDS in a fixed order (deterministic).
inline int gemm_waves(int G) { return G >= 256 ? 8 : 4; }

// One 32x32 output tile of one product (tile index `tile`).
template <int AK, int BK, int EPI, int SPLIT, int NW>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, int tile,
                                          float (*part)[16][64]) {
  const int G = g.G;
  const int tiles_n = G / 32;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31;         // row of the A fragment / column of the B fragment
  const int kq = (lane >> 5) * 4;  // k offset inside the 8-wide step
  const int arow = tm * 32 + r;
  const int bcol = tn * 32 + r;
  const int kspan = G / NW;
  const int kbeg = wave * kspan, kend = kbeg + kspan;
  float16_t acc = {};
  short4_t a0, a1, a2, b0, b1, b2;
  load_operand<AK, SPLIT, true>(g, arow, kbeg + kq, a0, a1, a2);
  load_operand<BK, SPLIT, false>(g, bcol, kbeg + kq, b0, b1, b2);
  for (int k0 = kbeg; k0 < kend; k0 += 8) {
    short4_t na0, na1, na2, nb0, nb1, nb2;
    const int kn = k0 + 8 < kend ? k0 + 8 : k0;  // prefetch the next step (last: reload)
    load_operand<AK, SPLIT, true>(g, arow, kn + kq, na0, na1, na2);
    load_operand<BK, SPLIT, false>(g, bcol, kn + kq, nb0, nb1, nb2);
    if (SPLIT == 3) {
      MFMA(a2, b0);
      MFMA(a1, b1);
      MFMA(a0, b2);
      MFMA(a1, b0);
      MFMA(a0, b1);
      MFMA(a0, b0);
    } else {
      if (g.passes == 4) MFMA(a1, b1);
      MFMA(a1, b0);
      MFMA(a0, b1);
      MFMA(a0, b0);
    }
    a0 = na0; a1 = na1; a2 = na2;
    b0 = nb0; b1 = nb1; b2 = nb2;
  }
  if (wave > 0) {
#pragma unroll
    for (int e = 0; e < 16; ++e) part[wave - 1][e][lane] = acc[e];
  }
  __syncthreads();
  if (wave == 0) {
  // D layout: register e -> row (e/4)*8 + (lane/32)*4 + e%4, col lane%32.
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    float v = acc[e];
#pragma unroll
    for (int w = 0; w < NW - 1; ++w) v += part[w][e][lane];
    const int row = tm * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
    const int col = tn * 32 + (lane & 31);
    const size_t o = (size_t)row * G + col;
    if (EPI == kEigen) v = (row == 0 && col == 0) ? 0.0f : -v / g.eig[o];
    g.out[o] = v;
  }
  }
}

template <int AK, int BK, int EPI, int SPLIT, int NW>
__global__ __launch_bounds__(64 * NW) void dct_gemm(GemmArgs g) {
  __shared__ float part[NW - 1][16][64];
  const int tiles = (g.G / 32) * (g.G / 32);
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    gemm_tile<AK, BK, EPI, SPLIT, NW>(g, t, part);
    __syncthreads();
  }
}

// Two independent products in one launch: the first G^2/1024 workgroups
// compute product 0, the rest product 1.
template <int AK0, int BK0, int AK1, int BK1, int NW>
__global__ __launch_bounds__(64 * NW) void dct_gemm_pair(GemmArgs g0, GemmArgs g1) {
  __shared__ float part[NW - 1][16][64];
  const int tiles = (g0.G / 32) * (g0.G / 32);
  for (int t = blockIdx.x; t < 2 * tiles; t += gridDim.x) {
    if (t < tiles) gemm_tile<AK0, BK0, kStore, 3, NW>(g0, t, part);
    else gemm_tile<AK1, BK1, kStore, 3, NW>(g1, t - tiles, part);
    __syncthreads();
  }
}
#undef MFMA

__global__ void deposit_kernel(int N, const float* __restrict__ x, const float* __restrict__ y,
                               double bin_w, double bin_h, int G, int* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double qc = fmin(fmax((double)x[i] / bin_w, 0.0), (double)(G - 1));
  const double qr = fmin(fmax((double)y[i] / bin_h, 0.0), (double)(G - 1));
  atomicAdd(&counts[(int)qr * G + (int)qc], 1);
}

// The same deposit over particles in Morton order (the Barnes-Hut sort's
// permutation): runs of equal bins inside a wave add their count with one
// atomic, which removes the contention of dense clusters.
__global__ void deposit_sorted_kernel(int N, const uint32_t* __restrict__ order,
                                      const float* __restrict__ x, const float* __restrict__ y,
                                      double bin_w, double bin_h, int G, int* __restrict__ counts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = t < N;
  int bin = -1;
  if (valid) {
    const int i = (int)order[t];
    const double qc = fmin(fmax((double)x[i] / bin_w, 0.0), (double)(G - 1));
    const double qr = fmin(fmax((double)y[i] / bin_h, 0.0), (double)(G - 1));
    bin = (int)qr * G + (int)qc;
  }
  const int prev = __shfl_up(bin, 1, 64);
  const bool head = valid && (lane == 0 || prev != bin);
  const unsigned long long heads = __ballot(head);
  const unsigned long long vmask = __ballot(valid);
  if (!head) return;
  const unsigned long long above = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
  const int last_valid = 63 - __clzll((long long)vmask);
  const int next = above ? __ffsll((long long)above) - 1 : last_valid + 1;
  atomicAdd(&counts[bin], next - lane);
}

__global__ void rho_kernel(int G, const int* __restrict__ counts, float cnt_w,
                           float* __restrict__ rho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < G * G) rho[i] = (float)counts[i] * cnt_w;
}

// Splits a row-major double matrix into three bf16 parts.
void split_host(const std::vector<double>& m, std::vector<uint16_t> out[3]) {
  for (int k = 0; k < 3; ++k) out[k].resize(m.size());
  for (size_t i = 0; i < m.size(); ++i) {
    const float v = (float)m[i];
    const uint16_t h = bf16_rne(v);
    const float r = v - bf16_to_f(h);
    const uint16_t md = bf16_rne(r);
    out[0][i] = h;
    out[1][i] = md;
    out[2][i] = bf16_rne(r - bf16_to_f(md));
  }
}

int upload3(hipStream_t st, const std::vector<double>& m, DevBuf* dst[3]) {
  std::vector<uint16_t> parts[3];
  split_host(m, parts);
  for (int k = 0; k < 3; ++k)
    FX_HIP(hipMemcpyAsync(dst[k]->p, parts[k].data(), 2 * m.size(), hipMemcpyHostToDevice, st));
  FX_HIP(hipStreamSynchronize(st));
  return FX_OK;
}

// spread: output tiles per workgroup (the grid is that many times smaller).
template <int AK, int BK, int EPI, int SPLIT>
void launch(const GemmArgs& g, hipStream_t st, int spread = 1) {
  const dim3 grid((unsigned)((g.G / 32) * (g.G / 32) / spread));
  if (gemm_waves(g.G) == 8)
    hipLaunchKernelGGL((dct_gemm<AK, BK, EPI, SPLIT, 8>), grid, dim3(512), 0, st, g);
  else
    hipLaunchKernelGGL((dct_gemm<AK, BK, EPI, SPLIT, 4>), grid, dim3(256), 0, st, g);
}

template <int AK0, int BK0, int AK1, int BK1>
void launch_pair(const GemmArgs& g0, const GemmArgs& g1, hipStream_t st, int spread = 1) {
  const dim3 grid((unsigned)(2 * (g0.G / 32) * (g0.G / 32) / spread));
  if (gemm_waves(g0.G) == 8)
    hipLaunchKernelGGL((dct_gemm_pair<AK0, BK0, AK1, BK1, 8>), grid, dim3(512), 0, st, g0, g1);
  else
    hipLaunchKernelGGL((dct_gemm_pair<AK0, BK0, AK1, BK1, 4>), grid, dim3(256), 0, st, g0, g1);
}

}  // namespace

int dct_alloc(Job& j) {
  DctState& d = j.dct;
  const int G = j.p.density_grid;
  d.G = G;
  const size_t GG = (size_t)G * G;
  int rc;
  fx_ctx* c = j.ctx;
  DevBuf* bufs16[] = {&d.c_hi, &d.c_mid, &d.c_lo, &d.ct_hi, &d.ct_mid, &d.ct_lo,
                      &d.dxt_hi, &d.dxt_mid, &d.dxt_lo, &d.dy_hi, &d.dy_mid, &d.dy_lo};
  for (DevBuf* b : bufs16)
    if ((rc = b->alloc(c, 2 * GG))) return rc;
  DevBuf* bufs32[] = {&d.counts, &d.t0, &d.t1, &d.t2, &d.phi, &d.ex, &d.ey, &d.eig};
  for (DevBuf* b : bufs32)
    if ((rc = b->alloc(c, 4 * GG))) return rc;
  const double pi = 3.14159265358979323846;
  std::vector<double> C(GG), Ct(GG), Dxt(GG), Dy(GG);
  for (int k = 0; k < G; ++k) {
    const double s = k == 0 ? std::sqrt(1.0 / G) : std::sqrt(2.0 / G);
    for (int m = 0; m < G; ++m) {
      // round through f32 so that the split matches the FP32 basis values
      const double v = (double)(float)(s * std::cos(pi * k * (2 * m + 1) / (2.0 * G)));
      C[(size_t)k * G + m] = v;
      Ct[(size_t)m * G + k] = v;
    }
  }
  const double bin_w = (double)j.die_w / G, bin_h = (double)j.die_h / G;
  const double wx = 1.0 / bin_w, wy = 1.0 / bin_h;
  // Centred difference of basis row k at column c (force_electrostatic.py:96-103).
  auto diff = [&](int k, int c, double w) {
    const double* row = &C[(size_t)k * G];
    if (c == 0) return -(row[1] - row[0]) * w;
    if (c == G - 1) return -(row[G - 1] - row[G - 2]) * w;
    return -(row[c + 1] - row[c - 1]) * w * 0.5;
  };
  for (int k = 0; k < G; ++k)
    for (int c2 = 0; c2 < G; ++c2) {
      Dxt[(size_t)c2 * G + k] = diff(k, c2, wx);  // [col][k]: transposed right operand
      Dy[(size_t)c2 * G + k] = diff(k, c2, wy);   // [row][k]: left operand
    }
  std::vector<float> eig(GG);
  for (int ky = 0; ky < G; ++ky)
    for (int kx = 0; kx < G; ++kx) {
      const double ex = wx * std::cos(pi * kx / G) - wx;
      const double ey = wy * std::cos(pi * ky / G) - wy;
      double ev = ex * ex + ey * ey;
      if (kx == 0 && ky == 0) ev = 1.0;
      eig[(size_t)ky * G + kx] = (float)ev;
    }
  DevBuf* pc[3] = {&d.c_hi, &d.c_mid, &d.c_lo};
  DevBuf* pct[3] = {&d.ct_hi, &d.ct_mid, &d.ct_lo};
  DevBuf* pdx[3] = {&d.dxt_hi, &d.dxt_mid, &d.dxt_lo};
  DevBuf* pdy[3] = {&d.dy_hi, &d.dy_mid, &d.dy_lo};
  if ((rc = upload3(j.stream, C, pc)) || (rc = upload3(j.stream, Ct, pct)) ||
      (rc = upload3(j.stream, Dxt, pdx)) || (rc = upload3(j.stream, Dy, pdy)))
    return rc;
  FX_HIP(hipMemcpyAsync(d.eig.p, eig.data(), 4 * GG, hipMemcpyHostToDevice, j.stream));
  FX_HIP(hipStreamSynchronize(j.stream));
  return FX_OK;
}

static void parts3(const DevBuf& a, const DevBuf& b, const DevBuf& c, const uint16_t* out[3]) {
  out[0] = a.as<uint16_t>();
  out[1] = b.as<uint16_t>();
  out[2] = c.as<uint16_t>();
}

// rho -> phi_hat (t1) and E (ex, ey); U = C^T phi_hat is left in t0.
int dct_field(Job& j, bool counts_cleared, hipStream_t stream) {
  DctState& d = j.dct;
  const int G = d.G;
  hipStream_t st = stream ? stream : j.stream;
  int rc;
  if (!counts_cleared && (rc = fill_i32(st, d.counts.as<int32_t>(), (int64_t)G * G, 0)))
    return rc;
  const double bin_w = (double)j.die_w / G, bin_h = (double)j.die_h / G;
  if (j.bh.sorted_valid)
    hipLaunchKernelGGL(deposit_sorted_kernel, dim3(div_up(j.N, 256)), dim3(256), 0, st, j.N,
                       j.bh.vals.as<uint32_t>(), j.x.as<float>(), j.y.as<float>(), bin_w, bin_h,
                       G, d.counts.as<int>());
  else
    hipLaunchKernelGGL(deposit_kernel, dim3(div_up(j.N, 256)), dim3(256), 0, st, j.N,
                       j.x.as<float>(), j.y.as<float>(), bin_w, bin_h, G, d.counts.as<int>());
  FX_HIP(hipGetLastError());
  GemmArgs a{};
  a.G = G;
  a.passes = j.p.dct_passes;
  const double w = 1.0 / (bin_w * bin_h);
  a.cnt_w = (float)w;
  a.mean = (float)((double)j.N / ((double)G * G) * (double)a.cnt_w);
  a.eig = d.eig.as<float>();
  const uint16_t *C[3], *Ct[3], *Dxt[3], *Dy[3];
  parts3(d.c_hi, d.c_mid, d.c_lo, C);
  parts3(d.ct_hi, d.ct_mid, d.ct_lo, Ct);
  parts3(d.dxt_hi, d.dxt_mid, d.dxt_lo, Dxt);
  parts3(d.dy_hi, d.dy_mid, d.dy_lo, Dy);
  const bool six = a.passes == 6;
  // On a second stream the products run beside the Barnes-Hut tree build and
  // walk: two output tiles per workgroup leave more of the GPU to those
  // (measured at G = 512: 2.7 % more iterations per second than one tile,
  // 4 % fewer with four).
  const int spread = (st != j.stream && G >= 64) ? 2 : 1;
  // T1 = C a
  GemmArgs g1 = a;
  for (int k = 0; k < 3; ++k) g1.a_p[k] = C[k];
  g1.b_cnt = d.counts.as<int>();
  g1.out = d.t2.as<float>();
  if (six) launch<kConst, kCounts, kStore, 3>(g1, st, spread);
  else launch<kConst, kCounts, kStore, 2>(g1, st, spread);
  // phi_hat = -(T1 C^T) / lambda, DC = 0   (right operand C^T stored transposed: C)
  GemmArgs g2 = a;
  g2.a_f = d.t2.as<float>();
  for (int k = 0; k < 3; ++k) g2.bt_p[k] = C[k];
  g2.out = d.t1.as<float>();
  if (six) launch<kData, kConst, kEigen, 3>(g2, st, spread);
  else launch<kData, kConst, kEigen, 2>(g2, st, spread);
  // U = C^T phi_hat and V = phi_hat C (right operand C stored transposed:
  // C^T), one launch
  GemmArgs g3 = a;
  for (int k = 0; k < 3; ++k) g3.a_p[k] = Ct[k];
  g3.b_f = d.t1.as<float>();
  g3.out = d.t0.as<float>();
  GemmArgs g5 = a;
  g5.a_f = d.t1.as<float>();
  for (int k = 0; k < 3; ++k) g5.bt_p[k] = Ct[k];
  g5.out = d.t2.as<float>();
  launch_pair<kConst, kData, kData, kConst>(g3, g5, st, spread);
  // Ex = U (C Dx) (centred x-differences of phi = U C) and
  // Ey = (Dy C^T) V (centred y-differences of phi = C^T V), one launch
  GemmArgs g4 = a;
  g4.a_f = d.t0.as<float>();
  for (int k = 0; k < 3; ++k) g4.bt_p[k] = Dxt[k];
  g4.out = d.ex.as<float>();
  GemmArgs g6 = a;
  for (int k = 0; k < 3; ++k) g6.a_p[k] = Dy[k];
  g6.b_f = d.t2.as<float>();
  g6.out = d.ey.as<float>();
  launch_pair<kData, kConst, kConst, kData>(g4, g6, st, spread);
  FX_HIP(hipGetLastError());
  return FX_OK;
}

// phi = U C (diagnostics only; the force uses ex, ey).
static int dct_phi(Job& j) {
  DctState& d = j.dct;
  GemmArgs g{};
  g.G = d.G;
  g.passes = 6;
  g.a_f = d.t0.as<float>();
  parts3(d.ct_hi, d.ct_mid, d.ct_lo, g.bt_p);
  g.out = d.phi.as<float>();
  launch<kData, kConst, kStore, 3>(g, j.stream);
  FX_HIP(hipGetLastError());
  return FX_OK;
}

}  // namespace fx

extern "C" int fx_job_density_fields(fx_job* job, float* rho, float* phi, float* ex, float* ey) {
  fx::clear_error();
  FX_CHECK(job, FX_E_INVALID, "NULL job");
  fx::Job& j = job->j;
  FX_CHECK(j.dct.G > 0, FX_E_STATE, "job has no density grid");
  int rc = fx::dct_field(j);
  if (!rc) rc = fx::dct_phi(j);
  if (rc) return rc;
  const int G = j.dct.G;
  const size_t GG = (size_t)G * G;
  fx::DevBuf drho;
  if ((rc = drho.alloc(j.ctx, 4 * GG))) return rc;
  const double bin_w = (double)j.die_w / G, bin_h = (double)j.die_h / G;
  hipLaunchKernelGGL(fx::rho_kernel, dim3(fx::div_up(GG, 256)), dim3(256), 0, j.stream, G,
                     j.dct.counts.as<int>(), (float)(1.0 / (bin_w * bin_h)), drho.as<float>());
  FX_HIP(hipGetLastError());
  if (rho) FX_HIP(hipMemcpyAsync(rho, drho.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  if (phi) FX_HIP(hipMemcpyAsync(phi, j.dct.phi.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  if (ex) FX_HIP(hipMemcpyAsync(ex, j.dct.ex.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  if (ey) FX_HIP(hipMemcpyAsync(ey, j.dct.ey.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  FX_HIP(hipStreamSynchronize(j.stream));
  return FX_OK;
}
