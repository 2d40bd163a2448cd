// DCT Poisson density force (neon_old_codebase/src/python/fluxion/
// force_electrostatic.py:47-157) on MFMA.
//
//   rho[r][c]  = count(r,c) / (bin_w*bin_h)           point-in-bin (47-60)
//   a          = rho - target                          (122-123)
//   rho_hat    = C a C^T                                dctn type 2 ortho (69)
//   phi_hat    = -rho_hat / lambda, phi_hat[0][0] = 0   (72-86)
//   phi        = C^T phi_hat C                          idctn (89)
// with C[k][m] = s_k cos(pi k (2m+1) / 2G), s_0 = sqrt(1/G), s_k = sqrt(2/G),
// and lambda = (wx cos(pi kx/G) - wx)^2 + (wy cos(pi ky/G) - wy)^2, lambda_00 = 1
// (the reference's eigenvalue expression, kept as written). The field
// E = -grad(phi) by centred differences and the per-gate gather are fused
// into the force-combine kernel.
//
// Each product is a GxGxG GEMM on v_mfma_f32_32x32x8_bf16 with FP32 operands
// split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi)); the product is
// hi*hi + hi*lo + lo*hi (+ lo*lo with 4 passes) accumulated in FP32, which
// recovers close to FP32 accuracy. One 64-lane wave owns a 32x32 output
// tile. The bin counts are deposited with int32 atomics (exact, order-free)
// and converted to a = count*w - target while loading the first GEMM's
// operand; the eigenvalue division is the second GEMM's epilogue.

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

__device__ __forceinline__ void split4(const float v[4], short4_t& hi, short4_t& lo) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint16_t h = bf16_rne(v[k]);
    hi[k] = (short)h;
    lo[k] = (short)bf16_rne(v[k] - bf16_to_f(h));
  }
}

enum OpKind { kConst = 0, kData = 1, kCounts = 2 };
enum Epilogue { kStore = 0, kEigen = 1 };

struct GemmArgs {
  int G;
  // left operand A[M][K]
  const uint16_t* a_hi;  // kConst: bf16 row-major [row][k]
  const uint16_t* a_lo;
  const float* a_f;      // kData: fp32 row-major [row][k]
  // right operand B[K][N]
  const uint16_t* bt_hi; // kConst: stored transposed, bt[col][k]
  const uint16_t* bt_lo;
  const float* b_f;      // kData: fp32 row-major [k][col]
  const int* b_cnt;      // kCounts: int32 row-major [k][col]
  float cnt_w, target;
  const float* eig;      // kEigen: lambda [row][col]
  float* out;
  int passes;
};

template <int AK, int BK, int EPI>
__global__ __launch_bounds__(64) void dct_gemm(GemmArgs g) {
  const int G = g.G;
  const int tiles_n = G / 32;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  const int lane = threadIdx.x;
  const int r = lane & 31;         // row of A fragment / col of B fragment
  const int kq = (lane >> 5) * 4;  // k offset inside the 8-wide step
  const int arow = tm * 32 + r;
  const int bcol = tn * 32 + r;
  float16_t acc = {};
  for (int k0 = 0; k0 < G; k0 += 8) {
    short4_t ahi, alo, bhi, blo;
    if (AK == kConst) {
      const uint2 h = *reinterpret_cast<const uint2*>(g.a_hi + (size_t)arow * G + k0 + kq);
      const uint2 l = *reinterpret_cast<const uint2*>(g.a_lo + (size_t)arow * G + k0 + kq);
      ahi = __builtin_bit_cast(short4_t, h);
      alo = __builtin_bit_cast(short4_t, l);
    } else {
      const float4 v4 = *reinterpret_cast<const float4*>(g.a_f + (size_t)arow * G + k0 + kq);
      const float v[4] = {v4.x, v4.y, v4.z, v4.w};
      split4(v, ahi, alo);
    }
    if (BK == kConst) {
      const uint2 h = *reinterpret_cast<const uint2*>(g.bt_hi + (size_t)bcol * G + k0 + kq);
      const uint2 l = *reinterpret_cast<const uint2*>(g.bt_lo + (size_t)bcol * G + k0 + kq);
      bhi = __builtin_bit_cast(short4_t, h);
      blo = __builtin_bit_cast(short4_t, l);
    } else {
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const size_t idx = (size_t)(k0 + kq + j) * G + bcol;
        if (BK == kCounts) v[j] = (float)g.b_cnt[idx] * g.cnt_w - g.target;
        else v[j] = g.b_f[idx];
      }
      split4(v, bhi, blo);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ahi, bhi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ahi, blo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(alo, bhi, acc, 0, 0, 0);
    if (g.passes == 4) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(alo, blo, acc, 0, 0, 0);
  }
  // D layout: register e -> row (e/4)*8 + (lane/32)*4 + e%4, col lane%32.
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int row = tm * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
    const int col = tn * 32 + (lane & 31);
    const size_t o = (size_t)row * G + col;
    float v = acc[e];
    if (EPI == kEigen) v = (row == 0 && col == 0) ? 0.0f : -v / g.eig[o];
    g.out[o] = v;
  }
}

__global__ void deposit_kernel(int N, const float* __restrict__ x, const float* __restrict__ y,
                               double bin_w, double bin_h, int G, int* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double qc = fmin(fmax((double)x[i] / bin_w, 0.0), (double)(G - 1));
  const double qr = fmin(fmax((double)y[i] / bin_h, 0.0), (double)(G - 1));
  atomicAdd(&counts[(int)qr * G + (int)qc], 1);
}

__global__ void fields_kernel(int G, const int* __restrict__ counts, float cnt_w,
                              const float* __restrict__ phi, float wx, float wy,
                              float* __restrict__ rho, float* __restrict__ ex,
                              float* __restrict__ ey) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G * G) return;
  const int r = i / G, c = i % G;
  if (rho) rho[i] = (float)counts[i] * cnt_w;
  if (ex) {
    if (c == 0) ex[i] = -(phi[i + 1] - phi[i]) * wx;
    else if (c == G - 1) ex[i] = -(phi[i] - phi[i - 1]) * wx;
    else ex[i] = -(phi[i + 1] - phi[i - 1]) * wx * 0.5f;
  }
  if (ey) {
    if (r == 0) ey[i] = -(phi[i + G] - phi[i]) * wy;
    else if (r == G - 1) ey[i] = -(phi[i] - phi[i - G]) * wy;
    else ey[i] = -(phi[i + G] - phi[i - G]) * wy * 0.5f;
  }
}

}  // namespace

int dct_alloc(Job& j) {
  DctState& d = j.dct;
  const int G = j.p.density_grid;
  d.G = G;
  const size_t GG = (size_t)G * G;
  int rc;
  fx_ctx* c = j.ctx;
  if ((rc = d.counts.alloc(c, 4 * GG)) || (rc = d.c_hi.alloc(c, 2 * GG)) ||
      (rc = d.c_lo.alloc(c, 2 * GG)) || (rc = d.ct_hi.alloc(c, 2 * GG)) ||
      (rc = d.ct_lo.alloc(c, 2 * GG)) || (rc = d.t0.alloc(c, 4 * GG)) ||
      (rc = d.t1.alloc(c, 4 * GG)) || (rc = d.phi.alloc(c, 4 * GG)) ||
      (rc = d.inv_eig.alloc(c, 4 * GG)))
    return rc;
  std::vector<uint16_t> ch(GG), cl(GG), th(GG), tl(GG);
  std::vector<float> eig(GG);
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < G; ++k) {
    const double s = k == 0 ? std::sqrt(1.0 / G) : std::sqrt(2.0 / G);
    for (int m = 0; m < G; ++m) {
      const float v = (float)(s * std::cos(pi * k * (2 * m + 1) / (2.0 * G)));
      const uint16_t h = bf16_rne(v);
      const uint16_t l = bf16_rne(v - bf16_to_f(h));
      ch[(size_t)k * G + m] = h;
      cl[(size_t)k * G + m] = l;
      th[(size_t)m * G + k] = h;
      tl[(size_t)m * G + k] = l;
    }
  }
  const double bin_w = (double)j.die_w / G, bin_h = (double)j.die_h / G;
  const double wx = 1.0 / bin_w, wy = 1.0 / bin_h;
  for (int ky = 0; ky < G; ++ky)
    for (int kx = 0; kx < G; ++kx) {
      const double ex = wx * std::cos(pi * kx / G) - wx;
      const double ey = wy * std::cos(pi * ky / G) - wy;
      double ev = ex * ex + ey * ey;
      if (kx == 0 && ky == 0) ev = 1.0;
      eig[(size_t)ky * G + kx] = (float)ev;
    }
  FX_HIP(hipMemcpyAsync(d.c_hi.p, ch.data(), 2 * GG, hipMemcpyHostToDevice, j.stream));
  FX_HIP(hipMemcpyAsync(d.c_lo.p, cl.data(), 2 * GG, hipMemcpyHostToDevice, j.stream));
  FX_HIP(hipMemcpyAsync(d.ct_hi.p, th.data(), 2 * GG, hipMemcpyHostToDevice, j.stream));
  FX_HIP(hipMemcpyAsync(d.ct_lo.p, tl.data(), 2 * GG, hipMemcpyHostToDevice, j.stream));
  FX_HIP(hipMemcpyAsync(d.inv_eig.p, eig.data(), 4 * GG, hipMemcpyHostToDevice, j.stream));
  FX_HIP(hipStreamSynchronize(j.stream));
  return FX_OK;
}

int dct_field(Job& j) {
  DctState& d = j.dct;
  const int G = d.G;
  hipStream_t st = j.stream;
  int rc;
  if ((rc = fill_i32(st, d.counts.as<int32_t>(), (int64_t)G * G, 0))) return rc;
  const double bin_w = (double)j.die_w / G, bin_h = (double)j.die_h / G;
  hipLaunchKernelGGL(deposit_kernel, dim3(div_up(j.N, 256)), dim3(256), 0, st, j.N,
                     j.x.as<float>(), j.y.as<float>(), bin_w, bin_h, G, d.counts.as<int>());
  FX_HIP(hipGetLastError());
  GemmArgs a{};
  a.G = G;
  a.passes = j.p.dct_passes;
  a.cnt_w = (float)(1.0 / (bin_w * bin_h));
  a.target = j.p.target_density;
  a.eig = d.inv_eig.as<float>();
  const dim3 grid((unsigned)((G / 32) * (G / 32))), block(64);
  // T1 = C * a
  GemmArgs g1 = a;
  g1.a_hi = d.c_hi.as<uint16_t>();
  g1.a_lo = d.c_lo.as<uint16_t>();
  g1.b_cnt = d.counts.as<int>();
  g1.out = d.t0.as<float>();
  hipLaunchKernelGGL((dct_gemm<kConst, kCounts, kStore>), grid, block, 0, st, g1);
  // phi_hat = -(T1 * C^T) / lambda   (B = C^T is stored transposed: C)
  GemmArgs g2 = a;
  g2.a_f = d.t0.as<float>();
  g2.bt_hi = d.c_hi.as<uint16_t>();
  g2.bt_lo = d.c_lo.as<uint16_t>();
  g2.out = d.t1.as<float>();
  hipLaunchKernelGGL((dct_gemm<kData, kConst, kEigen>), grid, block, 0, st, g2);
  // T2 = C^T * phi_hat
  GemmArgs g3 = a;
  g3.a_hi = d.ct_hi.as<uint16_t>();
  g3.a_lo = d.ct_lo.as<uint16_t>();
  g3.b_f = d.t1.as<float>();
  g3.out = d.t0.as<float>();
  hipLaunchKernelGGL((dct_gemm<kConst, kData, kStore>), grid, block, 0, st, g3);
  // phi = T2 * C   (B = C is stored transposed: C^T)
  GemmArgs g4 = a;
  g4.a_f = d.t0.as<float>();
  g4.bt_hi = d.ct_hi.as<uint16_t>();
  g4.bt_lo = d.ct_lo.as<uint16_t>();
  g4.out = d.phi.as<float>();
  hipLaunchKernelGGL((dct_gemm<kData, kConst, kStore>), grid, block, 0, st, g4);
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
  if (rc) return rc;
  const int G = j.dct.G;
  const size_t GG = (size_t)G * G;
  fx::DevBuf drho, dex, dey;
  if ((rc = drho.alloc(j.ctx, 4 * GG)) || (rc = dex.alloc(j.ctx, 4 * GG)) || (rc = dey.alloc(j.ctx, 4 * GG)))
    return rc;
  const double bin_w = (double)j.die_w / G, bin_h = (double)j.die_h / G;
  hipLaunchKernelGGL(fx::fields_kernel, dim3(fx::div_up(GG, 256)), dim3(256), 0, j.stream, G,
                     j.dct.counts.as<int>(), (float)(1.0 / (bin_w * bin_h)), j.dct.phi.as<float>(),
                     (float)(1.0 / bin_w), (float)(1.0 / bin_h), drho.as<float>(), dex.as<float>(),
                     dey.as<float>());
  FX_HIP(hipGetLastError());
  if (rho) FX_HIP(hipMemcpyAsync(rho, drho.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  if (phi) FX_HIP(hipMemcpyAsync(phi, j.dct.phi.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  if (ex) FX_HIP(hipMemcpyAsync(ex, dex.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  if (ey) FX_HIP(hipMemcpyAsync(ey, dey.p, 4 * GG, hipMemcpyDeviceToHost, j.stream));
  FX_HIP(hipStreamSynchronize(j.stream));
  return FX_OK;
}
