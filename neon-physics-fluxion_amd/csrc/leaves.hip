// Batched leaf placement: the loop of hapr_placement over leaves
// (HAPR_fluxion_v4/src/algorithm/placer.rs:66-69) running place_gates_in_leaf
// (placer.rs:210-287) for every leaf in ONE kernel launch.
//
// Layout: leaves are bin-packed into 256-thread workgroups by whole 64-lane
// waves (first-fit decreasing on ceil(n/64) waves, a leaf above 256 gates
// takes a whole workgroup and each thread owns up to 4 gates). Bins are
// dispatched longest first. A workgroup keeps its leaves' positions in LDS
// for all iterations; velocities live in registers; each thread owns gates
// and evaluates, per iteration, the in-leaf adjacency springs and the exact
// repulsion against every other gate of its leaf, reading positions by LDS
// broadcast (all lanes of a wave read the same address).
//
// Exact mode reproduces the reference's FP32 arithmetic bit for bit: the
// reference accumulates, for gate i, its springs in adjacency order, then
// -f(k,i) for k<i (rows before i) and +f(i,j) for j>i; since
// f(k,i) = -f(i,k) exactly in IEEE arithmetic, the per-gate sum over j in
// ascending order of f(i,j) is the same sequence of roundings. Division and
// square root are IEEE (correctly rounded), and the build disables FMA
// contraction. Fast mode uses v_rsq_f32 and FMA (within 1e-4 of exact).

#include <algorithm>
#include <cstring>
#include <numeric>

#include "internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxLeaf = 1024;
constexpr int kMaxPerThread = kMaxLeaf / kThreads;

template <bool kFast>
__global__ __launch_bounds__(kThreads) void place_leaves_kernel(
    const int4* __restrict__ bins, const int* __restrict__ leaf_base,
    const int* __restrict__ leaf_n, const float4* __restrict__ region,
    const int* __restrict__ lptr, const int* __restrict__ lnb, float attract, float repel,
    float damping, float dt, int iters, int init, float* __restrict__ px,
    float* __restrict__ py, float* __restrict__ pvx, float* __restrict__ pvy,
    float* __restrict__ pfx, float* __restrict__ pfy) {
  __shared__ float sx[kMaxLeaf];
  __shared__ float sy[kMaxLeaf];
  const int4 bin = bins[blockIdx.x];
  const int slots[4] = {bin.x, bin.y, bin.z, bin.w};
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int leaf = slots[wave];
  int first = wave, count = 0;
  if (leaf >= 0) {
    while (first > 0 && slots[first - 1] == leaf) --first;
    for (int s = first; s < 4 && slots[s] == leaf; ++s) ++count;
  }
  const int n = leaf >= 0 ? leaf_n[leaf] : 0;
  const int base = leaf >= 0 ? leaf_base[leaf] : 0;
  const int lo = first * 64;  // LDS offset of this leaf
  const int t = (wave - first) * 64 + lane;
  const int T = count * 64;
  float4 reg = leaf >= 0 ? region[leaf] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float half_w = reg.z / 2.0f, half_h = reg.w / 2.0f;
  const float min_x = reg.x - half_w, min_y = reg.y - half_h;
  const float max_x = min_x + reg.z, max_y = min_y + reg.w;

  float xr[kMaxPerThread], yr[kMaxPerThread], vxr[kMaxPerThread], vyr[kMaxPerThread];
  float fxr[kMaxPerThread], fyr[kMaxPerThread];
#pragma unroll
  for (int k = 0; k < kMaxPerThread; ++k) {
    const int i = t + k * T;
    xr[k] = yr[k] = vxr[k] = vyr[k] = fxr[k] = fyr[k] = 0.0f;
    if (leaf >= 0 && i < n) {
      if (init) {
        xr[k] = min_x + ((float)i / (float)n) * reg.z;
        yr[k] = min_y + ((float)((i * 7) % n) / (float)n) * reg.w;
      } else {
        xr[k] = px[base + i];
        yr[k] = py[base + i];
        vxr[k] = pvx[base + i];
        vyr[k] = pvy[base + i];
      }
      sx[lo + i] = xr[k];
      sy[lo + i] = yr[k];
    }
  }
  for (int it = 0; it < iters; ++it) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMaxPerThread; ++k) {
      const int i = t + k * T;
      if (leaf < 0 || i >= n) continue;
      const float xi = xr[k], yi = yr[k];
      float ax = 0.0f, ay = 0.0f;
      // Attraction (placer.rs:242-255), adjacency order.
      const int q1 = lptr[base + i + 1];
      for (int q = lptr[base + i]; q < q1; ++q) {
        const int lj = lnb[q];
        const float dx = sx[lo + lj] - xi;
        const float dy = sy[lo + lj] - yi;
        if (kFast) {
          ax = fmaf(attract, dx, ax);
          ay = fmaf(attract, dy, ay);
        } else {
          ax += attract * dx;
          ay += attract * dy;
        }
      }
      // Repulsion (placer.rs:257-272) against every other gate, j ascending.
      for (int j = 0; j < n; ++j) {
        const float dx = xi - sx[lo + j];
        const float dy = yi - sy[lo + j];
        if (kFast) {
          const float d2 = fmaf(dx, dx, dy * dy);
          const float r = fminf(__builtin_amdgcn_rsqf(d2), 10.0f);
          const float f = repel * (r * r * r);
          ax = fmaf(f, dx, ax);
          ay = fmaf(f, dy, ay);
        } else {
          const float dist_sq = dx * dx + dy * dy;
          const float dist = fmaxf(sqrtf(dist_sq), 0.1f);
          const float force = repel / (dist * dist);
          const float fx_ij = force * dx / dist;
          const float fy_ij = force * dy / dist;
          const float nax = ax + fx_ij, nay = ay + fy_ij;
          ax = (j != i) ? nax : ax;
          ay = (j != i) ? nay : ay;
        }
      }
      fxr[k] = ax;
      fyr[k] = ay;
    }
    __syncthreads();
    // Damped Euler step with clamp (placer.rs:274-280).
#pragma unroll
    for (int k = 0; k < kMaxPerThread; ++k) {
      const int i = t + k * T;
      if (leaf < 0 || i >= n) continue;
      vxr[k] = (vxr[k] + fxr[k] * dt) * damping;
      vyr[k] = (vyr[k] + fyr[k] * dt) * damping;
      float nx = xr[k] + vxr[k] * dt;
      float ny = yr[k] + vyr[k] * dt;
      if (nx < min_x) nx = min_x;
      if (nx > max_x) nx = max_x;
      if (ny < min_y) ny = min_y;
      if (ny > max_y) ny = max_y;
      xr[k] = nx;
      yr[k] = ny;
      sx[lo + i] = nx;
      sy[lo + i] = ny;
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxPerThread; ++k) {
    const int i = t + k * T;
    if (leaf < 0 || i >= n) continue;
    px[base + i] = xr[k];
    py[base + i] = yr[k];
    pvx[base + i] = vxr[k];
    pvy[base + i] = vyr[k];
    if (pfx) {
      pfx[base + i] = fxr[k];
      pfy[base + i] = fyr[k];
    }
  }
}

struct Bin {
  int slot[4] = {-1, -1, -1, -1};
  int used = 0;
  double cost = 0.0;
};

}  // namespace

extern "C" int fx_place_leaves_state(fx_ctx* ctx, const fx_leaf_batch* b, const fx_params* p,
                                     int32_t mode, int32_t init, int32_t iterations, float* x,
                                     float* y, float* vx, float* vy, float* fx, float* fy) {
  fx::clear_error();
  FX_CHECK(ctx && b && p && x && y, FX_E_INVALID, "fx_place_leaves: NULL argument");
  FX_CHECK(mode == FX_LEAF_EXACT || mode == FX_LEAF_FAST, FX_E_INVALID, "bad mode %d", mode);
  FX_CHECK(iterations >= 0, FX_E_INVALID, "negative iteration count");
  FX_CHECK(b->num_leaves >= 0 && b->num_gates >= 0, FX_E_INVALID, "negative sizes");
  FX_CHECK(!init || true, FX_E_INVALID, "");
  const int L = b->num_leaves;
  if (L == 0) return FX_OK;
  FX_CHECK(b->leaf_ptr && b->leaf_gates && b->region && b->adj_ptr && b->adj, FX_E_INVALID,
           "fx_place_leaves: NULL batch array");
  FX_CHECK(b->leaf_ptr[0] == 0, FX_E_INVALID, "leaf_ptr[0] must be 0");
  const int P = b->leaf_ptr[L];
  // Validate leaves and build the per-leaf local adjacency (placer.rs:232-235
  // global_to_local map and the filter of 247), in adjacency order.
  std::vector<int32_t> g2l(b->num_gates, -1);
  std::vector<int> leaf_base(L), leaf_n(L);
  std::vector<int> lptr(P + 1, 0);
  std::vector<int> lnb;
  lnb.reserve((size_t)P * 8);
  for (int c = 0; c < L; ++c) {
    const int s = b->leaf_ptr[c], e = b->leaf_ptr[c + 1];
    FX_CHECK(e >= s, FX_E_INVALID, "leaf_ptr not monotone at leaf %d", c);
    FX_CHECK(e - s <= kMaxLeaf, FX_E_INVALID, "leaf %d has %d gates (max %d)", c, e - s, kMaxLeaf);
    leaf_base[c] = s;
    leaf_n[c] = e - s;
    for (int i = s; i < e; ++i) {
      const int g = b->leaf_gates[i];
      FX_CHECK(g >= 0 && g < b->num_gates, FX_E_INVALID, "gate id %d out of range", g);
      FX_CHECK(g2l[g] < 0, FX_E_INVALID, "gate %d appears in more than one leaf slot", g);
      g2l[g] = i - s;
    }
    for (int i = s; i < e; ++i) {
      const int g = b->leaf_gates[i];
      for (int k = b->adj_ptr[g]; k < b->adj_ptr[g + 1]; ++k) {
        const int nb = b->adj[k];
        if (nb >= 0 && nb < b->num_gates && g2l[nb] >= 0) lnb.push_back(g2l[nb]);
      }
      lptr[i + 1] = (int)lnb.size();
    }
    for (int i = s; i < e; ++i) g2l[b->leaf_gates[i]] = -1;
  }
  // Bin packing by waves (first-fit decreasing), then longest bins first.
  std::vector<int> order(L);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return leaf_n[a] > leaf_n[c]; });
  std::vector<Bin> bins;
  std::vector<int> open[4];  // open bins indexed by free wave slots - 1
  for (int c : order) {
    const int n = leaf_n[c];
    if (n == 0) continue;
    const int w = std::min(4, (n + 63) / 64);
    const double cost = (double)n * (double)((n + 64 * w - 1) / (64 * w));
    int bi = -1;
    for (int f = w; f <= 4 && bi < 0; ++f) {
      if (!open[f - 1].empty()) {
        bi = open[f - 1].back();
        open[f - 1].pop_back();
      }
    }
    if (bi < 0) {
      bins.emplace_back();
      bi = (int)bins.size() - 1;
    }
    Bin& bn = bins[bi];
    for (int s = 0; s < w; ++s) bn.slot[bn.used + s] = c;
    bn.used += w;
    bn.cost = std::max(bn.cost, cost);
    if (bn.used < 4) open[4 - bn.used - 1].push_back(bi);
  }
  std::stable_sort(bins.begin(), bins.end(), [](const Bin& a, const Bin& c) { return a.cost > c.cost; });
  std::vector<int4> hbins(bins.size());
  for (size_t i = 0; i < bins.size(); ++i)
    hbins[i] = make_int4(bins[i].slot[0], bins[i].slot[1], bins[i].slot[2], bins[i].slot[3]);
  std::vector<float4> hreg(L);
  for (int c = 0; c < L; ++c)
    hreg[c] = make_float4(b->region[4 * c], b->region[4 * c + 1], b->region[4 * c + 2],
                          b->region[4 * c + 3]);
  std::vector<float> hx(P, 0.f), hy(P, 0.f), hvx(P, 0.f), hvy(P, 0.f);
  if (!init)
    for (int i = 0; i < P; ++i) {
      const int g = b->leaf_gates[i];
      hx[i] = x[g];
      hy[i] = y[g];
      if (vx && vy) {
        hvx[i] = vx[g];
        hvy[i] = vy[g];
      }
    }
  if (bins.empty()) return FX_OK;

  std::lock_guard<std::mutex> lock(ctx->mu);
  FX_HIP(hipSetDevice(ctx->device));
  fx::DevBuf dbins, dbase, dn, dreg, dlptr, dlnb, dx, dy, dvx, dvy, dfx, dfy;
  int rc;
  if ((rc = dbins.alloc(ctx, sizeof(int4) * hbins.size())) || (rc = dbase.alloc(ctx, 4 * L)) ||
      (rc = dn.alloc(ctx, 4 * L)) || (rc = dreg.alloc(ctx, sizeof(float4) * L)) ||
      (rc = dlptr.alloc(ctx, 4 * (P + 1))) || (rc = dlnb.alloc(ctx, 4 * (lnb.size() + 1))) ||
      (rc = dx.alloc(ctx, 4 * P)) || (rc = dy.alloc(ctx, 4 * P)) || (rc = dvx.alloc(ctx, 4 * P)) ||
      (rc = dvy.alloc(ctx, 4 * P)))
    return rc;
  if (fx && fy) {
    if ((rc = dfx.alloc(ctx, 4 * P)) || (rc = dfy.alloc(ctx, 4 * P))) return rc;
  }
  hipStream_t st = ctx->stream;
  FX_HIP(hipMemcpyAsync(dbins.p, hbins.data(), sizeof(int4) * hbins.size(), hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dbase.p, leaf_base.data(), 4 * L, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dn.p, leaf_n.data(), 4 * L, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dreg.p, hreg.data(), sizeof(float4) * L, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dlptr.p, lptr.data(), 4 * (P + 1), hipMemcpyHostToDevice, st));
  if (!lnb.empty())
    FX_HIP(hipMemcpyAsync(dlnb.p, lnb.data(), 4 * lnb.size(), hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dx.p, hx.data(), 4 * P, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dy.p, hy.data(), 4 * P, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dvx.p, hvx.data(), 4 * P, hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(dvy.p, hvy.data(), 4 * P, hipMemcpyHostToDevice, st));
  const dim3 grid((unsigned)hbins.size()), block(kThreads);
  FX_HIP(hipEventRecord(ctx->ev0, st));
  if (mode == FX_LEAF_FAST)
    hipLaunchKernelGGL(place_leaves_kernel<true>, grid, block, 0, st, dbins.as<int4>(),
                       dbase.as<int>(), dn.as<int>(), dreg.as<float4>(), dlptr.as<int>(),
                       dlnb.as<int>(), p->attract_weight, p->repel_weight, p->damping, p->dt,
                       iterations, init ? 1 : 0, dx.as<float>(), dy.as<float>(), dvx.as<float>(),
                       dvy.as<float>(), dfx.as<float>(), dfy.as<float>());
  else
    hipLaunchKernelGGL(place_leaves_kernel<false>, grid, block, 0, st, dbins.as<int4>(),
                       dbase.as<int>(), dn.as<int>(), dreg.as<float4>(), dlptr.as<int>(),
                       dlnb.as<int>(), p->attract_weight, p->repel_weight, p->damping, p->dt,
                       iterations, init ? 1 : 0, dx.as<float>(), dy.as<float>(), dvx.as<float>(),
                       dvy.as<float>(), dfx.as<float>(), dfy.as<float>());
  FX_HIP(hipGetLastError());
  FX_HIP(hipEventRecord(ctx->ev1, st));
  FX_HIP(hipMemcpyAsync(hx.data(), dx.p, 4 * P, hipMemcpyDeviceToHost, st));
  FX_HIP(hipMemcpyAsync(hy.data(), dy.p, 4 * P, hipMemcpyDeviceToHost, st));
  FX_HIP(hipMemcpyAsync(hvx.data(), dvx.p, 4 * P, hipMemcpyDeviceToHost, st));
  FX_HIP(hipMemcpyAsync(hvy.data(), dvy.p, 4 * P, hipMemcpyDeviceToHost, st));
  std::vector<float> hfx, hfy;
  if (fx && fy) {
    hfx.resize(P);
    hfy.resize(P);
    FX_HIP(hipMemcpyAsync(hfx.data(), dfx.p, 4 * P, hipMemcpyDeviceToHost, st));
    FX_HIP(hipMemcpyAsync(hfy.data(), dfy.p, 4 * P, hipMemcpyDeviceToHost, st));
  }
  FX_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  FX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->leaves_ms = ms;
  for (int i = 0; i < P; ++i) {
    const int g = b->leaf_gates[i];
    x[g] = hx[i];
    y[g] = hy[i];
    if (vx && vy) {
      vx[g] = hvx[i];
      vy[g] = hvy[i];
    }
    if (fx && fy) {
      fx[g] = hfx[i];
      fy[g] = hfy[i];
    }
  }
  return FX_OK;
}

extern "C" int fx_place_leaves(fx_ctx* ctx, const fx_leaf_batch* batch, const fx_params* params,
                               int32_t mode, float* x, float* y) {
  if (!params) return fx::fail(FX_E_INVALID, "fx_place_leaves: NULL params");
  return fx_place_leaves_state(ctx, batch, params, mode, 1, params->force_iterations, x, y,
                               nullptr, nullptr, nullptr, nullptr);
}

extern "C" double fx_leaves_last_kernel_ms(const fx_ctx* ctx) { return ctx ? ctx->leaves_ms : 0.0; }
