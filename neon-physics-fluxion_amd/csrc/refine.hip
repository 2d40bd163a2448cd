// Congestion-aware refinement: refine_placement
// (HAPR_fluxion_v4/src/algorithm/refiner.rs:34-95) on a device-resident job.
//
// Each iteration: the routing grid of the current positions (the bit-exact
// int32 count grid; h = v = count), its total overflow summed in FP32 in
// row-major order as RoutingGrid::total_overflow does (grid.rs:85-91; the
// host sums the downloaded grid, so the convergence and stall decisions are
// the reference's), then every gate in a hot cell (overflow > threshold,
// gates_in_hot_cells, congestion.rs:77-93) moves toward the least congested
// cell within 3 cells (move_gate_from_hot_cell, refiner.rs:98-125). The hot
// set is taken before any move and each move reads only its own gate and the
// grid, so all moves of an iteration are independent: one thread per gate,
// FP32 arithmetic as in the reference, results identical to its sequential
// loop.

#include <cfloat>
#include <cmath>
#include <vector>

#include "job.h"

namespace fx {
namespace {

constexpr int kBlock = 256;

// (q as usize).min(n-1) of RoutingGrid::cell_at (grid.rs:58-62).
__device__ __forceinline__ int grid_cell(float q, int n) {
  if (!(q > 0.0f)) return 0;
  if (q >= (float)n) return n - 1;
  return (int)q;
}

// f32::clamp (asserts lo <= hi; NaN passes through).
__device__ __forceinline__ float rust_clamp(float v, float lo, float hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}

__global__ __launch_bounds__(kBlock) void refine_move_kernel(
    int n, float* __restrict__ x, float* __restrict__ y, const int32_t* __restrict__ counts,
    int cols, int rows, float cw, float ch, float cap, float threshold, float max_move,
    float chip_w, float chip_h, unsigned long long* __restrict__ moved) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  bool did = false;
  if (g < n) {
    const float gx = x[g], gy = y[g];
    const int col = grid_cell(gx / cw, cols), row = grid_cell(gy / ch, rows);
    const float h = (float)counts[row * cols + col];
    // overflow_at: (h + v - capacity).max(0.0) > threshold  (grid.rs:74-78, 95-105)
    if (fmaxf((h + h) - cap, 0.0f) > threshold) {
      // neighbors(col, row, 3) in order, first minimum of h + v (refiner.rs:103-115)
      int bc = -1, br = -1;
      float best = 0.0f;
      for (int dr = -3; dr <= 3; ++dr)
        for (int dc = -3; dc <= 3; ++dc) {
          if (dr == 0 && dc == 0) continue;
          const int nc = col + dc, nr = row + dr;
          if (nc < 0 || nc >= cols || nr < 0 || nr >= rows) continue;
          const float c = (float)counts[nr * cols + nc];
          const float t = c + c;
          if (bc < 0 || t < best) {
            best = t;
            bc = nc;
            br = nr;
          }
        }
      if (bc >= 0) {
        // cell_center (grid.rs:129-133), then the clamped step (refiner.rs:117-124)
        const float tx = ((float)bc + 0.5f) * cw;
        const float ty = ((float)br + 0.5f) * ch;
        const float dx = rust_clamp(tx - gx, -max_move, max_move);
        const float dy = rust_clamp(ty - gy, -max_move, max_move);
        x[g] = rust_clamp(gx + dx, 0.0f, chip_w);
        y[g] = rust_clamp(gy + dy, 0.0f, chip_h);
        did = true;
      }
    }
  }
  const unsigned long long b = __ballot(did);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(moved, (unsigned long long)__popcll(b));
}

// RoutingGrid::total_overflow / max_overflow (grid.rs:80-91) in FP32, row-major.
void overflow_f32(const std::vector<int32_t>& counts, float cap, float* total, float* maxo) {
  float t = 0.0f, m = 0.0f;
  for (int32_t c : counts) {
    const float h = (float)c;
    const float o = std::fmax((h + h) - cap, 0.0f);
    t += o;
    m = std::fmax(m, o);
  }
  *total = t;
  *maxo = m;
}

// The grid of the job's positions on the device (*d) and on the host (h).
int grid_to_host(Job& j, int cols, int rows, std::vector<int32_t>& h, int32_t** d) {
  int rc = congestion_counts(j, cols, rows, d);
  if (rc) return rc;
  h.resize((size_t)cols * rows);
  FX_HIP(hipMemcpyAsync(h.data(), *d, 4 * h.size(), hipMemcpyDeviceToHost, j.stream));
  FX_HIP(hipStreamSynchronize(j.stream));
  return FX_OK;
}

}  // namespace
}  // namespace fx

extern "C" void fx_refine_defaults(fx_refine_config* c) {
  if (!c) return;
  // RefineConfig::default (refiner.rs:21-31)
  c->max_iterations = 100;
  c->move_distance = 5.0f;
  c->overflow_threshold = 0.0f;
  c->grid_cols = 256;
  c->grid_rows = 256;
}

extern "C" int fx_job_refine(fx_job* job, const fx_refine_config* cfg, fx_refine_result* out) {
  fx::clear_error();
  FX_CHECK(job && cfg && out, FX_E_INVALID, "fx_job_refine: NULL argument");
  FX_CHECK(cfg->grid_cols > 0 && cfg->grid_rows > 0 && cfg->grid_cols <= 8192 &&
               cfg->grid_rows <= 8192 && cfg->max_iterations >= 0,
           FX_E_INVALID, "fx_job_refine: bad configuration");
  fx::Job& j = job->j;
  const int cols = cfg->grid_cols, rows = cfg->grid_rows;
  const float cw = j.die_w / (float)cols, ch = j.die_h / (float)rows;  // grid.rs:27-28
  const float cap = (cw + ch) * 0.5f;                                   // grid.rs:31
  fx::DevBuf moved;
  int rc = moved.alloc(j.ctx, 8);
  if (rc) return rc;
  FX_HIP(hipMemsetAsync(moved.p, 0, 8, j.stream));
  std::vector<int32_t> h;
  float prev = FLT_MAX;
  int stall = 0, iter = 0, stop = FX_REFINE_MAX_ITERATIONS;
  float total = 0.0f, maxo = 0.0f;
  int32_t* d = nullptr;
  for (iter = 0; iter < cfg->max_iterations; ++iter) {
    if ((rc = fx::grid_to_host(j, cols, rows, h, &d))) return rc;
    fx::overflow_f32(h, cap, &total, &maxo);
    if (total <= cfg->overflow_threshold) {
      stop = FX_REFINE_CONVERGED;
      break;
    }
    if (std::fabs(prev - total) < 0.01f * std::fmax(prev, 1.0f)) {
      if (++stall >= 10) {
        stop = FX_REFINE_STALLED;
        break;
      }
    } else {
      stall = 0;
    }
    prev = total;
    hipLaunchKernelGGL(fx::refine_move_kernel, dim3(fx::div_up(j.N, fx::kBlock)),
                       dim3(fx::kBlock), 0, j.stream, j.N, j.x.as<float>(), j.y.as<float>(), d,
                       cols, rows, cw, ch, cap, cfg->overflow_threshold, cfg->move_distance,
                       j.die_w, j.die_h, moved.as<unsigned long long>());
    FX_HIP(hipGetLastError());
  }
  // Final grid (refiner.rs:84-86).
  if ((rc = fx::grid_to_host(j, cols, rows, h, &d))) return rc;
  fx::overflow_f32(h, cap, &total, &maxo);
  unsigned long long mv = 0;
  FX_HIP(hipMemcpyAsync(&mv, moved.p, 8, hipMemcpyDeviceToHost, j.stream));
  FX_HIP(hipStreamSynchronize(j.stream));
  double hp = 0.0;
  if ((rc = fx_hpwl(job, &hp))) return rc;
  out->iterations = iter;
  out->stop = stop;
  out->total_overflow = total;
  out->max_overflow = maxo;
  out->hpwl = hp;
  out->gate_moves = (int64_t)mv;
  return FX_OK;
}
