// optim.hip - the fit step's optimizer on gfx950: Adam over every parameter group of a field in one
// launch, with the log-scale clamp fused into the update.
//
// The update is torch.optim.Adam's (the reference's optimizer, fit.py:672-691, with its default
// betas (0.9, 0.999) and eps 1e-8), with the step count kept on the device so the launch can be
// captured in a HIP graph:
//   t += 1
//   m = m + (1 - b1) (g - m)          v = b2 v + (1 - b2) g^2
//   p = p - (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps)
// followed, for a clamped group, by p = min(max(p, lo), hi) with hi one value or one per row
// (fit.py:3033, :3359-3375: log-scales clamped to [ln 0.35, ln max(H, W)] of their image).

#include <cmath>
#include <cstdint>

#include "common.h"

namespace {

using namespace ssplat_detail;

constexpr int kMaxGroups = 8;
constexpr int kAdamThreads = 256;
constexpr int kMaxGroupBlocks = 4096;
constexpr int kAdamPer = 4;  // elements per thread and step

struct AdamGroupArgs {
  float* param;
  const float* grad;
  float* m;
  float* v;
  int64_t count;
  const float* hi_rows;  // per-row upper bound (clamp with rows) or NULL
  int64_t row_width;
  float lr;
  float lo, hi;
  int clamp;
};

struct AdamArgs {
  AdamGroupArgs g[kMaxGroups];
  int block_start[kMaxGroups + 1];  // each group's first workgroup (prefix of the block counts)
  int num_groups;
  const float* step;
  float beta1, beta2, eps;
};

__global__ void k_adam_tick(float* step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step += 1.0f;
}

__global__ __launch_bounds__(kAdamThreads) void k_adam(AdamArgs a) {
  // The bias corrections, once per workgroup (powf and sqrtf cost more than the update).
  __shared__ float s_bc[2];
  if (threadIdx.x == 0) {
    const float t = *a.step;
    s_bc[0] = 1.0f - powf(a.beta1, t);
    s_bc[1] = sqrtf(1.0f - powf(a.beta2, t));
  }
  __syncthreads();
  const float bc1 = s_bc[0];
  const float bc2_sqrt = s_bc[1];
  // Every workgroup serves one group (a uniform choice: the group's arguments stay scalar).
  int k = 0;
  while (k + 1 < a.num_groups && static_cast<int>(blockIdx.x) >= a.block_start[k + 1]) ++k;
  const AdamGroupArgs& g = a.g[k];
  float* __restrict__ param = g.param;
  const float* __restrict__ grad = g.grad;
  float* __restrict__ mom = g.m;
  float* __restrict__ var = g.v;
  const float step_size = g.lr / bc1;
  // kAdamPer elements per thread, blockDim apart: every load of a step is issued before the
  // first update.
  const int64_t span = int64_t(kAdamThreads) * kAdamPer;
  const int64_t stride = int64_t(a.block_start[k + 1] - a.block_start[k]) * span;
  for (int64_t j0 = int64_t(blockIdx.x - a.block_start[k]) * span + threadIdx.x; j0 < g.count;
       j0 += stride) {
    float gr[kAdamPer], m[kAdamPer], v[kAdamPer], p[kAdamPer], hi[kAdamPer];
#pragma unroll
    for (int u = 0; u < kAdamPer; ++u) {
      const int64_t j = j0 + int64_t(u) * kAdamThreads;
      const bool ok = j < g.count;
      gr[u] = ok ? grad[j] : 0.0f;
      m[u] = ok ? mom[j] : 0.0f;
      v[u] = ok ? var[j] : 0.0f;
      p[u] = ok ? param[j] : 0.0f;
      // Rows of at most 2^31 elements and tables below 2^32 (checked on the host): a 32-bit
      // division instead of a 64-bit one.
      hi[u] = (ok && g.clamp && g.hi_rows != nullptr)
                  ? g.hi_rows[static_cast<uint32_t>(j) / static_cast<uint32_t>(g.row_width)]
                  : g.hi;
    }
#pragma unroll
    for (int u = 0; u < kAdamPer; ++u) {
      const int64_t j = j0 + int64_t(u) * kAdamThreads;
      if (j >= g.count) continue;
      const float mu = fmaf(1.0f - a.beta1, gr[u] - m[u], m[u]);
      const float vu = fmaf(1.0f - a.beta2, gr[u] * gr[u], a.beta2 * v[u]);
      float pu = p[u] - step_size * mu / (sqrtf(vu) / bc2_sqrt + a.eps);
      if (g.clamp) pu = fminf(fmaxf(pu, g.lo), hi[u]);
      mom[j] = mu;
      var[j] = vu;
      param[j] = pu;
    }
  }
}

}  // namespace

extern "C" {

int ssplat_adam_step(const ssplat_adam_group* groups, int num_groups, float* step, double beta1,
                     double beta2, double eps, void* stream) {
  if (groups == nullptr || num_groups <= 0 || num_groups > kMaxGroups) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "num_groups must be in [1, %d]", kMaxGroups);
  }
  if (step == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "step is NULL");
  if (!(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0 && eps >= 0.0)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "betas must be in [0, 1) and eps >= 0");
  }
  AdamArgs a{};
  a.num_groups = num_groups;
  a.step = step;
  a.beta1 = static_cast<float>(beta1);
  a.beta2 = static_cast<float>(beta2);
  a.eps = static_cast<float>(eps);
  int blocks = 0;
  for (int k = 0; k < num_groups; ++k) {
    const ssplat_adam_group& s = groups[k];
    if (s.count < 0) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "group %d: count < 0", k);
    if (s.count > 0 && (s.param == nullptr || s.grad == nullptr || s.exp_avg == nullptr ||
                        s.exp_avg_sq == nullptr)) {
      return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "group %d: NULL tensor", k);
    }
    if (s.clamp && s.hi_rows != nullptr && s.row_width <= 0) {
      return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "group %d: row_width must be > 0", k);
    }
    if (s.count > (int64_t(1) << 32) - 1 || s.row_width > (int64_t(1) << 31)) {
      return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "group %d: more than 2^32 - 1 elements", k);
    }
    AdamGroupArgs& g = a.g[k];
    g.param = s.param;
    g.grad = s.grad;
    g.m = s.exp_avg;
    g.v = s.exp_avg_sq;
    g.count = s.count;
    g.lr = static_cast<float>(s.lr);
    g.clamp = s.clamp != 0;
    g.lo = static_cast<float>(s.lo);
    g.hi = static_cast<float>(s.hi);
    g.hi_rows = s.hi_rows;
    g.row_width = s.row_width > 0 ? s.row_width : 1;
    // One workgroup per kAdamThreads * kAdamPer elements, at most kMaxGroupBlocks per group
    // (grid-stride beyond).
    a.block_start[k] = blocks;
    const int64_t need = (s.count + kAdamThreads * kAdamPer - 1) / (kAdamThreads * kAdamPer);
    blocks += static_cast<int>(need < kMaxGroupBlocks ? need : kMaxGroupBlocks);
  }
  a.block_start[num_groups] = blocks;
  hipStream_t st = static_cast<hipStream_t>(stream);
  k_adam_tick<<<1, 64, 0, st>>>(step);
  SSPLAT_LAUNCH_CHECK("k_adam_tick");
  if (blocks > 0) {
    k_adam<<<blocks, kAdamThreads, 0, st>>>(a);
    SSPLAT_LAUNCH_CHECK("k_adam");
  }
  return success();
}

}  // extern "C"
