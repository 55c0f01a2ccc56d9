// field.hip - the renderer's input producer on gfx950: conics, integer support radii and opacities
// from a field's log-scales, rotations and opacity logits (the reference's GaussianField.conics /
// radii / opacity_values, gaussians.py:270-321), and the backward of conics and opacities to the
// log-scales, rotations and logits.  One launch each way instead of a chain of elementwise ops.
//
// Forward, per Gaussian (theta = rotation, sx2 = exp(l0)^2 + fv, sy2 = exp(l1)^2 + fv):
//   u = 1 / (sx2 + d), v = 1 / (sy2 + d), c = cos theta, s = sin theta
//   conic = (c^2 u + s^2 v,  c s (u - v),  s^2 u + c^2 v)
//   radius_x = max(ceil(sigma * sqrt(c^2 (sx2 + d) + s^2 (sy2 + d))), 1), radius_y likewise with
//   c and s exchanged; opacity = sigmoid(logit).
// The radii are evaluated with the reference's float32 operation order and no contraction, so
// they (and the tile lists built from them) are identical to the torch expression's.
//
// Backward, with gradients (ga, gb, gc) of the conic and go of the opacity:
//   gu = ga c^2 + gb c s + gc s^2,  gv = ga s^2 - gb c s + gc c^2
//   g_l0 = -2 gu u^2 exp(l0)^2,  g_l1 = -2 gv v^2 exp(l1)^2
//   g_theta = 2 c s (v - u)(ga - gc) + gb (c^2 - s^2)(u - v)
//   g_logit = go o (1 - o)

#include <cmath>
#include <climits>
#include <cstdint>

#include "common.h"

namespace {

using namespace ssplat_detail;

constexpr int kFieldThreads = 256;

struct ProduceArgs {
  const float* log_scales;       // (n, 2)
  const float* rotations;        // (n,)
  const float* logits;           // (n,) or NULL
  const float* filter_variance;  // (n,) or NULL
  int64_t n;
  float dilation;
  float sigma;
  float* conics;      // (n, 3) or NULL
  int64_t* radii;     // (n, 2) or NULL
  float* opacities;   // (n,) or NULL
};

__device__ __forceinline__ int64_t to_radius(float r) {
  if (!(r >= 1.0f)) return 1;
  if (r >= 9.2e18f) return INT64_MAX;
  return static_cast<int64_t>(r);
}

__global__ __launch_bounds__(kFieldThreads) void k_produce(ProduceArgs a) {
#pragma clang fp contract(off)
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float e0 = expf(a.log_scales[2 * i + 0]);
  const float e1 = expf(a.log_scales[2 * i + 1]);
  float sx2 = e0 * e0;
  float sy2 = e1 * e1;
  if (a.filter_variance) {
    const float fv = a.filter_variance[i];
    sx2 = sx2 + fv;
    sy2 = sy2 + fv;
  }
  const float rot = a.rotations[i];
  const float c = cosf(rot);
  const float s = sinf(rot);
  if (a.conics) {
    const float u = 1.0f / (sx2 + a.dilation);
    const float v = 1.0f / (sy2 + a.dilation);
    a.conics[3 * i + 0] = c * c * u + s * s * v;
    a.conics[3 * i + 1] = c * s * (u - v);
    a.conics[3 * i + 2] = s * s * u + c * c * v;
  }
  if (a.radii) {
    const float px = sx2 + a.dilation;
    const float py = sy2 + a.dilation;
    const float var_x = c * c * px + s * s * py;
    const float var_y = s * s * px + c * c * py;
    const float rx = ceilf(a.sigma * sqrtf(var_x));
    const float ry = ceilf(a.sigma * sqrtf(var_y));
    // float -> int64 like torch's .long() of an integral value, then clamped to >= 1; NaN gives 1
    // and values beyond int64 saturate (no undefined conversion).
    a.radii[2 * i + 0] = to_radius(rx);
    a.radii[2 * i + 1] = to_radius(ry);
  }
  if (a.opacities && a.logits) a.opacities[i] = 1.0f / (1.0f + expf(-a.logits[i]));
}

struct ProduceGradArgs {
  const float* log_scales;
  const float* rotations;
  const float* opacities;        // sigmoid values (n,) or NULL
  const float* filter_variance;  // (n,) or NULL
  int64_t n;
  float dilation;
  const float* g_conics;      // (n, 3)
  const float* g_opacities;   // (n,) or NULL
  float* g_log_scales;        // (n, 2)
  float* g_rotations;         // (n,)
  float* g_logits;            // (n,) or NULL
};

__global__ __launch_bounds__(kFieldThreads) void k_produce_grad(ProduceGradArgs a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float e0 = expf(a.log_scales[2 * i + 0]);
  const float e1 = expf(a.log_scales[2 * i + 1]);
  const float q0 = e0 * e0, q1 = e1 * e1;  // d(sx2)/dl0 = 2 q0
  float sx2 = q0, sy2 = q1;
  if (a.filter_variance) {
    sx2 += a.filter_variance[i];
    sy2 += a.filter_variance[i];
  }
  const float u = 1.0f / (sx2 + a.dilation);
  const float v = 1.0f / (sy2 + a.dilation);
  const float rot = a.rotations[i];
  const float c = cosf(rot);
  const float s = sinf(rot);
  const float ga = a.g_conics[3 * i + 0];
  const float gb = a.g_conics[3 * i + 1];
  const float gc = a.g_conics[3 * i + 2];
  const float cc = c * c, ss = s * s, cs = c * s;
  const float gu = ga * cc + gb * cs + gc * ss;
  const float gv = ga * ss - gb * cs + gc * cc;
  a.g_log_scales[2 * i + 0] = -2.0f * gu * u * u * q0;
  a.g_log_scales[2 * i + 1] = -2.0f * gv * v * v * q1;
  a.g_rotations[i] = 2.0f * cs * (v - u) * (ga - gc) + gb * (cc - ss) * (u - v);
  if (a.g_logits) {
    const float o = a.opacities[i];
    a.g_logits[i] = a.g_opacities ? a.g_opacities[i] * o * (1.0f - o) : 0.0f;
  }
}

int blocks_for(int64_t n) { return static_cast<int>((n + kFieldThreads - 1) / kFieldThreads); }

}  // namespace

extern "C" {

int ssplat_produce_field(const float* log_scales, const float* rotations,
                         const float* opacity_logits, const float* filter_variance, int64_t n,
                         double dilation, double sigma_cutoff, float* conics, int64_t* radii,
                         float* opacities, void* stream) {
  if (n < 0 || n > INT32_MAX) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "n out of range");
  if (!(dilation >= 0.0)) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "dilation must be >= 0");
  if (n > 0 && (log_scales == nullptr || rotations == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "log_scales and rotations must not be NULL");
  }
  if (opacities != nullptr && opacity_logits == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "opacities need opacity_logits");
  }
  if (n == 0) return success();
  ProduceArgs a{log_scales, rotations, opacity_logits, filter_variance, n,
                static_cast<float>(dilation), static_cast<float>(sigma_cutoff), conics, radii,
                opacities};
  k_produce<<<blocks_for(n), kFieldThreads, 0, static_cast<hipStream_t>(stream)>>>(a);
  SSPLAT_LAUNCH_CHECK("k_produce");
  return success();
}

int ssplat_produce_field_backward(const float* log_scales, const float* rotations,
                                  const float* opacities, const float* filter_variance,
                                  int64_t n, double dilation, const float* grad_conics,
                                  const float* grad_opacities, float* grad_log_scales,
                                  float* grad_rotations, float* grad_logits, void* stream) {
  if (n < 0 || n > INT32_MAX) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "n out of range");
  if (!(dilation >= 0.0)) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "dilation must be >= 0");
  if (n > 0 && (log_scales == nullptr || rotations == nullptr || grad_conics == nullptr ||
                grad_log_scales == nullptr || grad_rotations == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "field and gradient tensors must not be NULL");
  }
  if (grad_logits != nullptr && opacities == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "grad_logits needs the opacities");
  }
  if (n == 0) return success();
  ProduceGradArgs a{log_scales, rotations, opacities, filter_variance, n,
                    static_cast<float>(dilation), grad_conics, grad_opacities, grad_log_scales,
                    grad_rotations, grad_logits};
  k_produce_grad<<<blocks_for(n), kFieldThreads, 0, static_cast<hipStream_t>(stream)>>>(a);
  SSPLAT_LAUNCH_CHECK("k_produce_grad");
  return success();
}

}  // extern "C"
