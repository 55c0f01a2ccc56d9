// loss.hip - the fit step's image loss on gfx950: (1 - w) * L1 + w * (1 - SSIM) and its gradient
// with respect to the rendered image, for one image or a batch of images (fields) of any sizes.
//
// Definitions (the reference's fit loss, fit.py:800-823 with pixel_loss "l1", and its SSIM,
// metrics.py:88-151): per image of n = 3 H W values,
//   L1   = sum |p - t| / n
//   SSIM = sum S / n,  S = (2 mu_p mu_t + C1)(2 s_pt + C2) / ((mu_p^2 + mu_t^2 + C1)(s_p + s_t + C2))
// with mu, s from an 11-tap Gaussian window (sigma 1.5) applied per channel as two zero-padded
// 1-D passes, C1 = 0.01^2, C2 = 0.03^2.  A batch's loss is the sum of its images' losses.
//
// Gradient.  With A1 = 2 mu_p mu_t + C1, A2 = 2 s_pt + C2, B1 = mu_p^2 + mu_t^2 + C1,
// B2 = s_p + s_t + C2 and the window moments mu_p = G*p, m2 = G*p^2, mpt = G*(p t):
//   dS/dmu_p = 2 mu_t (A2 - A1) / (B1 B2) - 2 mu_p S (1/B1 - 1/B2)
//   dS/dm2   = -S / B2
//   dS/dmpt  = 2 A1 / (B1 B2)
// and, the window being symmetric, dSSIM/dp = (G*a + 2 p G*b + t G*c) / n for the three maps
// a, b, c above (zero outside the image).
//
// Kernels, over units of 16 columns by 32 rows (ssplat_loss_plan numbers the units of every
// field):
//   k_loss_stats : one 256-thread workgroup per unit: stages p and t with a 5-pixel halo in LDS,
//                  blurs p, p^2, p t (and t, t^2 without precomputed target moments)
//                  horizontally, then vertically, writes a, b, c and the unit's partial loss.
//   k_loss_grad  : two 256-thread workgroups per unit, one per 16 x 16 half: the same blurs of
//                  a, b, c; adds the SSIM gradient to the L1 gradient.
//   k_loss_l1    : the loss and gradient without SSIM.
//   k_loss_total : one workgroup sums the unit partials in a fixed order (deterministic loss).

#include <cmath>
#include <cstdint>

#include "common.h"

namespace {

using namespace ssplat_detail;

// Geometry.  A unit of work is 16 columns by 32 rows of one image (clipped by the image).  The
// stats pass blurs the unit's 42 staged rows horizontally (42 / 32 rows per output row; a
// 16 x 16 tile would blur 26 / 16) and its 32 output rows vertically.  The gradient pass works on
// the unit's two 16 x 16 halves in two workgroups (three staged maps of a 16 x 32 unit would not
// leave room in LDS for more than three workgroups per CU).
constexpr int kUnitW = 16;
constexpr int kUnitH = 32;
constexpr int kHalo = 5;                     // window radius
constexpr int kWin = 2 * kHalo + 1;          // 11 taps
constexpr int kSpanX = kUnitW + 2 * kHalo;   // 26 staged columns
constexpr int kLoadRow = 3 * kSpanX;         // 78 floats per staged row (3 channels)
constexpr int kLossThreads = 256;
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;

// Stats pass (the whole unit).
constexpr int kSpanY = kUnitH + 2 * kHalo;   // 42 staged rows
constexpr int kStageStride = 28;             // floats per staged row: 26 used, 16-byte rows
constexpr int kBlurStride = 18;              // floats per blurred row: 16 used, bank spread
constexpr int kHSeg = 8;                     // horizontal pass: outputs per task
constexpr int kVSeg = 8;                     // vertical pass: outputs per task
constexpr int kHTasks = 3 * kSpanY * (kUnitW / kHSeg);  // 252 (channel, row, segment)
constexpr int kVTasks = 3 * kUnitW * (kUnitH / kVSeg);  // 192 (channel, column, segment)
static_assert(kHTasks <= kLossThreads && kVTasks <= kLossThreads, "one task per thread");

// Gradient pass (one 16 x 16 half at a time).
constexpr int kGT = 16;                      // half edge
constexpr int kGSpan = kGT + 2 * kHalo;      // 26 staged rows / columns
constexpr int kGHSeg = 8;
constexpr int kGVSeg = 4;
constexpr int kGHTasks = 3 * kGSpan * (kGT / kGHSeg);  // 156
constexpr int kGVTasks = 3 * kGT * (kGT / kGVSeg);     // 192
constexpr int kGStageStride = 28;            // floats per staged row: 26 used, 16-byte rows
constexpr int kGBlurStride = 20;             // floats per blurred row: 16 used, 16-byte rows
static_assert(kGHTasks <= kLossThreads && kGVTasks <= kLossThreads, "one task per thread");

// What the stats pass blurs.  The target's window moments mu_t = G*t and G*t^2 do not change
// while a fit runs (the reference caches them in SSIMTargetStats, metrics.py:93-128), so a fit
// computes them once (kStatsTarget) and its steps read them
// (kStatsMoments), blurring three maps of the image instead of five.  The blurs are the same
// arithmetic in every mode, so the moments read equal the ones computed inline bit for bit.
constexpr int kStatsFull = 0;     // p, t, p^2, t^2, p t
constexpr int kStatsMoments = 1;  // p, p^2, p t; mu_t and G*t^2 read from tmom
constexpr int kStatsTarget = 2;   // t, t^2 only: writes mu_t and G*t^2 to tmom_out
template <int kMode>
constexpr int kQuantities = kMode == kStatsFull ? 5 : (kMode == kStatsMoments ? 3 : 2);

struct Window {
  float w[kWin];
};

struct LossArgs {
  FieldTable fields;    // units: first_tile = first unit, tiles_x = units per row, num_tiles
  const float* img;     // (pixels, 3)
  const float* target;  // (pixels, 3)
  float* grad;          // (pixels, 3): the gradient (backward)
  const float* scale;   // device scalar dL/dloss multiplying the gradient, or NULL (1)
  bool keep;            // forward: write a, b, c for a later backward
  float* abc;           // 3 maps of (pixels, 3): dS/dmu_p, dS/dm2, dS/dmpt
  const float* tmom;    // the target's moments, 2 maps of (pixels, 3): G*t, G*t^2 (or NULL)
  float* tmom_out;      // kStatsTarget: where they are written
  float* partial;       // one float per unit
  int64_t pixels;       // pixels over all fields
  float w_l1;           // 1 - ssim_weight
  float w_ssim;         // ssim_weight
  Window win;
};

struct UnitGeo {
  int x0, y0, W, H;
  int64_t first_pixel;
  float inv_n;  // 1 / (3 H W)
};

__device__ __forceinline__ UnitGeo unit_geo(const LossArgs& a, int unit) {
  const ssplat_field f = field_of_tile(a.fields, unit);
  const int local = unit - static_cast<int>(f.first_tile);
  UnitGeo g;
  g.x0 = (local % f.tiles_x) * kUnitW;
  g.y0 = (local / f.tiles_x) * kUnitH;
  g.W = f.width;
  g.H = f.height;
  g.first_pixel = f.first_pixel;
  g.inv_n = 1.0f / (3.0f * static_cast<float>(f.height) * static_cast<float>(f.width));
  return g;
}

// A float at a byte offset from a base: with a uniform base and a 32-bit offset the access
// takes the scalar-base form of the memory instructions (no 64-bit address arithmetic per lane).
__device__ __forceinline__ float load_at(const float* base, uint32_t byte_off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void store_at(float* base, uint32_t byte_off, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// Sum over the workgroup (returned to thread 0).  Every thread calls it.
__device__ float block_sum(float v) {
  __shared__ float s_red[kLossThreads / 64];
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  __syncthreads();  // earlier readers of s_red are done
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kLossThreads / 64; ++w) t += s_red[w];
  }
  return t;
}

// Staging of rows y0 - 5 .. y0 + kRows + 4, columns x0 - 5 .. x0 + 20 of kMaps 3-channel maps
// into LDS as [map][channel][row][column] (row stride kStride), zero outside the image.  Each
// staged row is 78 contiguous floats; thread t < 234 owns float t % 78 of rows t / 78,
// t / 78 + 3, ...  All loads are issued before the LDS stores (one memory latency per staging);
// a position outside the image loads element 0 of the map and stores 0.  With kL1, returns
// sum |map0 - map1| over the staged floats of rows y0 .. y0 + kRows - 1 and columns x0 .. x0 + 15
// inside the image.
template <int kMaps, int kRows, int kStride, bool kL1>
__device__ __forceinline__ float stage_halo(const float* const (&src)[kMaps], int x0, int y0,
                                            const UnitGeo& g, float* lds) {
  constexpr int kSpan = kRows + 2 * kHalo;
  constexpr int kPerPass = kLossThreads / kLoadRow;                 // 3 rows per pass
  constexpr int kPasses = (kSpan + kPerPass - 1) / kPerPass;
  constexpr int kMapFloats = 3 * kSpan * kStride;
  const int t = threadIdx.x;
  const bool active = t < kPerPass * kLoadRow;
  const int r0 = t / kLoadRow;
  const int c3 = t - r0 * kLoadRow;
  const int col = c3 / 3;
  const int ch = c3 - col * 3;
  const int gx = x0 - kHalo + col;
  const bool x_ok = active && gx >= 0 && gx < g.W;
  const bool x_center = col >= kHalo && col < kHalo + kUnitW;
  // Addresses: the staged window's first element (uniform; it may lie outside the image) plus a
  // 32-bit byte offset per lane, which is non-negative for every staged position.  A position outside
  // the image loads the unit's own first element (offset `safe`, inside the image) and stores 0.
  const int64_t origin = (g.first_pixel + int64_t(y0 - kHalo) * g.W + (x0 - kHalo)) * 3;
  const uint32_t row_bytes = static_cast<uint32_t>(g.W) * 12u;
  const uint32_t safe = kHalo * row_bytes + kHalo * 12u;
  uint32_t off = static_cast<uint32_t>(r0) * row_bytes + static_cast<uint32_t>(c3) * 4u;
  const int lds0 = (ch * kSpan + r0) * kStride + col;
  float v[kMaps][kPasses];
  bool ok[kPasses];
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    const int row = r0 + kPerPass * k;
    const int gy = y0 - kHalo + row;
    ok[k] = x_ok && row < kSpan && gy >= 0 && gy < g.H;
    const uint32_t at = ok[k] ? off : safe;
#pragma unroll
    for (int m = 0; m < kMaps; ++m) {
      const float x = load_at(src[m] + origin, at);
      v[m][k] = ok[k] ? x : 0.0f;
    }
    off += kPerPass * row_bytes;
  }
  float l1 = 0.0f;
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    const int row = r0 + kPerPass * k;
    if (!active || row >= kSpan) continue;
    const int off = lds0 + kPerPass * k * kStride;
#pragma unroll
    for (int m = 0; m < kMaps; ++m) lds[m * kMapFloats + off] = v[m][k];
    if (kL1 && ok[k] && x_center && row >= kHalo && row < kHalo + kRows) {
      l1 += fabsf(v[0][k] - v[1][k]);
    }
  }
  return l1;
}

// The window sums at one output: mu_p, mu_t, G*p^2, G*t^2, G*(p t).
struct Moments {
  float mu_p, mu_t, m_pp, m_tt, m_pt;
};

// SSIM statistics of every unit: L1 and SSIM partial sums, and (keep) the maps a, b, c of the
// gradient; or, for kStatsTarget, the target's moments.
template <int kMode>
__global__ __launch_bounds__(kLossThreads) void k_loss_stats(LossArgs a) {
  constexpr int kQ = kQuantities<kMode>;
  constexpr bool kTarget = kMode == kStatsTarget;
  constexpr int kMaps = kTarget ? 1 : 2;
  constexpr int kMapFloats = 3 * kSpanY * kStageStride;
  constexpr int kStaged = kMaps * kMapFloats;
  constexpr int kBlurred = kQ * 3 * kSpanY * kBlurStride;
  __shared__ __align__(16) float s_buf[kStaged > kBlurred ? kStaged : kBlurred];
  const int unit = xcd_remap(blockIdx.x, gridDim.x);
  const UnitGeo g = unit_geo(a, unit);
  float l1 = 0.0f, ssim_sum = 0.0f;
  if constexpr (kTarget) {
    const float* const src[1] = {a.target};
    stage_halo<1, kUnitH, kStageStride, false>(src, g.x0, g.y0, g, s_buf);
  } else {
    const float* const src[2] = {a.img, a.target};
    l1 = stage_halo<2, kUnitH, kStageStride, true>(src, g.x0, g.y0, g, s_buf);
  }
  __syncthreads();

  // Horizontal pass into registers: task = (channel, staged row, 8-column segment); each row's
  // 18 inputs are read as four float4 and a float2, and each product is formed once.
  float h[kQ][kHSeg];
  const bool h_task = threadIdx.x < kHTasks;
  const int h_ch = threadIdx.x / (kSpanY * 2);
  const int h_rem = threadIdx.x - h_ch * (kSpanY * 2);
  const int h_row = h_rem >> 1;
  const int h_c0 = (h_rem & 1) * kHSeg;
  if (h_task) {
    float in[kMaps][kHSeg + kWin - 1];
#pragma unroll
    for (int m = 0; m < kMaps; ++m) {
      const float* row = s_buf + m * kMapFloats + (h_ch * kSpanY + h_row) * kStageStride + h_c0;
      float tmp[20];
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) {
        const float4 x = *reinterpret_cast<const float4*>(row + 4 * q4);
        tmp[4 * q4] = x.x;
        tmp[4 * q4 + 1] = x.y;
        tmp[4 * q4 + 2] = x.z;
        tmp[4 * q4 + 3] = x.w;
      }
      const float2 x = *reinterpret_cast<const float2*>(row + 16);
      tmp[16] = x.x;
      tmp[17] = x.y;
#pragma unroll
      for (int i = 0; i < kHSeg + kWin - 1; ++i) in[m][i] = tmp[i];
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
#pragma unroll
      for (int o = 0; o < kHSeg; ++o) h[q][o] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kHSeg + kWin - 1; ++i) {
      float val[kQ];
      if constexpr (kMode == kStatsTarget) {
        const float t = in[0][i];
        val[0] = t;
        val[1] = t * t;
      } else {
        const float p = in[0][i], t = in[kMaps - 1][i];
        if constexpr (kMode == kStatsFull) {
          val[0] = p;
          val[1] = t;
          val[2] = p * p;
          val[3] = t * t;
          val[4] = p * t;
        } else {
          val[0] = p;
          val[1] = p * p;
          val[2] = p * t;
        }
      }
#pragma unroll
      for (int o = 0; o < kHSeg; ++o) {
        const int k = i - o;
        if (k < 0 || k >= kWin) continue;
        const float w = a.win.w[k];
#pragma unroll
        for (int q = 0; q < kQ; ++q) h[q][o] = fmaf(w, val[q], h[q][o]);
      }
    }
  }
  __syncthreads();  // every read of the staged maps is done
  float (*s_h)[3][kSpanY][kBlurStride] = reinterpret_cast<float (*)[3][kSpanY][kBlurStride]>(s_buf);
  if (h_task) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
#pragma unroll
      for (int o = 0; o < kHSeg; o += 2) {
        *reinterpret_cast<float2*>(&s_h[q][h_ch][h_row][h_c0 + o]) =
            make_float2(h[q][o], h[q][o + 1]);
      }
    }
  }
  __syncthreads();

  // Vertical pass: task = (channel, column, 8-row segment); then S and its derivatives.
  if (threadIdx.x < kVTasks) {
    const int ch = threadIdx.x >> 6;
    const int col = threadIdx.x & (kUnitW - 1);
    const int r0 = ((threadIdx.x >> 4) & 3) * kVSeg;
    const int x = g.x0 + col;
    // Output addresses: the unit's first element (uniform) plus a 32-bit byte offset per lane;
    // an output outside the image reads the unit's first element instead (offset 0).
    const int64_t origin = (g.first_pixel + int64_t(g.y0) * g.W + g.x0) * 3;
    const int64_t plane = a.pixels * 3;
    const uint32_t row_bytes = static_cast<uint32_t>(g.W) * 12u;
    const uint32_t off0 =
        static_cast<uint32_t>(r0) * row_bytes + static_cast<uint32_t>(col * 3 + ch) * 4u;
    // The target's moments at the outputs, loaded before the blur so it hides their latency.
    float om[2][kVSeg];
    if constexpr (kMode == kStatsMoments) {
      const float* m0_base = a.tmom + origin;
      const float* m1_base = a.tmom + plane + origin;
#pragma unroll
      for (int o = 0; o < kVSeg; ++o) {
        const int y = g.y0 + r0 + o;
        const bool ok = x < g.W && y < g.H;
        const uint32_t at = ok ? off0 + o * row_bytes : 0u;
        const float m0 = load_at(m0_base, at), m1 = load_at(m1_base, at);
        om[0][o] = m0;
        om[1][o] = m1;
      }
    }
    float st[kQ][kVSeg];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      float v[kVSeg + kWin - 1];
#pragma unroll
      for (int i = 0; i < kVSeg + kWin - 1; ++i) v[i] = s_h[q][ch][r0 + i][col];
#pragma unroll
      for (int o = 0; o < kVSeg; ++o) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < kWin; ++k) acc = fmaf(a.win.w[k], v[o + k], acc);
        st[q][o] = acc;
      }
    }
#pragma unroll
    for (int o = 0; o < kVSeg; ++o) {
      const int y = g.y0 + r0 + o;
      if (x >= g.W || y >= g.H) continue;
      const uint32_t at = off0 + o * row_bytes;
      if constexpr (kTarget) {
        store_at(a.tmom_out + origin, at, st[0][o]);
        store_at(a.tmom_out + plane + origin, at, st[1][o]);
      } else {
        // No contraction into fused multiply-adds here: both ways of obtaining the target's
        // moments then give the same S and derivatives bit for bit.
#pragma clang fp contract(off)
        Moments m;
        if constexpr (kMode == kStatsFull) {
          m = Moments{st[0][o], st[1][o], st[2][o], st[3][o], st[4][o]};
        } else {
          m = Moments{st[0][o], om[0][o], st[1][o], om[1][o], st[2][o]};
        }
        const float mu_p = m.mu_p, mu_t = m.mu_t;
        const float mu_p2 = mu_p * mu_p, mu_t2 = mu_t * mu_t, mu_pt = mu_p * mu_t;
        const float s_pp = m.m_pp - mu_p2;
        const float s_tt = m.m_tt - mu_t2;
        const float s_pt = m.m_pt - mu_pt;
        const float A1 = 2.0f * mu_pt + kC1;
        const float A2 = 2.0f * s_pt + kC2;
        const float B1 = mu_p2 + mu_t2 + kC1;
        const float B2 = s_pp + s_tt + kC2;
        // B1 >= C1 > 0 and B2 is C2 plus two variances: the hardware reciprocal is safe.
        const float inv_d = __builtin_amdgcn_rcpf(B1 * B2);
        const float S = A1 * A2 * inv_d;
        const float inv_b1 = B2 * inv_d, inv_b2 = B1 * inv_d;
        ssim_sum += S;
        if (a.keep) {
          store_at(a.abc + origin, at,
                   2.0f * mu_t * (A2 - A1) * inv_d - 2.0f * mu_p * S * (inv_b1 - inv_b2));
          store_at(a.abc + plane + origin, at, -S * inv_b2);
          store_at(a.abc + 2 * plane + origin, at, 2.0f * A1 * inv_d);
        }
      }
    }
  }
  if constexpr (!kTarget) {
    const float l1_tot = block_sum(l1);
    const float ssim_tot = block_sum(ssim_sum);
    if (threadIdx.x == 0) {
      a.partial[unit] = g.inv_n * (a.w_l1 * l1_tot - a.w_ssim * ssim_tot);
    }
  }
}

// L1 alone (ssim_weight 0): the unit's partial sum of |p - t|, or (backward) its gradient.
__global__ __launch_bounds__(kLossThreads) void k_loss_l1(LossArgs a, bool forward) {
  const int unit = xcd_remap(blockIdx.x, gridDim.x);
  const UnitGeo g = unit_geo(a, unit);
  const float scale = (!forward && a.scale != nullptr) ? *a.scale : 1.0f;
  float l1 = 0.0f;
  for (int e = threadIdx.x; e < kUnitH * kUnitW * 3; e += kLossThreads) {
    const int j = e / (kUnitW * 3);
    const int c3 = e - j * (kUnitW * 3);
    const int x = g.x0 + c3 / 3;
    const int y = g.y0 + j;
    if (x >= g.W || y >= g.H) continue;
    const int64_t at = (g.first_pixel + int64_t(y) * g.W + g.x0) * 3 + c3;
    const float d = a.img[at] - a.target[at];
    if (forward) {
      l1 += fabsf(d);
    } else {
      const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
      a.grad[at] = scale * (a.w_l1 * g.inv_n * sg);
    }
  }
  if (forward) {
    const float t = block_sum(l1);
    if (threadIdx.x == 0) a.partial[unit] = a.w_l1 * g.inv_n * t;
  }
}

// dL/dp = scale (w_l1 sign(p - t) / n - w_ssim (G*a + 2 p G*b + t G*c) / n); two workgroups per
// unit, one per 16 x 16 half.
__global__ __launch_bounds__(kLossThreads) void k_loss_grad(LossArgs a) {
  constexpr int kMapFloats = 3 * kGSpan * kGStageStride;
  constexpr int kStaged = 3 * kMapFloats;
  constexpr int kBlurred = 3 * 3 * kGSpan * kGBlurStride;
  __shared__ __align__(16) float s_buf[kStaged > kBlurred ? kStaged : kBlurred];
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const UnitGeo g = unit_geo(a, wg >> 1);
  const int y0 = g.y0 + (wg & 1) * kGT;
  if (y0 >= g.H) return;  // a unit clipped to one half (uniform)
  const float scale = a.scale != nullptr ? *a.scale : 1.0f;
  const float* const src[3] = {a.abc, a.abc + a.pixels * 3, a.abc + 2 * a.pixels * 3};
  const int h_ch = threadIdx.x / (kGSpan * 2);
  const int h_rem = threadIdx.x - h_ch * (kGSpan * 2);
  const int h_row = h_rem >> 1;
  const int h_c0 = (h_rem & 1) * kGHSeg;
  {
    stage_halo<3, kGT, kGStageStride, false>(src, g.x0, y0, g, s_buf);
    __syncthreads();
    float h[3][kGHSeg];
    if (threadIdx.x < kGHTasks) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        // The row's 18 inputs as four float4 and a float2 (16-byte staged rows).
        const float* row = s_buf + q * kMapFloats + (h_ch * kGSpan + h_row) * kGStageStride + h_c0;
        float v[kGHSeg + kWin - 1];
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
          const float4 x = *reinterpret_cast<const float4*>(row + 4 * q4);
          v[4 * q4] = x.x;
          v[4 * q4 + 1] = x.y;
          v[4 * q4 + 2] = x.z;
          v[4 * q4 + 3] = x.w;
        }
        const float2 x2 = *reinterpret_cast<const float2*>(row + 16);
        v[16] = x2.x;
        v[17] = x2.y;
#pragma unroll
        for (int o = 0; o < kGHSeg; ++o) {
          float acc = 0.0f;
#pragma unroll
          for (int k = 0; k < kWin; ++k) acc = fmaf(a.win.w[k], v[o + k], acc);
          h[q][o] = acc;
        }
      }
    }
    __syncthreads();  // every read of the staged maps is done
    float (*s_h)[3][kGSpan][kGBlurStride] =
        reinterpret_cast<float (*)[3][kGSpan][kGBlurStride]>(s_buf);
    if (threadIdx.x < kGHTasks) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int o = 0; o < kGHSeg; o += 4) {
          *reinterpret_cast<float4*>(&s_h[q][h_ch][h_row][h_c0 + o]) =
              make_float4(h[q][o], h[q][o + 1], h[q][o + 2], h[q][o + 3]);
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < kGVTasks) {
      const int ch = threadIdx.x >> 6;
      const int col = threadIdx.x & (kGT - 1);
      const int r0 = ((threadIdx.x >> 4) & 3) * kGVSeg;
      const int x = g.x0 + col;
      // Output addresses: the half's first element (uniform) plus a 32-bit byte offset per
      // lane; an output outside the image reads the half's first element instead (offset 0).
      const int64_t origin = (g.first_pixel + int64_t(y0) * g.W + g.x0) * 3;
      const uint32_t row_bytes = static_cast<uint32_t>(g.W) * 12u;
      const uint32_t off0 =
          static_cast<uint32_t>(r0) * row_bytes + static_cast<uint32_t>(col * 3 + ch) * 4u;
      float pv[kGVSeg], tv[kGVSeg];  // p and t at the outputs, loaded before the blur
#pragma unroll
      for (int o = 0; o < kGVSeg; ++o) {
        const int y = y0 + r0 + o;
        const bool ok = x < g.W && y < g.H;
        const uint32_t at = ok ? off0 + o * row_bytes : 0u;
        pv[o] = load_at(a.img + origin, at);
        tv[o] = load_at(a.target + origin, at);
      }
      float bl[3][kGVSeg];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        float v[kGVSeg + kWin - 1];
#pragma unroll
        for (int i = 0; i < kGVSeg + kWin - 1; ++i) v[i] = s_h[q][ch][r0 + i][col];
#pragma unroll
        for (int o = 0; o < kGVSeg; ++o) {
          float acc = 0.0f;
#pragma unroll
          for (int k = 0; k < kWin; ++k) acc = fmaf(a.win.w[k], v[o + k], acc);
          bl[q][o] = acc;
        }
      }
#pragma unroll
      for (int o = 0; o < kGVSeg; ++o) {
        const int y = y0 + r0 + o;
        if (x >= g.W || y >= g.H) continue;
        const uint32_t at = off0 + o * row_bytes;
        const float p = pv[o], t = tv[o];
        const float d = p - t;
        const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        const float dssim = bl[0][o] + 2.0f * p * bl[1][o] + t * bl[2][o];
        store_at(a.grad + origin, at, scale * (g.inv_n * (a.w_l1 * sg - a.w_ssim * dssim)));
      }
    }
  }
}

// loss = w_ssim * num_fields + sum of the unit partials, summed in float64 in a fixed order.
__global__ __launch_bounds__(1024) void k_loss_total(const float* __restrict__ partial,
                                                     int64_t tiles, double constant,
                                                     float* __restrict__ loss) {
  __shared__ double s_red[1024 / 64];
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < tiles; i += 1024) v += partial[i];
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = constant;
    for (int w = 0; w < 1024 / 64; ++w) t += s_red[w];
    *loss = static_cast<float>(t);
  }
}

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------

struct LossLayout {
  float* abc;
  float* partial;
  size_t bytes;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

LossLayout loss_layout(void* base, int64_t pixels, int64_t tiles) {
  LossLayout l{};
  const size_t abc_bytes = align256(size_t(pixels) * 9 * sizeof(float));
  const size_t part_bytes = align256(size_t(tiles) * sizeof(float));
  char* b = static_cast<char*>(base);
  l.abc = reinterpret_cast<float*>(b);
  l.partial = reinterpret_cast<float*>(b + abc_bytes);
  l.bytes = abc_bytes + part_bytes;
  return l;
}

Window gaussian_window() {
  // The reference's window (metrics.py:71-88), built in float32: exp(-x^2 / (2 sigma^2)) for
  // x = -5..5, sigma = 1.5, divided by its sum.
  Window w;
  float sum = 0.0f;
  for (int k = 0; k < kWin; ++k) {
    const float x = static_cast<float>(k - kHalo);
    w.w[k] = std::exp(-(x * x) / (2.0f * 1.5f * 1.5f));
    sum += w.w[k];
  }
  for (int k = 0; k < kWin; ++k) w.w[k] /= sum;
  return w;
}

// Validates the common arguments and fills `a` (everything but the per-call outputs).
int prepare_loss(const FieldTable& table, int64_t tiles, int64_t pixels, const float* img,
                 const float* target, double ssim_weight, void* ws, size_t ws_bytes,
                 LossArgs* a, const float* target_moments = nullptr) {
  if (img == nullptr || target == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "image and target must not be NULL");
  }
  if (!(std::isfinite(ssim_weight) && ssim_weight >= 0.0 && ssim_weight <= 1.0)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "ssim_weight must be in [0, 1]");
  }
  if (tiles <= 0 || tiles > INT32_MAX / 2 || pixels <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "the images must have pixels");
  }
  const LossLayout l = loss_layout(ws, pixels, tiles);
  if (ws == nullptr || reinterpret_cast<uintptr_t>(ws) % 256 != 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "workspace must be non-NULL and 256-byte aligned");
  }
  if (ws_bytes < l.bytes) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "workspace has %zu bytes, needs %zu", ws_bytes,
                     l.bytes);
  }
  *a = LossArgs{};
  a->fields = table;
  a->img = img;
  a->target = target;
  a->tmom = target_moments;
  a->abc = l.abc;
  a->partial = l.partial;
  a->pixels = pixels;
  a->w_ssim = static_cast<float>(ssim_weight);
  a->w_l1 = static_cast<float>(1.0 - ssim_weight);
  a->win = gaussian_window();
  return SSPLAT_OK;
}

// Loss value into *loss; with keep, also the SSIM derivative maps for run_loss_backward.
int run_loss_forward(LossArgs a, int num_fields, int64_t tiles, double ssim_weight, float* loss,
                     bool keep, hipStream_t s) {
  if (loss == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "loss is NULL");
  a.keep = keep;
  const int grid = static_cast<int>(tiles);
  if (ssim_weight > 0.0 && a.tmom != nullptr) {
    k_loss_stats<kStatsMoments><<<grid, kLossThreads, 0, s>>>(a);
  } else if (ssim_weight > 0.0) {
    k_loss_stats<kStatsFull><<<grid, kLossThreads, 0, s>>>(a);
  } else {
    k_loss_l1<<<grid, kLossThreads, 0, s>>>(a, true);
  }
  SSPLAT_LAUNCH_CHECK("k_loss_stats");
  k_loss_total<<<1, 1024, 0, s>>>(a.partial, tiles, ssim_weight * num_fields, loss);
  SSPLAT_LAUNCH_CHECK("k_loss_total");
  return SSPLAT_OK;
}

// The target's window moments G*t and G*t^2 into `moments` (2 maps of (pixels, 3)).
int run_target_moments(const FieldTable& table, int64_t tiles, int64_t pixels,
                       const float* target, float* moments, hipStream_t s) {
  if (target == nullptr || moments == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "target and moments must not be NULL");
  }
  if (tiles <= 0 || tiles > INT32_MAX / 2 || pixels <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "the images must have pixels");
  }
  LossArgs a{};
  a.fields = table;
  a.target = target;
  a.pixels = pixels;
  a.tmom_out = moments;
  a.win = gaussian_window();
  k_loss_stats<kStatsTarget><<<static_cast<int>(tiles), kLossThreads, 0, s>>>(a);
  SSPLAT_LAUNCH_CHECK("k_loss_stats");
  return SSPLAT_OK;
}

int run_loss_backward(LossArgs a, int64_t tiles, double ssim_weight, const float* scale,
                      float* grad, hipStream_t s) {
  if (grad == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "grad_image is NULL");
  a.grad = grad;
  a.scale = scale;
  const int grid = static_cast<int>(tiles);
  static_assert(kUnitH == 2 * kGT, "two gradient workgroups per unit");
  if (ssim_weight > 0.0) {
    k_loss_grad<<<2 * grid, kLossThreads, 0, s>>>(a);
  } else {
    k_loss_l1<<<grid, kLossThreads, 0, s>>>(a, false);
  }
  SSPLAT_LAUNCH_CHECK("k_loss_grad");
  return SSPLAT_OK;
}

// Units of one field: kUnitW columns by kUnitH rows, row-major.
int64_t plan_units(ssplat_field& f, int64_t first_unit, int64_t first_pixel) {
  const int64_t strips = (int64_t(f.width) + kUnitW - 1) / kUnitW;
  const int64_t segs = (int64_t(f.height) + kUnitH - 1) / kUnitH;
  f.first_pixel = first_pixel;
  f.first_tile = first_unit;
  f.tiles_x = static_cast<int32_t>(strips);
  f.num_tiles = static_cast<int32_t>(strips * segs);
  return strips * segs;
}

FieldTable single_image(int height, int width, int64_t* units) {
  ssplat_field f{};
  f.height = height;
  f.width = width;
  *units = plan_units(f, 0, 0);
  return FieldTable{nullptr, 0, f};
}

int check_size(int height, int width) {
  if (height <= 0 || width <= 0 || height > 65535 || width > 65535) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "height and width must be in [1, 65535]");
  }
  return SSPLAT_OK;
}

int check_table(const ssplat_field* fields, int num_fields) {
  if (fields == nullptr || num_fields <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "fields must be a non-empty device table");
  }
  return SSPLAT_OK;
}

}  // namespace

extern "C" {

int ssplat_loss_plan(ssplat_field* fields, int num_fields, int64_t* total_units,
                     int64_t* total_pixels, size_t* workspace_bytes) {
  int rc;
  if (fields == nullptr || num_fields <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "the field table must hold at least one field");
  }
  int64_t units = 0, px = 0;
  for (int i = 0; i < num_fields; ++i) {
    if ((rc = check_size(fields[i].height, fields[i].width))) return rc;
    units += plan_units(fields[i], units, px);
    px += int64_t(fields[i].height) * fields[i].width;
    if (units > INT32_MAX / 2) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "too many loss units");
  }
  if (total_units) *total_units = units;
  if (total_pixels) *total_pixels = px;
  if (workspace_bytes) *workspace_bytes = loss_layout(nullptr, px, units).bytes;
  return success();
}

int ssplat_loss_workspace_bytes(int64_t total_pixels, int64_t total_tiles, size_t* bytes) {
  if (bytes == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bytes is NULL");
  if (total_pixels <= 0 || total_tiles <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "pixels and tiles must be positive");
  }
  *bytes = loss_layout(nullptr, total_pixels, total_tiles).bytes;
  return success();
}

int ssplat_l1_ssim_loss(const float* image, const float* target, int height, int width,
                        double ssim_weight, float* loss, float* grad_image, void* workspace,
                        size_t workspace_bytes, void* stream) {
  int rc;
  if ((rc = check_size(height, width))) return rc;
  int64_t tiles;
  const FieldTable t = single_image(height, width, &tiles);
  LossArgs a;
  if ((rc = prepare_loss(t, tiles, int64_t(height) * width, image, target, ssim_weight,
                         workspace, workspace_bytes, &a))) {
    return rc;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((rc = run_loss_forward(a, 1, tiles, ssim_weight, loss, grad_image != nullptr, s))) return rc;
  if (grad_image && (rc = run_loss_backward(a, tiles, ssim_weight, nullptr, grad_image, s))) {
    return rc;
  }
  return success();
}

int ssplat_l1_ssim_loss_forward(const float* image, const float* target,
                                const float* target_moments, int height, int width,
                                double ssim_weight, float* loss, void* workspace,
                                size_t workspace_bytes, void* stream) {
  int rc;
  if ((rc = check_size(height, width))) return rc;
  int64_t tiles;
  const FieldTable t = single_image(height, width, &tiles);
  LossArgs a;
  if ((rc = prepare_loss(t, tiles, int64_t(height) * width, image, target, ssim_weight,
                         workspace, workspace_bytes, &a, target_moments))) {
    return rc;
  }
  if ((rc = run_loss_forward(a, 1, tiles, ssim_weight, loss, true,
                             static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

int ssplat_loss_target_moments(const float* target, int height, int width, float* moments,
                               void* stream) {
  int rc;
  if ((rc = check_size(height, width))) return rc;
  int64_t tiles;
  const FieldTable t = single_image(height, width, &tiles);
  if ((rc = run_target_moments(t, tiles, int64_t(height) * width, target, moments,
                               static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

int ssplat_batch_loss_target_moments(const ssplat_field* fields, int num_fields,
                                     int64_t total_tiles, int64_t total_pixels,
                                     const float* target, float* moments, void* stream) {
  int rc;
  if ((rc = check_table(fields, num_fields))) return rc;
  if ((rc = run_target_moments(FieldTable{fields, num_fields, ssplat_field{}}, total_tiles,
                               total_pixels, target, moments, static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

int ssplat_l1_ssim_loss_backward(const float* image, const float* target, int height, int width,
                                 double ssim_weight, const float* grad_loss, float* grad_image,
                                 void* workspace, size_t workspace_bytes, void* stream) {
  int rc;
  if ((rc = check_size(height, width))) return rc;
  int64_t tiles;
  const FieldTable t = single_image(height, width, &tiles);
  LossArgs a;
  if ((rc = prepare_loss(t, tiles, int64_t(height) * width, image, target, ssim_weight,
                         workspace, workspace_bytes, &a))) {
    return rc;
  }
  if ((rc = run_loss_backward(a, tiles, ssim_weight, grad_loss, grad_image,
                              static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

int ssplat_batch_l1_ssim_loss(const ssplat_field* fields, int num_fields, int64_t total_tiles,
                              int64_t total_pixels, const float* image, const float* target,
                              double ssim_weight, float* loss, float* grad_image,
                              void* workspace, size_t workspace_bytes, void* stream) {
  int rc;
  if ((rc = check_table(fields, num_fields))) return rc;
  LossArgs a;
  if ((rc = prepare_loss(FieldTable{fields, num_fields, ssplat_field{}}, total_tiles,
                         total_pixels, image, target, ssim_weight, workspace, workspace_bytes,
                         &a))) {
    return rc;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((rc = run_loss_forward(a, num_fields, total_tiles, ssim_weight, loss,
                             grad_image != nullptr, s))) {
    return rc;
  }
  if (grad_image &&
      (rc = run_loss_backward(a, total_tiles, ssim_weight, nullptr, grad_image, s))) {
    return rc;
  }
  return success();
}

int ssplat_batch_l1_ssim_loss_forward(const ssplat_field* fields, int num_fields,
                                      int64_t total_tiles, int64_t total_pixels,
                                      const float* image, const float* target,
                                      const float* target_moments, double ssim_weight,
                                      float* loss, void* workspace, size_t workspace_bytes,
                                      void* stream) {
  int rc;
  if ((rc = check_table(fields, num_fields))) return rc;
  LossArgs a;
  if ((rc = prepare_loss(FieldTable{fields, num_fields, ssplat_field{}}, total_tiles,
                         total_pixels, image, target, ssim_weight, workspace, workspace_bytes,
                         &a, target_moments))) {
    return rc;
  }
  if ((rc = run_loss_forward(a, num_fields, total_tiles, ssim_weight, loss, true,
                             static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

int ssplat_batch_l1_ssim_loss_backward(const ssplat_field* fields, int num_fields,
                                       int64_t total_tiles, int64_t total_pixels,
                                       const float* image, const float* target,
                                       double ssim_weight, const float* grad_loss,
                                       float* grad_image, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  int rc;
  if ((rc = check_table(fields, num_fields))) return rc;
  LossArgs a;
  if ((rc = prepare_loss(FieldTable{fields, num_fields, ssplat_field{}}, total_tiles,
                         total_pixels, image, target, ssim_weight, workspace, workspace_bytes,
                         &a))) {
    return rc;
  }
  if ((rc = run_loss_backward(a, total_tiles, ssim_weight, grad_loss, grad_image,
                              static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

}  // extern "C"
