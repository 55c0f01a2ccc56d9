// codec.hip - SSPL1 field decoding on gfx950 (the reference's codec.decode, codec.py:333-391,
// after the host has parsed the container and inflated its streams).
//
// The raw streams hold uniformly quantized integers (uint8/16/32 by bit width; the means stream
// optionally byte-planar).  Decoding is integer and byte work:
//   1. k_sspl1_unpack   - every value of the means and colors streams to a uint32 column;
//   2. the delta-coded columns (means under Morton coding, colors under color_delta) become
//      their prefix sums mod 2^bits: per-block sums (k_sspl1_partials), one scan of the block
//      sums per column (k_sspl1_top), block-local rescans with the block offsets
//      (k_sspl1_apply);
//   3. k_sspl1_dequant  - one thread per Gaussian: lo + (q / levels) * (hi - lo) in float64
//      with contraction off, rounded to float32 (the reference's NumPy arithmetic, so the
//      decoded values are the same bits); rotations over [0, pi) or [0, pi]; opacity
//      probabilities clipped to [1e-4, 1 - 1e-4] and turned into logits.

#include <cmath>
#include <cstdint>

#include "common.h"

namespace {

using namespace ssplat_detail;

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr int kBlockItems = kThreads * kPerThread;  // values per scan block
constexpr int kColumns = 5;                          // means x, y; colors r, g, b

// Value i of a stream of `width` bytes per value; byte-planar streams hold byte plane k of all
// `count` values at offset k * count.
__device__ __forceinline__ uint32_t read_value(const uint8_t* s, int width, bool planar,
                                               int64_t count, int64_t i) {
  if (width == 1) return s[i];
  uint32_t v = 0;
  if (planar) {
    for (int k = 0; k < width; ++k) v |= uint32_t(s[k * count + i]) << (8 * k);
  } else {
    for (int k = 0; k < width; ++k) v |= uint32_t(s[i * width + k]) << (8 * k);
  }
  return v;
}

struct Streams {
  const uint8_t* means;
  const uint8_t* scales;
  const uint8_t* rot;
  const uint8_t* colors;
  const uint8_t* opacity;  // NULL without opacities
};

struct Layout {
  int64_t n;
  int w_means, w_scales, w_rot, w_colors, w_opacity;  // bytes per value
  int bits_means, bits_scales, bits_rot, bits_colors, bits_opacity;
  bool morton, color_delta, planar, circular;
  double means_lo[2], means_span[2];
  double scale_lo[2], scale_span[2];
  double color_lo[3], color_span[3];
};

// Columns 0-1: means x, y (2n values of the means stream); 2-4: colors r, g, b (3n values).
__global__ __launch_bounds__(kThreads) void k_sspl1_unpack(Streams s, Layout l,
                                                           uint32_t* __restrict__ q) {
  const int64_t total = 5 * l.n;
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < total;
       k += int64_t(gridDim.x) * kThreads) {
    q[k] = k < 2 * l.n ? read_value(s.means, l.w_means, l.planar, 2 * l.n, k)
                       : read_value(s.colors, l.w_colors, false, 3 * l.n, k - 2 * l.n);
  }
}

// Sum of each kScanBlock-value block of every delta-coded column (uint64, exact).
__global__ __launch_bounds__(kThreads) void k_sspl1_partials(const uint32_t* __restrict__ q,
                                                             int64_t n,
                                                             uint64_t* __restrict__ bsums,
                                                             int blocks) {
  __shared__ uint64_t s_part[kThreads / 64];
  const int col = blockIdx.y;
  const uint32_t* c = q + col * n;
  const int64_t base = int64_t(blockIdx.x) * kBlockItems + int64_t(threadIdx.x) * kPerThread;
  uint64_t sum = 0;
  for (int k = 0; k < kPerThread; ++k) {
    if (base + k < n) sum += c[base + k];
  }
  for (int d = 32; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += s_part[w];
    bsums[int64_t(col) * blocks + blockIdx.x] = t;
  }
}

// Exclusive scan of each column's block sums, in place (one workgroup per column).
__global__ __launch_bounds__(kThreads) void k_sspl1_top(uint64_t* __restrict__ bsums,
                                                        int blocks) {
  __shared__ uint64_t s_carry;
  __shared__ uint64_t s_wave[kThreads / 64];
  uint64_t* b = bsums + int64_t(blockIdx.x) * blocks;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < blocks; base += kThreads) {
    const int i = base + threadIdx.x;
    const uint64_t v = i < blocks ? b[i] : 0;
    uint64_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint64_t before = s_carry;
    for (int w = 0; w < wave; ++w) before += s_wave[w];
    if (i < blocks) b[i] = before + x - v;
    __syncthreads();
    if (threadIdx.x == kThreads - 1) s_carry = before + x;
    __syncthreads();
  }
}

// Inclusive prefix sums mod 2^bits: each block rescans its values from its block offset.
__global__ __launch_bounds__(kThreads) void k_sspl1_apply(uint32_t* __restrict__ q, int64_t n,
                                                          const uint64_t* __restrict__ bsums,
                                                          int blocks, uint32_t mask_means,
                                                          uint32_t mask_colors, int first_col) {
  __shared__ uint64_t s_wave[kThreads / 64];
  const int col = first_col + blockIdx.y;
  uint32_t* c = q + col * n;
  const uint32_t mask = col < 2 ? mask_means : mask_colors;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = int64_t(blockIdx.x) * kBlockItems + int64_t(threadIdx.x) * kPerThread;
  uint32_t v[kPerThread];
  uint64_t sum = 0;
  for (int k = 0; k < kPerThread; ++k) {
    v[k] = base + k < n ? c[base + k] : 0u;
    sum += v[k];
  }
  uint64_t x = sum;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint64_t run = bsums[int64_t(col) * blocks + blockIdx.x] + x - sum;
  for (int w = 0; w < wave; ++w) run += s_wave[w];
  for (int k = 0; k < kPerThread; ++k) {
    run += v[k];
    if (base + k < n) c[base + k] = static_cast<uint32_t>(run) & mask;
  }
}

__device__ __forceinline__ float dequant(uint32_t q, int bits, double lo, double span) {
#pragma clang fp contract(off)
  const double levels = double((uint64_t(1) << bits) - 1);
  return static_cast<float>(lo + (double(q) / levels) * span);
}

__global__ __launch_bounds__(kThreads) void k_sspl1_dequant(Streams s, Layout l,
                                                            const uint32_t* __restrict__ q,
                                                            float* __restrict__ means,
                                                            float* __restrict__ log_scales,
                                                            float* __restrict__ rotations,
                                                            float* __restrict__ colors,
                                                            float* __restrict__ logits) {
#pragma clang fp contract(off)
  const int64_t n = l.n;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 2; ++a) {
    means[2 * i + a] = dequant(q[a * n + i], l.bits_means, l.means_lo[a], l.means_span[a]);
    const uint32_t qs = read_value(s.scales, l.w_scales, false, 2 * n, a * n + i);
    log_scales[2 * i + a] = dequant(qs, l.bits_scales, l.scale_lo[a], l.scale_span[a]);
  }
  const uint32_t qr = read_value(s.rot, l.w_rot, false, n, i);
  if (l.circular) {
    const double levels = double(uint64_t(1) << l.bits_rot);
    rotations[i] = static_cast<float>((double(qr) / levels) * M_PI);
  } else {
    rotations[i] = dequant(qr, l.bits_rot, 0.0, M_PI);
  }
  for (int c = 0; c < 3; ++c) {
    colors[3 * i + c] = dequant(q[(2 + c) * n + i], l.bits_colors, l.color_lo[c],
                                l.color_span[c]);
  }
  if (logits != nullptr && s.opacity != nullptr) {
    float p = dequant(read_value(s.opacity, l.w_opacity, false, n, i), l.bits_opacity, 0.0, 1.0);
    p = fminf(fmaxf(p, 1e-4f), static_cast<float>(1.0 - 1e-4));
    const float ratio = p / (1.0f - p);
    logits[i] = static_cast<float>(log(static_cast<double>(ratio)));
  }
}

int width_of(int bits) { return bits <= 8 ? 1 : (bits <= 16 ? 2 : 4); }

int scan_blocks(int64_t n) { return static_cast<int>((n + kBlockItems - 1) / kBlockItems); }

}  // namespace

extern "C" {

int ssplat_sspl1_workspace_bytes(int64_t n, size_t* bytes) {
  if (bytes == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bytes is NULL");
  if (n < 0 || n > (int64_t(1) << 31) - 1) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "n out of range");
  }
  const size_t q = (size_t(kColumns) * size_t(n) * sizeof(uint32_t) + 255) / 256 * 256;
  const size_t b = size_t(kColumns) * size_t(scan_blocks(n) + 1) * sizeof(uint64_t);
  *bytes = q + (b + 255) / 256 * 256;
  return success();
}

int ssplat_sspl1_decode(const ssplat_sspl1* layout, const uint8_t* means_raw,
                        const uint8_t* scales_raw, const uint8_t* rot_raw,
                        const uint8_t* colors_raw, const uint8_t* opacity_raw, float* means,
                        float* log_scales, float* rotations, float* colors,
                        float* opacity_logits, void* workspace, size_t workspace_bytes,
                        void* stream) {
  if (layout == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "layout is NULL");
  const ssplat_sspl1& h = *layout;
  if (h.n < 0 || h.n > (int64_t(1) << 31) - 1) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "n out of range");
  }
  const int bits[5] = {h.bits_means, h.bits_scales, h.bits_rot, h.bits_colors, h.bits_opacity};
  for (int k = 0; k < 5; ++k) {
    if (bits[k] < 1 || bits[k] > 32) {
      return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bit widths must be in [1, 32]");
    }
  }
  if (h.n == 0) return success();
  if (means_raw == nullptr || scales_raw == nullptr || rot_raw == nullptr ||
      colors_raw == nullptr || means == nullptr || log_scales == nullptr ||
      rotations == nullptr || colors == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "streams and outputs must not be NULL");
  }
  if (h.has_opacity && opacity_raw == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "the opacity stream is NULL");
  }
  size_t need = 0;
  int rc = ssplat_sspl1_workspace_bytes(h.n, &need);
  if (rc) return rc;
  if (workspace == nullptr || workspace_bytes < need ||
      reinterpret_cast<uintptr_t>(workspace) % 256 != 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT,
                     "workspace must be 256-byte aligned with at least %zu bytes", need);
  }
  Layout l{};
  l.n = h.n;
  l.bits_means = h.bits_means;
  l.bits_scales = h.bits_scales;
  l.bits_rot = h.bits_rot;
  l.bits_colors = h.bits_colors;
  l.bits_opacity = h.bits_opacity;
  l.w_means = width_of(h.bits_means);
  l.w_scales = width_of(h.bits_scales);
  l.w_rot = width_of(h.bits_rot);
  l.w_colors = width_of(h.bits_colors);
  l.w_opacity = width_of(h.bits_opacity);
  l.morton = h.morton != 0;
  l.color_delta = h.color_delta != 0;
  l.planar = h.means_byte_planar != 0;
  l.circular = h.rot_circular != 0;
  for (int a = 0; a < 2; ++a) {
    l.means_lo[a] = h.means_lo[a];
    l.means_span[a] = h.means_hi[a] - h.means_lo[a];
    l.scale_lo[a] = h.scale_lo[a];
    l.scale_span[a] = h.scale_hi[a] - h.scale_lo[a];
  }
  for (int c = 0; c < 3; ++c) {
    l.color_lo[c] = h.color_lo[c];
    l.color_span[c] = h.color_hi[c] - h.color_lo[c];
  }
  const Streams s{means_raw, scales_raw, rot_raw, colors_raw,
                  h.has_opacity ? opacity_raw : nullptr};
  uint32_t* q = static_cast<uint32_t*>(workspace);
  const size_t q_bytes = (size_t(kColumns) * size_t(h.n) * sizeof(uint32_t) + 255) / 256 * 256;
  uint64_t* bsums = reinterpret_cast<uint64_t*>(static_cast<char*>(workspace) + q_bytes);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t total = 5 * h.n;
  const int unpack_blocks = static_cast<int>(std::min<int64_t>((total + kThreads - 1) / kThreads,
                                                               int64_t(8192)));
  k_sspl1_unpack<<<unpack_blocks, kThreads, 0, st>>>(s, l, q);
  SSPLAT_LAUNCH_CHECK("k_sspl1_unpack");
  // Delta-coded columns: 0-1 (means) under Morton coding, 2-4 (colors) under color_delta.
  const int first = l.morton ? 0 : 2;
  const int last = l.color_delta ? 5 : 2;
  if (last > first) {
    const int blocks = scan_blocks(h.n);
    const dim3 grid(blocks, last - first);
    k_sspl1_partials<<<grid, kThreads, 0, st>>>(q + int64_t(first) * h.n, h.n,
                                                bsums + int64_t(first) * blocks, blocks);
    SSPLAT_LAUNCH_CHECK("k_sspl1_partials");
    k_sspl1_top<<<last - first, kThreads, 0, st>>>(bsums + int64_t(first) * blocks, blocks);
    SSPLAT_LAUNCH_CHECK("k_sspl1_top");
    const uint32_t mask_means =
        h.bits_means >= 32 ? 0xffffffffu : ((uint32_t(1) << h.bits_means) - 1u);
    const uint32_t mask_colors =
        h.bits_colors >= 32 ? 0xffffffffu : ((uint32_t(1) << h.bits_colors) - 1u);
    k_sspl1_apply<<<grid, kThreads, 0, st>>>(q, h.n, bsums, blocks, mask_means, mask_colors,
                                             first);
    SSPLAT_LAUNCH_CHECK("k_sspl1_apply");
  }
  k_sspl1_dequant<<<static_cast<int>((h.n + kThreads - 1) / kThreads), kThreads, 0, st>>>(
      s, l, q, means, log_scales, rotations, colors, h.has_opacity ? opacity_logits : nullptr);
  SSPLAT_LAUNCH_CHECK("k_sspl1_dequant");
  return success();
}

}  // extern "C"
