// ssplat.hip - sorting-free normalized / additive 2D Gaussian rasterizer for gfx950 (MI355X).
//
// Everything behind the C ABI of include/ssplat.h lives in this file:
//   * tile binning   : per-Gaussian clipped support rectangles -> per-tile pair counts (atomics)
//                      -> hand-written integer exclusive scan -> scatter into per-tile lists
//                      -> per-tile LDS bitonic sort, which yields ascending Gaussian ids per tile
//                      (the order the reference obtains from a stable radix sort).
//   * record packing : one 48-byte record per Gaussian (mean, conic, opacity, color, support
//                      rectangle) so a staged Gaussian costs three 16-byte loads.
//   * tiled forward  : one 16x16 tile per 256-thread workgroup, one 8x8 quadrant per wave64,
//                      records staged through LDS with coalesced loads, a ballot pre-filter per
//                      wave, and per-pixel register accumulation in ascending Gaussian id.
//   * tiled backward : recomputes the per-pixel numerator/denominator (nothing is saved by the
//                      forward), turns them into per-pixel adjoints in LDS, then sweeps the tile
//                      list again with each wave owning a quarter of the Gaussians over all 256
//                      pixels (four per lane), a transposed permlane/DPP wave reduction that
//                      leaves one gradient component per lane, one atomic set per
//                      (tile, Gaussian).
// No MFMA: the path is a gather/accumulate, not a contraction.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"

// ------------------------------------------------------------------------------------------
// Error reporting (shared with the other sources of the library through common.h)
// ------------------------------------------------------------------------------------------

namespace ssplat_detail {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int success() {
  g_last_error.clear();
  return SSPLAT_OK;
}

const char* last_error() { return g_last_error.c_str(); }

}  // namespace ssplat_detail

namespace {

using namespace ssplat_detail;

// ------------------------------------------------------------------------------------------
// Kernel timing (measurement hook): event pairs recorded on the launch stream around every
// render-kernel launch while enabled, summed and released by ssplat_kernel_timing_read.
// ------------------------------------------------------------------------------------------

struct KernelTimer {
  std::mutex mu;
  bool enabled = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs[2];  // [0] forward, [1] backward
};

KernelTimer& kernel_timer() {
  static KernelTimer t;
  return t;
}

// Records the start event of one timed launch (returns false when timing is off or fails).
bool timer_begin(int kernel, hipStream_t stream, hipEvent_t* start) {
  KernelTimer& t = kernel_timer();
  std::lock_guard<std::mutex> lock(t.mu);
  if (!t.enabled) return false;
  if (hipEventCreate(start) != hipSuccess) return false;
  if (hipEventRecord(*start, stream) != hipSuccess) {
    (void)hipEventDestroy(*start);
    return false;
  }
  (void)kernel;
  return true;
}

void timer_end(int kernel, hipStream_t stream, hipEvent_t start) {
  KernelTimer& t = kernel_timer();
  hipEvent_t stop;
  if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, stream) != hipSuccess) {
    (void)hipEventDestroy(start);
    return;
  }
  std::lock_guard<std::mutex> lock(t.mu);
  t.pairs[kernel].emplace_back(start, stop);
}

// ------------------------------------------------------------------------------------------
// Constants
// ------------------------------------------------------------------------------------------

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;  // counts per first-level scan block
constexpr int kMaxScanBlocks = kScanBlock;             // the top-level scan is one block
constexpr int64_t kMaxScanItems = int64_t(kScanBlock) * kMaxScanBlocks;
constexpr int kBinThreads = 256;
// Up to this many tile counters are scanned by one workgroup (one launch); larger grids use the
// three-launch scan.
constexpr int kSingleScanPerThread = 16;
constexpr int64_t kSingleScanItems = int64_t(1) << 16;
constexpr int kSortThreads = 256;
constexpr int kSortLds = 4096;      // lists up to this length sort in the block's LDS
constexpr int kSortWaveLds = 1024;  // lists up to this length sort in one wave's LDS slice
constexpr int kStage = 256;  // Gaussian records staged in LDS per batch
// The 16-pixel backward stages fewer records per batch: with the recomputing sweep's masks its
// LDS then stays within 20 KB, so eight workgroups (eight waves per SIMD) fit on a CU.
constexpr int kBwdStage16 = 128;
// The 16-pixel forward too: its filter pass walks the batch 64 records at a time, and tile lists
// are mostly shorter than 128.
constexpr int kFwdStage16 = 128;
// The backward sums a (tile, Gaussian) pair whose rectangle covers at most
// kLaneClassArea * lane_classes(tile) pixels of the tile by a quad of lanes (a quad sum instead
// of the wave's transposed reduction); larger pairs take a whole wave.  The small pairs are
// listed by area classes of kLaneClassArea pixels.  The limits were measured: 96 pixels with
// 16-pixel tiles, 128 with 32-pixel tiles.
constexpr int kLaneClassArea = 16;
constexpr int lane_classes(int tile) { return tile == 16 ? 6 : 8; }
// Deterministic gradient accumulation.  Gaussian g owns R(g) slots, one per tile of its clipped
// support rectangle (row-major); slots of consecutive Gaussians are consecutive.  The backward
// writes each (tile, Gaussian) pair's raw sums (nine values) and the call's tag to the pair's
// slot, and the reduction sums every Gaussian's slots in a fixed order, so the gradients do not
// depend on scheduling.  A slot whose tag is not the current call's (a tile the lists do not
// hold, e.g. culled) counts as zero.
constexpr int kSlotWords = 10;            // 9 partial gradients + the call's tag
// Affine colors add 6 color-gradient, 2 scale and 1 rotation partials: 18 values + padding +
// the tag (word 19).
constexpr int kSlotWordsAffine = 20;
constexpr size_t kSlotHeaderBytes = 256;  // slot buffer header: the call counter (uint32)
constexpr int kLaneSumMax = 64;           // Gaussians with up to this many slots: one lane each
// Work queues of a persistent render launch: up to kMaxQueuesPerXcd tile ranges per XCD.
constexpr int kMaxQueuesPerXcd = 8;
constexpr int kQueues = kXcds * kMaxQueuesPerXcd;
static_assert(kQueues <= 64, "one wave probes every queue");
constexpr int kWorkgroupsPerQueue = 64;  // resident workgroups sharing one queue (see TileQueue)
constexpr int kFieldWaves = 4;            // waves per 256-thread workgroup of per-Gaussian sweeps
// The render kernels run one workgroup per tile, one thread per pixel (tile edge 16: 256 threads,
// 4 waves; tile edge 32: 1024 threads, 16 waves), one wave per 8x8 quadrant.
// -0.5 * log2(e): exp(-q/2) == exp2(q * kNegHalfLog2e).
constexpr float kNegHalfLog2e = -0.72134752044448170368f;
constexpr float kInvNegHalfLog2e = -1.38629436111989061883f;  // 1 / kNegHalfLog2e = -2 ln 2
// Clamps that keep the integer support arithmetic exact in int64 (see support_rect).
constexpr float kCenterClamp = 1.0e12f;
constexpr int64_t kRadiusClamp = int64_t(1) << 36;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------------------------------------
// Shared device helpers
// ------------------------------------------------------------------------------------------

// Half-open support rectangle [x0, x1) x [y0, y1) of one Gaussian, clipped to the image.
struct Rect {
  int x0, x1, y0, y1;
};

// Clipped support rectangle: centre rounded half-to-even, +-radius, clipped to the image;
// fully outside -> empty (render.py:62-78).  Non-finite means are dropped (render_ext.cu:28-34).
// The centre is clamped to +-1e12 and the radii to +-2^36 so the int64 arithmetic cannot
// overflow; for every mean and radius inside those bounds the result equals the reference's.
__device__ __forceinline__ Rect support_rect(float mx, float my, int64_t rx, int64_t ry, int H,
                                             int W) {
  Rect r{0, 0, 0, 0};
  if (!isfinite(mx) || !isfinite(my)) return r;
  const float fx = fminf(fmaxf(rintf(mx), -kCenterClamp), kCenterClamp);
  const float fy = fminf(fmaxf(rintf(my), -kCenterClamp), kCenterClamp);
  const int64_t ix = static_cast<int64_t>(fx);
  const int64_t iy = static_cast<int64_t>(fy);
  rx = rx < -kRadiusClamp ? -kRadiusClamp : (rx > kRadiusClamp ? kRadiusClamp : rx);
  ry = ry < -kRadiusClamp ? -kRadiusClamp : (ry > kRadiusClamp ? kRadiusClamp : ry);
  const int64_t x0 = ix - rx > 0 ? ix - rx : 0;
  const int64_t x1 = ix + rx < W - 1 ? ix + rx : W - 1;
  const int64_t y0 = iy - ry > 0 ? iy - ry : 0;
  const int64_t y1 = iy + ry < H - 1 ? iy + ry : H - 1;
  if (x1 < x0 || y1 < y0) return r;
  r.x0 = static_cast<int>(x0);
  r.x1 = static_cast<int>(x1) + 1;
  r.y0 = static_cast<int>(y0);
  r.y1 = static_cast<int>(y1) + 1;
  return r;
}

// Conic quadratic in the reference's operation order, no FMA contraction (bit-exact with the
// oracle's float32 restatement; only the ellipse cull uses it).
__device__ __forceinline__ float conic_q_exact(float dx, float dy, float a, float b, float c) {
#pragma clang fp contract(off)
  return a * dx * dx + 2.0f * b * dx * dy + c * dy * dy;
}

// Exact continuous minimum of the conic quadratic over an axis-aligned rectangle (pixel-centre
// coordinates).  Same arithmetic as render_ext.cu:101-133 so the cull decision is bit-exact:
// a malformed conic keeps the pair (-FLT_MAX), a rectangle containing the mean gives 0, else the
// minimum over the four edges, each at its clamped 1-D minimiser.
__device__ __forceinline__ float tile_min_q(float mx, float my, float a, float b, float c,
                                            float x_lo, float x_hi, float y_lo, float y_hi) {
#pragma clang fp contract(off)
  const float det = a * c - b * b;
  if (!(a > 0.0f && c > 0.0f && det > 0.0f)) return -FLT_MAX;
  if (mx >= x_lo && mx <= x_hi && my >= y_lo && my <= y_hi) return 0.0f;
  const float slope_y = -(b / c);
  const float slope_x = -(b / a);
  float best = FLT_MAX;
  float dx = x_lo - mx;
  float dy = fminf(fmaxf(slope_y * dx, y_lo - my), y_hi - my);
  best = fminf(best, conic_q_exact(dx, dy, a, b, c));
  dx = x_hi - mx;
  dy = fminf(fmaxf(slope_y * dx, y_lo - my), y_hi - my);
  best = fminf(best, conic_q_exact(dx, dy, a, b, c));
  dy = y_lo - my;
  dx = fminf(fmaxf(slope_x * dy, x_lo - mx), x_hi - mx);
  best = fminf(best, conic_q_exact(dx, dy, a, b, c));
  dy = y_hi - my;
  dx = fminf(fmaxf(slope_x * dy, x_lo - mx), x_hi - mx);
  return fminf(best, conic_q_exact(dx, dy, a, b, c));
}

// Tile range of a clipped support rectangle packed as xr = x0 | x1 << 16, yr = y0 | y1 << 16,
// in a grid of tile edge 1 << lg; count = tw * th is the number of slots the Gaussian owns (0 for
// an empty rectangle).  Same tile bounds as the binning's enumeration.
struct RectTiles {
  int tx0, ty0, tw, th, count;
};

__device__ __forceinline__ RectTiles rect_tiles(uint32_t xr, uint32_t yr, int lg) {
  const int x0 = static_cast<int>(xr & 0xffff), x1 = static_cast<int>(xr >> 16);
  const int y0 = static_cast<int>(yr & 0xffff), y1 = static_cast<int>(yr >> 16);
  RectTiles r{0, 0, 0, 0, 0};
  if (x1 > x0 && y1 > y0) {
    r.tx0 = x0 >> lg;
    r.ty0 = y0 >> lg;
    r.tw = ((x1 - 1) >> lg) - r.tx0 + 1;
    r.th = ((y1 - 1) >> lg) - r.ty0 + 1;
    r.count = r.tw * r.th;
  }
  return r;
}

// Hardware 2^x (v_exp_f32).  The same instruction evaluates both the per-pixel weight and the
// fade tail, so a pixel whose quadratic equals sigma_cutoff^2 gets a faded weight of exactly 0.
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }

template <int kCtrl>
__device__ __forceinline__ float dpp_move(float v) {
  // Lanes without a source read 0 (old = 0, bound_ctrl off).
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}

// Inclusive prefix sums over the 64 lanes of a wave (lane l gets v(0) + ... + v(l)) by DPP:
// Hillis-Steele within each 16-lane row, then the rows' totals by row broadcasts.  No LDS
// traffic.  Must be called with all lanes active.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int32_t dpp_i32(int32_t v) {
  // Lanes without a source, and rows outside kRowMask, read 0.
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, kRowMask, 0xf, false);
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int64_t dpp_i64(int64_t v) {
  const uint32_t lo = static_cast<uint32_t>(dpp_i32<kCtrl, kRowMask>(static_cast<int32_t>(v)));
  const int32_t hi = dpp_i32<kCtrl, kRowMask>(static_cast<int32_t>(v >> 32));
  return static_cast<int64_t>((uint64_t(uint32_t(hi)) << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  auto move = [](auto ctrl, auto mask, T x) -> T {
    if constexpr (sizeof(T) == 8) {
      return dpp_i64<decltype(ctrl)::value, decltype(mask)::value>(x);
    } else {
      return dpp_i32<decltype(ctrl)::value, decltype(mask)::value>(x);
    }
  };
  using C1 = std::integral_constant<int, 0x111>;   // row_shr:1
  using C2 = std::integral_constant<int, 0x112>;   // row_shr:2
  using C4 = std::integral_constant<int, 0x114>;   // row_shr:4
  using C8 = std::integral_constant<int, 0x118>;   // row_shr:8
  using B15 = std::integral_constant<int, 0x142>;  // row_bcast:15 (lane 15 of a row -> next row)
  using B31 = std::integral_constant<int, 0x143>;  // row_bcast:31 (lane 31 -> rows 2 and 3)
  using All = std::integral_constant<int, 0xf>;
  v += move(C1{}, All{}, v);
  v += move(C2{}, All{}, v);
  v += move(C4{}, All{}, v);
  v += move(C8{}, All{}, v);
  v += move(B15{}, std::integral_constant<int, 0xa>{}, v);  // rows 1 and 3
  v += move(B31{}, std::integral_constant<int, 0xc>{}, v);  // rows 2 and 3
  return v;
}

// Sum over the 64 lanes of a wave; the total lands in lane 63.  Must be called with all lanes
// active.
__device__ __forceinline__ float wave_sum_lane63(float v) {
  v += dpp_move<0xb1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x114>(v);  // row_shr:4
  v += dpp_move<0x118>(v);  // row_shr:8
  v += dpp_move<0x142>(v);  // row_bcast:15
  v += dpp_move<0x143>(v);  // row_bcast:31
  return v;
}

// x + y after exchanging the upper 32 lanes of x with the lower 32 lanes of y
// (v_permlane32_swap): lanes 0-31 get x(l) + x(l+32), lanes 32-63 get y(l-32) + y(l).
__device__ __forceinline__ float swap32_add(float x, float y) {
  const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false,
                                                  false);
  return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}

// Same with the odd 16-lane rows of x exchanged with the even rows of y (v_permlane16_swap):
// row 0 gets x's rows 0+1, row 1 y's rows 0+1, row 2 x's rows 2+3, row 3 y's rows 2+3.
__device__ __forceinline__ float swap16_add(float x, float y) {
  const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false,
                                                  false);
  return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}

// Transposed reduction of eight per-lane values over a wave64: on return, lane 8k+7 holds the
// wave sum of v[k].  Each step halves the lanes a value spans while pairing values, so eight
// values cost 18 VALU operations instead of 8 independent 6-step trees.  `hi8` is
// (lane & 8) != 0.  Must be called with all lanes active.
__device__ __forceinline__ float wave_sum8_transposed(const float (&v)[8], bool hi8) {
  const float a0 = swap32_add(v[0], v[4]);  // lanes 0-31: v0, 32-63: v4
  const float a1 = swap32_add(v[1], v[5]);
  const float a2 = swap32_add(v[2], v[6]);
  const float a3 = swap32_add(v[3], v[7]);
  const float b0 = swap16_add(a0, a2);      // rows: v0, v2, v4, v6
  const float b1 = swap16_add(a1, a3);      // rows: v1, v3, v5, v7
  const float t0 = b0 + dpp_move<0x128>(b0);  // row_ror:8
  const float t1 = b1 + dpp_move<0x128>(b1);
  float c = hi8 ? t1 : t0;                   // lanes 0-7 of a row: b0's value, 8-15: b1's
  c += dpp_move<0xb1>(c);                    // quad_perm [1,0,3,2]
  c += dpp_move<0x4e>(c);                    // quad_perm [2,3,0,1]
  c += dpp_move<0x114>(c);                   // row_shr:4 -> lanes 7 and 15 of each row
  return c;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t count, int32_t value) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    p[i] = value;
  }
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t count, float value) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    p[i] = value;
  }
}

// Sticky status: a non-OK source is copied, an OK source leaves the destination untouched, so a
// caller can clear the word once and check it after many binnings.
__global__ void k_merge_status(const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && *src != SSPLAT_OK) *dst = *src;
}

int fill_i32(int32_t* p, int64_t count, int32_t value, hipStream_t stream) {
  if (count <= 0) return SSPLAT_OK;
  const int64_t blocks64 = (count + 255) / 256;
  const int blocks = static_cast<int>(blocks64 < 4096 ? blocks64 : 4096);
  k_fill_i32<<<blocks, 256, 0, stream>>>(p, count, value);
  SSPLAT_LAUNCH_CHECK("fill_i32");
  return SSPLAT_OK;
}

int fill_f32(float* p, int64_t count, float value, hipStream_t stream) {
  if (count <= 0) return SSPLAT_OK;
  const int64_t blocks64 = (count + 255) / 256;
  const int blocks = static_cast<int>(blocks64 < 4096 ? blocks64 : 4096);
  k_fill_f32<<<blocks, 256, 0, stream>>>(p, count, value);
  SSPLAT_LAUNCH_CHECK("fill_f32");
  return SSPLAT_OK;
}

// ------------------------------------------------------------------------------------------
// Tile binning
// ------------------------------------------------------------------------------------------

struct BinArgs {
  const float* means;
  const float* conics;
  const int64_t* radii;
  FieldTable fields;
  int n, ts, cull;
  float cull_thr;  // cutoff^2 + 1e-5 * max(cutoff^2, 1), computed in float32 on the host
};

// Enumerates every live (tile, Gaussian) pair of this workgroup's 256 Gaussians.  Each wave owns
// 64 consecutive Gaussians; their pairs are numbered by a wave-level prefix sum and dealt out
// to the lanes round-robin, so one large Gaussian does not serialize its wave.  Each lane takes
// kEnumUnroll pairs per step (pair k, k + 64, ...): their owner searches (a fixed six-step
// binary search over the wave's 64 prefix sums in LDS) are independent and overlap, and the
// step's pairs go to `emit` together so that its memory operations overlap too.
// emit(tile[kEnumUnroll], gid[kEnumUnroll], live[kEnumUnroll]).
constexpr int kEnumUnroll = 4;

template <class Emit>
__device__ __forceinline__ void enumerate_pairs(const BinArgs& a, Emit&& emit) {
  __shared__ int64_t s_incl[kBinThreads];
  __shared__ int4 s_rect[kBinThreads];
  __shared__ int4 s_tile[kBinThreads];
  __shared__ int2 s_grid[kBinThreads];
  __shared__ float4 s_geo[kBinThreads];
  __shared__ float s_b[kBinThreads];
  const int lane = threadIdx.x & 63;
  const int wbase = threadIdx.x & ~63;
  const int i = blockIdx.x * kBinThreads + threadIdx.x;
  int64_t cnt = 0;
  Rect r{0, 0, 0, 0};
  int tx0 = 0, ty0 = 0, tw = 1;
  int first_tile = 0, tiles_x = 1;
  float mx = 0.0f, my = 0.0f;
  if (i < a.n) {
    const ssplat_field f = field_of_gaussian(a.fields, i);
    first_tile = static_cast<int>(f.first_tile);
    tiles_x = f.tiles_x;
    mx = a.means[2 * i + 0];
    my = a.means[2 * i + 1];
    r = support_rect(mx, my, a.radii[2 * i + 0], a.radii[2 * i + 1], f.height, f.width);
    if (r.x1 > r.x0 && r.y1 > r.y0) {
      tx0 = r.x0 / a.ts;
      ty0 = r.y0 / a.ts;
      const int tx1 = (r.x1 - 1) / a.ts;
      const int ty1 = (r.y1 - 1) / a.ts;
      tw = tx1 - tx0 + 1;
      cnt = int64_t(tw) * (ty1 - ty0 + 1);
    }
  }
  const int64_t incl = wave_inclusive_scan(cnt);
  const int64_t total =
      (int64_t(__builtin_amdgcn_readlane(static_cast<int32_t>(incl >> 32), 63)) << 32) |
      uint32_t(__builtin_amdgcn_readlane(static_cast<int32_t>(incl), 63));
  s_incl[threadIdx.x] = incl;
  s_rect[threadIdx.x] = make_int4(r.x0, r.x1, r.y0, r.y1);
  s_tile[threadIdx.x] = make_int4(tx0, ty0, tw, i);
  s_grid[threadIdx.x] = make_int2(first_tile, tiles_x);
  if (a.cull) {
    float ca = 0.0f, cb = 0.0f, cc = 0.0f;
    if (i < a.n) {
      ca = a.conics[3 * i + 0];
      cb = a.conics[3 * i + 1];
      cc = a.conics[3 * i + 2];
    }
    s_geo[threadIdx.x] = make_float4(mx, my, ca, cc);
    s_b[threadIdx.x] = cb;
  }
  __syncthreads();
  for (int64_t base = 0; base < total; base += 64 * kEnumUnroll) {
    int64_t kk[kEnumUnroll];
    int lo[kEnumUnroll];
#pragma unroll
    for (int u = 0; u < kEnumUnroll; ++u) {
      const int64_t k = base + 64 * u + lane;
      kk[u] = k < total ? k : total - 1;
      lo[u] = 0;
    }
    // Owner = the first of the 64 Gaussians whose inclusive prefix exceeds k.
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
#pragma unroll
      for (int u = 0; u < kEnumUnroll; ++u) {
        if (s_incl[wbase + lo[u] + step - 1] <= kk[u]) lo[u] += step;
      }
    }
    int tiles[kEnumUnroll], gidv[kEnumUnroll];
    bool live[kEnumUnroll];
#pragma unroll
    for (int u = 0; u < kEnumUnroll; ++u) {
      const int o = wbase + lo[u];
      const int64_t excl = lo[u] > 0 ? s_incl[o - 1] : 0;
      // A pair's index within its Gaussian is below tw * th <= 65535^2, so 32-bit unsigned.
      const uint32_t local = static_cast<uint32_t>(kk[u] - excl);
      const int4 tl = s_tile[o];
      const uint32_t w = static_cast<uint32_t>(tl.z);
      const uint32_t row = local / w;
      const int tile_x = tl.x + static_cast<int>(local - row * w);
      const int tile_y = tl.y + static_cast<int>(row);
      bool keep = base + 64 * u + lane < total;
      if (a.cull) {
        const int4 rc = s_rect[o];
        const float4 g = s_geo[o];
        const int rx0 = max(rc.x, tile_x * a.ts);
        const int rx1 = min(rc.y - 1, (tile_x + 1) * a.ts - 1);
        const int ry0 = max(rc.z, tile_y * a.ts);
        const int ry1 = min(rc.w - 1, (tile_y + 1) * a.ts - 1);
        const float mq = tile_min_q(g.x, g.y, g.z, s_b[o], g.w, static_cast<float>(rx0),
                                    static_cast<float>(rx1), static_cast<float>(ry0),
                                    static_cast<float>(ry1));
        keep = keep && !(mq > a.cull_thr);
      }
      const int2 grid = s_grid[o];
      tiles[u] = grid.x + tile_y * grid.y + tile_x;
      gidv[u] = tl.w;
      live[u] = keep;
    }
    emit(tiles, gidv, live);
  }
}

// Exclusive scan of one value per thread over a kThreads-thread block; returns the block total
// through *total.  Reusable back to back (ends with a barrier).
template <int kThreads>
__device__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
  constexpr int kWaves = kThreads / 64;
  __shared__ int64_t s_wave[kWaves];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t x = wave_inclusive_scan(v);
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  if (wave == 0) {
    const int64_t s = wave_inclusive_scan(lane < kWaves ? s_wave[lane] : int64_t(0));
    if (lane < kWaves) s_wave[lane] = s;
  }
  __syncthreads();
  const int64_t prefix = wave > 0 ? s_wave[wave - 1] : 0;
  *total = s_wave[kWaves - 1];
  __syncthreads();
  return prefix + x - v;
}

// Clears the tile counters and the binning status word.
__global__ void k_bin_reset(int32_t* __restrict__ counts, int64_t items,
                            int32_t* __restrict__ status) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < items; i += stride) {
    counts[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *status = SSPLAT_OK;
}

__global__ __launch_bounds__(kBinThreads) void k_bin_count(BinArgs a, int32_t* __restrict__ counts) {
  enumerate_pairs(a, [&](const int (&tiles)[kEnumUnroll], const int (&)[kEnumUnroll],
                         const bool (&live)[kEnumUnroll]) {
#pragma unroll
    for (int u = 0; u < kEnumUnroll; ++u) {
      if (live[u]) atomicAdd(&counts[tiles[u]], 1);
    }
  });
}

// One-workgroup exclusive scan of counts[0, items) into offsets, for grids of up to
// kSingleScanItems counters: each thread scans kSingleScanPerThread consecutive counters of a
// kScanThreads * kSingleScanPerThread chunk (16-byte loads and stores), with a running carry
// between chunks.  Saturates at INT32_MAX and flags *status like k_scan_apply.  The counters are
// left as they are: the scatter counts them back down to zero.
__global__ __launch_bounds__(kScanThreads) void k_scan_single(const int32_t* __restrict__ counts,
                                                              int64_t items,
                                                              int32_t* __restrict__ offsets,
                                                              int32_t* __restrict__ status,
                                                              uint32_t* __restrict__ bump,
                                                              int32_t* __restrict__ queue) {
  if (bump != nullptr && threadIdx.x == 0) *bump += 1u;  // a counter the caller advances
  // The next launch's persistent work-queue counters, zeroed for it.
  if (queue != nullptr && threadIdx.x < kQueues) queue[threadIdx.x] = 0;
  constexpr int kPer = kSingleScanPerThread;
  constexpr int kChunk = kScanThreads * kPer;
  static_assert(kPer % 4 == 0, "counters are loaded four at a time");
  // The workspace (counts) is 256-byte aligned; the caller's offsets may not be.
  const bool offsets_aligned = (reinterpret_cast<uintptr_t>(offsets) & 15) == 0;
  int64_t carry = 0;
  bool saturated = false;
  for (int64_t base = 0; base < items; base += kChunk) {
    const int64_t i0 = base + int64_t(threadIdx.x) * kPer;
    int32_t v[kPer];
    int64_t sum = 0;
    if (i0 + kPer <= items) {
      // counts is 256-byte aligned and i0 a multiple of 4: aligned 16-byte loads.
#pragma unroll
      for (int k = 0; k < kPer; k += 4) {
        const int4 q = *reinterpret_cast<const int4*>(counts + i0 + k);
        v[k] = q.x;
        v[k + 1] = q.y;
        v[k + 2] = q.z;
        v[k + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kPer; ++k) v[k] = i0 + k < items ? counts[i0 + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) sum += v[k];
    int64_t total;
    int64_t run = carry + block_exclusive_scan<kScanThreads>(sum, &total);
    int32_t o[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      saturated |= (i0 + k < items) & (run > INT32_MAX);
      o[k] = static_cast<int32_t>(run > INT32_MAX ? INT32_MAX : run);
      run += v[k];
    }
    if (offsets_aligned && i0 + kPer <= items) {
#pragma unroll
      for (int k = 0; k < kPer; k += 4) {
        *reinterpret_cast<int4*>(offsets + i0 + k) = make_int4(o[k], o[k + 1], o[k + 2], o[k + 3]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        if (i0 + k < items) offsets[i0 + k] = o[k];
      }
    }
    carry += total;
  }
  if (saturated) *status = SSPLAT_ERR_CAPACITY;
}

// Scatter cursors rebuilt from a finished offsets table: cursor[t] = offsets[t+1] - offsets[t]
// (for a fill issued on its own, after the counting call's counters may have been reused).
__global__ void k_cursors_from_offsets(const int32_t* __restrict__ offsets, int num_tiles,
                                       int32_t* __restrict__ cursors) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num_tiles) return;
  const int32_t d = offsets[t + 1] - offsets[t];
  cursors[t] = d > 0 ? d : 0;
}

__global__ __launch_bounds__(kBinThreads) void k_bin_scatter(BinArgs a,
                                                             const int32_t* __restrict__ offsets,
                                                             int32_t* __restrict__ cursors,
                                                             int32_t* __restrict__ gids,
                                                             int64_t capacity,
                                                             int32_t* __restrict__ status) {
  // Each tile's cursor holds its pair count and is counted down to zero: the pairs land in
  // [offsets[t], offsets[t+1]) in arbitrary order (the per-tile sort orders them) and the counters
  // end at zero.
  enumerate_pairs(a, [&](const int (&tiles)[kEnumUnroll], const int (&gid)[kEnumUnroll],
                         const bool (&live)[kEnumUnroll]) {
    // All of the step's cursor updates are issued before their results are used.
    int32_t slot[kEnumUnroll], first[kEnumUnroll];
#pragma unroll
    for (int u = 0; u < kEnumUnroll; ++u) {
      slot[u] = live[u] ? atomicSub(&cursors[tiles[u]], 1) : 0;
      first[u] = live[u] ? offsets[tiles[u]] : 0;
    }
#pragma unroll
    for (int u = 0; u < kEnumUnroll; ++u) {
      if (!live[u]) continue;
      const int64_t pos = int64_t(first[u]) + slot[u] - 1;
      if (pos >= 0 && pos < capacity) {
        gids[pos] = gid[u];
      } else {
        *status = SSPLAT_ERR_CAPACITY;
      }
    }
  });
}

// First level: per-block sums of `items` int32 counts.
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(const int32_t* __restrict__ counts,
                                                                int64_t items,
                                                                int64_t* __restrict__ block_sums) {
  const int64_t base = int64_t(blockIdx.x) * kScanBlock + int64_t(threadIdx.x) * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < items) s += counts[base + k];
  }
  int64_t total;
  block_exclusive_scan<kScanThreads>(s, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Second level: exclusive scan of the block sums, in place (one block, <= kMaxScanBlocks sums).
__global__ __launch_bounds__(kScanThreads) void k_scan_top(int64_t* __restrict__ block_sums,
                                                           int blocks,
                                                           uint32_t* __restrict__ bump,
                                                           int32_t* __restrict__ queue) {
  if (bump != nullptr && threadIdx.x == 0) *bump += 1u;  // a counter the caller advances
  if (queue != nullptr && threadIdx.x < kQueues) queue[threadIdx.x] = 0;
  const int base = threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < blocks ? block_sums[base + k] : 0;
    s += v[k];
  }
  int64_t total;
  int64_t run = block_exclusive_scan<kScanThreads>(s, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < blocks) block_sums[base + k] = run;
    run += v[k];
  }
}

// Third level: exclusive offsets of every count, written as int32.  A running total that does not
// fit int32 is saturated at INT32_MAX and flags the workspace status (the tile index stores int32
// offsets, like the reference), so a host reading tile_offsets[T] sees the overflow.
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int32_t* __restrict__ counts,
                                                             int64_t items,
                                                             const int64_t* __restrict__ block_sums,
                                                             int32_t* __restrict__ offsets,
                                                             int32_t* __restrict__ status) {
  const int64_t base = int64_t(blockIdx.x) * kScanBlock + int64_t(threadIdx.x) * kScanItems;
  int32_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < items ? counts[base + k] : 0;
    s += v[k];
  }
  int64_t total;
  int64_t run = block_exclusive_scan<kScanThreads>(s, &total) + block_sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < items) {
      if (run > INT32_MAX) *status = SSPLAT_ERR_CAPACITY;
      offsets[base + k] = static_cast<int32_t>(run > INT32_MAX ? INT32_MAX : run);
    }
    run += v[k];
  }
}

// Ascending sort of a[0, L) by a bitonic network in its "flip" form: every comparator puts the
// smaller value at the lower index, so virtual +inf padding above L never moves and comparators
// reaching past L are skipped.  kGlobal: a lives in global memory (lists longer than the LDS
// buffer), which needs device-scope fences between the stages.
template <bool kGlobal>
__device__ void bitonic_sort(int32_t* a, int L) {
  int p2 = 1;
  while (p2 < L) p2 <<= 1;
  const int pairs = p2 >> 1;
  for (int k = 2; k <= p2; k <<= 1) {
    const int half = k >> 1;
    for (int p = threadIdx.x; p < pairs; p += blockDim.x) {
      const int i = (p / half) * k + (p % half);
      const int j = i ^ (k - 1);
      if (j < L) {
        const int32_t x = a[i];
        const int32_t y = a[j];
        if (x > y) {
          a[i] = y;
          a[j] = x;
        }
      }
    }
    if (kGlobal) __threadfence();
    __syncthreads();
    for (int m = k >> 2; m >= 1; m >>= 1) {
      for (int p = threadIdx.x; p < pairs; p += blockDim.x) {
        const int i = (p / m) * 2 * m + (p % m);
        const int j = i + m;
        if (j < L) {
          const int32_t x = a[i];
          const int32_t y = a[j];
          if (x > y) {
            a[i] = y;
            a[j] = x;
          }
        }
      }
      if (kGlobal) __threadfence();
      __syncthreads();
    }
  }
}

// Orders the LDS accesses of one wave: its lanes run in lockstep and LDS executes a wave's
// operations in issue order, so a release/acquire pair (no s_barrier) is enough between the
// compare-exchange stages of a wave-private sort.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Ascending bitonic sort of up to 64 values held one per lane (INT32_MAX padding sorts last).
// Must be called with all 64 lanes active.
// The value of lane (lane ^ kJ), by DPP (kJ <= 8) or a permlane swap (16, 32): no LDS traffic.
// Must be called with all 64 lanes active.
template <int kJ>
__device__ __forceinline__ int32_t lane_xor(int32_t v, int lane) {
  if constexpr (kJ == 1) {
    return __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
  } else if constexpr (kJ == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
  } else if constexpr (kJ == 4) {
    // Mirror within 8 lanes (i -> i ^ 7), then within 4 (i -> i ^ 3): i -> i ^ 4.
    const int32_t m = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    return __builtin_amdgcn_update_dpp(0, m, 0x1b, 0xf, 0xf, false);  // quad_perm [3,2,1,0]
  } else if constexpr (kJ == 8) {
    return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xf, 0xf, false);  // row_ror:8
  } else if constexpr (kJ == 16) {
    // v_permlane16_swap(v, v): s[0] holds rows (0, 0, 2, 2) of v, s[1] rows (1, 1, 3, 3).
    const auto s = __builtin_amdgcn_permlane16_swap(uint32_t(v), uint32_t(v), false, false);
    return static_cast<int32_t>((lane & 16) ? s[0] : s[1]);
  } else {
    static_assert(kJ == 32, "lane_xor covers 1, 2, 4, 8, 16 and 32");
    // v_permlane32_swap(v, v): s[0] holds halves (0, 0) of v, s[1] halves (1, 1).
    const auto s = __builtin_amdgcn_permlane32_swap(uint32_t(v), uint32_t(v), false, false);
    return static_cast<int32_t>((lane & 32) ? s[0] : s[1]);
  }
}

// One half-cleaner step at distance kJ < 64 over element e = 64 r + lane of each register.
template <int kR, int kK, int kJ>
__device__ __forceinline__ void sort_step_lanes(int32_t (&v)[kR], int lane) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int32_t p = lane_xor<kJ>(v[r], lane);
    const bool up = ((64 * r + lane) & kK) == 0;  // ascending block
    const bool keep_min = ((lane & kJ) == 0) == up;
    v[r] = keep_min ? min(v[r], p) : max(v[r], p);
  }
}

// A half-cleaner step at distance kJ >= 64: registers r and r + kJ / 64 of one lane.
template <int kR, int kK, int kJ>
__device__ __forceinline__ void sort_step_regs(int32_t (&v)[kR]) {
  constexpr int kD = kJ / 64;
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if ((r & kD) == 0) {
      const bool up = ((64 * r) & kK) == 0;
      const int32_t lo = min(v[r], v[r + kD]), hi = max(v[r], v[r + kD]);
      v[r] = up ? lo : hi;
      v[r + kD] = up ? hi : lo;
    }
  }
}

template <int kR, int kK, int kJ>
__device__ __forceinline__ void sort_merge(int32_t (&v)[kR], int lane) {
  if constexpr (kJ >= 64) {
    sort_step_regs<kR, kK, kJ>(v);
  } else {
    sort_step_lanes<kR, kK, kJ>(v, lane);
  }
  if constexpr (kJ > 1) sort_merge<kR, kK, kJ / 2>(v, lane);
}

template <int kR, int kK>
__device__ __forceinline__ void sort_stages(int32_t (&v)[kR], int lane) {
  sort_merge<kR, kK, kK / 2>(v, lane);
  if constexpr (kK < 64 * kR) sort_stages<kR, 2 * kK>(v, lane);
}

// Ascending bitonic sort of 64 kR values held kR per lane (element 64 r + lane in v[r]), all in
// registers: lane exchanges by DPP and permlane swaps.  INT32_MAX padding sorts last.  Must be
// called with all 64 lanes active.
template <int kR>
__device__ __forceinline__ void wave_sort_regs(int32_t (&v)[kR], int lane) {
  sort_stages<kR, 2>(v, lane);
}

// Sorts a tile list of L <= 64 kR entries by one wave, in registers.
template <int kR>
__device__ __forceinline__ void sort_tile_in_registers(int32_t* __restrict__ g, int L, int lane) {
  int32_t v[kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) v[r] = 64 * r + lane < L ? g[64 * r + lane] : INT32_MAX;
  wave_sort_regs<kR>(v, lane);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (64 * r + lane < L) g[64 * r + lane] = v[r];
  }
}

// Flip-form bitonic sort of a[0, L) by the 64 lanes of one wave (a in the wave's LDS slice).
__device__ void wave_sort_lds(int32_t* a, int L, int lane) {
  int p2 = 1;
  while (p2 < L) p2 <<= 1;
  const int pairs = p2 >> 1;
  for (int k = 2; k <= p2; k <<= 1) {
    const int half = k >> 1;
    for (int p = lane; p < pairs; p += 64) {
      const int i = (p / half) * k + (p % half);
      const int j = i ^ (k - 1);
      if (j < L) {
        const int32_t x = a[i];
        const int32_t y = a[j];
        if (x > y) {
          a[i] = y;
          a[j] = x;
        }
      }
    }
    wave_lds_sync();
    for (int m = k >> 2; m >= 1; m >>= 1) {
      for (int p = lane; p < pairs; p += 64) {
        const int i = (p / m) * 2 * m + (p % m);
        const int j = i + m;
        if (j < L) {
          const int32_t x = a[i];
          const int32_t y = a[j];
          if (x > y) {
            a[i] = y;
            a[j] = x;
          }
        }
      }
      wave_lds_sync();
    }
  }
}

__device__ __forceinline__ int tile_segment(const int32_t* offsets, int t, int64_t capacity,
                                            int64_t* start) {
  int64_t s = offsets[t];
  int64_t e = offsets[t + 1];
  s = s < 0 ? 0 : (s > capacity ? capacity : s);
  e = e < s ? s : (e > capacity ? capacity : e);
  *start = s;
  return static_cast<int>(e - s);
}

// Sorts every tile's gid list and merges the binning status.  One wave per tile: lists of <= 64 entries sort in registers
// (shuffle bitonic, no barriers), lists of <= kSortWaveLds entries in the wave's LDS slice; the
// rare longer lists are sorted by the whole block afterwards (LDS up to kSortLds entries, global
// memory beyond).
__global__ __launch_bounds__(kSortThreads) void k_sort_tiles(int32_t* __restrict__ gids,
                                                             const int32_t* __restrict__ offsets,
                                                             int num_tiles, int64_t capacity,
                                                             const int32_t* __restrict__ ws_status,
                                                             int32_t* __restrict__ status) {
  constexpr int kWaves = kSortThreads / 64;
  // The binning's status is final (the scatter ran before this launch): merge it into the
  // caller's sticky status word.
  if (status != nullptr && blockIdx.x == 0 && threadIdx.x == 0 && *ws_status != SSPLAT_OK) {
    *status = *ws_status;
  }
  __shared__ int32_t buf[kSortLds];
  __shared__ int s_big[kWaves];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  for (int base = blockIdx.x * kWaves; base < num_tiles; base += gridDim.x * kWaves) {
    const int t = base + wave;
    int64_t s = 0;
    const int L = t < num_tiles ? tile_segment(offsets, t, capacity, &s) : 0;
    if (L <= 64) {
      if (L > 1) sort_tile_in_registers<1>(gids + s, L, lane);
    } else if (L <= 128) {
      sort_tile_in_registers<2>(gids + s, L, lane);
    } else if (L <= 256) {
      sort_tile_in_registers<4>(gids + s, L, lane);
    } else if (L <= kSortWaveLds) {
      int32_t* a = buf + wave * kSortWaveLds;
      for (int i = lane; i < L; i += 64) a[i] = gids[s + i];
      wave_lds_sync();
      wave_sort_lds(a, L, lane);
      for (int i = lane; i < L; i += 64) gids[s + i] = a[i];
    }
    if (lane == 0) s_big[wave] = L > kSortWaveLds ? t : -1;
    __syncthreads();
    for (int w = 0; w < kWaves; ++w) {
      const int tb = s_big[w];
      if (tb < 0) continue;
      int64_t sb = 0;
      const int Lb = tile_segment(offsets, tb, capacity, &sb);
      if (Lb <= kSortLds) {
        for (int i = threadIdx.x; i < Lb; i += blockDim.x) buf[i] = gids[sb + i];
        __syncthreads();
        bitonic_sort<false>(buf, Lb);
        for (int i = threadIdx.x; i < Lb; i += blockDim.x) gids[sb + i] = buf[i];
      } else {
        bitonic_sort<true>(gids + sb, Lb);
      }
      __syncthreads();
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// Strip binning: the tile lists without a global atomic per (tile, Gaussian) pair
// ------------------------------------------------------------------------------------------
//
// The global tile order (every field's tiles, row-major, fields in order) is cut into strips of
// 2^lg consecutive tiles (lg = 4, 5 for large grids, down to 0 for small ones), and the Gaussians into chunks of consecutive ids, one
// chunk per level-1 workgroup.  Two launches:
//
//   k_strip_level1, one workgroup per chunk: lists every (strip, Gaussian) pair of the chunk
//     whose rectangle meets the strip, with the Gaussian's tile mask inside the strip (its pairs
//     there, after the ellipse cull), first in LDS, then grouped by strip in the chunk's own
//     region of the bucket.  The chunk's row of a chunks x (strips + 1) matrix gets, per strip,
//     the chunk's first entry of the strip and the chunk's pairs in the strips before it.  LDS
//     atomics only.
//   k_strip_level2, one workgroup per strip: reads its matrix column, whose sum over the chunks
//     is the number of pairs before the strip (its first list position), gathers its entries
//     from every chunk's region, counts its pairs per tile, writes its tiles' offsets, places the
//     pairs per tile in LDS and sorts every tile's pairs into its list, which gives the
//     reference's ascending Gaussian ids.
//
// A chunk whose entries outgrow its region (Gaussians that meet very many strips) stores none;
// level 2 then takes each of its Gaussians as a candidate and computes the masks itself.
// Grids with more than kStripMax strips of 32 tiles use the per-pair kernels above.

constexpr int kStripMax = 7680;         // strips of one binning (level-1 LDS bins)
constexpr int kStripTarget = 512;       // narrower strips below this many
constexpr int kStripL1Threads = 1024;   // level-1 workgroup
constexpr int kStripL2Threads = 512;    // level-2 workgroup
constexpr int kStripMaxChunks = 1024;   // chunks of one binning (level-2 LDS table)
constexpr int kStripChunkUnit = 1024;   // chunks hold a multiple of this many Gaussians
constexpr int kStripIdBits = 19;        // level 1 packs (strip, id in chunk) into 32 bits
constexpr int kStripMaxChunk = 1 << kStripIdBits;
constexpr int kStripL1Lds = 65024;      // level-1 dynamic LDS: counters plus the entry list
constexpr int kStripEntriesPerGaussian = 8;  // a chunk's region: 8 entries per Gaussian
constexpr int kStripCache = 1024;       // level-2 candidates in LDS per round
constexpr int kStripPairsLds = 4096;    // pairs a strip places in LDS (more: in the lists)
constexpr int kStripL1PerThread = 8;    // strips a level-1 thread scans
static_assert(kStripMax <= kStripL1Threads * kStripL1PerThread, "level 1 scans every strip");
static_assert(kStripMaxChunks <= kStripL2Threads * 2, "level 2 scans 2 chunks per thread");
static_assert(kStripMax < (1 << (32 - kStripIdBits)), "strip numbers fit the packed entry");
static_assert(kStripPairsLds >= kSortLds, "long lists sort in the placement buffer");

struct StripArgs {
  int lg;                // log2 of the tiles per strip
  int strips;            // ceil(T / 2^lg)
  int num_tiles;         // T
  int chunks;            // ceil(n / chunk), at least 1
  int chunk;             // Gaussians per chunk (a multiple of kStripChunkUnit)
  int chunk_cap;         // entries of a chunk's region
  int list_cap;          // entries of level 1's LDS list
  int2* offs;            // chunks x (strips + 1): per strip, the chunk's first entry of the
                         // strip and its pairs before the strip; [strips] = its totals
  int2* bucket;          // chunks x chunk_cap entries (gid, tile mask)
};

// One Gaussian's tile rectangle in its field's tile grid (inclusive; empty when tx1 < tx0), its
// clipped pixel rectangle and mean (the cull's inputs), and its field's tile range.
struct GaussTiles {
  int tx0, tx1, ty0, ty1;
  int first_tile, tiles_x, num_tiles;
  int x0, x1, y0, y1;  // half-open clipped support rectangle
  float mx, my;
};

__device__ __forceinline__ GaussTiles empty_tiles() {
  GaussTiles t;
  t.tx0 = 0, t.tx1 = -1, t.ty0 = 0, t.ty1 = -1;
  t.first_tile = 0, t.tiles_x = 1, t.num_tiles = 0;
  t.x0 = 0, t.x1 = 0, t.y0 = 0, t.y1 = 0;
  t.mx = 0.0f, t.my = 0.0f;
  return t;
}

// Same rectangle and tile bounds as enumerate_pairs.
__device__ __forceinline__ GaussTiles gauss_tiles(const BinArgs& a, int g) {
  GaussTiles t = empty_tiles();
  const ssplat_field f = field_of_gaussian(a.fields, g);
  t.first_tile = static_cast<int>(f.first_tile);
  t.tiles_x = f.tiles_x;
  t.num_tiles = f.num_tiles;
  t.mx = a.means[2 * g + 0];
  t.my = a.means[2 * g + 1];
  const Rect r = support_rect(t.mx, t.my, a.radii[2 * g + 0], a.radii[2 * g + 1], f.height,
                              f.width);
  if (r.x1 > r.x0 && r.y1 > r.y0) {
    t.x0 = r.x0, t.x1 = r.x1, t.y0 = r.y0, t.y1 = r.y1;
    t.tx0 = r.x0 / a.ts;
    t.tx1 = (r.x1 - 1) / a.ts;
    t.ty0 = r.y0 / a.ts;
    t.ty1 = (r.y1 - 1) / a.ts;
  }
  return t;
}

// The tiles of strip s (global tiles [s << lg, (s + 1) << lg)) that hold a (tile, Gaussian) pair
// of this Gaussian: the tiles of its rectangle inside the strip, less the culled ones (the
// per-pair kernels' test, operation for operation).  Bit k stands for tile (s << lg) + k.
__device__ __forceinline__ uint32_t strip_mask(const BinArgs& a, const GaussTiles& t, float ca,
                                               float cb, float cc, int s, int lg) {
  if (t.tx1 < t.tx0) return 0u;
  const int s0 = s << lg;
  // The strip's tiles inside this Gaussian's field, as field-local indices [l0, l1).
  const int l0 = max(s0, t.first_tile) - t.first_tile;
  const int l1 = min(s0 + (1 << lg), t.first_tile + t.num_tiles) - t.first_tile;
  if (l1 <= l0) return 0u;
  const int rlo = max(l0 / t.tiles_x, t.ty0);
  const int rhi = min((l1 - 1) / t.tiles_x, t.ty1);
  uint32_t m = 0u;
  for (int row = rlo; row <= rhi; ++row) {
    const int rb = row * t.tiles_x;  // field-local index of the row's first tile
    const int xa = max(t.tx0, max(l0 - rb, 0));
    const int xb = min(t.tx1, min(l1 - 1 - rb, t.tiles_x - 1));
    if (xb < xa) continue;
    const int k0 = t.first_tile + rb + xa - s0;  // in [0, 2^lg), and k0 + xb - xa < 2^lg
    if (!a.cull) {
      m |= static_cast<uint32_t>(((uint64_t(1) << (xb - xa + 1)) - 1) << k0);
    } else {
      const int ry0 = max(t.y0, row * a.ts);
      const int ry1 = min(t.y1 - 1, (row + 1) * a.ts - 1);
      for (int tx = xa; tx <= xb; ++tx) {
        const int rx0 = max(t.x0, tx * a.ts);
        const int rx1 = min(t.x1 - 1, (tx + 1) * a.ts - 1);
        const float mq = tile_min_q(t.mx, t.my, ca, cb, cc, static_cast<float>(rx0),
                                    static_cast<float>(rx1), static_cast<float>(ry0),
                                    static_cast<float>(ry1));
        if (!(mq > a.cull_thr)) m |= 1u << (k0 + tx - xa);
      }
    }
  }
  return m;
}

// Gaussian g's mask for strip s, from its inputs.
__device__ __forceinline__ uint32_t entry_mask(const BinArgs& a, int g, int s, int lg) {
  const GaussTiles t = gauss_tiles(a, g);
  float ca = 0.0f, cb = 0.0f, cc = 0.0f;
  if (a.cull) {
    ca = a.conics[3 * g + 0];
    cb = a.conics[3 * g + 1];
    cc = a.conics[3 * g + 2];
  }
  return strip_mask(a, t, ca, cb, cc, s, lg);
}

// The owner lane's copy of every GaussTiles field (all lanes active).
__device__ __forceinline__ GaussTiles shfl_tiles(const GaussTiles& t, int o) {
  GaussTiles r;
  r.tx0 = __shfl(t.tx0, o), r.tx1 = __shfl(t.tx1, o);
  r.ty0 = __shfl(t.ty0, o), r.ty1 = __shfl(t.ty1, o);
  r.first_tile = __shfl(t.first_tile, o), r.tiles_x = __shfl(t.tiles_x, o);
  r.num_tiles = __shfl(t.num_tiles, o);
  r.x0 = __shfl(t.x0, o), r.x1 = __shfl(t.x1, o), r.y0 = __shfl(t.y0, o), r.y1 = __shfl(t.y1, o);
  r.mx = __shfl(t.mx, o), r.my = __shfl(t.my, o);
  return r;
}

// Gaussian g for level 1 (empty past n) and its conic when kFull and culling.
template <bool kFull>
__device__ __forceinline__ void strip_gaussian(const BinArgs& a, int64_t g, GaussTiles* t,
                                               float* ca, float* cb, float* cc) {
  *t = empty_tiles();
  *ca = 0.0f, *cb = 0.0f, *cc = 0.0f;
  if (g < a.n) {
    const int i = static_cast<int>(g);
    *t = gauss_tiles(a, i);
    if (kFull && a.cull) {
      *ca = a.conics[3 * i + 0];
      *cb = a.conics[3 * i + 1];
      *cc = a.conics[3 * i + 2];
    }
  }
}

// One pass of level 1 over 1024 Gaussians of a chunk (ids g0 + sub + thread): the Gaussians' tile
// rows are numbered by a wave prefix sum and dealt to the lanes round-robin, so a tall Gaussian
// does not serialize its wave.  A lane handling row `row` of a Gaussian takes the strips of that
// row's tiles, less the first when the previous row's last strip is the same (a strip met by two
// rows is listed once: tile indices grow with the row, so a strip meeting rows r and r + 2 also
// meets r + 1).  step(s_lo, ne, owner tiles, ca, cb, cc, id in chunk) runs on every lane of each
// step (ne = 0 on a lane without a row): the lane's strips are s_lo .. s_lo + ne - 1.
template <class Step>
__device__ __forceinline__ void strip_steps(const BinArgs& a, int lg, int64_t g0, int sub,
                                            Step&& step) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  GaussTiles t;
  float ca, cb, cc;
  strip_gaussian<true>(a, g0 + sub + tid, &t, &ca, &cb, &cc);
  const int local0 = sub + (tid & ~63);  // the wave's first Gaussian, relative to the chunk
  const int units = t.tx1 >= t.tx0 ? t.ty1 - t.ty0 + 1 : 0;
  const int incl = wave_inclusive_scan(units);
  const int total = __builtin_amdgcn_readlane(incl, 63);
  for (int base = 0; base < total; base += 64) {
    const int k = min(base + lane, total - 1);
    int o = 0;
#pragma unroll
    for (int sz = 32; sz >= 1; sz >>= 1) {
      if (__shfl(incl, o + sz - 1) <= k) o += sz;
    }
    const int before = __shfl(incl, o > 0 ? o - 1 : 0);
    const int excl = o > 0 ? before : 0;
    const GaussTiles ot = shfl_tiles(t, o);
    const float oa = __shfl(ca, o), ob = __shfl(cb, o), oc = __shfl(cc, o);
    int s_lo = 0, s_hi = -1;
    if (base + lane < total) {
      const int row = ot.ty0 + (k - excl);
      const int rowbase = ot.first_tile + row * ot.tiles_x;
      s_lo = (rowbase + ot.tx0) >> lg;
      s_hi = (rowbase + ot.tx1) >> lg;
      if (row > ot.ty0 && s_lo == ((rowbase - ot.tiles_x + ot.tx1) >> lg)) ++s_lo;
    }
    step(s_lo, s_hi >= s_lo ? s_hi - s_lo + 1 : 0, ot, oa, ob, oc, local0 + o);
  }
}

// Level 1: one workgroup per chunk.  Every (strip, Gaussian) entry gets its mask and is appended
// to the chunk's LDS list (one LDS atomic per wave and step for the list, per entry for the
// strip's entry and pair counters); one block scan of the packed counters (pairs in the high
// word, entries in the low) gives the chunk's offsets, written as its matrix row; the list is
// then copied into the chunk's region, grouped by strip.  A chunk with more entries than its
// list holds places them by a second pass; one with more than its region holds stores none
// (level 2 recomputes its masks).
__global__ __launch_bounds__(kStripL1Threads) void k_strip_level1(BinArgs a, StripArgs sa,
                                                                  int32_t* __restrict__ status) {
  extern __shared__ int32_t s_lds[];
  int32_t* s_ent = s_lds;                   // strips + 1: entry counts, offsets, then cursors
  int32_t* s_pairs = s_lds + sa.strips + 1;  // strips + 1: pair counts, then their prefix
  int2* s_list = reinterpret_cast<int2*>(s_lds + 2 * ((sa.strips + 2) & ~1));
  __shared__ int s_count;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int c = blockIdx.x;
  if (c == 0 && tid == 0) *status = SSPLAT_OK;
  for (int s = tid; s <= sa.strips; s += kStripL1Threads) {
    s_ent[s] = 0;
    s_pairs[s] = 0;
  }
  if (tid == 0) s_count = 0;
  __syncthreads();
  const int64_t g0 = int64_t(c) * sa.chunk;
  const int lg = sa.lg;
  for (int sub = 0; sub < sa.chunk; sub += kStripL1Threads) {
    strip_steps(a, lg, g0, sub,
                [&](int s_lo, int ne, const GaussTiles& ot, float oa, float ob, float oc, int id) {
                  const int ne_incl = wave_inclusive_scan(ne);
                  const int step_total = __builtin_amdgcn_readlane(ne_incl, 63);
                  int wbase = 0;
                  if (lane == 0 && step_total > 0) wbase = atomicAdd(&s_count, step_total);
                  wbase = __builtin_amdgcn_readfirstlane(wbase) + ne_incl - ne;
                  for (int i = 0; i < ne; ++i) {
                    const int s = s_lo + i;
                    const uint32_t m = strip_mask(a, ot, oa, ob, oc, s, lg);
                    atomicAdd(&s_ent[s], 1);
                    if (m != 0u) atomicAdd(&s_pairs[s], __popc(m));
                    if (wbase + i < sa.list_cap) {
                      const uint32_t key = (static_cast<uint32_t>(s) << kStripIdBits) | uint32_t(id);
                      s_list[wbase + i] = make_int2(static_cast<int>(key), static_cast<int>(m));
                    }
                  }
                });
  }
  __syncthreads();
  // Exclusive scans of the entry and pair counts in one: pairs << 32 | entries (the entries of a
  // chunk stay below 2^32, so the low word never carries).
  constexpr int kPer = kStripL1PerThread;
  const int b0 = tid * kPer;
  uint64_t x[kPer];
  uint64_t sum = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    x[k] = b0 + k < sa.strips
               ? (uint64_t(uint32_t(s_pairs[b0 + k])) << 32) | uint32_t(s_ent[b0 + k])
               : 0ull;
    sum += x[k];
  }
  int64_t total64;
  uint64_t run = static_cast<uint64_t>(
      block_exclusive_scan<kStripL1Threads>(static_cast<int64_t>(sum), &total64));
  const uint64_t total = static_cast<uint64_t>(total64);
  const int64_t entries = static_cast<int64_t>(total & 0xffffffffull);
  const int64_t pairs = static_cast<int64_t>(total >> 32);
  int2* row = sa.offs + int64_t(c) * (sa.strips + 1);
  auto clamp32 = [](uint64_t v) { return static_cast<int32_t>(v < INT32_MAX ? v : INT32_MAX); };
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (b0 + k < sa.strips) {
      const int32_t e = clamp32(run & 0xffffffffull);
      s_ent[b0 + k] = e;  // the strip's cursor in the region
      row[b0 + k] = make_int2(e, clamp32(run >> 32));
    }
    run += x[k];
  }
  if (tid == 0) row[sa.strips] = make_int2(clamp32(entries), clamp32(pairs));
  if (entries > sa.chunk_cap) return;  // level 2 recomputes this chunk's masks
  __syncthreads();
  int2* region = sa.bucket + int64_t(c) * sa.chunk_cap;
  if (entries <= sa.list_cap) {
    for (int i = tid; i < entries; i += kStripL1Threads) {
      const int2 e = s_list[i];
      const int s = static_cast<int>(static_cast<uint32_t>(e.x) >> kStripIdBits);
      const int g = static_cast<int>(g0 + (e.x & (kStripMaxChunk - 1)));
      region[atomicAdd(&s_ent[s], 1)] = make_int2(g, e.y);
    }
    return;
  }
  for (int sub = 0; sub < sa.chunk; sub += kStripL1Threads) {
    strip_steps(a, lg, g0, sub,
                [&](int s_lo, int ne, const GaussTiles& ot, float oa, float ob, float oc, int id) {
                  for (int i = 0; i < ne; ++i) {
                    const int s = s_lo + i;
                    const uint32_t m = strip_mask(a, ot, oa, ob, oc, s, lg);
                    region[atomicAdd(&s_ent[s], 1)] =
                        make_int2(static_cast<int>(g0 + id), static_cast<int>(m));
                  }
                });
  }
}

// Set bits of a wave-wide mask below this lane (v_mbcnt).
__device__ __forceinline__ int lanes_below(uint64_t b) {
  return static_cast<int>(__builtin_amdgcn_mbcnt_hi(
      static_cast<uint32_t>(b >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u)));
}

// OR of a value over the 64 lanes of a wave (all lanes active).
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  v |= __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
  v |= __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
  v |= __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v |= __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v |= __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v |= __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Sorts the list g[0, L) ascending by one wave in registers (L <= 1024); longer lists are left
// to the caller.
__device__ __forceinline__ void wave_sort_list(int32_t* g, int L, int lane) {
  if (L <= 1) return;
  if (L <= 64) {
    sort_tile_in_registers<1>(g, L, lane);
  } else if (L <= 128) {
    sort_tile_in_registers<2>(g, L, lane);
  } else if (L <= 256) {
    sort_tile_in_registers<4>(g, L, lane);
  } else if (L <= 512) {
    sort_tile_in_registers<8>(g, L, lane);
  } else if (L <= 1024) {
    sort_tile_in_registers<16>(g, L, lane);
  }
}

constexpr int kWaveSortMax = 1024;  // the longest list one wave sorts in registers

constexpr int kStripCountOnly = 0;  // level 2 writes the strip's offsets
constexpr int kStripBuild = 1;      // offsets and lists
constexpr int kStripFill = 2;       // lists, placed by the caller's offsets

// Sorts the list src[0, L) (L <= 64 kR) ascending by one wave in registers and writes the first
// n <= L values to dst.
template <int kR>
__device__ __forceinline__ void sort_list_to(const int32_t* src, int32_t* dst, int L, int n,
                                             int lane) {
  int32_t v[kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) v[r] = 64 * r + lane < L ? src[64 * r + lane] : INT32_MAX;
  wave_sort_regs<kR>(v, lane);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (64 * r + lane < n) dst[64 * r + lane] = v[r];
  }
}

// One wave: src[0, L) sorted into dst[0, n) (L <= kWaveSortMax; returns false for longer lists).
__device__ __forceinline__ bool wave_sort_to(const int32_t* src, int32_t* dst, int L, int n,
                                             int lane) {
  if (L <= 64) {
    sort_list_to<1>(src, dst, L, n, lane);
  } else if (L <= 128) {
    sort_list_to<2>(src, dst, L, n, lane);
  } else if (L <= 256) {
    sort_list_to<4>(src, dst, L, n, lane);
  } else if (L <= 512) {
    sort_list_to<8>(src, dst, L, n, lane);
  } else if (L <= kWaveSortMax) {
    sort_list_to<16>(src, dst, L, n, lane);
  } else {
    return false;
  }
  return true;
}

// Level 2: one workgroup per strip.  Its matrix column gives, per chunk, the strip's entries in
// the chunk's region (or, for an overflowed chunk, its Gaussians as candidates, whose masks are
// computed here) and the chunk's pairs before the strip, whose sum is the strip's first list
// position.  Candidates are in chunk order (ascending ids between chunks, any order inside one).
//
// The candidates pass through LDS kStripCache at a time, and each wave compacts, for each of its
// tiles, the candidates holding that tile into the tile's LDS segment (a ballot and a bit count
// per 64 candidates; the counts come out exact even when a segment is full); a wave scan of the
// counts gives the tiles' offsets; each tile's pairs are then sorted by one wave in registers
// straight into its list.  When a tile's pairs outgrow its segment, the pairs are placed per tile
// in LDS in any order (in the lists themselves when the strip has more than kStripPairsLds) and
// each list is sorted.  Fill mode places by the caller's offsets (clamped to gid_capacity) and
// flags *status when a tile holds more pairs than its range.  Eight waves a SIMD: every
// workgroup of a 2048x2048 grid at 16-pixel tiles is resident at once.
template <int kMode>
__global__ __launch_bounds__(kStripL2Threads) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_strip_level2(BinArgs a, StripArgs sa, int32_t* offsets, int32_t* gids,
                    int64_t gid_capacity, int32_t* __restrict__ ws_status,
                    int32_t* __restrict__ status) {
  constexpr int kWaves = kStripL2Threads / 64;
  __shared__ int s_cnt[32];
  __shared__ int s_cur[32];
  __shared__ int s_loc[32];   // the tile's first pair in the strip
  __shared__ int s_len[32];   // pairs written to its list
  __shared__ int64_t s_out[32];
  __shared__ int s_full;      // a tile's pairs outgrew its segment
  // Per chunk: the exclusive prefix of its candidates and the first entry of this strip in its
  // region (-1: overflowed chunk, candidates are its Gaussians).
  __shared__ int32_t s_tab[2 * kStripMaxChunks];
  __shared__ int2 s_cache[kStripCache];
  __shared__ int32_t s_buf[kStripPairsLds];  // the strip's pairs, per tile
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int S = 1 << sa.lg;
  const int seg = kStripPairsLds >> sa.lg;  // a tile's LDS segment in the compaction path
  const int s = xcd_remap(blockIdx.x, gridDim.x);
  const int s0 = s << sa.lg;
  if (tid < 32) {
    s_cnt[tid] = 0;
    s_cur[tid] = 0;
  }
  if (tid == 0) s_full = 0;
  // The level-1 status is final here: merge it into the caller's sticky status word.
  if (kMode != kStripCountOnly && status != nullptr && blockIdx.x == 0 && tid == 0 &&
      *ws_status != SSPLAT_OK) {
    *status = *ws_status;
  }

  // Candidates per chunk and their exclusive prefix; the pairs before the strip (2 chunks a
  // thread); the last strip also sums every chunk's pairs.
  constexpr int kPer = 2;
  const bool last = s == sa.strips - 1;
  int nc[kPer];
  int64_t sum = 0, before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int c = tid * kPer + k;
    nc[k] = 0;
    if (c < sa.chunks) {
      const int2* row = sa.offs + int64_t(c) * (sa.strips + 1);
      const int2 lo = row[s], hi = row[s + 1], total = row[sa.strips];
      if (total.x > sa.chunk_cap) {
        const int64_t left = int64_t(a.n) - int64_t(c) * sa.chunk;
        nc[k] = static_cast<int>(left < sa.chunk ? left : sa.chunk);
        s_tab[2 * c + 1] = -1;
      } else {
        nc[k] = hi.x - lo.x;
        s_tab[2 * c + 1] = lo.x;
      }
      before += lo.y;
      all += total.y;
    }
    sum += nc[k];
  }
  int64_t count, first, total_pairs = 0;
  int64_t run = block_exclusive_scan<kStripL2Threads>(sum, &count);  // ends with a barrier
  block_exclusive_scan<kStripL2Threads>(before, &first);
  if (last) block_exclusive_scan<kStripL2Threads>(all, &total_pairs);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int c = tid * kPer + k;
    if (c < sa.chunks) s_tab[2 * c] = static_cast<int32_t>(run);
    run += nc[k];
  }
  __syncthreads();

  // Candidate j: its chunk by binary search over the prefixes, then its entry (or Gaussian).
  auto candidate = [&](int64_t j, int* g, uint32_t* m) {
    *g = -1;
    *m = 0u;
    if (j >= count) return;
    int lo = 0, hi = sa.chunks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_tab[2 * mid] <= j) {
        lo = mid;
      } else {
        hi = mid - 1;
      }
    }
    const int i = static_cast<int>(j - s_tab[2 * lo]);
    const int entry0 = s_tab[2 * lo + 1];
    if (entry0 < 0) {
      const int id = static_cast<int>(int64_t(lo) * sa.chunk + i);  // below n
      *g = id;
      *m = entry_mask(a, id, s, sa.lg);
    } else {
      const int2 e = sa.bucket[int64_t(lo) * sa.chunk_cap + entry0 + i];
      if (e.x >= 0 && e.x < a.n) {
        *g = e.x;
        *m = static_cast<uint32_t>(e.y);
      }
    }
  };

  // Pairs per tile: the candidates pass through the LDS cache kStripCache at a time, and each
  // wave compacts, for each of its tiles, the cached candidates holding that tile into the
  // tile's LDS segment after the earlier rounds' (a ballot and a bit count per 64 candidates;
  // the counts come out exact even when a segment is full).  `cached`: the cache ends up holding
  // every candidate.
  const bool cached = count <= kStripCache;
  for (int64_t r0 = 0; r0 < count; r0 += kStripCache) {
    const int m = static_cast<int>(count - r0 < kStripCache ? count - r0 : kStripCache);
    if (r0 > 0) __syncthreads();  // the previous round's compaction is done with the cache
    for (int i = tid; i < m; i += kStripL2Threads) {
      int g;
      uint32_t mk;
      candidate(r0 + i, &g, &mk);
      s_cache[i] = make_int2(g, static_cast<int>(mk));
    }
    __syncthreads();
    // Compaction: a wave per tile, candidates in order (64 a step).  Tile k's count is read and
    // written by its wave only (lane 0 writes it, the wave's later LDS reads are ordered after).
    const int rounds = (m + 63) >> 6;
    for (int k = wave; k < S; k += kWaves) {
      int32_t* dst = s_buf + k * seg;
      int n = s_cnt[k];
      for (int r = 0; r < rounds; ++r) {
        const int j = (r << 6) + lane;
        const int2 e = j < m ? s_cache[j] : make_int2(-1, 0);
        const bool hit = (static_cast<uint32_t>(e.y) >> k) & 1u;
        const uint64_t b = __ballot(hit);
        const int pos = n + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                static_cast<uint32_t>(b >> 32),
                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u)));
        if (hit && pos < seg) dst[pos] = e.x;
        n += static_cast<int>(__popcll(b));
      }
      if (lane == 0) {
        s_cnt[k] = n;
        // A tile beyond its segment, or beyond what one wave sorts: the general path below.
        if (n > seg || n > kWaveSortMax) s_full = 1;
      }
    }
  }
  __syncthreads();

  // Offsets: a wave scan of the per-tile counts from the strip's first pair.
  if (wave == 0) {
    const int c = lane < S ? s_cnt[lane] : 0;
    const int incl = wave_inclusive_scan(c);
    const int excl = incl - c;
    if (lane < 32) s_loc[lane] = excl;
    if (kMode != kStripFill && last && lane == 0) {
      offsets[sa.num_tiles] =
          static_cast<int32_t>(total_pairs > INT32_MAX ? INT32_MAX : total_pairs);
      if (total_pairs > INT32_MAX || (kMode == kStripBuild && total_pairs > gid_capacity)) {
        *ws_status = SSPLAT_ERR_CAPACITY;
        if (status != nullptr) *status = SSPLAT_ERR_CAPACITY;
      }
    }
    if (lane < S) {
      const int tile = s0 + lane;
      if (kMode == kStripFill) {
        int64_t lo = 0, hi = 0;
        if (tile < sa.num_tiles) {
          lo = offsets[tile];
          hi = offsets[tile + 1];
        }
        const int64_t cap = gid_capacity;
        const int64_t b = lo < 0 ? 0 : (lo > cap ? cap : lo);
        const int64_t end = hi < b ? b : (hi > cap ? cap : hi);
        s_out[lane] = b;
        s_len[lane] = static_cast<int>(c < end - b ? c : end - b);
        if (c > end - b && status != nullptr) *status = SSPLAT_ERR_CAPACITY;
      } else {
        const int64_t pos = first + excl;
        if (tile < sa.num_tiles) {
          offsets[tile] = static_cast<int32_t>(pos > INT32_MAX ? INT32_MAX : pos);
        }
        const int64_t cap = kMode == kStripBuild ? gid_capacity : 0;
        s_out[lane] = pos;
        s_len[lane] = static_cast<int>(pos >= cap ? 0 : (c < cap - pos ? c : cap - pos));
      }
    }
  }
  if (kMode == kStripCountOnly) return;
  __syncthreads();

  if (s_full == 0) {
    // Each tile's compacted pairs sorted by one wave into its list.
    for (int k = wave; k < S; k += kWaves) {
      wave_sort_to(s_buf + k * seg, gids + s_out[k], s_cnt[k], s_len[k], lane);
    }
    return;
  }

  // The strip's pairs fit the LDS buffer: place there, sort into the lists.  Otherwise place
  // straight into the lists (up to each list's length) and sort them in place.
  const bool in_lds = s_loc[S - 1] + s_cnt[S - 1] <= kStripPairsLds;
  // Placement: ranks per tile by ballots, one LDS atomic per wave and tile.
  for (int64_t base = 0; base < count; base += kStripL2Threads) {
    const int64_t j = base + tid;
    int g;
    uint32_t m;
    if (cached) {
      const int2 e = j < count ? s_cache[j] : make_int2(-1, 0);
      g = e.x;
      m = static_cast<uint32_t>(e.y);
    } else {
      candidate(j, &g, &m);
    }
    uint32_t wm = wave_or(m);
    while (wm != 0u) {
      const int k = __ffs(wm) - 1;
      wm &= wm - 1;
      const uint64_t b = __ballot((m >> k) & 1u);
      const int leader = __ffsll(static_cast<unsigned long long>(b)) - 1;
      int r0 = 0;
      if (lane == leader) r0 = atomicAdd(&s_cur[k], static_cast<int>(__popcll(b)));
      r0 = __builtin_amdgcn_readlane(r0, leader);
      if ((m >> k) & 1u) {
        const int rank =
            r0 + static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(
                                                                static_cast<uint32_t>(b), 0u)));
        if (in_lds) {
          s_buf[s_loc[k] + rank] = g;
        } else if (rank < s_len[k]) {
          gids[s_out[k] + rank] = g;
        }
      }
    }
  }
  // The barrier's workgroup-scope fences order the LDS and global stores before the reads of
  // other waves of this workgroup (a device-scope fence would write back this XCD's whole L2,
  // per workgroup).
  __syncthreads();

  // Sort every tile's pairs into its list: a wave per tile in registers, the block for the rare
  // lists above 1024 entries (in LDS up to kSortLds, else in global memory).
  for (int k = wave; k < S; k += kWaves) {
    if (in_lds) {
      wave_sort_to(s_buf + s_loc[k], gids + s_out[k], s_cnt[k], s_len[k], lane);
    } else {
      wave_sort_list(gids + s_out[k], s_len[k], lane);
    }
  }
  for (int k = 0; k < S; ++k) {
    const int L = in_lds ? s_cnt[k] : s_len[k];
    if (L <= 1024) continue;
    __syncthreads();
    int32_t* g = gids + s_out[k];
    if (in_lds) {
      bitonic_sort<false>(s_buf + s_loc[k], L);
      const int n = s_len[k];
      for (int i = tid; i < n; i += kStripL2Threads) g[i] = s_buf[s_loc[k] + i];
    } else if (L <= kSortLds) {
      for (int i = tid; i < L; i += kStripL2Threads) s_buf[i] = g[i];
      __syncthreads();
      bitonic_sort<false>(s_buf, L);
      for (int i = tid; i < L; i += kStripL2Threads) g[i] = s_buf[i];
    } else {
      bitonic_sort<true>(g, L);
    }
  }
}

// ------------------------------------------------------------------------------------------
// Rendering
// ------------------------------------------------------------------------------------------

// One packed Gaussian: four 16-byte chunks.
//   c0 = (mx, my, qa, qb)       conic in the exponent's units: the visit computes
//   c1 = (qc, xr, yr, gid)        e = qa dx^2 + qb dx dy + qc dy^2 and weights with exp2(e)
//   c2 = (o r, o g, o b, o)     opacity-premultiplied color and opacity (accumulation)
//   c3 = (r, g, b, slot)        plain color (gradient sweep); first slot relative to the wave
// Without support fade the record holds (qa, qb, qc) = -log2(e)/2 * (a, 2b, c) (the fold saves
// a multiply per visit); with fade it holds (a, 2b, c) and the kernels scale the quadratic so
// that a pixel at exactly sigma_cutoff evaluates the same 2^x as the fade tail.
// xr = x0 | x1 << 16 and yr = y0 | y1 << 16 hold the half-open clipped support rectangle.
struct __align__(16) GaussRec {
  float4 c0, c1, c2, c3;
};
static_assert(sizeof(GaussRec) == 64, "GaussRec must be 64 bytes");

// Per-wave slot bookkeeping written by the record packing (forward and backward alike):
// rect_wave[w] = slots owned by the 64 Gaussians of wave w (the scan input; rect_wave[num_waves]
// = 0), each record's c3.w = the Gaussian's first slot relative to its wave.
struct SlotPrep {
  int32_t* rect_wave;
  int num_waves;
  int lg;
  int32_t* rect_status;
};

template <bool kFold>
__global__ __launch_bounds__(256) void k_prepare(
    const float* __restrict__ means, const float* __restrict__ conics,
    const float* __restrict__ colors, const int64_t* __restrict__ radii,
    const float* __restrict__ opacities, int n, FieldTable fields, GaussRec* __restrict__ recs,
    SlotPrep sp, int32_t* __restrict__ queue) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  if (i == 0) {
    sp.rect_wave[sp.num_waves] = 0;
    *sp.rect_status = SSPLAT_OK;
  }
  // The forward's persistent work-queue counters, zeroed for it.
  if (i < kQueues) queue[i] = 0;
  // Every input is loaded up front (one memory round trip per thread).
  const bool live = i < n;
  const int j = live ? i : 0;
  const float mx = live ? means[2 * j + 0] : 0.0f;
  const float my = live ? means[2 * j + 1] : 0.0f;
  const int64_t rx = live ? radii[2 * j + 0] : 0;
  const int64_t ry = live ? radii[2 * j + 1] : 0;
  const float qa = live ? conics[3 * j + 0] : 0.0f;
  const float qb = live ? conics[3 * j + 1] : 0.0f;
  const float qc = live ? conics[3 * j + 2] : 0.0f;
  const float cr = live ? colors[3 * j + 0] : 0.0f;
  const float cg = live ? colors[3 * j + 1] : 0.0f;
  const float cb = live ? colors[3 * j + 2] : 0.0f;
  const float op = (live && opacities) ? opacities[j] : 1.0f;
  uint32_t xr = 0, yr = 0;
  if (live) {
    const ssplat_field f = field_of_gaussian(fields, i);
    const Rect r = support_rect(mx, my, rx, ry, f.height, f.width);
    if (r.x1 > r.x0 && r.y1 > r.y0) {
      xr = uint32_t(r.x0) | (uint32_t(r.x1) << 16);
      yr = uint32_t(r.y0) | (uint32_t(r.y1) << 16);
    }
  }
  // Slots of this wave's Gaussians: inclusive prefix of R over the 64 lanes.
  const int slots = rect_tiles(xr, yr, sp.lg).count;
  const int incl = wave_inclusive_scan(slots);
  if (lane == 63 && (i >> 6) < sp.num_waves) sp.rect_wave[i >> 6] = incl;
  if (!live) return;
  const float scale = kFold ? kNegHalfLog2e : 1.0f;
  GaussRec g;
  g.c0 = make_float4(mx, my, scale * qa, scale * (2.0f * qb));
  g.c1 = make_float4(scale * qc, __uint_as_float(xr), __uint_as_float(yr), __int_as_float(i));
  g.c2 = make_float4(op * cr, op * cg, op * cb, op);
  g.c3 = make_float4(cr, cg, cb, __int_as_float(incl - slots));
  recs[i] = g;
}

// Affine colors (render.py:110-123): per Gaussian three 16-byte chunks
//   a0 = (cos t / sx, sin t / sx, cos t / sy, sin t / sy)   local x = a0.x dx + a0.y dy,
//                                                          local y = a0.z dy - a0.w dx
//   a1 = (G00, G01, G02, G10), a2 = (G11, G12, 1/sx, 1/sy)  color_k = c_k + G0k x + G1k y
// with sx, sy clamped below at 1e-6 like the reference; a2.z (a2.w) is stored negated where
// the clamp holds, whose scale then gets no gradient.
constexpr int kAffineChunks = 3;

// The thread's index in its workgroup; with kOpaque, opaque to the optimizer.  In the persistent
// kernels' tile loop the compiler otherwise hoists what derives from the thread index (lane, wave,
// pixel offsets, LDS addresses) out of the loop and holds it in registers across every tile: the
// persistent backward then needs 101 VGPRs instead of 59.  Opaque, those values are recomputed
// where used.  The grid kernels run one tile and gain nothing from the recomputation.
template <bool kOpaque>
__device__ __forceinline__ int thread_index() {
  int v = static_cast<int>(threadIdx.x);
  if constexpr (kOpaque) asm volatile("" : "+v"(v));
  return v;
}

struct RenderArgs {
  const GaussRec* recs;
  const float4* arecs;      // affine records (kAffineChunks per Gaussian), or NULL
  const int32_t* rect_off;  // first slot of each wave of 64 Gaussians (backward only)
  const int32_t* gids;
  const int32_t* offsets;
  int64_t num_gids;
  int n;  // number of records; list entries outside [0, n) stage as empty records
  FieldTable fields;
  float eps;
  float tail_exp;    // sigma_cutoff^2 * kNegHalfLog2e
  float fade_alpha;  // support-fade strength in [0, 1]; the faded tail is alpha * exp(-cutoff^2/2)
};

// Clamped [start, end) of a tile's list.
__device__ __forceinline__ void tile_range(const RenderArgs& a, int tile, int64_t* start,
                                           int64_t* end) {
  int64_t s = a.offsets[tile];
  int64_t e = a.offsets[tile + 1];
  s = s < 0 ? 0 : (s > a.num_gids ? a.num_gids : s);
  e = e < s ? s : (e > a.num_gids ? a.num_gids : e);
  *start = s;
  *end = e;
}

// The first kChunks 16-byte chunks of up to kStage records: the forward stages chunks 0-2, the
// backward all four.
template <int kChunks, int kS = kStage>
struct StageLds {
  static constexpr int kSize = kS;
  float4 c[kChunks][kS];
};

// The tile a gradient sweep stages records for, to find each pair's slot.
struct SlotTile {
  int tile_x, tile_y;  // the tile's position in its field's tile grid
  int lg;              // log2 of the tile edge
  int64_t capacity;    // slots the buffer holds
};

// Coalesced staging of up to kStage records of the tile list into LDS (kChunks 16-byte loads
// per thread).  An id outside [0, n) (a list slot a capacity-limited binning left unwritten)
// stages an empty rectangle, which no pixel and no quadrant test ever selects.
// kAffine: the affine chunks follow in st.c[4 .. 4 + kAffineChunks).  kSlot: the gradient
// sweep's records carry in c3.w the pair's slot, the Gaussian's first slot plus this tile's
// row-major index in its rectangle's tiles, or -1 where the pair has none (a list entry outside
// the rectangle, a slot beyond the capacity, an empty record).
template <int kLoad, int kChunks, int kS, bool kAffine = false, bool kSlot = false,
          bool kOpaque = false>
__device__ __forceinline__ void stage_batch(const RenderArgs& a, StageLds<kChunks, kS>& st,
                                            int64_t batch, int m, SlotTile stile = {}) {
  static_assert(kLoad <= kChunks, "staging more chunks than the buffer holds");
  static_assert(!kAffine || (kLoad == 4 && kChunks == 4 + kAffineChunks), "affine layout");
  const int t = thread_index<kOpaque>();
  if (t < m) {
    const int g = a.gids[batch + t];
    if (g >= 0 && g < a.n) {
      const float4* src = reinterpret_cast<const float4*>(a.recs + g);
      float4 v[kLoad];
#pragma unroll
      for (int c = 0; c < kLoad; ++c) v[c] = src[c];
      if constexpr (kSlot) {
        const RectTiles rt = rect_tiles(__float_as_uint(v[1].y), __float_as_uint(v[1].z), stile.lg);
        const int kx = stile.tile_x - rt.tx0, ky = stile.tile_y - rt.ty0;
        int slot = -1;
        if (kx >= 0 && kx < rt.tw && ky >= 0 && ky < rt.th) {
          const int64_t s64 = int64_t(__float_as_int(v[3].w)) + a.rect_off[g >> 6] +
                              ky * rt.tw + kx;
          if (s64 < stile.capacity) slot = static_cast<int>(s64);
        }
        v[3].w = __int_as_float(slot);
      }
#pragma unroll
      for (int c = 0; c < kLoad; ++c) st.c[c][t] = v[c];
      if constexpr (kAffine) {
#pragma unroll
        for (int c = 0; c < kAffineChunks; ++c) {
          st.c[4 + c][t] = a.arecs[int64_t(g) * kAffineChunks + c];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < kChunks; ++c) st.c[c][t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if constexpr (kSlot) st.c[3][t].w = __int_as_float(-1);
    }
  }
}

// Requires each value to be in a VGPR at this point of the program (an empty asm statement that
// reads them); it emits no instruction of its own.
template <typename T>
__device__ __forceinline__ int in_register(T v) {
  asm volatile("" ::"v"(v));
  return 0;
}
template <typename... T>
__device__ __forceinline__ void in_registers(T... v) {
  const int all[] = {in_register(v)...};
  (void)all;
}

// Pixels of the 8x8 quadrant at (qx0, qy0) inside the half-open rectangle packed in (xr, yr), as
// a 64-bit mask with bit 8*ly + lx for the pixel (qx0 + lx, qy0 + ly).
__device__ __forceinline__ uint64_t quadrant_mask(uint32_t xr, uint32_t yr, int qx0, int qy0) {
  const int lx = max(static_cast<int>(xr & 0xffff) - qx0, 0);
  const int hx = min(static_cast<int>(xr >> 16) - qx0, 8);
  const int ly = max(static_cast<int>(yr & 0xffff) - qy0, 0);
  const int hy = min(static_cast<int>(yr >> 16) - qy0, 8);
  if (hx <= lx || hy <= ly) return 0;
  const uint32_t row = ((1u << hx) - (1u << lx)) * 0x01010101u;  // the column run in every row
  const uint64_t rows = (uint64_t(row) << 32) | row;
  const uint64_t below_hy = hy == 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * hy)) - 1;
  const uint64_t below_ly = (uint64_t(1) << (8 * ly)) - 1;
  return rows & below_hy & ~below_ly;
}

// Accumulates numerator (r,g,b) and denominator of one pixel per lane over the whole tile list,
// in ascending Gaussian id.  Each wave owns the 8x8 quadrant (qx0, qy0).  For every staged record
// the wave first computes (one record per lane) the mask of quadrant pixels inside its rectangle
// into its LDS slice; a ballot keeps the records that meet the quadrant, and each lane reads its
// bit of their masks.
// kDenOnly: only the denominator D (acc.w) is accumulated (the backward's normalizer
// recomputation); it is the same sum, in the same order, as the forward's D.  staged: the
// caller has staged the tile's only batch already (with a barrier after it).
template <bool kFade, bool kDenOnly, int kChunks, int kS, bool kAffine = false, bool kOpaque = false>
__device__ __forceinline__ void accumulate_quadrant(const RenderArgs& a, StageLds<kChunks, kS>& st,
                                                    uint32_t* s_mask, int tile, int qx0, int qy0,
                                                    int px, int py, float tail, float4& acc,
                                                    bool staged = false) {
  const int lane = thread_index<kOpaque>() & 63;
  const float fx = static_cast<float>(px);
  const float fy = static_cast<float>(py);
  int64_t start, end;
  tile_range(a, tile, &start, &end);
  float nr = 0.0f, ng = 0.0f, nb = 0.0f, d = 0.0f;
  for (int64_t batch = start; batch < end; batch += kS) {
    const int m = static_cast<int>(end - batch < kS ? end - batch : kS);
    if (!staged) {
      if constexpr (kAffine && !kDenOnly) {
        stage_batch<4, kChunks, kS, true, false, kOpaque>(a, st, batch, m);
      } else {
        stage_batch<3, kChunks, kS, false, false, kOpaque>(a, st, batch, m);
      }
      __syncthreads();
    }
    uint64_t hit[kS / 64];
#pragma unroll
    for (int k = 0; k < kS / 64; ++k) {
      const int j = lane + 64 * k;
      uint64_t qm = 0;
      if (j < m) {
        const float4 v = st.c[1][j];
        qm = quadrant_mask(__float_as_uint(v.y), __float_as_uint(v.z), qx0, qy0);
      }
      s_mask[2 * j] = static_cast<uint32_t>(qm);  // pixels of lanes 0-31
      s_mask[2 * j + 1] = static_cast<uint32_t>(qm >> 32);  // lanes 32-63
      hit[k] = __ballot(qm != 0);
    }
    wave_lds_sync();  // the masks are read back by every lane of this wave
    const uint32_t lane_bit_index = lane & 31;
    const uint32_t* my_half = s_mask + (lane >> 5);
    // One record's contribution to this lane's pixel, if the pixel is inside its rectangle.
    auto visit = [&](int j, bool inside, const float4& v0, const float4& v1, const float4& v2) {
      if (inside) {
        const float dx = fx - v0.x;
        const float dy = fy - v0.y;
        const float t = fmaf(v0.z, dx, v0.w * dy);
        const float e = fmaf(t, dx, v1.x * dy * dy);
        const float raw = exp2_hw(kFade ? e * kNegHalfLog2e : e);
        const float base = kFade ? fmaxf(raw - tail, 0.0f) : raw;
        if constexpr (kAffine && !kDenOnly) {
          // w (c_k + G0k x + G1k y) with the pixel's local coordinates x, y.
          const float4 v3 = st.c[3][j];
          const float4 a0 = st.c[4][j];
          const float4 a1 = st.c[5][j];
          const float4 a2 = st.c[6][j];
          const float lxc = fmaf(a0.x, dx, a0.y * dy);
          const float lyc = fmaf(a0.z, dy, -(a0.w * dx));
          const float w = base * v2.w;
          nr = fmaf(w, fmaf(a1.x, lxc, fmaf(a1.w, lyc, v3.x)), nr);
          ng = fmaf(w, fmaf(a1.y, lxc, fmaf(a2.x, lyc, v3.y)), ng);
          nb = fmaf(w, fmaf(a1.z, lxc, fmaf(a2.y, lyc, v3.z)), nb);
        } else if constexpr (!kDenOnly) {
          nr = fmaf(base, v2.x, nr);
          ng = fmaf(base, v2.y, ng);
          nb = fmaf(base, v2.z, nb);
        }
        d = fmaf(base, v2.w, d);
      }
    };
#pragma unroll
    for (int k = 0; k < kS / 64; ++k) {
      uint64_t mask = hit[k];
      // Two records per step: both are read from LDS before the first is evaluated, so the
      // second's reads overlap the first's arithmetic.  They are added in ascending order.
      while (mask) {
        const int b1 = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        const bool two = mask != 0;
        const int b2 = two ? __ffsll(static_cast<unsigned long long>(mask)) - 1 : b1;
        mask &= mask - 1;
        const int j1 = 64 * k + b1, j2 = 64 * k + b2;
        const float4 p0 = st.c[0][j1], p1 = st.c[1][j1], p2 = st.c[2][j1];
        const float4 q0 = st.c[0][j2], q1 = st.c[1][j2], q2 = st.c[2][j2];
        // The lane's bit of each record's quadrant mask (all ones for a covered quadrant),
        // read with the records so one wait covers every read.
        const uint32_t m1 = my_half[2 * j1], m2 = my_half[2 * j2];
        // Every value both visits read is in registers here (the compiler would otherwise
        // sink each record's reads into its masked visit and wait once per record).
        in_registers(p0.x, p0.y, p0.z, p0.w, p1.x, p2.w, m1, q0.x, q0.y, q0.z, q0.w, q1.x, q2.w,
                     m2);
        if constexpr (!kDenOnly) in_registers(p2.x, p2.y, p2.z, q2.x, q2.y, q2.z);
        const bool in1 = __builtin_amdgcn_ubfe(m1, lane_bit_index, 1) != 0;
        const bool in2 = __builtin_amdgcn_ubfe(m2, lane_bit_index, 1) != 0;
        visit(j1, in1, p0, p1, p2);
        if (two) visit(j2, in2, q0, q1, q2);
      }
    }
    __syncthreads();
  }
  acc = make_float4(nr, ng, nb, d);
}

// Work queue of a persistent render launch: the tiles split into `queues` contiguous ranges,
// queues / kXcds consecutive ranges per XCD, each with an atomic counter in the workspace that the
// launch's preceding kernel zeroes; a claim takes one tile.  A workgroup claims from its home
// range (its XCD from the round-robin dispatch, a range of that XCD from its index) and, once that
// is exhausted, from the next range with tiles left.  About kWorkgroupsPerQueue
// workgroups share a range: they work on neighbouring tiles at any moment, which share most of
// their records (one XCD's L2 serves them), while the claims on any one counter (atomics on one
// address serialize) stay few.  Stealing balances the XCDs, whose ranges can differ in work
// (a batch's fields are concatenated by size).
struct TileQueue {
  int32_t* counters;  // kQueues counters, zero at launch
  int num_tiles;
  int queues;         // ranges in use: kXcds times 1 .. kMaxQueuesPerXcd
};

__device__ __forceinline__ int queue_bound(const TileQueue& q, int r) {
  return static_cast<int>(int64_t(q.num_tiles) * r / q.queues);
}

// Thread 0's claim state in a persistent workgroup: the range it drains, the tile of its current
// claim, and the counter value of the following claim in that range, fetched when the current
// one is taken so its round trip overlaps the tile's work.
struct TileClaim {
  int range;
  int pending;
};

__device__ __forceinline__ TileClaim first_claim(const TileQueue& q) {
  const int b = static_cast<int>(blockIdx.x);
  const int per_xcd = q.queues / kXcds;
  TileClaim c;
  c.range = (b & (kXcds - 1)) * per_xcd + (b / kXcds) % per_xcd;
  c.pending = atomicAdd(&q.counters[c.range], 1);
  return c;
}

// Tile of claim value v in range r, or -1 past the range.
__device__ __forceinline__ int claimed_tile(const TileQueue& q, int r, int v) {
  const int lo = queue_bound(q, r), hi = queue_bound(q, r + 1);
  return v < hi - lo ? lo + v : -1;
}

// The next tile of the workgroup, broadcast through LDS (-1 once every range is exhausted).
// Thread 0 takes its pending claim.  When its range is exhausted, the first wave reads every
// range's counter at once (device-scope atomic loads: no cache serves a stale value, and no
// read-modify-write competes with the claims) and thread 0 claims
// in the next range (in range order) with tiles left, repeating only if other workgroups took
// them first.
// Starts with a barrier, so the previous tile's threads are done with LDS and the broadcast word.
__device__ __forceinline__ int claim_tile(const TileQueue& q, TileClaim& c) {
  __shared__ int s_tile;
  __syncthreads();
  if (thread_index<true>() == 0) {
    const int t = claimed_tile(q, c.range, c.pending);
    if (t >= 0) c.pending = atomicAdd(&q.counters[c.range], 1);  // the claim after this one
    s_tile = t;
  }
  __syncthreads();
  // Every thread reads the same broadcast word: readfirstlane tells the compiler the tile is
  // uniform, so what derives from it (origin, field, list bounds) stays in scalar registers.
  int tile = __builtin_amdgcn_readfirstlane(s_tile);
  while (tile < 0) {
    if (thread_index<true>() < 64) {
      const int lane = thread_index<true>();
      int left = 0;
      if (lane < q.queues) {
        const int size = queue_bound(q, lane + 1) - queue_bound(q, lane);
        const int taken =
            __hip_atomic_load(&q.counters[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        left = taken < size ? size - taken : 0;
      }
      // The first range with tiles left after the current one, in range order (workgroups
      // that run dry at different ranges spread over different victims), by a wave minimum.
      const int cur = __shfl(c.range, 0, 64);  // thread 0's range (the claim state is its own)
      const int dist = (lane - cur + q.queues) % q.queues;
      int best = left > 0 ? (dist << 6) | lane : INT32_MAX;
      for (int d = 32; d > 0; d >>= 1) best = min(best, __shfl_xor(best, d, 64));
      if (lane == 0) {
        int t = -2;  // -2: every range exhausted
        if (best != INT32_MAX) {
          const int r = best & 63;
          t = claimed_tile(q, r, atomicAdd(&q.counters[r], 1));  // -1: raced, probe again
          if (t >= 0) {
            c.range = r;
            c.pending = atomicAdd(&q.counters[r], 1);
          }
        }
        s_tile = t;
      }
    }
    __syncthreads();
    tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();  // every thread has read the word before it can change again
    if (tile == -2) return -1;
  }
  return tile;
}

// One tile of the forward: the body of k_forward and of its persistent form.
template <int kT, bool kFade, bool kNormalize, bool kAffine, bool kOpaque = false>
__device__ __forceinline__ void forward_tile(const RenderArgs& a, int tile, float* __restrict__ out,
                                             float* __restrict__ den) {
  constexpr int kQ = kT / 8;  // quadrants per tile row
  constexpr int kS = kT == 16 ? kFwdStage16 : kStage;  // records per staged batch
  __shared__ StageLds<kAffine ? 4 + kAffineChunks : 3, kS> st;
  __shared__ uint32_t s_mask[kQ * kQ][2 * kS];  // per-wave quadrant masks
  const ssplat_field f = field_of_tile(a.fields, tile);
  const int local = tile - static_cast<int>(f.first_tile);
  const int tile_x = local % f.tiles_x;
  const int tile_y = local / f.tiles_x;
  const int wave = thread_index<kOpaque>() >> 6;
  const int lane = thread_index<kOpaque>() & 63;
  const int qx0 = tile_x * kT + (wave % kQ) * 8;
  const int qy0 = tile_y * kT + (wave / kQ) * 8;
  const int px = qx0 + (lane & 7);
  const int py = qy0 + (lane >> 3);
  const float tail = kFade ? a.fade_alpha * exp2_hw(a.tail_exp) : 0.0f;
  float4 acc;
  accumulate_quadrant<kFade, false, kAffine ? 4 + kAffineChunks : 3, kS, kAffine, kOpaque>(
      a, st, s_mask[wave], tile, qx0, qy0, px, py, tail, acc);
  if (px < f.width && py < f.height) {
    const int64_t flat = f.first_pixel + int64_t(py) * f.width + px;
    if (den) den[flat] = acc.w;
    if (kNormalize) {
      const float denom = acc.w + a.eps;
      out[flat * 3 + 0] = acc.x / denom;
      out[flat * 3 + 1] = acc.y / denom;
      out[flat * 3 + 2] = acc.z / denom;
    } else {
      out[flat * 3 + 0] = acc.x;
      out[flat * 3 + 1] = acc.y;
      out[flat * 3 + 2] = acc.z;
    }
  }
}

// Grid launch: one workgroup per tile, XCD-aware tile order.
template <int kT, bool kFade, bool kNormalize, bool kAffine = false>
__global__ __launch_bounds__(kT * kT) void k_forward(RenderArgs a, float* __restrict__ out,
                                                     float* __restrict__ den) {
  forward_tile<kT, kFade, kNormalize, kAffine>(a, xcd_remap(blockIdx.x, gridDim.x), out, den);
}

// Persistent launch: as many workgroups as fit on the device, each taking tiles from the queue.
// With 32-pixel tiles the kernel is held to eight waves per SIMD (two workgroups per CU); with
// 16-pixel tiles the LDS of seven workgroups fills the CU, so seven waves per SIMD is the most.
template <int kT, bool kFade, bool kNormalize>
__global__ __launch_bounds__(kT * kT) __attribute__((amdgpu_waves_per_eu(kT == 32 ? 8 : 7, 8)))
void k_forward_persistent(RenderArgs a, TileQueue q,
                                                                float* __restrict__ out,
                                                                float* __restrict__ den) {
  TileClaim c{};
  if (threadIdx.x == 0) c = first_claim(q);
  for (int tile = claim_tile(q, c); tile >= 0; tile = claim_tile(q, c)) {
    forward_tile<kT, kFade, kNormalize, false, true>(a, tile, out, den);
  }
}

struct GradArgs {
  const float* grad_out;
  const float* den;  // forward coverage (read when normalizing)
  const float* out;  // forward image (read when normalizing)
  float* slots;      // kSlotWords floats per slot
  int64_t slot_capacity;
  const uint32_t* epoch;  // the call's tag
  bool has_opacity;
};

// Where the normalized backward gets the forward's D (den) and out:
constexpr int kNormSaved = 0;      // both read from the forward's saved outputs
constexpr int kNormRecompD = 1;    // out read, D recomputed by a denominator-only sweep
constexpr int kNormRecompAll = 2;  // N and D recomputed by the forward's sweep, out = N/(D+eps)

// One tile of the backward: the body of k_backward and of its persistent form.
template <int kT, bool kFade, bool kNormalize, int kNorm, bool kAffine, bool kOpaque = false>
__device__ __forceinline__ void backward_tile(const RenderArgs& a, const GradArgs& g, int tile) {
  constexpr bool kRecompute = kNormalize && kNorm != kNormSaved;
  constexpr int kQ = kT / 8;  // quadrants per tile row
  constexpr int kWaves = kQ * kQ;
  constexpr int kS = kT == 16 ? kBwdStage16 : kStage;  // records per staged batch
  constexpr int kMaskF4 = kWaves * 2 * kS / 4;          // the sweep's masks, in float4 units
  constexpr int kChunks = kAffine ? 4 + kAffineChunks : 4;
  constexpr int kWords = kAffine ? kSlotWordsAffine : kSlotWords;
  __shared__ StageLds<kChunks, kS> st;
  // The per-pixel adjoints; when recomputing, the same space first holds the sweep's per-wave
  // quadrant masks (two uint32 per staged record and wave).
  __shared__ float4 s_scratch[kRecompute && kMaskF4 > kT * kT ? kMaskF4 : kT * kT];
  float4* s_adj = s_scratch;
  const ssplat_field f = field_of_tile(a.fields, tile);
  const int local = tile - static_cast<int>(f.first_tile);
  const int tile_x = local % f.tiles_x;
  const int tile_y = local / f.tiles_x;
  const int wave = thread_index<kOpaque>() >> 6;
  const int lane = thread_index<kOpaque>() & 63;
  const int lx = lane & 7;
  const int ly = lane >> 3;
  const int tx0 = tile_x * kT;
  const int ty0 = tile_y * kT;
  const float tail = kFade ? a.fade_alpha * exp2_hw(a.tail_exp) : 0.0f;

  // A tile whose list fits one staged batch (most tiles of dense fields) stages it once, with
  // the gradient sweep's chunks and slots, before the normalizer sweep, which reads the same
  // records; the gradient sweep then skips its staging.
  constexpr int kLg = kT == 16 ? 4 : 5;
  int64_t start, end;
  tile_range(a, tile, &start, &end);
  const bool single = kRecompute && end - start <= kS;
  if (single) {
    stage_batch<4, kChunks, kS, kAffine, true, kOpaque>(
        a, st, start, static_cast<int>(end - start),
        SlotTile{tile_x, tile_y, kLg, g.slot_capacity});
    __syncthreads();
  }

  // Per-pixel adjoints (render_ext.cu:496-510):
  //   dL/dN_k = G_k/(D+eps), dL/dD = -sum_k G_k out_k/(D+eps)  (additive: G_k and 0).
  if (kRecompute) {
    // D (and N) recomputed for one pixel per lane by the forward's own sweep and order, so they
    // equal the forward's values bit for bit.
    const int qx0 = tx0 + (wave % kQ) * 8;
    const int qy0 = ty0 + (wave / kQ) * 8;
    const int px = qx0 + lx;
    const int py = qy0 + ly;
    uint32_t* s_mask = reinterpret_cast<uint32_t*>(s_scratch) + wave * (2 * kS);
    float4 acc;
    accumulate_quadrant<kFade, kNorm == kNormRecompD, kChunks, kS, kAffine, kOpaque>(
        a, st, s_mask, tile, qx0, qy0, px, py, tail, acc, single);
    // Every wave is past the sweep's last barrier: the mask space can take the adjoints.
    float4 adj = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (px < f.width && py < f.height) {
      const int64_t flat = f.first_pixel + int64_t(py) * f.width + px;
      const float gr = g.grad_out[flat * 3 + 0];
      const float gg = g.grad_out[flat * 3 + 1];
      const float gb = g.grad_out[flat * 3 + 2];
      const float denom = acc.w + a.eps;
      float o_r, o_g, o_b;
      if (kNorm == kNormRecompD) {
        o_r = g.out[flat * 3 + 0];
        o_g = g.out[flat * 3 + 1];
        o_b = g.out[flat * 3 + 2];
      } else {
        o_r = acc.x / denom;
        o_g = acc.y / denom;
        o_b = acc.z / denom;
      }
      const float dot = gr * o_r + gg * o_g + gb * o_b;
      adj = make_float4(gr / denom, gg / denom, gb / denom, -dot / denom);
    }
    s_adj[(py - ty0) * kT + (px - tx0)] = adj;
  } else {
    const int px = tx0 + static_cast<int>(thread_index<kOpaque>() % kT);
    const int py = ty0 + static_cast<int>(thread_index<kOpaque>() / kT);
    float4 adj = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (px < f.width && py < f.height) {
      const int64_t flat = f.first_pixel + int64_t(py) * f.width + px;
      const float gr = g.grad_out[flat * 3 + 0];
      const float gg = g.grad_out[flat * 3 + 1];
      const float gb = g.grad_out[flat * 3 + 2];
      if (kNormalize) {
        const float denom = g.den[flat] + a.eps;
        const float dot = gr * g.out[flat * 3 + 0] + gg * g.out[flat * 3 + 1] +
                          gb * g.out[flat * 3 + 2];
        adj = make_float4(gr / denom, gg / denom, gb / denom, -dot / denom);
      } else {
        adj = make_float4(gr, gg, gb, 0.0f);
      }
    }
    s_adj[thread_index<kOpaque>()] = adj;
  }
  __syncthreads();

  // Gradient sweep: each lane owns the pixel (lx, ly) of every quadrant; with 16-pixel tiles a
  // wave record reads the four quadrants' adjoints into registers once.
  const bool has_opacity = g.has_opacity;
  const float tag = __uint_as_float(*g.epoch);
  // Lane roles of the transposed reduction: lane 8k+7 owns gradient component k
  // (0-1 mean x/y, 2-4 conic a/b/c, 5-7 color r/g/b); lane 63 also the opacity and the tag.
  const bool hi8 = (lane & 8) != 0;
  const int comp = lane >> 3;
  const bool result_lane = (lane & 7) == 7;
  const float comp_scale = (comp == 2 || comp == 4) ? -0.5f : (comp == 3 ? -1.0f : 1.0f);
  // The batch's small and large records (see below) and their counts.
  __shared__ uint16_t s_small[kAffine ? 1 : kS];
  __shared__ uint16_t s_large[kAffine ? 1 : kS];
  __shared__ int s_nsmall, s_nlarge;
  // One small staged record (index j, -1: none) by one quad of lanes: lane `sub` of the quad
  // takes the record's pixels of this tile with row-major index sub, sub + 4, ... (the visit's
  // arithmetic, summed in registers); a fixed two-step DPP sum gives every lane of the quad the
  // same totals and the quad's lanes store the pair's slot.  All lanes of the wave call it
  // together.
  auto quad_record = [&](int j, int sub) {
    float sx = 0.0f, sy = 0.0f, sxx = 0.0f, sxy = 0.0f, syy = 0.0f;
    float gcr = 0.0f, gcg = 0.0f, gcb = 0.0f, gop = 0.0f;
    int slot = -1;
    if (j >= 0) {
      const float4 v0 = st.c[0][j];
      const float4 v1 = st.c[1][j];
      const float4 v3 = st.c[3][j];
      slot = __float_as_int(v3.w);
      const uint32_t xr = __float_as_uint(v1.y), yr = __float_as_uint(v1.z);
      const int ix0 = max(static_cast<int>(xr & 0xffff), tx0);
      const int ix1 = min(static_cast<int>(xr >> 16), tx0 + kT);
      const int iy0 = max(static_cast<int>(yr & 0xffff), ty0);
      const int w = ix1 - ix0;
      const int n = w * (min(static_cast<int>(yr >> 16), ty0 + kT) - iy0);
      const float mx = v0.x, my = v0.y, qa = v0.z, qb = v0.w, qc = v1.x;
      const float cr = v3.x, cg = v3.y, cbl = v3.z;
      // Pixel `sub` of the row-major order, then steps of 4 pixels: dr = 4 / w rows and dc
      // columns, or, past the row's end, one row more and w columns fewer.  The pixel's
      // coordinates are kept as floats (small integers, exact) and its adjoints' LDS address in
      // bytes; the quotients by w (1 <= w) are spelled out for the numerators 0-4.
      const int dr = w >= 5 ? 0 : (w == 1 ? 4 : (w == 2 ? 2 : 1));
      const int dc = 4 - dr * w;
      const int r0 = w >= 4 ? 0 : (w == 1 ? sub : (w == 2 ? sub >> 1 : (sub == 3 ? 1 : 0)));
      float fx = static_cast<float>(ix0 + sub - r0 * w);
      float fy = static_cast<float>(iy0 + r0);
      const float fx_end = static_cast<float>(ix1 - dc);  // at or past it, the step wraps
      // The steps stay registers (opaque values: otherwise the select of two conversions is
      // folded into a conversion of a select, two more instructions per pixel).
      float sx_n = static_cast<float>(dc), sx_w = static_cast<float>(dc - w);
      float sy_n = static_cast<float>(dr), sy_w = static_cast<float>(dr + 1);
      asm volatile("" : "+v"(sx_n), "+v"(sx_w), "+v"(sy_n), "+v"(sy_w));
      constexpr int kRowBytes = kT * static_cast<int>(sizeof(float4));
      const int sa_n = dr * kRowBytes + dc * static_cast<int>(sizeof(float4));
      const int sa_w = sa_n + kRowBytes - w * static_cast<int>(sizeof(float4));
      int addr = ((iy0 + r0 - ty0) * kT + (ix0 + sub - r0 * w - tx0)) *
                 static_cast<int>(sizeof(float4));
      const char* adj_base = reinterpret_cast<const char*>(s_adj);
      for (int k = sub; k < n; k += 4) {
        const float4 ad = *reinterpret_cast<const float4*>(adj_base + addr);
        const float dx = fx - mx;
        const float dy = fy - my;
        const bool wrap = fx >= fx_end;
        fx += wrap ? sx_w : sx_n;
        fy += wrap ? sy_w : sy_n;
        addr += wrap ? sa_w : sa_n;
        const float t = fmaf(qa, dx, qb * dy);
        const float e = fmaf(t, dx, qc * dy * dy);
        const float raw = exp2_hw(kFade ? e * kNegHalfLog2e : e);
        const float dw = fmaf(ad.x, cr, fmaf(ad.y, cg, fmaf(ad.z, cbl, ad.w)));
        float base, dqr;
        if (kFade) {
          base = fmaxf(raw - tail, 0.0f);
          dqr = raw > tail ? dw * raw : 0.0f;
        } else {
          base = raw;
          dqr = dw * raw;
        }
        gcr = fmaf(ad.x, base, gcr);
        gcg = fmaf(ad.y, base, gcg);
        gcb = fmaf(ad.z, base, gcb);
        gop = fmaf(dw, base, gop);
        sx = fmaf(dqr, dx, sx);
        sy = fmaf(dqr, dy, sy);
        const float dqx = dqr * dx;
        sxx = fmaf(dqx, dx, sxx);
        sxy = fmaf(dqx, dy, sxy);
        syy = fmaf(dqr * dy, dy, syy);
      }
    }
    // Quad sums: (v0 + v1) + (v2 + v3) on every lane of the quad, bit for bit.
    auto quad_sum = [](float& v) {
      v += dpp_move<0xb1>(v);  // quad_perm [1,0,3,2]
      v += dpp_move<0x4e>(v);  // quad_perm [2,3,0,1]
    };
    quad_sum(sx), quad_sum(sy), quad_sum(sxx), quad_sum(sxy), quad_sum(syy);
    quad_sum(gcr), quad_sum(gcg), quad_sum(gcb), quad_sum(gop);
    if (slot >= 0) {
      // The slot's first 32 bytes by the quad's four lanes in one store, then opacity and tag.
      float2* dst = reinterpret_cast<float2*>(g.slots + int64_t(slot) * kSlotWords);
      const float lo = sub == 0 ? sx : (sub == 1 ? sxx : (sub == 2 ? syy : gcg));
      const float hi = sub == 0 ? sy : (sub == 1 ? sxy : (sub == 2 ? gcr : gcb));
      dst[sub] = make_float2(lo, hi);
      if (sub == 0) dst[4] = make_float2(has_opacity ? gop : 0.0f, tag);
    }
  };
  for (int64_t batch = start; batch < end; batch += kS) {
    const int m = static_cast<int>(end - batch < kS ? end - batch : kS);
    if (!single) {
      stage_batch<4, kChunks, kS, kAffine, true, kOpaque>(a, st, batch, m,
                                                         SlotTile{tile_x, tile_y, kLg,
                                                                  g.slot_capacity});
      __syncthreads();
    }
    // One staged record by the whole wave: each lane visits its pixel of every quadrant the
    // rectangle meets, then the wave's transposed reduction and one slot store.
    // With 16-pixel tiles `adj` holds the lane's adjoints of the four quadrants.
    auto wave_record = [&](int j, const float4 (&adj)[4]) {
      const float4 v0 = st.c[0][j];
      const float4 v1 = st.c[1][j];
      const float4 v2 = st.c[2][j];
      const float4 v3 = st.c[3][j];
      const uint32_t xr = __builtin_amdgcn_readfirstlane(__float_as_uint(v1.y));
      const uint32_t yr = __builtin_amdgcn_readfirstlane(__float_as_uint(v1.z));
      const int x0 = xr & 0xffff, x1 = xr >> 16, y0 = yr & 0xffff, y1 = yr >> 16;
      const float mx = v0.x, my = v0.y, qa = v0.z, qb = v0.w, qc = v1.x;
      const float cr = v3.x, cg = v3.y, cbl = v3.z;
      // The accumulators; the first quadrant's visit (16-pixel tiles) or an explicit clear sets
      // them.
      float sx, sy, sxx, sxy, syy, gcr, gcg, gcb, gop;
      // Affine colors: the frame and color gradients, and their partials (unused otherwise).
      float4 a0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), a1 = a0, a2 = a0;
      if constexpr (kAffine) {
        a0 = st.c[4][j];
        a1 = st.c[5][j];
        a2 = st.c[6][j];
      }
      float g00, g01, g02, g10, g11, g12, sa, sb, cxx, cyy, cxy, cyx;
      // The clear's moves stay on its own path (a volatile statement is not hoisted above the
      // branch that selects it; a plain zero would be set ahead of the branch on every pair).
      auto zero = [](float& v) { asm volatile("v_mov_b32 %0, 0" : "=v"(v)); };
      auto clear = [&]() {
        zero(sx), zero(sy), zero(sxx), zero(sxy), zero(syy);
        zero(gcr), zero(gcg), zero(gcb), zero(gop);
        if constexpr (kAffine) {
          zero(g00), zero(g01), zero(g02), zero(g10), zero(g11), zero(g12);
          zero(sa), zero(sb), zero(cxx), zero(cyy), zero(cxy), zero(cyx);
        }
      };
      // One visit of pixel (px, py) with its adjoint.  kFirst: the visit sets the accumulators
      // instead of adding to them and runs on every lane, a lane outside the rectangle evaluating
      // its weight at exp2(-inf) = 0 so that its terms are zero (for finite adjoints).  Otherwise
      // only lanes inside the rectangle run it.
      auto visit = [&](auto first, int px, int py, const float4& ad, bool inside) {
        constexpr bool kFirst = decltype(first)::value;
        auto mad = [](float& acc, float x, float y) { acc = kFirst ? x * y : fmaf(x, y, acc); };
        auto add = [](float& acc, float x) { acc = kFirst ? x : acc + x; };
        const float dx = static_cast<float>(px) - mx;
        const float dy = static_cast<float>(py) - my;
        const float t = fmaf(qa, dx, qb * dy);
        const float e = fmaf(t, dx, qc * dy * dy);
        const float arg = kFade ? e * kNegHalfLog2e : e;
        const float raw = exp2_hw(kFirst && !inside ? -INFINITY : arg);
        float lxc = 0.0f, lyc = 0.0f;
        float pr = cr, pg = cg, pb = cbl;  // the pixel's color
        if constexpr (kAffine) {
          lxc = fmaf(a0.x, dx, a0.y * dy);
          lyc = fmaf(a0.z, dy, -(a0.w * dx));
          pr = fmaf(a1.x, lxc, fmaf(a1.w, lyc, cr));
          pg = fmaf(a1.y, lxc, fmaf(a2.x, lyc, cg));
          pb = fmaf(a1.z, lxc, fmaf(a2.y, lyc, cbl));
        }
        const float dw = fmaf(ad.x, pr, fmaf(ad.y, pg, fmaf(ad.z, pb, ad.w)));
        float base, dqr;
        if (kFade) {
          base = fmaxf(raw - tail, 0.0f);
          dqr = raw > tail ? dw * raw : 0.0f;
        } else {
          base = raw;
          dqr = dw * raw;
        }
        mad(gcr, ad.x, base);
        mad(gcg, ad.y, base);
        mad(gcb, ad.z, base);
        mad(gop, dw, base);
        mad(sx, dqr, dx);
        mad(sy, dqr, dy);
        const float dqx = dqr * dx;
        mad(sxx, dqx, dx);
        mad(sxy, dqx, dy);
        mad(syy, dqr * dy, dy);
        if constexpr (kAffine) {
          // Color-gradient partials sum base * local * dL/dN; the local-coordinate adjoints
          // base * (dL/dN . G0) and base * (dL/dN . G1) feed the mean, scale and rotation
          // gradients (times the opacity, applied per pair).
          const float bl = base * lxc, bly = base * lyc;
          mad(g00, ad.x, bl);
          mad(g01, ad.y, bl);
          mad(g02, ad.z, bl);
          mad(g10, ad.x, bly);
          mad(g11, ad.y, bly);
          mad(g12, ad.z, bly);
          const float ax = base * fmaf(ad.x, a1.x, fmaf(ad.y, a1.y, ad.z * a1.z));
          const float ay = base * fmaf(ad.x, a1.w, fmaf(ad.y, a2.x, ad.z * a2.y));
          add(sa, ax);
          add(sb, ay);
          mad(cxx, ax, lxc);
          mad(cyy, ay, lyc);
          mad(cxy, ax, lyc);
          mad(cyx, ay, lxc);
        }
      };
      using First = std::true_type;
      using Next = std::false_type;
      const unsigned rw = static_cast<unsigned>(x1 - x0), rh = static_cast<unsigned>(y1 - y0);
      if constexpr (kT == 16) {
        // One straight-line block per quadrant: the per-lane rectangle test alone decides which
        // lanes accumulate (a quadrant the rectangle misses has no lane inside and is skipped by
        // the exec-mask branch), so the accumulators stay in the same registers throughout.
        // The first quadrant, when the rectangle meets it (a uniform test), sets the
        // accumulators on every lane; otherwise they are cleared.
        const bool meets0 = x0 < tx0 + 8 && x1 > tx0 && y0 < ty0 + 8 && y1 > ty0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int px = tx0 + (s & 1) * 8 + lx;
          const int py = ty0 + (s >> 1) * 8 + ly;
          const bool inside =
              (static_cast<unsigned>(px - x0) < rw) & (static_cast<unsigned>(py - y0) < rh);
          if (s == 0) {
            if (meets0) {
              visit(First{}, px, py, adj[s], inside);
            } else {
              clear();
            }
          } else if (inside) {
            visit(Next{}, px, py, adj[s], inside);
          }
        }
      } else {
        clear();
        // Only the quadrants the rectangle meets (uniform bounds), adjoints read from LDS.
        const int cx0 = max(x0, tx0) - tx0, cx1 = min(x1, tx0 + kT) - tx0;
        const int cy0 = max(y0, ty0) - ty0, cy1 = min(y1, ty0 + kT) - ty0;
        if (cx1 > cx0 && cy1 > cy0) {
          for (int qy = cy0 >> 3; qy <= (cy1 - 1) >> 3; ++qy) {
            for (int qx = cx0 >> 3; qx <= (cx1 - 1) >> 3; ++qx) {
              const int px = tx0 + qx * 8 + lx;
              const int py = ty0 + qy * 8 + ly;
              if ((static_cast<unsigned>(px - x0) < rw) & (static_cast<unsigned>(py - y0) < rh)) {
                visit(Next{}, px, py, s_adj[(qy * 8 + ly) * kT + qx * 8 + lx], true);
              }
            }
          }
        }
      }
      // Reduce the eight partials to lanes 8k+7 (k = component), the opacity partial to lane 63.
      const float vals[8] = {sx, sy, sxx, sxy, syy, gcr, gcg, gcb};
      const float red = wave_sum8_transposed(vals, hi8);
      float gop_sum, red2 = 0.0f;
      // The slot's components: without affine colors the pair's raw sums (Sx, Sy, Sxx, Sxy, Syy,
      // Cr, Cg, Cb; the opacity and conic factors are per Gaussian and k_grad_reduce applies them
      // once to the sums over its tiles); with affine colors the pair's gradients.
      float gval = red;
      float extra[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // affine: scale x, scale y, rotation
      const float op = v2.w;
      if constexpr (kAffine) {
        const float partner = dpp_move<0x128>(red);  // lane 7 <-> lane 15: Sx and Sy meet
        const float vals2[8] = {g00, g01, g02, g10, g11, g12, sa, sb};
        const float vals3[8] = {cxx, cyy, cxy, cyx, gop, 0.0f, 0.0f, 0.0f};
        red2 = wave_sum8_transposed(vals2, hi8);
        const float red3 = wave_sum8_transposed(vals3, hi8);
        gop_sum = has_opacity ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red3), 39))
                              : 0.0f;
        // dL/dq = -op/2 * dw * raw per pixel; chain to the mean and conic entries:
        //   g_mx = op (a Sx + b Sy), g_my = op (b Sx + c Sy), g_a = -op/2 Sxx, g_b = -op Sxy,
        //   g_c = -op/2 Syy, g_color = op C.
        // The record's quadratic is the conic times -log2(e)/2 when folded; undo that here.
        const float unfold = kFade ? 1.0f : kInvNegHalfLog2e;
        const float ca = qa * unfold, cc = qc * unfold, cb = (0.5f * unfold) * qb;
        const float own_scale = comp == 0 ? ca : (comp == 1 ? cc : comp_scale);
        const float cross = comp < 2 ? partner * cb : 0.0f;
        gval = op * fmaf(red, own_scale, cross);
        // Local coordinates x = (c dx + s dy) isx, y = (c dy - s dx) isy (isx = 1/sx), d = p - mean,
        // with A, B the sums of the local-coordinate adjoints and Cuv = sum (adjoint u) * v:
        //   dL/dmean = -op (A (a0.x, a0.y) + B (-a0.w, a0.z));
        //   dL/dsx = -op Cxx isx (zero where the 1e-6 clamp holds), likewise sy;
        //   dL/dtheta = op (Cxy isx / isy - Cyx isy / isx).
        const float A = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red2), 55));
        const float B = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red2), 63));
        const float Cxx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red3), 7));
        const float Cyy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red3), 15));
        const float Cxy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red3), 23));
        const float Cyx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(red3), 31));
        if (comp == 0) gval -= op * fmaf(a0.x, A, -(a0.w * B));
        if (comp == 1) gval -= op * fmaf(a0.y, A, a0.z * B);
        // a2.z / a2.w carry 1/s, negated where the clamp holds (s < 1e-6).
        const float isx = fabsf(a2.z), isy = fabsf(a2.w);
        extra[0] = a2.z < 0.0f ? 0.0f : -op * Cxx * isx;
        extra[1] = a2.w < 0.0f ? 0.0f : -op * Cyy * isy;
        extra[2] = op * (Cxy * (isx / isy) - Cyx * (isy / isx));
      } else {
        gop_sum = has_opacity ? wave_sum_lane63(gop) : 0.0f;
      }
      // The pair's slot, found at staging (-1: none).
      const int slot = __builtin_amdgcn_readfirstlane(__float_as_int(v3.w));
      if (slot >= 0) {
        float* dst = g.slots + int64_t(slot) * kWords;
        if (result_lane) dst[comp] = gval;
        if constexpr (kAffine) {
          if (result_lane && comp < 6) dst[9 + comp] = op * red2;
          if (lane == 63) {
            dst[8] = gop_sum;
            dst[15] = extra[0];
            dst[16] = extra[1];
            dst[17] = extra[2];
            dst[kWords - 1] = tag;
          }
        } else {
          if (lane == 63) *reinterpret_cast<float2*>(dst + 8) = make_float2(gop_sum, tag);
        }
      }
    };
    // The lane's adjoints of the four quadrants (16-pixel tiles), read from LDS into registers
    // once per batch, where a wave starts its large records (not live during the lane path).
    auto load_adj = [&](float4 (&adj)[4]) {
      if constexpr (kT == 16) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          adj[s] = s_adj[((s >> 1) * 8 + ly) * kT + (s & 1) * 8 + lx];
        }
      }
    };
    if constexpr (kAffine) {
      float4 adj[4];
      load_adj(adj);
      for (int j = wave; j < m; j += kWaves) wave_record(j, adj);
    } else {
      // Records whose rectangle covers at most kClasses * kLaneClassArea pixels of the tile are
      // summed by a quad of lanes each over those pixels (adjoints from LDS): a wave takes 16 of
      // them at a time, with only a quad sum to reduce; wave 0 sorts the batch into those and
      // the larger ones (empty intersections have no slot and are dropped).  The small ones are
      // listed by area class (1-16, 17-32, ... pixels), so the 16 records a wave takes together
      // need about the same number of steps (a wave runs as many as its largest record needs).
      // Both kinds are dealt round-robin to the waves: 16 small records a wave at a time, then
      // one large record at a time.
      constexpr int kClasses = lane_classes(kT);
      if (wave == 0) {
        constexpr int kPerLane = kS / 64;
        int cls[kPerLane];  // 0: no pixels, 1..kClasses: small, kClasses + 1: large
#pragma unroll
        for (int r = 0; r < kPerLane; ++r) {
          const int j = r * 64 + lane;
          int area = 0;
          if (j < m) {
            const float4 v1 = st.c[1][j];
            const uint32_t xr = __float_as_uint(v1.y), yr = __float_as_uint(v1.z);
            const int w = min(static_cast<int>(xr >> 16), tx0 + kT) -
                          max(static_cast<int>(xr & 0xffff), tx0);
            const int h = min(static_cast<int>(yr >> 16), ty0 + kT) -
                          max(static_cast<int>(yr & 0xffff), ty0);
            area = w > 0 && h > 0 ? w * h : 0;
          }
          cls[r] = min((area + kLaneClassArea - 1) / kLaneClassArea, kClasses + 1);
        }
        int ns = 0, nl = 0;
#pragma unroll
        for (int c = 1; c <= kClasses; ++c) {
#pragma unroll
          for (int r = 0; r < kPerLane; ++r) {
            const bool in = cls[r] == c;
            const uint64_t b = __ballot(in);
            if (in) s_small[ns + lanes_below(b)] = static_cast<uint16_t>(r * 64 + lane);
            ns += static_cast<int>(__popcll(b));
          }
        }
#pragma unroll
        for (int r = 0; r < kPerLane; ++r) {
          const bool in = cls[r] == kClasses + 1;
          const uint64_t b = __ballot(in);
          if (in) s_large[nl + lanes_below(b)] = static_cast<uint16_t>(r * 64 + lane);
          nl += static_cast<int>(__popcll(b));
        }
        if (lane == 0) {
          s_nsmall = ns;
          s_nlarge = nl;
        }
      }
      __syncthreads();
      const int ns = s_nsmall, nl = s_nlarge;
      for (int base = wave * 16; base < ns; base += kWaves * 16) {
        const int i = base + (lane >> 2);
        quad_record(i < ns ? static_cast<int>(s_small[i]) : -1, lane & 3);
      }
      if (wave < nl) {
        float4 adj[4];
        load_adj(adj);
        for (int i = wave; i < nl; i += kWaves) wave_record(s_large[i], adj);
      }
    }
    __syncthreads();
  }
}

// Eight waves per SIMD (at most 64 VGPRs): with 16-pixel tiles that is eight resident
// workgroups per CU, whose different phases (list sweeps, reductions) overlap.  The affine
// sweep's 21 accumulators need more registers: four waves per SIMD (up to 128 VGPRs).
template <int kT, bool kFade, bool kNormalize, int kNorm, bool kAffine = false>
__global__ __launch_bounds__(kT * kT) __attribute__((amdgpu_waves_per_eu(
    kAffine ? 4 : 8, kAffine ? 4 : 8))) void k_backward(RenderArgs a, GradArgs g) {
  backward_tile<kT, kFade, kNormalize, kNorm, kAffine>(a, g, xcd_remap(blockIdx.x, gridDim.x));
}

// The tile loop holds the kernel arguments in scalar registers across tiles.  With 32-pixel tiles
// eight waves per SIMD (two workgroups per CU) costs a few scalar spills into vector lanes; with
// 16-pixel tiles the recomputing sweeps need 69-71 VGPRs, and eight waves per SIMD would spill
// vector registers to memory, so seven is the floor.
template <int kT, bool kFade, bool kNormalize, int kNorm>
__global__ __launch_bounds__(kT * kT) __attribute__((amdgpu_waves_per_eu(kT == 32 ? 8 : 7, 8)))
void k_backward_persistent(RenderArgs a, GradArgs g, TileQueue q) {
  TileClaim c{};
  if (threadIdx.x == 0) c = first_claim(q);
  for (int tile = claim_tile(q, c); tile >= 0; tile = claim_tile(q, c)) {
    backward_tile<kT, kFade, kNormalize, kNorm, false, true>(a, g, tile);
  }
}

struct ReduceArgs {
  const GaussRec* recs;
  int n;
  int lg;
  int num_waves;
  const int32_t* rect_off;
  const int32_t* rect_status;
  const float* slots;
  int64_t slot_capacity;
  const uint32_t* epoch;
  float* grad_means;
  float* grad_conics;
  float* grad_colors;
  float* grad_opacities;    // NULL when the field has no opacities
  float* grad_color_grads;  // affine only: (n, 2, 3)
  float* grad_scales;       // affine only: (n, 2)
  float* grad_rotations;    // affine only: (n,)
  int32_t* status;          // may be NULL
  float unfold;             // the records' conic factor undone: 1, or 1 / (-log2(e)/2) if folded
};

// Adds one slot's kVals partial gradients when its tag (word kWords - 1) is the call's: a slot
// the call did not write counts as zero, and the select keeps stale values out of the sum.
template <int kWords, int kVals>
__device__ __forceinline__ void add_slot_values(const float2 (&v)[kWords / 2], uint32_t tag,
                                                float (&acc)[kVals]) {
  const bool live = __float_as_uint(v[kWords / 2 - 1].y) == tag;
#pragma unroll
  for (int c = 0; c < kVals; ++c) {
    const float x = (c & 1) ? v[c >> 1].y : v[c >> 1].x;
    acc[c] += live ? x : 0.0f;
  }
}

template <int kWords, int kVals>
__device__ __forceinline__ void add_slot(const float* s, uint32_t tag, float (&acc)[kVals]) {
  float2 v[kWords / 2];
#pragma unroll
  for (int w = 0; w < kWords / 2; ++w) v[w] = *reinterpret_cast<const float2*>(s + 2 * w);
  add_slot_values<kWords, kVals>(v, tag, acc);
}

// Gradients of every Gaussian from its slots, in an order fixed by the Gaussian alone: up to
// kLaneSumMax slots are summed by one lane in ascending slot order; more are summed by the whole
// wave, lane l taking slots l, l + 64, ... in order, then a fixed DPP tree.  Every Gaussian's
// outputs are written (zero without slots).  kVals = 9 (means, conics, colors, opacity) or 18
// (plus color gradients, scales and rotation with affine colors).
template <int kWords, int kVals>
__global__ __launch_bounds__(256) void k_grad_reduce(ReduceArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const uint32_t tag = *a.epoch;
  int count = 0;
  int64_t base = 0;
  if (i < a.n) {
    const GaussRec& r = a.recs[i];
    const float4 c1 = r.c1;
    count = rect_tiles(__float_as_uint(c1.y), __float_as_uint(c1.z), a.lg).count;
    base = int64_t(a.rect_off[i >> 6]) + __float_as_int(r.c3.w);
    // Slots beyond the capacity were never written (the status reports the overflow).
    const int64_t room = a.slot_capacity - base;
    count = room <= 0 ? 0 : (room < count ? static_cast<int>(room) : count);
  }
  float acc[kVals];
#pragma unroll
  for (int c = 0; c < kVals; ++c) acc[c] = 0.0f;
  // Lane sums, each in ascending slot order.  A Gaussian's slots follow the previous
  // Gaussian's, so the wave's slots form one run.  With a few slots per lane the run is copied
  // to LDS kChunk slots at a time with coalesced loads (lanes reading their own slots straight
  // from memory would touch a different cache line per lane and load); with many, each lane
  // streams its own slots (a chunk would then hold the slots of only a few lanes).
  constexpr int kChunk = 2560 / kWords;      // slots per chunk: 10 KB per wave
  constexpr int kPer = kChunk * kWords / 128;  // float2 loads per lane and chunk
  constexpr int kChunkedMaxPerLane = 12;       // slots per small lane up to which chunks pay
  constexpr int kInFlight = kWords <= kSlotWords ? 8 : 4;  // loads issued together per lane
  __shared__ float2 s_chunk[256 / 64][kChunk * kWords / 2];
  float2* chunk = s_chunk[threadIdx.x >> 6];
  const bool small = count > 0 && count <= kLaneSumMax;
  // Slot runs ascend with the lane, so the wave's run is [first small lane's base, last small
  // lane's end).
  const uint64_t small_lanes = __ballot(small);
  int64_t lo = 0, hi = 0;
  if (small_lanes != 0) {
    const int first = __ffsll(static_cast<unsigned long long>(small_lanes)) - 1;
    const int last = 63 - __clzll(static_cast<long long>(small_lanes));
    const int64_t end = base + count;
    lo = (int64_t(__builtin_amdgcn_readlane(static_cast<int>(base >> 32), first)) << 32) |
         uint32_t(__builtin_amdgcn_readlane(static_cast<int>(base), first));
    hi = (int64_t(__builtin_amdgcn_readlane(static_cast<int>(end >> 32), last)) << 32) |
         uint32_t(__builtin_amdgcn_readlane(static_cast<int>(end), last));
  }
  const bool chunked = hi - lo <= int64_t(kChunkedMaxPerLane) * __popcll(small_lanes);
  if (!chunked && small) {
    const float* sl = a.slots + base * kWords;
    for (int k = 0; k < count; k += kInFlight) {
      float2 v[kInFlight][kWords / 2];
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) {
        const bool live = k + u < count;
#pragma unroll
        for (int w = 0; w < kWords / 2; ++w) {
          v[u][w] = live ? *reinterpret_cast<const float2*>(sl + (k + u) * kWords + 2 * w)
                         : make_float2(0.0f, 0.0f);
        }
      }
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) {
        if (k + u < count) add_slot_values<kWords, kVals>(v[u], tag, acc);
      }
    }
  }
  for (int64_t c0 = lo; chunked && c0 < hi; c0 += kChunk) {
    const bool need = small && base < c0 + kChunk && base + count > c0;
    if (__ballot(need) == 0) continue;  // a run of big Gaussians' slots
    const int len2 = static_cast<int>(hi - c0 < kChunk ? hi - c0 : kChunk) * (kWords / 2);
    const float2* src = reinterpret_cast<const float2*>(a.slots + c0 * kWords);
    float2 tmp[kPer];  // every load of the chunk in flight at once
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int t = lane + 64 * u;
      tmp[u] = t < len2 ? src[t] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      if (lane + 64 * u < len2) chunk[lane + 64 * u] = tmp[u];
    }
    wave_lds_sync();
    if (need) {
      const int k0 = static_cast<int>((base > c0 ? base : c0) - c0);
      const int64_t end = base + count < c0 + kChunk ? base + count : c0 + kChunk;
      const int k1 = static_cast<int>(end - c0);
      for (int k = k0; k < k1; ++k) {
        float2 v[kWords / 2];
#pragma unroll
        for (int w = 0; w < kWords / 2; ++w) v[w] = chunk[k * (kWords / 2) + w];
        add_slot_values<kWords, kVals>(v, tag, acc);
      }
    }
    wave_lds_sync();  // every lane is done with the chunk before the next copy
  }
  uint64_t big = __ballot(count > kLaneSumMax);
  while (big) {
    const int b = __ffsll(static_cast<unsigned long long>(big)) - 1;
    big &= big - 1;
    const int cb = __builtin_amdgcn_readlane(count, b);
    const int64_t bb = (int64_t(__builtin_amdgcn_readlane(static_cast<int>(base >> 32), b)) << 32) |
                       uint32_t(__builtin_amdgcn_readlane(static_cast<int>(base), b));
    float part[kVals];
#pragma unroll
    for (int c = 0; c < kVals; ++c) part[c] = 0.0f;
    for (int k = lane; k < cb; k += 64) {
      add_slot<kWords, kVals>(a.slots + (bb + k) * kWords, tag, part);
    }
#pragma unroll
    for (int c = 0; c < kVals; ++c) {
      const float total = __int_as_float(
          __builtin_amdgcn_readlane(__float_as_int(wave_sum_lane63(part[c])), 63));
      if (lane == b) acc[c] = total;
    }
  }
  if (i < a.n) {
    if constexpr (kVals == 9) {
      // The slots hold each pair's raw sums; the per-Gaussian factors (dL/dq = -op/2 dw raw per
      // pixel, chained to the mean and conic entries) apply to their totals:
      //   g_mx = op (a Sx + b Sy), g_my = op (b Sx + c Sy), g_a = -op/2 Sxx, g_b = -op Sxy,
      //   g_c = -op/2 Syy, g_color = op C.
      const GaussRec& r = a.recs[i];
      const float4 c0 = r.c0;
      const float op = r.c2.w;
      const float ca = c0.z * a.unfold, cc = r.c1.x * a.unfold, cb = (0.5f * a.unfold) * c0.w;
      const float sx = acc[0], sy = acc[1];
      acc[0] = op * fmaf(sx, ca, sy * cb);
      acc[1] = op * fmaf(sy, cc, sx * cb);
      acc[2] = op * (-0.5f * acc[2]);
      acc[3] = op * -acc[3];
      acc[4] = op * (-0.5f * acc[4]);
      acc[5] *= op;
      acc[6] *= op;
      acc[7] *= op;
    }
    a.grad_means[2 * i + 0] = acc[0];
    a.grad_means[2 * i + 1] = acc[1];
    a.grad_conics[3 * i + 0] = acc[2];
    a.grad_conics[3 * i + 1] = acc[3];
    a.grad_conics[3 * i + 2] = acc[4];
    a.grad_colors[3 * i + 0] = acc[5];
    a.grad_colors[3 * i + 1] = acc[6];
    a.grad_colors[3 * i + 2] = acc[7];
    if (a.grad_opacities != nullptr) a.grad_opacities[i] = acc[8];
    if constexpr (kVals > 9) {
#pragma unroll
      for (int c = 0; c < 6; ++c) a.grad_color_grads[6 * i + c] = acc[9 + c];
      a.grad_scales[2 * i + 0] = acc[15];
      a.grad_scales[2 * i + 1] = acc[16];
      a.grad_rotations[i] = acc[17];
    }
  }
  if (i == 0 && a.status != nullptr &&
      (*a.rect_status != SSPLAT_OK || int64_t(a.rect_off[a.num_waves]) > a.slot_capacity)) {
    *a.status = SSPLAT_ERR_CAPACITY;
  }
}

// Affine records (see RenderArgs) from scales (n,2), rotations (n,) and color gradients
// (n,2,3), in float32 like the reference's torch expression (render.py:110-123).
__global__ __launch_bounds__(256) void k_prepare_affine(const float* __restrict__ scales,
                                                        const float* __restrict__ rotations,
                                                        const float* __restrict__ color_grads,
                                                        int n, float4* __restrict__ arecs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float sx = scales[2 * i + 0], sy = scales[2 * i + 1];
  const float isx = 1.0f / fmaxf(sx, 1e-6f);
  const float isy = 1.0f / fmaxf(sy, 1e-6f);
  const float t = rotations[i];
  const float ct = cosf(t), st = sinf(t);
  const float* G = color_grads + 6 * int64_t(i);
  arecs[3 * int64_t(i) + 0] = make_float4(ct * isx, st * isx, ct * isy, st * isy);
  arecs[3 * int64_t(i) + 1] = make_float4(G[0], G[1], G[2], G[3]);
  arecs[3 * int64_t(i) + 2] = make_float4(G[4], G[5], sx < 1e-6f ? -isx : isx,
                                          sy < 1e-6f ? -isy : isy);
}

// Number of slots (support-rectangle tiles) of every Gaussian into rect_wave, per wave of 64,
// for a slot count without record packing.
__global__ __launch_bounds__(256) void k_rect_count(const float* __restrict__ means,
                                                    const int64_t* __restrict__ radii, int n,
                                                    FieldTable fields, SlotPrep sp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  if (i == 0) {
    sp.rect_wave[sp.num_waves] = 0;
    *sp.rect_status = SSPLAT_OK;
  }
  uint32_t xr = 0, yr = 0;
  if (i < n) {
    const ssplat_field f = field_of_gaussian(fields, i);
    const Rect r = support_rect(means[2 * i + 0], means[2 * i + 1], radii[2 * i + 0],
                                radii[2 * i + 1], f.height, f.width);
    if (r.x1 > r.x0 && r.y1 > r.y0) {
      xr = uint32_t(r.x0) | (uint32_t(r.x1) << 16);
      yr = uint32_t(r.y0) | (uint32_t(r.y1) << 16);
    }
  }
  const int incl = wave_inclusive_scan(rect_tiles(xr, yr, sp.lg).count);
  if (lane == 63 && (i >> 6) < sp.num_waves) sp.rect_wave[i >> 6] = incl;
}

// *total = the slot count after the scan (int64), or -1 when it overflowed int32.
__global__ void k_rect_total(const int32_t* __restrict__ rect_off, int num_waves,
                             const int32_t* __restrict__ rect_status, int64_t* __restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *total = *rect_status == SSPLAT_OK ? int64_t(rect_off[num_waves]) : int64_t(-1);
  }
}

// Per Gaussian: the sum of its weight exp(-q/2) (with the support fade applied) over the pixels
// of its clipped support rectangle, without opacity (render.py:417-442).  Each wave takes one
// Gaussian at a time (grid-stride over the Gaussians) and sweeps its rectangle in 8x8 pixel
// blocks, one pixel per lane; every lane sums its pixels in sweep order and a fixed DPP tree
// adds the lanes, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_activity(const float* __restrict__ means,
                                                  const float* __restrict__ conics,
                                                  const int64_t* __restrict__ radii, int n,
                                                  FieldTable fields, float fade_alpha,
                                                  float tail_exp, float* __restrict__ activity) {
  const int lane = threadIdx.x & 63;
  const int lx = lane & 7, ly = lane >> 3;
  const int waves = static_cast<int>(gridDim.x) * (kFieldWaves);
  const bool fade = fade_alpha > 0.0f;
  const float tail = fade ? fade_alpha * exp2_hw(tail_exp) : 0.0f;
  const int first = __builtin_amdgcn_readfirstlane(
      static_cast<int>(blockIdx.x) * kFieldWaves + static_cast<int>(threadIdx.x >> 6));
  for (int g = first; g < n; g += waves) {
    const ssplat_field f = field_of_gaussian(fields, g);
    const float mx = means[2 * g + 0];
    const float my = means[2 * g + 1];
    const Rect r = support_rect(mx, my, radii[2 * g + 0], radii[2 * g + 1], f.height, f.width);
    // The quadratic in exp2 units, as the render kernels evaluate it: (a, 2b, c) scaled by
    // -log2(e)/2 (with fade, the scale is applied after so that a pixel at exactly the cutoff
    // gives the tail's value).
    const float scale = fade ? 1.0f : kNegHalfLog2e;
    const float qa = scale * conics[3 * g + 0];
    const float qb = scale * (2.0f * conics[3 * g + 1]);
    const float qc = scale * conics[3 * g + 2];
    float sum = 0.0f;
    for (int by = r.y0; by < r.y1; by += 8) {
      for (int bx = r.x0; bx < r.x1; bx += 8) {
        const int px = bx + lx, py = by + ly;
        if (px < r.x1 && py < r.y1) {
          const float dx = static_cast<float>(px) - mx;
          const float dy = static_cast<float>(py) - my;
          const float e = fmaf(fmaf(qa, dx, qb * dy), dx, qc * dy * dy);
          const float raw = exp2_hw(fade ? e * kNegHalfLog2e : e);
          sum += fade ? fmaxf(raw - tail, 0.0f) : raw;
        }
      }
    }
    sum = wave_sum_lane63(sum);
    if (lane == 63) activity[g] = sum;
  }
}

// ------------------------------------------------------------------------------------------
// Host-side argument handling
// ------------------------------------------------------------------------------------------

// Geometry of one call: the field table (one by-value field, or a device table) and totals.
struct Plan {
  FieldTable table;
  int64_t n;        // Gaussians over all fields
  int num_tiles;    // tiles over all fields
  int ts;
  bool persistent;  // render kernels as one persistent launch (SSPLAT_TILE_PERSISTENT)
};

// Splits a tile_size argument into the tile edge and the schedule flag.
int check_tile_size(int tile_size, int* ts, bool* persistent) {
  *persistent = (tile_size & SSPLAT_TILE_PERSISTENT) != 0;
  *ts = tile_size & ~SSPLAT_TILE_PERSISTENT;
  if (*ts <= 0 || *ts * *ts > 1024) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT,
                     "tile_size must be positive with tile_size^2 <= 1024");
  }
  return SSPLAT_OK;
}

int check_image(int H, int W) {
  if (H <= 0 || W <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "height and width must be positive");
  }
  if (H > 65535 || W > 65535) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "height and width must be <= 65535");
  }
  return SSPLAT_OK;
}

int check_totals(int64_t n, int64_t tiles) {
  if (n < 0 || n > INT32_MAX) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "n must be in [0, 2^31)");
  if (tiles <= 0 || tiles + 1 > kMaxScanItems) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile grid of %lld tiles is out of range",
                     static_cast<long long>(tiles));
  }
  return SSPLAT_OK;
}

int make_single_plan(int64_t n, int H, int W, int tile_size, Plan* p) {
  int rc;
  int ts;
  bool persistent;
  if (n < 0 || n > INT32_MAX) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "n must be in [0, 2^31)");
  if ((rc = check_image(H, W)) || (rc = check_tile_size(tile_size, &ts, &persistent))) return rc;
  const int64_t tx = (int64_t(W) + ts - 1) / ts;
  const int64_t ty = (int64_t(H) + ts - 1) / ts;
  if ((rc = check_totals(n, tx * ty))) return rc;
  ssplat_field f{};
  f.height = H;
  f.width = W;
  f.first_gaussian = 0;
  f.num_gaussians = n;
  f.first_pixel = 0;
  f.first_tile = 0;
  f.tiles_x = static_cast<int32_t>(tx);
  f.num_tiles = static_cast<int32_t>(tx * ty);
  p->table = FieldTable{nullptr, 0, f};
  p->n = n;
  p->num_tiles = f.num_tiles;
  p->ts = ts;
  p->persistent = persistent;
  return SSPLAT_OK;
}

int make_batch_plan(const ssplat_field* fields, int num_fields, int64_t total_gaussians,
                    int64_t total_tiles, int tile_size, Plan* p) {
  int rc;
  int ts;
  bool persistent;
  if (fields == nullptr || num_fields <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "the field table must hold at least one field");
  }
  if ((rc = check_tile_size(tile_size, &ts, &persistent)) ||
      (rc = check_totals(total_gaussians, total_tiles))) {
    return rc;
  }
  p->table = FieldTable{fields, num_fields, ssplat_field{}};
  p->n = total_gaussians;
  p->num_tiles = static_cast<int>(total_tiles);
  p->ts = ts;
  p->persistent = persistent;
  return SSPLAT_OK;
}

struct Workspace {
  GaussRec* recs;
  int32_t* counts;       // num_tiles + 1 per-tile counters (binning), reused as scatter cursors
  int64_t* bsums;        // first-level scan block sums
  int32_t* status;       // internal status word of the last binning
  int32_t* rect_wave;    // slots per wave of 64 Gaussians (num_waves + 1, the last 0)
  int32_t* rect_off;     // their exclusive scan: each wave's first slot; [num_waves] = total
  int64_t* rect_bsums;   // first-level block sums of that scan
  int32_t* rect_status;  // int32 overflow of that scan
  int32_t* queue_fwd;    // kQueues work-queue counters of a persistent forward
  int32_t* queue_bwd;    // and of a persistent backward
  StripArgs strip;       // strip binning buffers (strip.strips == 0: the per-pair kernels bin)
  int num_waves;
  size_t render_bytes;   // the part the render calls use (the strip buffers follow it)
  size_t bytes;
};

// Level-1 LDS of the two counter arrays (strips + 1 each, padded to 8 bytes) and its list.
int strip_counter_bytes(int strips) { return 2 * ((strips + 2) & ~1) * int(sizeof(int32_t)); }

// Strips of a tile grid: 16 tiles each, or 32 when 16 would make more than kStripMax strips
// (none when 32 does too: the per-pair kernels bin such grids); fewer tiles a strip, down to
// one, when 16 would leave fewer than kStripTarget strips, so a small grid still spreads its
// pairs over many level-2 workgroups.  Chunks: kStripChunkUnit Gaussians, or the multiple of
// it that keeps them within kStripMaxChunks.
void strip_geometry(int64_t n, int num_tiles, StripArgs* sa) {
  sa->lg = 0;
  sa->strips = 0;
  sa->num_tiles = num_tiles;
  auto strips_at = [&](int l) { return (int64_t(num_tiles) + (1 << l) - 1) >> l; };
  int lg = 4;
  if (strips_at(4) > kStripMax) {
    if (strips_at(5) > kStripMax) return;
    lg = 5;
  } else {
    while (lg > 0 && strips_at(lg) < kStripTarget) --lg;
  }
  sa->lg = lg;
  sa->strips = static_cast<int>(strips_at(lg));
  const int64_t per = int64_t(kStripMaxChunks) * kStripChunkUnit;
  const int64_t units = (n + per - 1) / per;
  const int64_t chunk = (units > 1 ? units : 1) * kStripChunkUnit;
  if (chunk > kStripMaxChunk) {  // beyond the packed entry's ids: the per-pair kernels bin
    sa->lg = 0;
    sa->strips = 0;
    return;
  }
  sa->chunk = static_cast<int>(chunk);
  const int64_t chunks = (n + sa->chunk - 1) / sa->chunk;
  sa->chunks = static_cast<int>(chunks > 1 ? chunks : 1);
  sa->chunk_cap = kStripEntriesPerGaussian * sa->chunk;
  // The level-1 list takes what its LDS leaves after the two counter arrays.
  sa->list_cap = (kStripL1Lds - strip_counter_bytes(sa->strips)) / int(sizeof(int2));
}

Workspace layout(void* base, int64_t n, int num_tiles) {
  Workspace w{};
  size_t off = 0;
  const size_t recs = off;
  off = align_up(off + size_t(n) * sizeof(GaussRec), 256);
  const size_t counts = off;
  off = align_up(off + size_t(num_tiles + 1) * sizeof(int32_t), 256);
  const size_t bsums = off;
  const size_t blocks = (size_t(num_tiles) + 1 + kScanBlock - 1) / kScanBlock;
  off = align_up(off + blocks * sizeof(int64_t), 256);
  const size_t status = off;
  off = align_up(off + sizeof(int32_t), 256);
  const size_t waves = (size_t(n) + 63) / 64;
  const size_t rect_wave = off;
  off = align_up(off + (waves + 1) * sizeof(int32_t), 256);
  const size_t rect_off = off;
  off = align_up(off + (waves + 1) * sizeof(int32_t), 256);
  const size_t rect_bsums = off;
  off = align_up(off + ((waves + 1 + kScanBlock - 1) / kScanBlock) * sizeof(int64_t), 256);
  const size_t rect_status = off;
  off = align_up(off + sizeof(int32_t), 256);
  const size_t queues = off;
  off = align_up(off + 2 * kQueues * sizeof(int32_t), 256);
  const size_t render_bytes = off;
  // Strip binning: the chunks' offset matrix and regions.
  StripArgs sa{};
  strip_geometry(n, num_tiles, &sa);
  const bool strips = sa.strips > 0;
  const size_t offs = off;
  if (strips) off = align_up(off + size_t(sa.chunks) * (sa.strips + 1) * sizeof(int2), 256);
  const size_t bucket = off;
  if (strips) off = align_up(off + size_t(sa.chunks) * sa.chunk_cap * sizeof(int2), 256);
  char* b = static_cast<char*>(base);
  w.recs = reinterpret_cast<GaussRec*>(b + recs);
  w.counts = reinterpret_cast<int32_t*>(b + counts);
  w.bsums = reinterpret_cast<int64_t*>(b + bsums);
  w.status = reinterpret_cast<int32_t*>(b + status);
  w.rect_wave = reinterpret_cast<int32_t*>(b + rect_wave);
  w.rect_off = reinterpret_cast<int32_t*>(b + rect_off);
  w.rect_bsums = reinterpret_cast<int64_t*>(b + rect_bsums);
  w.rect_status = reinterpret_cast<int32_t*>(b + rect_status);
  w.queue_fwd = reinterpret_cast<int32_t*>(b + queues);
  w.queue_bwd = w.queue_fwd + kQueues;
  w.strip = sa;
  w.strip.offs = reinterpret_cast<int2*>(b + offs);
  w.strip.bucket = reinterpret_cast<int2*>(b + bucket);
  w.num_waves = static_cast<int>(waves);
  w.render_bytes = render_bytes;
  w.bytes = off;
  return w;
}

// binning: the call bins (it needs the strip buffers too, not only the render part).
int check_workspace(void* ws, size_t ws_bytes, const Plan& p, Workspace* w,
                    bool binning = false) {
  *w = layout(ws, p.n, p.num_tiles);
  if (ws == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "workspace is NULL");
  if (reinterpret_cast<uintptr_t>(ws) % 256 != 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  }
  const size_t need = binning ? w->bytes : w->render_bytes;
  if (ws_bytes < need) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "workspace has %zu bytes, needs %zu", ws_bytes,
                     need);
  }
  return SSPLAT_OK;
}

int make_bin_args(const float* means, const float* conics, const int64_t* radii, const Plan& p,
                  int cull, double sigma_cutoff, BinArgs* a) {
  if (p.n > 0 && (means == nullptr || radii == nullptr || (cull && conics == nullptr))) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "means/conics/radii must not be NULL");
  }
  if (!std::isfinite(sigma_cutoff)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "sigma_cutoff must be finite");
  }
  a->means = means;
  a->conics = conics;
  a->radii = radii;
  a->fields = p.table;
  a->n = static_cast<int>(p.n);
  a->ts = p.ts;
  a->cull = cull ? 1 : 0;
  // Same float32 arithmetic as the reference: cutoff2 = float(sigma^2), guard 1e-5*max(cutoff2,1)
  // (render_ext.cu:187-188, :920).
  const float cutoff2 = static_cast<float>(sigma_cutoff * sigma_cutoff);
  volatile float guard = 1e-5f * (cutoff2 > 1.0f ? cutoff2 : 1.0f);
  volatile float thr = cutoff2 + guard;
  a->cull_thr = thr;
  return SSPLAT_OK;
}

// Exclusive int32 scan of counts[0, items) into offsets (saturating; *status flags an overflow):
// one launch up to kSingleScanItems items, three beyond.
int run_scan(const int32_t* counts, int64_t items, int32_t* offsets, int64_t* bsums,
             int32_t* status, hipStream_t stream, uint32_t* bump = nullptr,
             int32_t* queue = nullptr) {
  if (items > kMaxScanItems) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "scan of %lld items is out of range",
                     static_cast<long long>(items));
  }
  if (items <= kSingleScanItems) {
    k_scan_single<<<1, kScanThreads, 0, stream>>>(counts, items, offsets, status, bump, queue);
    SSPLAT_LAUNCH_CHECK("k_scan_single");
    return SSPLAT_OK;
  }
  const int blocks = static_cast<int>((items + kScanBlock - 1) / kScanBlock);
  k_scan_partials<<<blocks, kScanThreads, 0, stream>>>(counts, items, bsums);
  SSPLAT_LAUNCH_CHECK("k_scan_partials");
  k_scan_top<<<1, kScanThreads, 0, stream>>>(bsums, blocks, bump, queue);
  SSPLAT_LAUNCH_CHECK("k_scan_top");
  k_scan_apply<<<blocks, kScanThreads, 0, stream>>>(counts, items, bsums, offsets, status);
  SSPLAT_LAUNCH_CHECK("k_scan_apply");
  return SSPLAT_OK;
}

// Counts and scans: offsets[0..T] written, the counters left holding the per-tile counts (the
// scatter's cursors).  Three launches (reset, count, one-workgroup scan) up to kSingleScanItems
// counters, five beyond.
int run_count(const BinArgs& a, const Plan& p, const Workspace& w, int32_t* offsets,
              hipStream_t stream) {
  const int64_t items = int64_t(p.num_tiles) + 1;
  {
    const int64_t blocks64 = (items + 255) / 256;
    const int blocks = static_cast<int>(blocks64 < 4096 ? blocks64 : 4096);
    k_bin_reset<<<blocks, 256, 0, stream>>>(w.counts, items, w.status);
    SSPLAT_LAUNCH_CHECK("k_bin_reset");
  }
  if (a.n > 0) {
    k_bin_count<<<(a.n + kBinThreads - 1) / kBinThreads, kBinThreads, 0, stream>>>(a, w.counts);
    SSPLAT_LAUNCH_CHECK("k_bin_count");
  }
  return run_scan(w.counts, items, offsets, w.bsums, w.status, stream);
}

// Scatter + per-tile sort (which also merges the status).  rebuild_cursors: the counters do not
// hold run_count's per-tile counts (a fill issued on its own), so they are rebuilt from offsets.
int run_fill(const BinArgs& a, const Plan& p, const Workspace& w, const int32_t* offsets,
             int32_t* gids, int64_t capacity, int32_t* status, bool rebuild_cursors,
             hipStream_t stream) {
  if (rebuild_cursors && p.num_tiles > 0) {
    k_cursors_from_offsets<<<(p.num_tiles + 255) / 256, 256, 0, stream>>>(offsets, p.num_tiles,
                                                                          w.counts);
    SSPLAT_LAUNCH_CHECK("k_cursors_from_offsets");
  }
  if (a.n > 0) {
    k_bin_scatter<<<(a.n + kBinThreads - 1) / kBinThreads, kBinThreads, 0, stream>>>(
        a, offsets, w.counts, gids, capacity, w.status);
    SSPLAT_LAUNCH_CHECK("k_bin_scatter");
  }
  const int sort_blocks = (p.num_tiles + kSortThreads / 64 - 1) / (kSortThreads / 64);
  if (a.n > 0 && sort_blocks > 0) {
    k_sort_tiles<<<sort_blocks, kSortThreads, 0, stream>>>(gids, offsets, p.num_tiles, capacity,
                                                            w.status, status);
    SSPLAT_LAUNCH_CHECK("k_sort_tiles");
  } else if (status) {
    k_merge_status<<<1, 64, 0, stream>>>(w.status, status);
    SSPLAT_LAUNCH_CHECK("k_merge_status");
  }
  return SSPLAT_OK;
}

// Strip binning (two launches): mode kStripCountOnly writes offsets, kStripBuild offsets and
// gids, kStripFill gids placed by the given offsets.
int run_strip_bin(const BinArgs& a, const Workspace& w, int mode, int32_t* offsets,
                  int32_t* gids, int64_t capacity, int32_t* status, hipStream_t stream) {
  const StripArgs& sa = w.strip;
  const size_t lds = kStripL1Lds;
  k_strip_level1<<<sa.chunks, kStripL1Threads, lds, stream>>>(a, sa, w.status);
  SSPLAT_LAUNCH_CHECK("k_strip_level1");
  if (mode == kStripCountOnly) {
    k_strip_level2<kStripCountOnly><<<sa.strips, kStripL2Threads, 0, stream>>>(
        a, sa, offsets, gids, capacity, w.status, status);
  } else if (mode == kStripBuild) {
    k_strip_level2<kStripBuild><<<sa.strips, kStripL2Threads, 0, stream>>>(
        a, sa, offsets, gids, capacity, w.status, status);
  } else {
    k_strip_level2<kStripFill><<<sa.strips, kStripL2Threads, 0, stream>>>(
        a, sa, offsets, gids, capacity, w.status, status);
  }
  SSPLAT_LAUNCH_CHECK("k_strip_level2");
  return SSPLAT_OK;
}

int tile_index_count(const float* means, const float* conics, const int64_t* radii,
                     const Plan& p, int cull, double sigma_cutoff, int32_t* offsets, void* ws,
                     size_t ws_bytes, void* stream) {
  int rc;
  if (offsets == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_offsets is NULL");
  Workspace w;
  if ((rc = check_workspace(ws, ws_bytes, p, &w, true))) return rc;
  BinArgs a;
  if ((rc = make_bin_args(means, conics, radii, p, cull, sigma_cutoff, &a))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (w.strip.strips > 0) {
    rc = run_strip_bin(a, w, kStripCountOnly, offsets, nullptr, 0, nullptr, s);
  } else {
    rc = run_count(a, p, w, offsets, s);
  }
  if (rc) return rc;
  return success();
}

int tile_index_fill(const float* means, const float* conics, const int64_t* radii,
                    const Plan& p, int cull, double sigma_cutoff, const int32_t* offsets,
                    int32_t* gids, int64_t capacity, int32_t* status, void* ws, size_t ws_bytes,
                    bool rebuild_cursors, void* stream) {
  int rc;
  if (offsets == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_offsets is NULL");
  if (capacity < 0 || (capacity > 0 && gids == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_gids is NULL or capacity negative");
  }
  Workspace w;
  if ((rc = check_workspace(ws, ws_bytes, p, &w, true))) return rc;
  BinArgs a;
  if ((rc = make_bin_args(means, conics, radii, p, cull, sigma_cutoff, &a))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (w.strip.strips > 0) {
    rc = run_strip_bin(a, w, kStripFill, const_cast<int32_t*>(offsets), gids, capacity, status, s);
  } else {
    rc = run_fill(a, p, w, offsets, gids, capacity, status, rebuild_cursors, s);
  }
  if (rc) return rc;
  return success();
}

// Both binning phases in one call: offsets and lists, no host synchronization.
int build_tile_index(const float* means, const float* conics, const int64_t* radii,
                     const Plan& p, int cull, double sigma_cutoff, int32_t* offsets,
                     int32_t* gids, int64_t capacity, int32_t* status, void* ws, size_t ws_bytes,
                     void* stream) {
  int rc;
  if (offsets == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_offsets is NULL");
  if (capacity < 0 || (capacity > 0 && gids == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_gids is NULL or capacity negative");
  }
  Workspace w;
  if ((rc = check_workspace(ws, ws_bytes, p, &w, true))) return rc;
  if (w.strip.strips == 0) {
    rc = tile_index_count(means, conics, radii, p, cull, sigma_cutoff, offsets, ws, ws_bytes,
                          stream);
    if (rc) return rc;
    return tile_index_fill(means, conics, radii, p, cull, sigma_cutoff, offsets, gids, capacity,
                           status, ws, ws_bytes, false, stream);
  }
  BinArgs a;
  if ((rc = make_bin_args(means, conics, radii, p, cull, sigma_cutoff, &a))) return rc;
  if ((rc = run_strip_bin(a, w, kStripBuild, offsets, gids, capacity, status,
                          static_cast<hipStream_t>(stream)))) {
    return rc;
  }
  return success();
}

// Support-fade strength, clamped to [0, 1] like the reference (render.py:94-98).
int check_fade(double alpha, double* out) {
  if (!std::isfinite(alpha)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "support_fade_alpha must be finite");
  }
  *out = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
  return SSPLAT_OK;
}

int check_render_inputs(const float* means, const float* conics, const float* colors,
                        const int64_t* radii, const Plan& p, const int32_t* gids,
                        const int32_t* offsets, int64_t num_gids, int normalize, double eps) {
  if (p.n > 0 && (means == nullptr || conics == nullptr || colors == nullptr || radii == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "means/conics/colors/radii must not be NULL");
  }
  if (offsets == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_offsets is NULL");
  if (num_gids < 0 || (num_gids > 0 && gids == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "tile_gids is NULL or num_gids negative");
  }
  if (p.ts != 16 && p.ts != 32) {
    return set_error(SSPLAT_ERR_UNSUPPORTED,
                     "the render kernels are built for tile_size 16 and 32, got %d", p.ts);
  }
  if (normalize && !(std::isfinite(eps) && eps > 0.0)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "normalization_eps must be finite and > 0");
  }
  return SSPLAT_OK;
}

RenderArgs make_render_args(const Workspace& w, const Plan& p, const int32_t* gids,
                            const int32_t* offsets, int64_t num_gids, double eps,
                            double fade_alpha, double sigma_cutoff) {
  RenderArgs r;
  r.recs = w.recs;
  r.arecs = nullptr;
  r.rect_off = nullptr;
  r.gids = gids;
  r.offsets = offsets;
  r.num_gids = num_gids;
  r.n = static_cast<int>(p.n);
  r.fields = p.table;
  r.eps = static_cast<float>(eps);
  r.tail_exp = static_cast<float>(sigma_cutoff * sigma_cutoff) * kNegHalfLog2e;
  r.fade_alpha = static_cast<float>(fade_alpha);
  return r;
}

int tile_lg(int ts) { return ts == 32 ? 5 : 4; }

SlotPrep make_slot_prep(const Plan& p, const Workspace& w) {
  return SlotPrep{w.rect_wave, w.num_waves, tile_lg(p.ts), w.rect_status};
}

// Packs the records with each Gaussian's slot bookkeeping; without support fade the conic is
// folded with -log2(e)/2.  The forward and the backward pack identical records.
int run_prepare(const float* means, const float* conics, const float* colors,
                const int64_t* radii, const float* opacities, const Plan& p, const Workspace& w,
                bool support_fade, hipStream_t stream) {
  if (p.n == 0) return SSPLAT_OK;
  const int blocks = static_cast<int>((p.n + 255) / 256);
  const SlotPrep sp = make_slot_prep(p, w);
  if (support_fade) {
    k_prepare<false><<<blocks, 256, 0, stream>>>(means, conics, colors, radii, opacities,
                                                 static_cast<int>(p.n), p.table, w.recs, sp,
                                                 w.queue_fwd);
  } else {
    k_prepare<true><<<blocks, 256, 0, stream>>>(means, conics, colors, radii, opacities,
                                                static_cast<int>(p.n), p.table, w.recs, sp,
                                                w.queue_fwd);
  }
  SSPLAT_LAUNCH_CHECK("k_prepare");
  return SSPLAT_OK;
}

// Kernel selection for one tile edge: support fade, normalization and the backward's normalizer
// source are template parameters (no per-pixel branches on them).
template <int kT, bool kAffine = false>
void launch_forward(bool fade, bool normalize, int tiles, const RenderArgs& ra, float* out,
                    float* den, hipStream_t s) {
  const dim3 grid(tiles), block(kT * kT);
  if (fade && normalize) {
    k_forward<kT, true, true, kAffine><<<grid, block, 0, s>>>(ra, out, den);
  } else if (fade) {
    k_forward<kT, true, false, kAffine><<<grid, block, 0, s>>>(ra, out, den);
  } else if (normalize) {
    k_forward<kT, false, true, kAffine><<<grid, block, 0, s>>>(ra, out, den);
  } else {
    k_forward<kT, false, false, kAffine><<<grid, block, 0, s>>>(ra, out, den);
  }
}

template <int kT, bool kFade, bool kAffine>
void launch_backward_normalized(int norm, int tiles, const RenderArgs& ra, const GradArgs& ga,
                                hipStream_t s) {
  const dim3 grid(tiles), block(kT * kT);
  if (norm == kNormSaved) {
    k_backward<kT, kFade, true, kNormSaved, kAffine><<<grid, block, 0, s>>>(ra, ga);
  } else if (norm == kNormRecompD) {
    k_backward<kT, kFade, true, kNormRecompD, kAffine><<<grid, block, 0, s>>>(ra, ga);
  } else {
    k_backward<kT, kFade, true, kNormRecompAll, kAffine><<<grid, block, 0, s>>>(ra, ga);
  }
}

template <int kT, bool kAffine = false>
void launch_backward(bool fade, bool normalize, int norm, int tiles, const RenderArgs& ra,
                     const GradArgs& ga, hipStream_t s) {
  const dim3 grid(tiles), block(kT * kT);
  if (!normalize) {
    if (fade) {
      k_backward<kT, true, false, kNormSaved, kAffine><<<grid, block, 0, s>>>(ra, ga);
    } else {
      k_backward<kT, false, false, kNormSaved, kAffine><<<grid, block, 0, s>>>(ra, ga);
    }
  } else if (fade) {
    launch_backward_normalized<kT, true, kAffine>(norm, tiles, ra, ga, s);
  } else {
    launch_backward_normalized<kT, false, kAffine>(norm, tiles, ra, ga, s);
  }
}

// Workgroups of a persistent launch: as many as the device holds at once, at most one a tile.
template <typename Kernel>
int persistent_grid(Kernel kernel, int block, int tiles, int* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  if (e != hipSuccess) {
    return set_error(SSPLAT_ERR_HIP, "persistent launch sizing: %s", hipGetErrorString(e));
  }
  const int64_t resident = int64_t(cus) * (per_cu > 0 ? per_cu : 1);
  *grid = static_cast<int>(resident < tiles ? resident : tiles);
  if (*grid < 1) *grid = 1;
  return SSPLAT_OK;
}

// Ranges of a persistent launch of `grid` workgroups: about kWorkgroupsPerQueue workgroups a
// range, 1 .. kMaxQueuesPerXcd ranges per XCD.
int queue_count(int grid) {
  int per_xcd = (grid / kXcds + kWorkgroupsPerQueue / 2) / kWorkgroupsPerQueue;
  per_xcd = per_xcd < 1 ? 1 : (per_xcd > kMaxQueuesPerXcd ? kMaxQueuesPerXcd : per_xcd);
  return kXcds * per_xcd;
}

template <int kT, bool kFade, bool kNormalize>
int launch_forward_persistent_one(TileQueue q, const RenderArgs& ra, float* out, float* den,
                                  hipStream_t s) {
  auto kernel = k_forward_persistent<kT, kFade, kNormalize>;
  int grid = 0;
  const int rc = persistent_grid(kernel, kT * kT, q.num_tiles, &grid);
  if (rc) return rc;
  q.queues = queue_count(grid);
  kernel<<<grid, kT * kT, 0, s>>>(ra, q, out, den);
  return SSPLAT_OK;
}

template <int kT>
int launch_forward_persistent(bool fade, bool normalize, const TileQueue& q, const RenderArgs& ra,
                              float* out, float* den, hipStream_t s) {
  if (fade && normalize) return launch_forward_persistent_one<kT, true, true>(q, ra, out, den, s);
  if (fade) return launch_forward_persistent_one<kT, true, false>(q, ra, out, den, s);
  if (normalize) return launch_forward_persistent_one<kT, false, true>(q, ra, out, den, s);
  return launch_forward_persistent_one<kT, false, false>(q, ra, out, den, s);
}

template <int kT, bool kFade, bool kNormalize, int kNorm>
int launch_backward_persistent_one(TileQueue q, const RenderArgs& ra, const GradArgs& ga,
                                   hipStream_t s) {
  auto kernel = k_backward_persistent<kT, kFade, kNormalize, kNorm>;
  int grid = 0;
  const int rc = persistent_grid(kernel, kT * kT, q.num_tiles, &grid);
  if (rc) return rc;
  q.queues = queue_count(grid);
  kernel<<<grid, kT * kT, 0, s>>>(ra, ga, q);
  return SSPLAT_OK;
}

template <int kT, bool kFade>
int launch_backward_persistent_norm(int norm, const TileQueue& q, const RenderArgs& ra,
                                    const GradArgs& ga, hipStream_t s) {
  if (norm == kNormSaved) return launch_backward_persistent_one<kT, kFade, true, kNormSaved>(q, ra, ga, s);
  if (norm == kNormRecompD) {
    return launch_backward_persistent_one<kT, kFade, true, kNormRecompD>(q, ra, ga, s);
  }
  return launch_backward_persistent_one<kT, kFade, true, kNormRecompAll>(q, ra, ga, s);
}

template <int kT>
int launch_backward_persistent(bool fade, bool normalize, int norm, const TileQueue& q,
                               const RenderArgs& ra, const GradArgs& ga, hipStream_t s) {
  if (!normalize) {
    return fade ? launch_backward_persistent_one<kT, true, false, kNormSaved>(q, ra, ga, s)
                : launch_backward_persistent_one<kT, false, false, kNormSaved>(q, ra, ga, s);
  }
  return fade ? launch_backward_persistent_norm<kT, true>(norm, q, ra, ga, s)
              : launch_backward_persistent_norm<kT, false>(norm, q, ra, ga, s);
}

// Affine-color inputs (render.py:110-123): color gradients (n,2,3), scales (n,2), rotations (n,)
// and, in the backward, their gradient outputs.  The affine records live in the workspace
// after the plain layout (ssplat_affine_workspace_bytes).
struct Affine {
  const float* color_grads;
  const float* scales;
  const float* rotations;
  float* grad_color_grads;
  float* grad_scales;
  float* grad_rotations;
};

size_t affine_record_bytes(int64_t n) {
  return align_up(size_t(n) * kAffineChunks * sizeof(float4), 256);
}

// Checks the affine inputs and packs their records; *arecs gets their address.
int prepare_affine(const Plan& p, const Affine& af, void* ws, size_t ws_bytes,
                   const Workspace& w, bool backward, const float4** arecs, hipStream_t s) {
  if (p.ts != 16) {
    return set_error(SSPLAT_ERR_UNSUPPORTED, "affine colors render with tile_size 16 only");
  }
  if (p.n > 0 && (af.color_grads == nullptr || af.scales == nullptr ||
                  af.rotations == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT,
                     "affine color rendering requires scales and rotations");
  }
  if (backward && p.n > 0 && (af.grad_color_grads == nullptr || af.grad_scales == nullptr ||
                              af.grad_rotations == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "affine gradient outputs must not be NULL");
  }
  if (ws_bytes < w.bytes + affine_record_bytes(p.n)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "workspace has %zu bytes, needs %zu", ws_bytes,
                     w.bytes + affine_record_bytes(p.n));
  }
  float4* a = reinterpret_cast<float4*>(static_cast<char*>(ws) + w.bytes);
  if (p.n > 0) {
    k_prepare_affine<<<static_cast<int>((p.n + 255) / 256), 256, 0, s>>>(
        af.scales, af.rotations, af.color_grads, static_cast<int>(p.n), a);
    SSPLAT_LAUNCH_CHECK("k_prepare_affine");
  }
  *arecs = a;
  return SSPLAT_OK;
}

int forward_tiled(const Plan& p, const float* means, const float* conics, const float* colors,
                  const int64_t* radii, const float* opacities, const int32_t* gids,
                  const int32_t* offsets, int64_t num_gids, int normalize, double eps,
                  double support_fade_alpha, double sigma_cutoff, float* out, float* den,
                  void* ws, size_t ws_bytes, void* stream, const Affine* af = nullptr) {
  int rc;
  double fade_alpha;
  if ((rc = check_fade(support_fade_alpha, &fade_alpha))) return rc;
  const bool support_fade = fade_alpha > 0.0;
  if ((rc = check_render_inputs(means, conics, colors, radii, p, gids, offsets, num_gids,
                                normalize, eps))) {
    return rc;
  }
  if (out == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "out is NULL");
  Workspace w;
  if ((rc = check_workspace(ws, ws_bytes, p, &w))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float4* arecs = nullptr;
  if (af != nullptr && (rc = prepare_affine(p, *af, ws, ws_bytes, w, false, &arecs, s))) {
    return rc;
  }
  if ((rc = run_prepare(means, conics, colors, radii, opacities, p, w, support_fade, s))) {
    return rc;
  }
  RenderArgs ra = make_render_args(w, p, gids, offsets, num_gids, eps, fade_alpha, sigma_cutoff);
  ra.arecs = arecs;
  hipEvent_t t0;
  const bool timed = timer_begin(0, s, &t0);
  // A field without Gaussians has no record packing to zero the queue: grid launch.
  const bool persistent = p.persistent && af == nullptr && p.n > 0;
  const TileQueue q{w.queue_fwd, p.num_tiles, kXcds};
  if (af != nullptr) {
    launch_forward<16, true>(support_fade, normalize != 0, p.num_tiles, ra, out, den, s);
  } else if (persistent && p.ts == 16) {
    if ((rc = launch_forward_persistent<16>(support_fade, normalize != 0, q, ra, out, den, s))) {
      return rc;
    }
  } else if (persistent) {
    if ((rc = launch_forward_persistent<32>(support_fade, normalize != 0, q, ra, out, den, s))) {
      return rc;
    }
  } else if (p.ts == 16) {
    launch_forward<16>(support_fade, normalize != 0, p.num_tiles, ra, out, den, s);
  } else {
    launch_forward<32>(support_fade, normalize != 0, p.num_tiles, ra, out, den, s);
  }
  SSPLAT_LAUNCH_CHECK("k_forward");
  if (timed) timer_end(0, s, t0);
  return success();
}

int check_slots(void* grad_slots, int64_t slot_capacity) {
  if (slot_capacity < 0) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "slot_capacity < 0");
  if (grad_slots == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "grad_slots is NULL");
  if (reinterpret_cast<uintptr_t>(grad_slots) % 256 != 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "grad_slots must be 256-byte aligned");
  }
  return SSPLAT_OK;
}

int backward_tiled(const Plan& p, const float* grad_out, const float* means, const float* conics,
                   const float* colors, const int64_t* radii, const float* opacities,
                   const float* den, const float* out, const int32_t* gids,
                   const int32_t* offsets, int64_t num_gids, int normalize, double eps,
                   double support_fade_alpha, double sigma_cutoff, float* grad_means,
                   float* grad_conics, float* grad_colors, float* grad_opacities,
                   void* grad_slots, int64_t slot_capacity, int32_t* status, void* ws,
                   size_t ws_bytes, void* stream, const Affine* af = nullptr,
                   bool after_forward = false) {
  int rc;
  double fade_alpha;
  if ((rc = check_fade(support_fade_alpha, &fade_alpha))) return rc;
  const bool support_fade = fade_alpha > 0.0;
  if ((rc = check_render_inputs(means, conics, colors, radii, p, gids, offsets, num_gids,
                                normalize, eps))) {
    return rc;
  }
  if (grad_out == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "grad_out is NULL");
  if (normalize && den != nullptr && out == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT,
                     "a normalized backward given den also needs the forward's out");
  }
  // Normalized: saved den and out, saved out with D recomputed, or both recomputed.
  const int norm = den != nullptr ? kNormSaved : (out != nullptr ? kNormRecompD : kNormRecompAll);
  if (p.n > 0 && (grad_means == nullptr || grad_conics == nullptr || grad_colors == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "gradient outputs must not be NULL");
  }
  if ((rc = check_slots(grad_slots, slot_capacity))) return rc;
  Workspace w;
  if ((rc = check_workspace(ws, ws_bytes, p, &w))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p.n == 0) return success();
  float* gop = (opacities != nullptr) ? grad_opacities : nullptr;
  uint32_t* epoch = static_cast<uint32_t*>(grad_slots);
  float* slots = reinterpret_cast<float*>(static_cast<char*>(grad_slots) + kSlotHeaderBytes);
  const float4* arecs = nullptr;
  if (af != nullptr && (rc = prepare_affine(p, *af, ws, ws_bytes, w, true, &arecs, s))) {
    return rc;
  }
  // After a forward of the same inputs into this workspace, its records (and their slot
  // bookkeeping) are still there and are used as they are.
  if (!after_forward &&
      (rc = run_prepare(means, conics, colors, radii, opacities, p, w, support_fade, s))) {
    return rc;
  }
  // The slot offsets; the scan also advances the slot buffer's call counter (the call's tag).
  if ((rc = run_scan(w.rect_wave, int64_t(w.num_waves) + 1, w.rect_off, w.rect_bsums,
                     w.rect_status, s, epoch, w.queue_bwd))) {
    return rc;
  }
  RenderArgs ra = make_render_args(w, p, gids, offsets, num_gids, eps, fade_alpha, sigma_cutoff);
  ra.rect_off = w.rect_off;
  ra.arecs = arecs;
  const GradArgs ga{grad_out, den, out, slots, slot_capacity, epoch, gop != nullptr};
  hipEvent_t t0;
  const bool timed = timer_begin(1, s, &t0);
  const TileQueue q{w.queue_bwd, p.num_tiles, kXcds};
  if (af != nullptr) {
    launch_backward<16, true>(support_fade, normalize != 0, norm, p.num_tiles, ra, ga, s);
  } else if (p.persistent && p.ts == 16) {
    if ((rc = launch_backward_persistent<16>(support_fade, normalize != 0, norm, q, ra, ga, s))) {
      return rc;
    }
  } else if (p.persistent) {
    if ((rc = launch_backward_persistent<32>(support_fade, normalize != 0, norm, q, ra, ga, s))) {
      return rc;
    }
  } else if (p.ts == 16) {
    launch_backward<16>(support_fade, normalize != 0, norm, p.num_tiles, ra, ga, s);
  } else {
    launch_backward<32>(support_fade, normalize != 0, norm, p.num_tiles, ra, ga, s);
  }
  SSPLAT_LAUNCH_CHECK("k_backward");
  if (timed) timer_end(1, s, t0);
  const ReduceArgs red{w.recs,      static_cast<int>(p.n), tile_lg(p.ts), w.num_waves,
                       w.rect_off,  w.rect_status,         slots,         slot_capacity,
                       epoch,       grad_means,            grad_conics,   grad_colors,
                       gop,         af ? af->grad_color_grads : nullptr,
                       af ? af->grad_scales : nullptr,      af ? af->grad_rotations : nullptr,
                       status,      support_fade ? 1.0f : kInvNegHalfLog2e};
  const int blocks = static_cast<int>((p.n + 255) / 256);
  if (af != nullptr) {
    k_grad_reduce<kSlotWordsAffine, 18><<<blocks, 256, 0, s>>>(red);
  } else {
    k_grad_reduce<kSlotWords, 9><<<blocks, 256, 0, s>>>(red);
  }
  SSPLAT_LAUNCH_CHECK("k_grad_reduce");
  return success();
}

int gaussian_activity(const Plan& p, const float* means, const float* conics,
                      const int64_t* radii, double support_fade_alpha, double sigma_cutoff,
                      float* activity, void* stream) {
  int rc;
  double fade_alpha;
  if ((rc = check_fade(support_fade_alpha, &fade_alpha))) return rc;
  if (!std::isfinite(sigma_cutoff)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "sigma_cutoff must be finite");
  }
  if (p.n > 0 && (means == nullptr || conics == nullptr || radii == nullptr ||
                  activity == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "means/conics/radii/activity must not be NULL");
  }
  if (p.n == 0) return success();
  const int64_t blocks64 = (p.n + kFieldWaves - 1) / kFieldWaves;
  const int blocks = static_cast<int>(blocks64 < 2048 ? blocks64 : 2048);
  k_activity<<<blocks, 64 * kFieldWaves, 0, static_cast<hipStream_t>(stream)>>>(
      means, conics, radii, static_cast<int>(p.n), p.table, static_cast<float>(fade_alpha),
      static_cast<float>(sigma_cutoff * sigma_cutoff) * kNegHalfLog2e, activity);
  SSPLAT_LAUNCH_CHECK("k_activity");
  return success();
}

int rect_pair_count(const Plan& p, const float* means, const int64_t* radii, int64_t* total,
                    void* ws, size_t ws_bytes, void* stream) {
  int rc;
  if (total == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "total is NULL");
  if (p.ts != 16 && p.ts != 32) {
    return set_error(SSPLAT_ERR_UNSUPPORTED,
                     "the render kernels are built for tile_size 16 and 32, got %d", p.ts);
  }
  if (p.n > 0 && (means == nullptr || radii == nullptr)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "means/radii must not be NULL");
  }
  Workspace w;
  if ((rc = check_workspace(ws, ws_bytes, p, &w))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int blocks = static_cast<int>((p.n + 255) / 256);
  k_rect_count<<<blocks > 0 ? blocks : 1, 256, 0, s>>>(means, radii, static_cast<int>(p.n),
                                                       p.table, make_slot_prep(p, w));
  SSPLAT_LAUNCH_CHECK("k_rect_count");
  if ((rc = run_scan(w.rect_wave, int64_t(w.num_waves) + 1, w.rect_off, w.rect_bsums,
                     w.rect_status, s))) {
    return rc;
  }
  k_rect_total<<<1, 64, 0, s>>>(w.rect_off, w.num_waves, w.rect_status, total);
  SSPLAT_LAUNCH_CHECK("k_rect_total");
  return success();
}

// ------------------------------------------------------------------------------------------
// Allocation-free forward: the workspace of ssplat_forward_tiled_noalloc is a static device
// array (one copy per GPU, held until the process ends) instead of a caller's buffer.
// ------------------------------------------------------------------------------------------

constexpr int64_t kNoAllocMaxGaussians = int64_t(1) << 18;  // 262,144 packed records
constexpr int64_t kNoAllocMaxTiles = int64_t(1) << 16;      // e.g. 4096 x 4096 at 16 px
// The forward's layout at both maxima (records 16 MiB plus the tile counters and the slot
// bookkeeping), rounded up.
constexpr size_t kNoAllocScratchBytes = size_t(17) << 20;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up once per device with that device
// current.  The array's host-side name is a stand-in, not a device address: only this pointer
// may be handed to kernels.
int scratch_address(void** out) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    return set_error(SSPLAT_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  }
  if (dev < 0 || dev >= kMaxDevices) {
    return set_error(SSPLAT_ERR_UNSUPPORTED, "device %d is beyond the %d devices supported", dev,
                     kMaxDevices);
  }
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      return set_error(SSPLAT_ERR_HIP, "hipGetSymbolAddress(scratch): %s",
                       hipGetErrorString(e));
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return SSPLAT_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------

extern "C" {

int ssplat_abi_version(void) { return SSPLAT_ABI_VERSION; }

const char* ssplat_last_error(void) { return last_error(); }

int ssplat_kernel_timing_enable(int enable) {
  KernelTimer& t = kernel_timer();
  std::lock_guard<std::mutex> lock(t.mu);
  t.enabled = enable != 0;
  return success();
}

int ssplat_kernel_timing_read(int kernel, int64_t* launches, double* total_ms) {
  if (kernel != SSPLAT_KERNEL_FORWARD && kernel != SSPLAT_KERNEL_BACKWARD) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "kernel must be 0 (forward) or 1 (backward)");
  }
  if (launches == nullptr || total_ms == nullptr) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "launches and total_ms must not be NULL");
  }
  KernelTimer& t = kernel_timer();
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs;
  {
    std::lock_guard<std::mutex> lock(t.mu);
    pairs.swap(t.pairs[kernel]);
  }
  double sum = 0.0;
  int rc = SSPLAT_OK;
  for (auto& pr : pairs) {
    float ms = 0.0f;
    if (rc == SSPLAT_OK) {
      hipError_t e = hipEventSynchronize(pr.second);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, pr.first, pr.second);
      if (e != hipSuccess) {
        rc = set_error(SSPLAT_ERR_HIP, "kernel timing: %s", hipGetErrorString(e));
      }
    }
    sum += ms;
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  *launches = static_cast<int64_t>(pairs.size());
  *total_ms = sum;
  return rc == SSPLAT_OK ? success() : rc;
}

int ssplat_workspace_bytes(int64_t n, int height, int width, int tile_size, size_t* bytes) {
  if (bytes == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bytes is NULL");
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  *bytes = layout(nullptr, p.n, p.num_tiles).bytes;
  return success();
}

int ssplat_tile_index_count(const float* means, const float* conics, const int64_t* radii,
                            int64_t n, int height, int width, int tile_size, int ellipse_cull,
                            double sigma_cutoff, int32_t* tile_offsets, void* workspace,
                            size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return tile_index_count(means, conics, radii, p, ellipse_cull, sigma_cutoff, tile_offsets,
                          workspace, workspace_bytes, stream);
}

int ssplat_tile_index_fill(const float* means, const float* conics, const int64_t* radii,
                           int64_t n, int height, int width, int tile_size, int ellipse_cull,
                           double sigma_cutoff, const int32_t* tile_offsets, int32_t* tile_gids,
                           int64_t gid_capacity, int32_t* status, void* workspace,
                           size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return tile_index_fill(means, conics, radii, p, ellipse_cull, sigma_cutoff, tile_offsets,
                         tile_gids, gid_capacity, status, workspace, workspace_bytes, true, stream);
}

int ssplat_build_tile_index(const float* means, const float* conics, const int64_t* radii,
                            int64_t n, int height, int width, int tile_size, int ellipse_cull,
                            double sigma_cutoff, int32_t* tile_offsets, int32_t* tile_gids,
                            int64_t gid_capacity, int32_t* status, void* workspace,
                            size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return build_tile_index(means, conics, radii, p, ellipse_cull, sigma_cutoff, tile_offsets,
                          tile_gids, gid_capacity, status, workspace, workspace_bytes, stream);
}

int ssplat_forward_tiled(const float* means, const float* conics, const float* colors,
                         const int64_t* radii, const float* opacities, int64_t n,
                         const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids,
                         int height, int width, int tile_size, int normalize, double eps,
                         double support_fade_alpha, double sigma_cutoff, float* out,
                         float* den, void* workspace, size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return forward_tiled(p, means, conics, colors, radii, opacities, tile_gids, tile_offsets,
                       num_gids, normalize, eps, support_fade_alpha, sigma_cutoff, out, den,
                       workspace, workspace_bytes, stream);
}

int ssplat_forward_tiled_noalloc(const float* means, const float* conics, const float* colors,
                                 const int64_t* radii, const float* opacities, int64_t n,
                                 const int32_t* tile_gids, const int32_t* tile_offsets,
                                 int64_t num_gids, int height, int width, int tile_size,
                                 int normalize, double eps, double support_fade_alpha,
                                 double sigma_cutoff, float* out, float* den, void* stream) {
  Plan p;
  int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  const size_t need = layout(nullptr, p.n, p.num_tiles).render_bytes;
  if (p.n > kNoAllocMaxGaussians || p.num_tiles > kNoAllocMaxTiles ||
      need > kNoAllocScratchBytes) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT,
                     "ssplat_forward_tiled_noalloc: %lld Gaussians and %d tiles need %zu bytes of "
                     "scratch; the static scratch holds %zu (at most %lld Gaussians and %lld "
                     "tiles)",
                     static_cast<long long>(p.n), p.num_tiles, need, kNoAllocScratchBytes,
                     static_cast<long long>(kNoAllocMaxGaussians),
                     static_cast<long long>(kNoAllocMaxTiles));
  }
  void* ws = nullptr;
  if ((rc = scratch_address(&ws))) return rc;
  return forward_tiled(p, means, conics, colors, radii, opacities, tile_gids, tile_offsets,
                       num_gids, normalize, eps, support_fade_alpha, sigma_cutoff, out, den, ws,
                       kNoAllocScratchBytes, stream);
}

int ssplat_backward_tiled(const float* grad_out, const float* means, const float* conics,
                          const float* colors, const int64_t* radii, const float* opacities,
                          const float* den, const float* out, int64_t n,
                          const int32_t* tile_gids, const int32_t* tile_offsets,
                          int64_t num_gids, int height, int width, int tile_size, int normalize,
                          double eps, double support_fade_alpha, double sigma_cutoff,
                          float* grad_means, float* grad_conics, float* grad_colors,
                          float* grad_opacities, void* grad_slots, int64_t slot_capacity,
                          int32_t* status, void* workspace, size_t workspace_bytes,
                          void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return backward_tiled(p, grad_out, means, conics, colors, radii, opacities, den, out, tile_gids,
                        tile_offsets, num_gids, normalize, eps, support_fade_alpha, sigma_cutoff,
                        grad_means, grad_conics, grad_colors, grad_opacities, grad_slots,
                        slot_capacity, status, workspace, workspace_bytes, stream);
}

int ssplat_backward_tiled_after_forward(
    const float* grad_out, const float* means, const float* conics, const float* colors,
    const int64_t* radii, const float* opacities, const float* den, const float* out, int64_t n,
    const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids, int height,
    int width, int tile_size, int normalize, double eps, double support_fade_alpha,
    double sigma_cutoff, float* grad_means, float* grad_conics, float* grad_colors,
    float* grad_opacities, void* grad_slots, int64_t slot_capacity, int32_t* status,
    void* workspace, size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return backward_tiled(p, grad_out, means, conics, colors, radii, opacities, den, out, tile_gids,
                        tile_offsets, num_gids, normalize, eps, support_fade_alpha, sigma_cutoff,
                        grad_means, grad_conics, grad_colors, grad_opacities, grad_slots,
                        slot_capacity, status, workspace, workspace_bytes, stream, nullptr, true);
}

int ssplat_forward_tiled_affine(const float* means, const float* conics, const float* colors,
                                const int64_t* radii, const float* opacities,
                                const float* color_grads, const float* scales,
                                const float* rotations, int64_t n, const int32_t* tile_gids,
                                const int32_t* tile_offsets, int64_t num_gids, int height,
                                int width, int tile_size, int normalize, double eps,
                                double support_fade_alpha, double sigma_cutoff, float* out,
                                float* den, void* workspace, size_t workspace_bytes,
                                void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  const Affine af{color_grads, scales, rotations, nullptr, nullptr, nullptr};
  return forward_tiled(p, means, conics, colors, radii, opacities, tile_gids, tile_offsets,
                       num_gids, normalize, eps, support_fade_alpha, sigma_cutoff, out, den,
                       workspace, workspace_bytes, stream, &af);
}

int ssplat_backward_tiled_affine(
    const float* grad_out, const float* means, const float* conics, const float* colors,
    const int64_t* radii, const float* opacities, const float* color_grads, const float* scales,
    const float* rotations, const float* den, const float* out, int64_t n,
    const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids, int height,
    int width, int tile_size, int normalize, double eps, double support_fade_alpha,
    double sigma_cutoff, float* grad_means, float* grad_conics, float* grad_colors,
    float* grad_opacities, float* grad_color_grads, float* grad_scales, float* grad_rotations,
    void* grad_slots, int64_t slot_capacity, int32_t* status, void* workspace,
    size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  const Affine af{color_grads, scales, rotations, grad_color_grads, grad_scales, grad_rotations};
  return backward_tiled(p, grad_out, means, conics, colors, radii, opacities, den, out, tile_gids,
                        tile_offsets, num_gids, normalize, eps, support_fade_alpha, sigma_cutoff,
                        grad_means, grad_conics, grad_colors, grad_opacities, grad_slots,
                        slot_capacity, status, workspace, workspace_bytes, stream, &af);
}

int ssplat_affine_workspace_bytes(int64_t n, int height, int width, int tile_size,
                                  size_t* bytes) {
  if (bytes == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bytes is NULL");
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  *bytes = layout(nullptr, p.n, p.num_tiles).bytes + affine_record_bytes(p.n);
  return success();
}

int ssplat_affine_grad_slots_bytes(int64_t slot_capacity, size_t* bytes) {
  if (bytes == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bytes is NULL");
  if (slot_capacity < 0 || slot_capacity > (int64_t(1) << 40)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "slot_capacity out of range");
  }
  *bytes = align_up(kSlotHeaderBytes + size_t(slot_capacity) * kSlotWordsAffine * sizeof(float),
                    256);
  return success();
}

int ssplat_grad_slots_bytes(int64_t slot_capacity, size_t* bytes) {
  if (bytes == nullptr) return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "bytes is NULL");
  if (slot_capacity < 0 || slot_capacity > (int64_t(1) << 40)) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "slot_capacity out of range");
  }
  *bytes = align_up(kSlotHeaderBytes + size_t(slot_capacity) * kSlotWords * sizeof(float), 256);
  return success();
}

int ssplat_gaussian_activity(const float* means, const float* conics, const int64_t* radii,
                             int64_t n, int height, int width, double support_fade_alpha,
                             double sigma_cutoff, float* activity, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, 16, &p);
  if (rc) return rc;
  return gaussian_activity(p, means, conics, radii, support_fade_alpha, sigma_cutoff, activity,
                           stream);
}

int ssplat_rect_pair_count(const float* means, const int64_t* radii, int64_t n, int height,
                           int width, int tile_size, int64_t* total, void* workspace,
                           size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_single_plan(n, height, width, tile_size, &p);
  if (rc) return rc;
  return rect_pair_count(p, means, radii, total, workspace, workspace_bytes, stream);
}

int ssplat_batch_plan(ssplat_field* fields, int num_fields, int tile_size,
                      int64_t* total_gaussians, int64_t* total_tiles, int64_t* total_pixels,
                      size_t* workspace_bytes) {
  int rc;
  if (fields == nullptr || num_fields <= 0) {
    return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "the field table must hold at least one field");
  }
  int edge;
  bool persistent;
  if ((rc = check_tile_size(tile_size, &edge, &persistent))) return rc;
  tile_size = edge;  // the schedule flag does not change the layout
  int64_t g = 0, t = 0, px = 0;
  for (int i = 0; i < num_fields; ++i) {
    ssplat_field& f = fields[i];
    if ((rc = check_image(f.height, f.width))) return rc;
    if (f.num_gaussians < 0) {
      return set_error(SSPLAT_ERR_INVALID_ARGUMENT, "field %d has a negative Gaussian count", i);
    }
    const int64_t tx = (int64_t(f.width) + tile_size - 1) / tile_size;
    const int64_t ty = (int64_t(f.height) + tile_size - 1) / tile_size;
    f.first_gaussian = g;
    f.first_pixel = px;
    f.first_tile = t;
    f.tiles_x = static_cast<int32_t>(tx);
    f.num_tiles = static_cast<int32_t>(tx * ty);
    g += f.num_gaussians;
    t += tx * ty;
    px += int64_t(f.height) * f.width;
    if ((rc = check_totals(g, t))) return rc;
  }
  if (total_gaussians) *total_gaussians = g;
  if (total_tiles) *total_tiles = t;
  if (total_pixels) *total_pixels = px;
  if (workspace_bytes) *workspace_bytes = layout(nullptr, g, static_cast<int>(t)).bytes;
  return success();
}

int ssplat_batch_build_tile_index(const ssplat_field* fields, int num_fields,
                                  int64_t total_gaussians, int64_t total_tiles, int tile_size,
                                  const float* means, const float* conics, const int64_t* radii,
                                  int ellipse_cull, double sigma_cutoff, int32_t* tile_offsets,
                                  int32_t* tile_gids, int64_t gid_capacity, int32_t* status,
                                  void* workspace, size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_batch_plan(fields, num_fields, total_gaussians, total_tiles, tile_size, &p);
  if (rc) return rc;
  return build_tile_index(means, conics, radii, p, ellipse_cull, sigma_cutoff, tile_offsets,
                          tile_gids, gid_capacity, status, workspace, workspace_bytes, stream);
}

int ssplat_batch_forward_tiled(const ssplat_field* fields, int num_fields,
                               int64_t total_gaussians, int64_t total_tiles, int tile_size,
                               const float* means, const float* conics, const float* colors,
                               const int64_t* radii, const float* opacities,
                               const int32_t* tile_gids, const int32_t* tile_offsets,
                               int64_t num_gids, int normalize, double eps,
                               double support_fade_alpha, double sigma_cutoff, float* out,
                               float* den, void* workspace,
                               size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_batch_plan(fields, num_fields, total_gaussians, total_tiles, tile_size, &p);
  if (rc) return rc;
  return forward_tiled(p, means, conics, colors, radii, opacities, tile_gids, tile_offsets,
                       num_gids, normalize, eps, support_fade_alpha, sigma_cutoff, out, den,
                       workspace, workspace_bytes, stream);
}

int ssplat_batch_backward_tiled(const ssplat_field* fields, int num_fields,
                                int64_t total_gaussians, int64_t total_tiles, int tile_size,
                                const float* grad_out, const float* means, const float* conics,
                                const float* colors, const int64_t* radii,
                                const float* opacities, const float* den, const float* out,
                                const int32_t* tile_gids,
                                const int32_t* tile_offsets, int64_t num_gids, int normalize,
                                double eps, double support_fade_alpha, double sigma_cutoff,
                                float* grad_means, float* grad_conics, float* grad_colors,
                                float* grad_opacities, void* grad_slots, int64_t slot_capacity,
                                int32_t* status, void* workspace, size_t workspace_bytes,
                                void* stream) {
  Plan p;
  const int rc = make_batch_plan(fields, num_fields, total_gaussians, total_tiles, tile_size, &p);
  if (rc) return rc;
  return backward_tiled(p, grad_out, means, conics, colors, radii, opacities, den, out, tile_gids,
                        tile_offsets, num_gids, normalize, eps, support_fade_alpha, sigma_cutoff,
                        grad_means, grad_conics, grad_colors, grad_opacities, grad_slots,
                        slot_capacity, status, workspace, workspace_bytes, stream);
}

int ssplat_batch_backward_tiled_after_forward(
    const ssplat_field* fields, int num_fields, int64_t total_gaussians, int64_t total_tiles,
    int tile_size, const float* grad_out, const float* means, const float* conics,
    const float* colors, const int64_t* radii, const float* opacities, const float* den,
    const float* out, const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids,
    int normalize, double eps, double support_fade_alpha, double sigma_cutoff,
    float* grad_means, float* grad_conics, float* grad_colors, float* grad_opacities,
    void* grad_slots, int64_t slot_capacity, int32_t* status, void* workspace,
    size_t workspace_bytes, void* stream) {
  Plan p;
  const int rc = make_batch_plan(fields, num_fields, total_gaussians, total_tiles, tile_size, &p);
  if (rc) return rc;
  return backward_tiled(p, grad_out, means, conics, colors, radii, opacities, den, out, tile_gids,
                        tile_offsets, num_gids, normalize, eps, support_fade_alpha, sigma_cutoff,
                        grad_means, grad_conics, grad_colors, grad_opacities, grad_slots,
                        slot_capacity, status, workspace, workspace_bytes, stream, nullptr, true);
}

}  // extern "C"
