/*
 * ssplat.h - C ABI of the MI355X (gfx950) sorting-free 2D Gaussian rasterizer.
 *
 * The library renders StructSplat's normalized weighted sum
 *     out(p) = sum_i c_i o_i G_i(p) / (sum_i o_i G_i(p) + eps)
 * and its additive variant out(p) = sum_i c_i o_i G_i(p), forward and backward, over the
 * clipped per-Gaussian support rectangles of the reference renderer.
 *
 * Conventions shared by every entry point:
 *   - All array arguments are DEVICE pointers to contiguous row-major memory owned by the
 *     caller.  The library never allocates device memory; scratch space is a caller-owned
 *     workspace whose size comes from ssplat_workspace_bytes().
 *   - `stream` is a hipStream_t (passed as void* so this header needs no HIP headers).  All
 *     work is enqueued on it; no entry point synchronizes with the host, so every call can be
 *     recorded into a HIP graph.
 *   - Return value: SSPLAT_OK (0) or an SSPLAT_ERR_* code; ssplat_last_error() returns a
 *     thread-local message for the last failing call on the calling thread.
 *   - Field layout (the reference's tensor contract, render_ext.cpp:94-125):
 *       means (N,2) f32, conics (N,3) f32 = (a, b, c) of inv(Sigma) = [[a,b],[b,c]],
 *       colors (N,3) f32, radii (N,2) int64 support half-widths, opacities (N,) f32 or NULL
 *       (NULL means opacity 1).  Images are (H,W,3) f32; coverage/den is (H,W) f32.
 *   - Tile index: tile_offsets (T+1,) int32 and tile_gids (P,) int32 where
 *       T = ceil(W/ts) * ceil(H/ts) and tile t owns tile_gids[tile_offsets[t] : tile_offsets[t+1]]
 *       in ascending Gaussian id.
 */
#ifndef SSPLAT_H_
#define SSPLAT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSPLAT_ABI_VERSION 6  /* 6: loss units (ssplat_loss_plan), target moments */

/*
 * Render schedule flag, OR-ed into the tile_size argument of the render entry points (the plain
 * and batched forwards and backwards, ssplat_workspace_bytes, ssplat_batch_plan): the forward and
 * backward kernels then run as one persistent launch (as many workgroups as the device holds at
 * once, each taking tiles from per-XCD work queues in the workspace, then from the other XCDs'
 * queues) instead of one workgroup per tile.  Results are bit-identical to the grid launch's.
 * The binning entry points take the plain tile edge.  No reference counterpart.
 */
#define SSPLAT_TILE_PERSISTENT 0x100

enum ssplat_status {
  SSPLAT_OK = 0,
  SSPLAT_ERR_INVALID_ARGUMENT = 1, /* bad shape, size, tile size or pointer */
  SSPLAT_ERR_CAPACITY = 2,         /* tile_gids or gradient slot capacity too small */
  SSPLAT_ERR_HIP = 3,              /* a HIP runtime call or kernel launch failed */
  SSPLAT_ERR_UNSUPPORTED = 4       /* valid request the library does not implement */
};

/* Version of this ABI (SSPLAT_ABI_VERSION of the build). */
int ssplat_abi_version(void);

/* Message describing the last failing call on this thread ("" after a success). */
const char* ssplat_last_error(void);

/*
 * Kernel timing, a measurement hook with no reference counterpart.  While enabled, every
 * forward / backward call records a HIP event pair on its stream around its render-kernel
 * launch (k_forward / k_backward, not the record packing or the gradient reduction).  ssplat_kernel_timing_read
 * waits for the recorded pairs of one kernel, returns their count and summed milliseconds, and
 * releases them.  Not for use while a stream is being captured into a HIP graph.
 */
enum ssplat_kernel { SSPLAT_KERNEL_FORWARD = 0, SSPLAT_KERNEL_BACKWARD = 1 };
int ssplat_kernel_timing_enable(int enable);
int ssplat_kernel_timing_read(int kernel, int64_t* launches, double* total_ms);

/*
 * Workspace bytes needed by the calls below for a field of n Gaussians on an HxW image binned
 * with tile_size.  One workspace may be reused by consecutive calls on one stream.
 * No reference counterpart: the reference allocates its scratch with the torch caching
 * allocator inside each call (render_ext.cu:889-919).
 */
int ssplat_workspace_bytes(int64_t n, int height, int width, int tile_size, size_t* bytes);

/*
 * Tile binning, phase 1: per-tile pair counts and their exclusive scan.
 * Writes tile_offsets[0..T]; tile_offsets[T] is the number of (tile, Gaussian) pairs P.
 * Replaces the counting half of build_tile_index (render_ext.cpp:139-167,
 * render_ext.cu:858-911: tile_count_kernel + cub::DeviceScan + the host readback).
 * ellipse_cull drops pairs whose support-fade weight is exactly zero on the whole tile
 * (render_ext.cu:97-133, 178-191); callers pass it only together with support fade.
 */
int ssplat_tile_index_count(const float* means, const float* conics, const int64_t* radii,
                            int64_t n, int height, int width, int tile_size, int ellipse_cull,
                            double sigma_cutoff, int32_t* tile_offsets, void* workspace,
                            size_t workspace_bytes, void* stream);

/*
 * Tile binning, phase 2: fills tile_gids from the offsets written by phase 1 on the same
 * workspace, in ascending Gaussian id within every tile (the stable order of the reference's
 * CUB radix sort, render_ext.cu:927-944).  gid_capacity is the length of tile_gids; when it is
 * smaller than tile_offsets[T] (or the pair count exceeds int32), pairs beyond it are dropped and
 * *status (a device int32, may be NULL) is set to SSPLAT_ERR_CAPACITY.  A successful fill leaves
 * *status untouched (sticky), so the caller clears it once and may check it after many calls.
 * Replaces expand_pairs_kernel + SortPairs + tile_offsets_from_keys_kernel (render_ext.cu:135-218,
 * :912-951).
 */
int ssplat_tile_index_fill(const float* means, const float* conics, const int64_t* radii,
                           int64_t n, int height, int width, int tile_size, int ellipse_cull,
                           double sigma_cutoff, const int32_t* tile_offsets, int32_t* tile_gids,
                           int64_t gid_capacity, int32_t* status, void* workspace,
                           size_t workspace_bytes, void* stream);

/*
 * Both phases in one call with no host synchronization (capturable in a HIP graph).
 * Equivalent of build_tile_index(means, conics, radii, H, W, tile_size, ellipse_cull,
 * sigma_cutoff) (render_ext.cpp:139-167) with a caller-chosen gid capacity.
 */
int ssplat_build_tile_index(const float* means, const float* conics, const int64_t* radii,
                            int64_t n, int height, int width, int tile_size, int ellipse_cull,
                            double sigma_cutoff, int32_t* tile_offsets, int32_t* tile_gids,
                            int64_t gid_capacity, int32_t* status, void* workspace,
                            size_t workspace_bytes, void* stream);

/*
 * Render calls read tile_gids[tile_offsets[t] : tile_offsets[t+1]] clamped to [0, num_gids),
 * where num_gids is the length of tile_gids, so a stale offset table cannot read past it.
 * The render kernels are built for tile_size 16 (the reference default, cuda_render.py:258)
 * and 32; other tile sizes return SSPLAT_ERR_UNSUPPORTED.  Both give the same images (every
 * pixel sums the same Gaussians in the same id order); 32 gives fewer (tile, Gaussian) pairs.
 *
 * support_fade_alpha in [0, 1] (clamped) subtracts alpha * exp(-cutoff^2/2) from every weight,
 * clamped at 0 (render.py:94-107): 0 renders the plain Gaussian, 1 is the reference's
 * support_fade (weight exactly 0 at sigma_cutoff), values in between are the reference's
 * fractional fade strength.
 *
 * Tiled forward.  Writes out (H,W,3) and, when den is not NULL, the raw coverage
 * den (H,W) = sum_i o_i G_i (no eps).  Every pixel is accumulated in ascending Gaussian id,
 * so the result is bitwise deterministic.
 * Replaces forward_tiled (render_ext.cpp:188-211, render_ext.cu:349-431, :1011-1063) and the
 * coverage forward used by render_cuda_exact_with_coverage (cuda_render.py:229-253).
 */
int ssplat_forward_tiled(const float* means, const float* conics, const float* colors,
                         const int64_t* radii, const float* opacities, int64_t n,
                         const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids,
                         int height, int width, int tile_size, int normalize, double eps,
                         double support_fade_alpha, double sigma_cutoff, float* out, float* den,
                         void* workspace, size_t workspace_bytes, void* stream);

/*
 * ssplat_forward_tiled without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Its scratch (the packed records and their bookkeeping) is a
 * static device array of 17 MiB, one copy per GPU, held until the process ends; a call takes at
 * most 262,144 Gaussians and 65,536 tiles (SSPLAT_ERR_INVALID_ARGUMENT beyond).  It calls no
 * allocating routine.  Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams or threads without ordering between them) race and must not be made.
 * Results are those of ssplat_forward_tiled, bit for bit.
 */
int ssplat_forward_tiled_noalloc(const float* means, const float* conics, const float* colors,
                                 const int64_t* radii, const float* opacities, int64_t n,
                                 const int32_t* tile_gids, const int32_t* tile_offsets,
                                 int64_t num_gids, int height, int width, int tile_size,
                                 int normalize, double eps, double support_fade_alpha,
                                 double sigma_cutoff, float* out, float* den, void* stream);

/*
 * Tiled backward.  Given grad_out (H,W,3) = dL/dout, writes dL/dmeans (N,2), dL/dconics (N,3),
 * dL/dcolors (N,3) and, when opacities and grad_opacities are both non-NULL, dL/dopacities (N,).
 * A normalized backward needs the per-pixel adjoints dL/dN_k = G_k/(den+eps) and
 * dL/dden = -sum_k G_k out_k/(den+eps) of the forward's den (H,W) and out (H,W,3):
 *   den and out given   : both read (the reference's saved tensors);
 *   out given, den NULL : den, the normalizer, is recomputed by a denominator-only sweep of
 *                         the tile lists instead of being stored;
 *   both NULL           : den and the numerator are recomputed, out = N/(den+eps).
 * The recomputed values equal the forward's bit for bit, so all three give the same gradients.
 * An additive call reads neither.  Every gradient output is written.
 *
 * The gradients are deterministic: they do not depend on scheduling, so repeated calls on the
 * same inputs give the same bits.  Gaussian g owns one "slot" per tile of its clipped support
 * rectangle, R(g) = tw * th slots (tile bounds x0 / tile_size .. (x1 - 1) / tile_size, likewise
 * in y; 0 for an empty rectangle), numbered consecutively over the Gaussians.  Every
 * (tile, Gaussian) pair of the lists stores its partial gradients in its slot of grad_slots, and
 * each Gaussian's slots are then summed in a fixed order; a slot the lists do not reach (a pair
 * left out by the ellipse cull) counts as zero.  grad_slots is a 256-byte aligned buffer of
 * ssplat_grad_slots_bytes(slot_capacity) bytes, ZERO-FILLED ONCE when allocated (its header
 * counts the calls that used it, and tags each slot with the call that wrote it); it may then
 * be reused by any number of calls on one stream.  slot_capacity must be at least the total
 * slot count sum_g R(g): that equals the pair count P of lists built without ellipse cull, and
 * ssplat_rect_pair_count computes it in general.  With too small a capacity the slots beyond it
 * are lost and *status (a device int32, may be NULL) is set to SSPLAT_ERR_CAPACITY (sticky,
 * as in ssplat_tile_index_fill).
 * Replaces backward_tiled (render_ext.cpp:269-301, render_ext.cu:696-847, :1169-1231), whose
 * float atomics make the gradients depend on scheduling.
 */
int ssplat_backward_tiled(const float* grad_out, const float* means, const float* conics,
                          const float* colors, const int64_t* radii, const float* opacities,
                          const float* den, const float* out,
                          int64_t n, const int32_t* tile_gids, const int32_t* tile_offsets,
                          int64_t num_gids, int height, int width, int tile_size, int normalize,
                          double eps, double support_fade_alpha, double sigma_cutoff,
                          float* grad_means, float* grad_conics, float* grad_colors,
                          float* grad_opacities, void* grad_slots, int64_t slot_capacity,
                          int32_t* status, void* workspace, size_t workspace_bytes,
                          void* stream);

/*
 * ssplat_backward_tiled for a backward that follows ssplat_forward_tiled of the same inputs
 * (means, conics, colors, radii, opacities, tile lists, image, tile size, fade) into the same
 * workspace, with nothing else written to that workspace in between: the forward's packed
 * records are used as they are instead of being packed again (one launch and ~13N bytes of
 * reads and 64N of writes less).  With any other workspace content the gradients are wrong;
 * use ssplat_backward_tiled then.  No reference counterpart.
 */
int ssplat_backward_tiled_after_forward(
    const float* grad_out, const float* means, const float* conics, const float* colors,
    const int64_t* radii, const float* opacities, const float* den, const float* out, int64_t n,
    const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids, int height,
    int width, int tile_size, int normalize, double eps, double support_fade_alpha,
    double sigma_cutoff, float* grad_means, float* grad_conics, float* grad_colors,
    float* grad_opacities, void* grad_slots, int64_t slot_capacity, int32_t* status,
    void* workspace, size_t workspace_bytes, void* stream);

/*
 * Affine (gradient) colors: Gaussian i's color at pixel p is
 *   c_i + G_i0 * lx + G_i1 * ly,   lx = ( cos t dx + sin t dy) / max(sx, 1e-6),
 *                                  ly = (-sin t dx + cos t dy) / max(sy, 1e-6),
 * (dx, dy) = p - mean_i, t = rotations[i], (sx, sy) = scales[i] (linear, not log), G_i the
 * (2,3) row of color_grads (N,2,3).  The weights, normalization and fade are those of
 * ssplat_forward_tiled.  Replaces the color_grads path of render_normalized_weighted_sum /
 * render_additive_weighted_sum (render.py:110-123, :190-197, :300-307; the reference has no
 * kernel for it and runs it as a torch loop over pixel chunks).  tile_size 16 only
 * (SSPLAT_ERR_UNSUPPORTED otherwise); the workspace holds
 * ssplat_affine_workspace_bytes(n, height, width, tile_size) bytes.
 */
int ssplat_forward_tiled_affine(const float* means, const float* conics, const float* colors,
                                const int64_t* radii, const float* opacities,
                                const float* color_grads, const float* scales,
                                const float* rotations, int64_t n, const int32_t* tile_gids,
                                const int32_t* tile_offsets, int64_t num_gids, int height,
                                int width, int tile_size, int normalize, double eps,
                                double support_fade_alpha, double sigma_cutoff, float* out,
                                float* den, void* workspace, size_t workspace_bytes,
                                void* stream);

/*
 * Backward of ssplat_forward_tiled_affine: as ssplat_backward_tiled, and also writes
 * dL/dcolor_grads (N,2,3), dL/dscales (N,2) and dL/drotations (N,).  A scale below 1e-6
 * (clamped in the forward) gets gradient 0.  grad_slots holds
 * ssplat_affine_grad_slots_bytes(slot_capacity) bytes (80 bytes a slot), zero-filled once and
 * not shared with the plain backward's slot buffers.  Deterministic like ssplat_backward_tiled.
 */
int ssplat_backward_tiled_affine(
    const float* grad_out, const float* means, const float* conics, const float* colors,
    const int64_t* radii, const float* opacities, const float* color_grads, const float* scales,
    const float* rotations, const float* den, const float* out, int64_t n,
    const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids, int height,
    int width, int tile_size, int normalize, double eps, double support_fade_alpha,
    double sigma_cutoff, float* grad_means, float* grad_conics, float* grad_colors,
    float* grad_opacities, float* grad_color_grads, float* grad_scales, float* grad_rotations,
    void* grad_slots, int64_t slot_capacity, int32_t* status, void* workspace,
    size_t workspace_bytes, void* stream);

/* Workspace bytes of the affine entry points (the plain layout plus 48 bytes a Gaussian). */
int ssplat_affine_workspace_bytes(int64_t n, int height, int width, int tile_size,
                                  size_t* bytes);

/* Bytes of an affine gradient slot buffer holding slot_capacity slots (80 bytes each plus a
   header).  No reference counterpart. */
int ssplat_affine_grad_slots_bytes(int64_t slot_capacity, size_t* bytes);

/*
 * Per-Gaussian activity: activity[i] = sum over the pixels of Gaussian i's clipped support
 * rectangle of its weight exp(-q/2) (support fade applied; opacity not included), the pruning
 * statistic of gaussian_activity (render.py:417-442, a torch loop over pixel slices in the
 * reference).  Each Gaussian's sum is formed in a fixed order (bitwise repeatable).
 */
int ssplat_gaussian_activity(const float* means, const float* conics, const int64_t* radii,
                             int64_t n, int height, int width, double support_fade_alpha,
                             double sigma_cutoff, float* activity, void* stream);

/* Bytes of a gradient slot buffer holding slot_capacity slots (40 bytes each plus a header).
   No reference counterpart (the reference accumulates with float atomics). */
int ssplat_grad_slots_bytes(int64_t slot_capacity, size_t* bytes);

/*
 * Total slot count sum_g R(g) of a field (see ssplat_backward_tiled) into *total, a device
 * int64 (-1 if it exceeds int32), without host synchronization.  tile_size is 16 or 32.
 * No reference counterpart.
 */
int ssplat_rect_pair_count(const float* means, const int64_t* radii, int64_t n, int height,
                           int width, int tile_size, int64_t* total, void* workspace,
                           size_t workspace_bytes, void* stream);

/*
 * Batches: many independent fields (images of their own size, each with its own Gaussians)
 * rendered by one launch per kernel.  The fields' Gaussians are concatenated in field order
 * (means, conics, colors, radii, opacities: rows first_gaussian .. first_gaussian+num_gaussians),
 * their images are concatenated in field order (out/grad_out: pixels first_pixel ..), and their
 * tile grids likewise (tile_offsets has total_tiles + 1 entries).
 * No reference counterpart: the reference fits many images with one process per image
 * (batch.py:163-245); this replaces the process farm with one ragged launch.
 */
typedef struct ssplat_field {
  int32_t height;         /* set by the caller */
  int32_t width;          /* set by the caller */
  int64_t first_gaussian; /* filled by ssplat_batch_plan */
  int64_t num_gaussians;  /* set by the caller */
  int64_t first_pixel;    /* filled by ssplat_batch_plan */
  int64_t first_tile;     /* filled by ssplat_batch_plan */
  int32_t tiles_x;        /* filled by ssplat_batch_plan */
  int32_t num_tiles;      /* filled by ssplat_batch_plan */
} ssplat_field;

/*
 * Host-only: completes a HOST table of fields (offsets of every field in the concatenated
 * Gaussians, pixels and tiles) and returns the totals and the workspace size of the batch.  The
 * caller copies the completed table to device memory for the batch calls below.
 */
int ssplat_batch_plan(ssplat_field* fields, int num_fields, int tile_size,
                      int64_t* total_gaussians, int64_t* total_tiles, int64_t* total_pixels,
                      size_t* workspace_bytes);

/* Batched ssplat_build_tile_index; `fields` is the DEVICE copy of the planned table. */
int ssplat_batch_build_tile_index(const ssplat_field* fields, int num_fields,
                                  int64_t total_gaussians, int64_t total_tiles, int tile_size,
                                  const float* means, const float* conics, const int64_t* radii,
                                  int ellipse_cull, double sigma_cutoff, int32_t* tile_offsets,
                                  int32_t* tile_gids, int64_t gid_capacity, int32_t* status,
                                  void* workspace, size_t workspace_bytes, void* stream);

/* Batched ssplat_forward_tiled; out (and den) hold every field's image, concatenated. */
int ssplat_batch_forward_tiled(const ssplat_field* fields, int num_fields,
                               int64_t total_gaussians, int64_t total_tiles, int tile_size,
                               const float* means, const float* conics, const float* colors,
                               const int64_t* radii, const float* opacities,
                               const int32_t* tile_gids, const int32_t* tile_offsets,
                               int64_t num_gids, int normalize, double eps,
                               double support_fade_alpha, double sigma_cutoff, float* out,
                               float* den, void* workspace,
                               size_t workspace_bytes, void* stream);

/* Batched ssplat_backward_tiled; den and out are the batched forward's concatenated outputs.
   Slots are numbered over the concatenated Gaussians; each Gaussian's sum is the same as in a
   single-field call (the order depends on the Gaussian alone). */
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
                                void* stream);

/* ssplat_batch_backward_tiled after ssplat_batch_forward_tiled of the same inputs into the same
   workspace (see ssplat_backward_tiled_after_forward).  No reference counterpart. */
int ssplat_batch_backward_tiled_after_forward(
    const ssplat_field* fields, int num_fields, int64_t total_gaussians, int64_t total_tiles,
    int tile_size, const float* grad_out, const float* means, const float* conics,
    const float* colors, const int64_t* radii, const float* opacities, const float* den,
    const float* out, const int32_t* tile_gids, const int32_t* tile_offsets, int64_t num_gids,
    int normalize, double eps, double support_fade_alpha, double sigma_cutoff,
    float* grad_means, float* grad_conics, float* grad_colors, float* grad_opacities,
    void* grad_slots, int64_t slot_capacity, int32_t* status, void* workspace,
    size_t workspace_bytes, void* stream);

/*
 * SSPL1 decoding (the reference's codec.decode, codec.py:333-391).  The host parses the
 * container (magic, JSON header, framed streams) and inflates the streams; this decodes the
 * raw streams, copied to the device, into the field arrays: means (n,2), log_scales (n,2),
 * rotations (n,), colors (n,3) and, with has_opacity, opacity logits (n,), all float32 and
 * bit-identical to the reference's NumPy decode (float64 dequantization rounded to float32; the
 * logits' final log is correctly rounded).  Streams: means 2n values (x then y; byte-planar when
 * means_byte_planar), log-scales 2n, rotations n, colors 3n (r, g, b), opacities n; each value
 * uint8/16/32 for bit widths <= 8/16/32.  Morton-coded means and delta-coded colors are prefix
 * sums mod 2^bits.  The workspace holds ssplat_sspl1_workspace_bytes(n) bytes, 256-aligned.
 */
typedef struct ssplat_sspl1 {
  int64_t n;
  int32_t bits_means, bits_scales, bits_rot, bits_colors, bits_opacity;
  int32_t morton, color_delta, means_byte_planar, rot_circular, has_opacity;
  double means_lo[2], means_hi[2];
  double scale_lo[2], scale_hi[2];
  double color_lo[3], color_hi[3];
} ssplat_sspl1;

int ssplat_sspl1_workspace_bytes(int64_t n, size_t* bytes);
int ssplat_sspl1_decode(const ssplat_sspl1* layout, const uint8_t* means_raw,
                        const uint8_t* scales_raw, const uint8_t* rot_raw,
                        const uint8_t* colors_raw, const uint8_t* opacity_raw, float* means,
                        float* log_scales, float* rotations, float* colors,
                        float* opacity_logits, void* workspace, size_t workspace_bytes,
                        void* stream);

/*
 * Fit loss: (1 - w) * L1 + w * (1 - SSIM) of an RGB image against its target, and its gradient
 * with respect to the image (the reference computes it with torch in fit.py:800-823 and
 * metrics.py:88-151; no native counterpart).  L1 is the mean absolute difference over the 3 H W
 * values; SSIM the mean of the SSIM map with an 11-tap Gaussian window (sigma 1.5) applied per
 * channel as two zero-padded 1-D passes, C1 = 0.01^2, C2 = 0.03^2.  image, target and
 * grad_image are (H, W, 3) float32; *loss is one device float.  grad_image may be NULL (loss
 * only).  ssim_weight w is in [0, 1].  The loss is summed in a fixed order (deterministic).
 */
int ssplat_loss_workspace_bytes(int64_t total_pixels, int64_t total_tiles, size_t* bytes);

/*
 * Host-only: completes a HOST table of fields for the loss calls (the loss works in units, strips
 * 16 columns wide and up to 128 rows tall): first_pixel, first_tile (the field's first unit),
 * tiles_x (its strips) and num_tiles (its units); height and width are the caller's.  Returns the
 * total units (the total_tiles of the loss calls, and of ssplat_loss_workspace_bytes for one
 * image), the total pixels and the loss workspace size.
 */
int ssplat_loss_plan(ssplat_field* fields, int num_fields, int64_t* total_units,
                     int64_t* total_pixels, size_t* workspace_bytes);
int ssplat_l1_ssim_loss(const float* image, const float* target, int height, int width,
                        double ssim_weight, float* loss, float* grad_image, void* workspace,
                        size_t workspace_bytes, void* stream);

/*
 * The target's window moments for a fit whose target does not change: moments receives two
 * (H, W, 3) float32 maps, G*t then G*t^2 (the 11-tap window per channel, zero-padded): the
 * target-side statistics the reference caches across a fit in SSIMTargetStats
 * (metrics.py:93-128).
 */
int ssplat_loss_target_moments(const float* target, int height, int width, float* moments,
                               void* stream);

/*
 * The same loss in two calls for autograd: _forward writes *loss and keeps the SSIM derivative
 * maps in the workspace; _backward (same image, target, weight and workspace, nothing else using
 * the workspace in between) writes grad_image = (*grad_loss) * dloss/dimage, with grad_loss a
 * device float (NULL: 1).  target_moments, when not NULL, holds ssplat_loss_target_moments of
 * this target: the forward then blurs three maps of the image instead of five, with the same
 * result bit for bit.
 */
int ssplat_l1_ssim_loss_forward(const float* image, const float* target,
                                const float* target_moments, int height, int width,
                                double ssim_weight, float* loss, void* workspace,
                                size_t workspace_bytes, void* stream);
int ssplat_l1_ssim_loss_backward(const float* image, const float* target, int height, int width,
                                 double ssim_weight, const float* grad_loss, float* grad_image,
                                 void* workspace, size_t workspace_bytes, void* stream);

/* Batched loss: the sum of every field's loss over concatenated images (a field table planned
   by ssplat_loss_plan, copied to the device; total_tiles = its units), its two-call form and the
   target moments of every field (2 maps of (total_pixels, 3)). */
int ssplat_batch_l1_ssim_loss(const ssplat_field* fields, int num_fields, int64_t total_tiles,
                              int64_t total_pixels, const float* image, const float* target,
                              double ssim_weight, float* loss, float* grad_image,
                              void* workspace, size_t workspace_bytes, void* stream);
int ssplat_batch_loss_target_moments(const ssplat_field* fields, int num_fields,
                                     int64_t total_tiles, int64_t total_pixels,
                                     const float* target, float* moments, void* stream);
int ssplat_batch_l1_ssim_loss_forward(const ssplat_field* fields, int num_fields,
                                      int64_t total_tiles, int64_t total_pixels,
                                      const float* image, const float* target,
                                      const float* target_moments, double ssim_weight,
                                      float* loss, void* workspace, size_t workspace_bytes,
                                      void* stream);
int ssplat_batch_l1_ssim_loss_backward(const ssplat_field* fields, int num_fields,
                                       int64_t total_tiles, int64_t total_pixels,
                                       const float* image, const float* target,
                                       double ssim_weight, const float* grad_loss,
                                       float* grad_image, void* workspace,
                                       size_t workspace_bytes, void* stream);

/*
 * Fit optimizer: one Adam step over up to 8 parameter groups in one launch (the reference's
 * torch.optim.Adam, fit.py:672-691, default betas (0.9, 0.999), eps 1e-8), with the step count
 * in device memory (*step is incremented first, so the call can be captured in a HIP graph):
 *   m = m + (1 - b1)(g - m),  v = b2 v + (1 - b2) g^2,
 *   p = p - lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps).
 * A group with clamp set is then clamped to [lo, hi], or to [lo, hi_rows[i / row_width]] when
 * hi_rows is given (the log-scale clamp of fit.py:3033, :3359-3375).  No reference native
 * counterpart.
 */
typedef struct ssplat_adam_group {
  float* param;
  const float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  int64_t count;          /* elements of param (and grad, exp_avg, exp_avg_sq) */
  double lr;
  int32_t clamp;          /* 0: no clamp */
  double lo;
  double hi;              /* used when hi_rows is NULL */
  const float* hi_rows;   /* per-row upper bound, or NULL */
  int64_t row_width;      /* elements per row of param for hi_rows */
} ssplat_adam_group;

int ssplat_adam_step(const ssplat_adam_group* groups, int num_groups, float* step, double beta1,
                     double beta2, double eps, void* stream);

/*
 * Renderer inputs from a field (the reference's GaussianField.conics / radii / opacity_values,
 * gaussians.py:270-321, in one launch): conics (n,3), int64 radii (n,2) and opacities (n,) =
 * sigmoid(opacity_logits) from log_scales (n,2), rotations (n,) and optional opacity_logits and
 * filter_variance (n,).  Any output may be NULL.  The radii follow the reference's float32
 * operation order, so they equal the torch expression's bit for bit.
 */
int ssplat_produce_field(const float* log_scales, const float* rotations,
                         const float* opacity_logits, const float* filter_variance, int64_t n,
                         double dilation, double sigma_cutoff, float* conics, int64_t* radii,
                         float* opacities, void* stream);

/*
 * Backward of ssplat_produce_field: gradients of the conics (and opacities) to the log_scales,
 * rotations (and opacity logits; opacities are the forward's sigmoid values).  Outputs are
 * overwritten.  The radii carry no gradient (gaussians.py:306).
 */
int ssplat_produce_field_backward(const float* log_scales, const float* rotations,
                                  const float* opacities, const float* filter_variance,
                                  int64_t n, double dilation, const float* grad_conics,
                                  const float* grad_opacities, float* grad_log_scales,
                                  float* grad_rotations, float* grad_logits, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SSPLAT_H_ */
