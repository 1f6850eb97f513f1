/*
 * diorama-mi355x: the C ABI of the MI355X (gfx950) Gaussian-splatting renderer
 * and trainer. Plain pointers and sizes only; every object is an opaque handle.
 *
 * Each group replaces one operator class of the reference (paths relative to the
 * reference repository) and keeps its method names, argument meaning, buffer
 * layouts and error behaviour. Work is enqueued on the context's HIP stream in
 * call order, the analogue of recording into a GPUCommandEncoder and submitting:
 * `encode` calls return without waiting for the GPU; downloads, reads and
 * `after_submit` wait.
 *
 * Errors: functions return DIO_OK (0) or a negative code, and dio_last_error()
 * gives the message (thread-local). Contract violations the reference throws for
 * (src/gaussianRenderer.ts:99-101, src/tileRasterizer.ts:136-138,202,
 * src/trainer.ts:133-135,295) return DIO_ERR_INVALID with the same message.
 * Handles are not thread-safe; one host thread drives a context.
 *
 * Buffer layouts (src/gaussians.ts:32-70):
 *   Gaussian  16 floats: position xyz, opacity, scale xyz, pad, quaternion wxyz, rgb, pad
 *   Splat2D   12 floats: mean xy (pixels, y down), depth, radius (0 = culled),
 *             conic abc, pad, rgb, opacity
 *   SH rest   45 floats: coefficients 1..15, rgb each, coefficient-major
 *   Camera    view[16], proj[16] column-major (src/mat4.ts:1-3), viewport[2] pixels
 *   images    rgba8 as one u32 per pixel (r in the low byte), row-major
 *   pixels    colour buffers: 4 floats per pixel (rgb, final transmittance)
 */
#ifndef DIORAMA_H
#define DIORAMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIO_OK 0
#define DIO_ERR_INVALID (-1)   /* contract violation (the reference throws an Error) */
#define DIO_ERR_INTERNAL (-2)  /* workspace or internal invariant */
#define DIO_ERR_HIP (-3)       /* a HIP runtime call failed */
#define DIO_ERR_MEMORY (-4)    /* device allocation failed */

typedef struct dio_ctx dio_ctx;
typedef struct dio_buf dio_buf;
typedef struct dio_prefix_sum dio_prefix_sum;
typedef struct dio_radix_sort dio_radix_sort;
typedef struct dio_tile_rasterizer dio_tile_rasterizer;
typedef struct dio_renderer dio_renderer;
typedef struct dio_training_loss dio_training_loss;
typedef struct dio_trainer dio_trainer;

/* CameraData (src/gaussianRenderer.ts:11-15). */
/* Launch settings that change speed, never results (see dio_trainer_autotune). */
typedef struct dio_tuning {
  /* pixels per tile side of the rasterizer: 16 (the reference's tiling,
   * src/tileRasterizer.ts:8), 8 (a workgroup per 8x8 quadrant of the 16-px lists) or
   * 32 (pairs binned per 32-px bin, each 16-px tile culling the bin's list by its
   * exact 16-px rectangle; splat batch 64 or 128). Images and gradients are the same
   * for all three; with 32, pixelLast counts positions in the bin's list. Batch
   * trainers take 16. */
  uint32_t tile_size;
  uint32_t splat_batch; /* pairs a wave stages per round: 64, 128 or 256 */
  uint32_t digit_bits;  /* radix digit width of the binning sorts: 4 or 8 */
  /* 16-px tiles only: 0 launches a workgroup per tile; 1 launches as many
   * workgroups as the GPU holds at once, each taking the next tile (heaviest first)
   * from a counter until none is left */
  uint32_t persistent;
  /* 16-px tiles, one workgroup each, single-view trainers: 0 replays each tile's
   * list in one backward workgroup; 256, 512 or 1024 splits the lists at the pair
   * positions that are multiples of it, the forward pass keeping every pixel's
   * transmittance and colour there, and the backward replays the parts (longest
   * first) in workgroups of their own. Images are the same; gradients agree to
   * float rounding. */
  uint32_t segment;
  /* single-view trainers: 0 sorts the splats by depth before writing their pairs
   * (src/tileRasterizer.ts:127-192); 1 writes the pairs in splat order, sorts them
   * stably by tile and then each tile's list by depth; 2 bins the pairs by tile
   * from per-tile counts, with no sort over the pairs, and then sorts each tile's
   * list by (depth, splat) (frames of more than 8192 tiles or bins run 1 instead).
   * The lists, and so images and gradients, are the same, also on a frame whose
   * pairs overflow the pair buffers (all keep the nearest pair slots). */
  uint32_t per_tile_depth;
} dio_tuning; /* a trainer starts at {16, 64, 8, 0, 512, 2}; a batch trainer at {16, 64, 8, 0, 0, 0} */

typedef struct dio_tuning_result {
  dio_tuning tuning;
  double ms_per_step;
} dio_tuning_result;

typedef struct dio_camera {
  float view[16];
  float proj[16];
  float viewport[2];
} dio_camera;

/* ------------------------------------------------------------ runtime -- */
/* Replaces initGpu (src/gpu.ts:8-47): one device, one in-order stream. */
const char* dio_last_error(void);
const char* dio_version(void);
int dio_ctx_create(int device, dio_ctx** out);
int dio_ctx_synchronize(dio_ctx* ctx);
/* Waits for the stream and destroys it. Buffers and objects of the context may
 * still be destroyed afterwards (nothing else may be done with them); the
 * context's own memory goes with the last of them. */
void dio_ctx_destroy(dio_ctx* ctx);
/* Timestamp queries (the 'timestamp-query' feature, src/main.ts:40): record a
 * marker in one of DIO_TIMESTAMP_SLOTS slots on the context's stream, then read
 * the milliseconds between two recorded slots (waits for the later one). */
#define DIO_TIMESTAMP_SLOTS 64
int dio_ctx_timestamp(dio_ctx* ctx, uint32_t slot);
int dio_ctx_elapsed(dio_ctx* ctx, uint32_t from_slot, uint32_t to_slot, double* ms);

/* Device buffers (GPUBuffer + queue.writeBuffer / readBuffer, src/readback.ts). */
int dio_buf_create(dio_ctx* ctx, size_t bytes, dio_buf** out);
int dio_buf_upload(dio_buf* buf, size_t offset, const void* src, size_t bytes);
int dio_buf_download(const dio_buf* buf, size_t offset, void* dst, size_t bytes);
int dio_buf_clear(dio_buf* buf); /* encoder.clearBuffer */
/* encoder.copyBufferToBuffer (src/readback.ts:11), enqueued on dst's stream. */
int dio_buf_copy(dio_buf* dst, size_t dst_offset, const dio_buf* src, size_t src_offset, size_t bytes);
size_t dio_buf_size(const dio_buf* buf);
void dio_buf_destroy(dio_buf* buf);

/* ------------------------------------------------------------ PrefixSum -- */
/* src/prefixSum.ts:10-103: exclusive u32 scan of `data` in place. `count`
 * (nullable) holds in element 0 how many values to scan (at most `capacity`).
 * After encode, element 0 of total() is the sum of the scanned values. */
int dio_prefix_sum_create(dio_ctx* ctx, dio_buf* data, uint32_t capacity, dio_buf* count, dio_prefix_sum** out);
dio_buf* dio_prefix_sum_total(dio_prefix_sum* scan);
int dio_prefix_sum_encode(dio_prefix_sum* scan);
void dio_prefix_sum_destroy(dio_prefix_sum* scan);
/* No counterpart in the reference: PrefixSum's create + encode + total in one call
 * that allocates nothing, for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops). It scans min(*count, capacity) values of
 * `data` in place (all `capacity` when `count` is null) and writes the grand total
 * to element 0 of `total`, as dio_prefix_sum_encode does, with its workspace in a
 * static device array of the library (128 KiB per GPU, kept until the process
 * ends). At most 2^26 values; larger capacities fail with DIO_ERR_INVALID. Calls
 * that overlap on one GPU share that array and race: serialise them. */
int dio_prefix_sum_encode_static(dio_ctx* ctx, dio_buf* data, uint32_t capacity, const dio_buf* count, dio_buf* total);

/* ------------------------------------------------------------ RadixSort -- */
/* src/radixSort.ts:27-142: stable sort of u32 keys (high_bits = 0) or of 64-bit
 * keys stored as (low, high) u32 pairs, looking at the low `high_bits` bits of the
 * high word (at most 32). `count` (nullable) holds how many keys to sort. The
 * result lands back in keys() and values(). */
int dio_radix_sort_create(dio_ctx* ctx, uint32_t capacity, uint32_t high_bits, dio_buf* count, dio_radix_sort** out);
dio_buf* dio_radix_sort_keys(dio_radix_sort* sort);
dio_buf* dio_radix_sort_values(dio_radix_sort* sort);
uint32_t dio_radix_sort_capacity(const dio_radix_sort* sort);
int dio_radix_sort_encode(dio_radix_sort* sort);
void dio_radix_sort_destroy(dio_radix_sort* sort);

/* ------------------------------------------------------- TileRasterizer -- */
/* src/tileRasterizer.ts:29-317: bins `count` splats of `splats` into 16x16 tiles
 * (sorted front to back), blends every pixel, and keeps what the backward pass
 * needs. initial_pair_capacity 0 means max(65536, 8 * count) (:97). The camera
 * buffer of the reference is not needed: the viewport is an argument. */
int dio_tile_rasterizer_create(dio_ctx* ctx, dio_buf* splats, uint32_t count, uint32_t initial_pair_capacity,
                               dio_tile_rasterizer** out);
/* encode (:127-192). `target` (nullable) receives rgba8 pixels. */
int dio_tile_rasterizer_encode(dio_tile_rasterizer* tr, dio_buf* target, uint32_t width, uint32_t height,
                               const float background[3]);
/* encodeBackward (:200-222): pixel_grads vec4 per pixel -> splat_grads 9 floats per splat. */
int dio_tile_rasterizer_encode_backward(dio_tile_rasterizer* tr, uint32_t width, uint32_t height, dio_buf* pixel_grads,
                                        dio_buf* splat_grads);
/* afterSubmit (:229-244): reads back the frame's pair count and grows the pair
 * buffers to 2^ceil(log2(1.25 * count)) when they were too small. Waits. */
int dio_tile_rasterizer_after_submit(dio_tile_rasterizer* tr);
/* No counterpart in the reference: the operator's launch settings (dio_tuning;
 * it starts at {16, 64, 8, 0, 0, 0}), for callers that tune them. Images and
 * gradients do not depend on them. Takes effect at the next encode, which a
 * backward pass then needs. */
int dio_tile_rasterizer_set_tuning(dio_tile_rasterizer* tr, const dio_tuning* tuning);
uint32_t dio_tile_rasterizer_last_pair_count(const dio_tile_rasterizer* tr);
uint32_t dio_tile_rasterizer_pair_capacity(const dio_tile_rasterizer* tr);
/* debugBuffers: each sorted pair's splat, and each tile's [start, end) run. */
dio_buf* dio_tile_rasterizer_pair_splats(dio_tile_rasterizer* tr);
dio_buf* dio_tile_rasterizer_ranges(dio_tile_rasterizer* tr);
dio_buf* dio_tile_rasterizer_pixel_color(dio_tile_rasterizer* tr);
dio_buf* dio_tile_rasterizer_pixel_last(dio_tile_rasterizer* tr);
void dio_tile_rasterizer_destroy(dio_tile_rasterizer* tr);

/* ----------------------------------------------------- GaussianRenderer -- */
/* src/gaussianRenderer.ts:55-311. initial_pair_capacity 0 = the rasterizer's default. */
int dio_renderer_create(dio_ctx* ctx, uint32_t initial_pair_capacity, dio_renderer** out);
/* setGaussians (:93-110): copies count x 16 floats; sh_rest (nullable) count x 45. */
int dio_renderer_set_gaussians(dio_renderer* r, const float* gaussians, uint32_t count, const float* sh_rest,
                               uint32_t sh_degree);
/* adoptGaussians (:118-122): takes ownership of both buffers. */
int dio_renderer_adopt_gaussians(dio_renderer* r, uint32_t count, dio_buf* gaussians, dio_buf* sh, uint32_t sh_degree);
int dio_renderer_set_sh_degree(dio_renderer* r, uint32_t degree);
/* encode (:169-192) = encode_preprocess + encode_tiles; the split is the
 * afterPreprocess hook point (src/insideView.ts:221). */
int dio_renderer_encode(dio_renderer* r, dio_buf* target, const dio_camera* camera, const float background[3]);
int dio_renderer_encode_preprocess(dio_renderer* r, const dio_camera* camera);
int dio_renderer_encode_tiles(dio_renderer* r, dio_buf* target, const float background[3]);
/* encodeDraw (:198-211): ready-made splats, skipping the preprocess. */
int dio_renderer_encode_draw(dio_renderer* r, dio_buf* target, dio_buf* splats, uint32_t count, uint32_t width,
                             uint32_t height, const float background[3]);
/* encodeBackward (:217-227): results in the gradient buffers below. */
int dio_renderer_encode_backward(dio_renderer* r, dio_buf* pixel_grads);
int dio_renderer_after_submit(dio_renderer* r);
uint32_t dio_renderer_count(const dio_renderer* r);
dio_buf* dio_renderer_splat_buffer(dio_renderer* r);
dio_buf* dio_renderer_gaussian_buffer(dio_renderer* r);
dio_buf* dio_renderer_sh_buffer(dio_renderer* r);
dio_tile_rasterizer* dio_renderer_tiles(dio_renderer* r);
dio_buf* dio_renderer_pixel_buffer(dio_renderer* r);
dio_buf* dio_renderer_pixel_last_buffer(dio_renderer* r);
/* gradients (:29-36): 16 floats per Gaussian, 45 SH floats, 9 splat floats. */
dio_buf* dio_renderer_gradient_gaussians(dio_renderer* r);
dio_buf* dio_renderer_gradient_sh(dio_renderer* r);
dio_buf* dio_renderer_gradient_splats(dio_renderer* r);
void dio_renderer_destroy(dio_renderer* r);

/* --------------------------------------------------------- TrainingLoss -- */
/* src/trainingLoss.ts:16-130: 0.8 * L1 + 0.2 * (1 - SSIM) and dL/dpixel. */
int dio_training_loss_create(dio_ctx* ctx, uint32_t width, uint32_t height, dio_training_loss** out);
/* rendered: vec4 per pixel (rgb used); target: rgba8 per pixel. */
int dio_training_loss_encode(dio_training_loss* loss, dio_buf* rendered, dio_buf* target);
/* read (:110-125): sum of the per-pixel shares + 0.2. Waits. */
int dio_training_loss_read(dio_training_loss* loss, double* out);
dio_buf* dio_training_loss_pixel_grads(dio_training_loss* loss);
dio_buf* dio_training_loss_pixel_loss(dio_training_loss* loss);
void dio_training_loss_destroy(dio_training_loss* loss);

/* ---------------------------------------------------------------- Trainer -- */
/* TrainingSettings / DEFAULT_SETTINGS (src/trainer.ts:37-82). */
typedef struct dio_training_settings {
  float extent;
  uint32_t position_lr_steps;
  float position_lr[2];
  float opacity_lr, scale_lr, rotation_lr, color_lr, sh_lr;
  uint32_t sh_degree_interval;
  uint32_t densify_from, densify_until, densify_interval;
  float densify_grad_threshold;
  float percent_dense;
  float min_opacity;
  uint32_t opacity_reset_interval;
  float background[3];
} dio_training_settings;

/* Densification record (src/trainer.ts:86-95). */
typedef struct dio_densification {
  uint32_t step, before, after, kept, cloned, split, removed;
} dio_densification;

void dio_default_settings(float extent, dio_training_settings* out);
/* Trainer constructor (:127-168): raw params (count x 16: position, opacity logit,
 * log scale, pad, rotation, degree-0 coefficient, pad) and SH (count x 45);
 * one rgba8 image per view, all views the same size. */
int dio_trainer_create(dio_ctx* ctx, const float* params, const float* sh, uint32_t count, const dio_camera* cameras,
                       const uint32_t* const* images, uint32_t view_count, const dio_training_settings* settings,
                       dio_trainer** out);
/* Multi-view batching (no counterpart in the reference; for callers training on
 * many views per step): `batch` training views of any sizes per step are rendered,
 * scored and back-propagated in the same launches and their gradients summed into
 * one Adam step, the step being one captured HIP graph. Every other dio_trainer_*
 * call works as for dio_trainer_create; step counts, learning-rate schedule and
 * densification cadence count batch steps; dio_trainer_last_loss is the mean of
 * the batch's per-view losses. */
int dio_trainer_create_batched(dio_ctx* ctx, const float* params, const float* sh, uint32_t count,
                               const dio_camera* cameras, const uint32_t* const* images, uint32_t view_count,
                               uint32_t batch, const dio_training_settings* settings, dio_trainer** out);
uint32_t dio_trainer_batch_size(const dio_trainer* t); /* 0 for a plain trainer */
int dio_trainer_step(dio_trainer* t);      /* optimize() then maintain() */
int dio_trainer_optimize(dio_trainer* t);  /* render, loss, backward, stats, Adam; no wait */
int dio_trainer_maintain(dio_trainer* t);  /* densify / reset when due; waits when densifying */
int dio_trainer_densify(dio_trainer* t, int prune_large);
int dio_trainer_reset_opacity(dio_trainer* t);
int dio_trainer_export_raw(dio_trainer* t, float* params, float* sh); /* count x 16, count x 45 */
int dio_trainer_last_loss(dio_trainer* t, double* out);
/* evaluate (:371-386): mean PSNR/SSIM of renders clamped and rounded to 8 bits
 * against rgba8 photos in host memory, scored on the device (src/metrics.ts:5-72);
 * ssim is NaN without with_ssim. */
int dio_trainer_evaluate(dio_trainer* t, const dio_camera* cameras, const uint32_t* const* images, uint32_t view_count,
                         int with_ssim, double* psnr, double* ssim);
/* measurePsnr (src/insideView.ts:233-267): the mean PSNR of `view_count` views
 * whose photos are already device buffers (rgba8, at least width x height x 4
 * bytes), for training curves without host transfers. */
int dio_trainer_measure_psnr(dio_trainer* t, const dio_camera* cameras, const dio_buf* const* images,
                             uint32_t view_count, double* psnr);
uint32_t dio_trainer_count(const dio_trainer* t);
uint32_t dio_trainer_iteration(const dio_trainer* t);
void dio_trainer_set_iteration(dio_trainer* t, uint32_t iteration);
uint32_t dio_trainer_sh_degree(const dio_trainer* t);
/* shDegree is a public field of the reference Trainer (src/trainer.ts:105). */
int dio_trainer_set_sh_degree(dio_trainer* t, uint32_t degree);
/* The training view (index into the views given at creation) of the last plain
 * step: the one nextView() returned (src/trainer.ts:395-404). */
uint32_t dio_trainer_last_view(const dio_trainer* t);
uint32_t dio_trainer_densification_count(const dio_trainer* t);
int dio_trainer_densification(const dio_trainer* t, uint32_t index, dio_densification* out);
dio_buf* dio_trainer_params(dio_trainer* t);
dio_buf* dio_trainer_moments(dio_trainer* t);
dio_buf* dio_trainer_stats(dio_trainer* t);
dio_renderer* dio_trainer_renderer(dio_trainer* t);
void dio_trainer_destroy(dio_trainer* t);

/* --------------------------------------------------------- measurement -- */
/* ------------------------------------------------------------- tuning -- */
/* No counterpart in the reference (the reference's settings are fixed); for
 * callers that tune a trainer to their GPU and views. */
int dio_trainer_set_tuning(dio_trainer* t, const dio_tuning* tuning);
int dio_trainer_get_tuning(const dio_trainer* t, dio_tuning* out);
/* The autotuner: for every candidate (tile size 16, 8, 32; splat batch 64, 128,
 * 256; digit width 4, 8; persistent 0, 1 with 16-px tiles; segment 0, 256, 512
 * with 16-px tiles and a workgroup per tile; per-tile depth order 0, 1 for
 * single-view trainers; see dio_tuning) run one
 * untimed optimize() step, then time `steps` more
 * (ordinary training steps; a batch trainer replays its step graph on the same
 * views); a single-view trainer then times its four fastest again over 3 x steps
 * (their listed times are those). Keep the fastest setting, and list up to
 * `capacity` results. */
int dio_trainer_autotune(dio_trainer* t, uint32_t steps, dio_tuning_result* results, uint32_t capacity,
                         uint32_t* count);

/* ------------------------------------------------ learned densification -- */
/* No counterpart in the reference; for callers that densify with a learned
 * proposal. A 12 -> 64 -> 64 -> 4 MLP (5,252 parameters, BF16 weights, FP32
 * accumulation) reads 12 statistics per Gaussian and scores keep / clone / split
 * / prune. Weights cross the boundary as 5,252 floats: W1[12][64], b1[64],
 * W2[64][64], b2[64], W3[64][4], b3[4], input index first. */
typedef struct dio_densify_mlp dio_densify_mlp;
#define DIO_MLP_PARAMS 5252
#define DIO_MLP_INPUTS 12
int dio_mlp_create(dio_ctx* ctx, uint32_t seed, dio_densify_mlp** out); /* He-uniform, zero biases */
int dio_mlp_get_weights(dio_densify_mlp* mlp, float* weights);        /* the BF16 weights as floats */
int dio_mlp_set_weights(dio_densify_mlp* mlp, const float* weights);  /* rounded to BF16; resets Adam */
/* features: count x 12 floats; logits: count x 4 floats (keep, clone, split, prune). */
int dio_mlp_forward(dio_densify_mlp* mlp, dio_buf* features, uint32_t count, dio_buf* logits);
/* Gradients of the mean class-weighted softmax cross-entropy over all `count` samples (labels:
 * u32 class 0 keep, 1 clone, 2 split, 3 prune) and that mean loss. */
int dio_mlp_gradients(dio_densify_mlp* mlp, dio_buf* features, dio_buf* labels, uint32_t count, float* grads,
                      double* loss);
/* `steps` Adam steps, each on `batch` samples drawn (hashed from `seed`) from the
 * `count` samples; last_loss (nullable) gets the last step's mean loss. */
int dio_mlp_train(dio_densify_mlp* mlp, dio_buf* features, dio_buf* labels, uint32_t count, uint32_t steps,
                  uint32_t batch, float lr, uint32_t seed, double* last_loss);
/* The inputs' standardisation, applied before the first layer in inference and
 * training: x' = (x - mean) * scale (mean 0, scale 1 by default). */
int dio_mlp_set_input_norm(dio_densify_mlp* mlp, const float mean[12], const float scale[12]);
/* Per-class loss weights (keep, clone, split, prune) for dio_mlp_gradients and
 * dio_mlp_train; all 1 by default. */
int dio_mlp_set_class_weights(dio_densify_mlp* mlp, const float weights[4]);
void dio_mlp_destroy(dio_densify_mlp* mlp);
/* mode 0: the plain rule (the default); 1: the plain rule, logging samples (the
 * 12 inputs and the rule's class of Gaussians drawn at each densification,
 * kept when at the next densification the Gaussian was pruned or its mean
 * view-space gradient fell); 2: the MLP's class where its probability reaches
 * `confidence`, else the rule's. budget > 0 caps the count after densifying,
 * growth ranked by mean gradient (rule) or P(clone) + P(split) (MLP). */
int dio_trainer_set_densify_mode(dio_trainer* t, uint32_t mode, dio_densify_mlp* mlp, uint32_t budget,
                                 float confidence, uint32_t samples_per_densify);
/* The logged samples (non-owning views) and how many there are. */
int dio_trainer_samples(dio_trainer* t, dio_buf** features, dio_buf** labels, uint32_t* count);
/* The 12 inputs of the current Gaussians (a trainer-owned count x 12 buffer):
 * log(1 + mean and largest view-space gradient / 2e-4), opacity, log(largest and
 * smallest scale / extent), log(largest / smallest scale), SH DC luminance,
 * log(1 + views), log(1 + age in views seen), log(1 + largest radius),
 * log(mean depth), gradient-direction variance. */
int dio_trainer_features(dio_trainer* t, dio_buf** features);
/* Per Gaussian (largest gradient length, summed depth, summed gradient x, y)
 * since the last densification; null in mode 0. */
dio_buf* dio_trainer_extra_stats(dio_trainer* t);
/* Mode 1 only (else null): per Gaussian, since the last densification, the sums
 * over the pixels it blended into of blend weight (alpha T) x the pixel's loss
 * share, and of blend weight (2 floats each). Their ratio is the loss over the
 * Gaussian's footprint, the outcome signal of the logged samples. */
dio_buf* dio_trainer_footprint(dio_trainer* t);
/* A batch trainer's last step (no reference counterpart; for checks against a
 * per-view oracle): the training-view index of each of its `*count` slots (up to
 * `capacity` written to `views`); the splat buffer in which slot s holds its
 * view's splats at records s * (*stride) ... s * (*stride) + count - 1; and the
 * pixel colour buffer (vec4 per pixel) holding the slots' images one after
 * another, row by row, in slot order. */
int dio_trainer_batch_frames(dio_trainer* t, uint32_t* views, uint32_t capacity, uint32_t* count, dio_buf** splats,
                             uint32_t* stride, dio_buf** pixels);

/* Times `repeats` launches of one named stage between HIP events recorded on the
 * context's stream, for the benchmark's roofline figures. Stages: "preprocess",
 * "scan", "sort", "rasterize", "rasterize_backward", "loss", "backward_adam" (the
 * step's fused preprocess backward + statistics + Adam), and the separate
 * "preprocess_backward" and "adam" of densification steps; a batch trainer has
 * the one stage "batch_step" (its captured step graph). Returns the mean milliseconds per launch of the last optimize()
 * step's inputs re-run `repeats` times. */
int dio_trainer_time_stage(dio_trainer* t, const char* stage, uint32_t repeats, double* ms);

/* Stage clock of the plain trainer. The reference times no stage: it requests
 * 'timestamp-query' (src/main.ts:40) but records none, and its checks wall-clock
 * whole submissions with performance.now() (e.g. src/checks/backward.ts:160-163).
 * While enabled, every plain optimize() step records HIP
 * events between its launches on the context's stream, and their elapsed times are
 * summed per stage: the stages run in their real order, on a real step's data.
 * Densification steps are not counted. Enabling or disabling clears the totals;
 * a batch trainer is an error. */
int dio_trainer_stage_clock(dio_trainer* t, int enable);
/* Waits for the clocked steps and returns the summed milliseconds per stage in
 * the order of dio_trainer_stage_name (ms may be NULL; otherwise it holds
 * `capacity` >= *stages values), the number of stages, of steps clocked, and
 * their summed pair counts (each step's count is copied out after its last
 * stage). Any output pointer may be NULL. */
int dio_trainer_stage_totals(dio_trainer* t, double* ms, uint32_t capacity, uint32_t* stages, uint32_t* steps,
                             uint64_t* pairs);
/* Name of stage k of a clocked step, or NULL past the last: "preprocess",
 * "depth_sort", "scan", "duplicate", "tile_sort", "ranges", "rasterize", "loss"
 * (with the pair-count readback copy), "clear_grads", "rasterize_backward",
 * "backward_adam". */
const char* dio_trainer_stage_name(uint32_t stage);

/* --------------------------------------------------- COLMAP reconstructions --
 * One registered photo of a COLMAP sparse reconstruction: the TrainingCamera of
 * src/sceneCameras.ts:6-19 (OpenCV axes in world space: x right, y down, z
 * forward; focal lengths in pixels). */
typedef struct dio_colmap_camera {
  char name[256];
  uint32_t width, height;
  double fx, fy;
  double position[3];
  double right[3], down[3], forward[3];
} dio_colmap_camera;
/* parseColmap (src/colmap.ts:18-129): cameras.bin, images.bin and points3D.bin
 * (binary, little-endian; undistorted SIMPLE_PINHOLE or PINHOLE cameras) into the
 * photos sorted by name and the points' float32 xyz and rgb bytes. *camera_count
 * and *point_count receive the counts; with cameras, positions and colors all
 * null nothing else is written, so a caller can size its arrays first. Unknown
 * camera models, photos of missing cameras and short files fail with
 * DIO_ERR_INVALID and the reader's messages. */
int dio_colmap_parse(const void* cameras_bin, size_t cameras_bytes, const void* images_bin, size_t images_bytes,
                     const void* points_bin, size_t points_bytes, dio_colmap_camera* cameras, uint32_t camera_capacity,
                     uint32_t* camera_count, float* positions, uint8_t* colors, uint32_t point_capacity,
                     uint32_t* point_count);
/* trainingCameraData (src/loadDataset.ts:80-86): the renderer's camera for a photo
 * rendered at width x height: cameraView (src/sceneCameras.ts:54-67) and
 * perspective(verticalFov, width / height, 0.01, 1000). */
int dio_colmap_training_camera(const dio_colmap_camera* camera, uint32_t width, uint32_t height, dio_camera* out);

/* ---------------------------------------------------- starting Gaussians --
 * meanNeighborDistance2 (src/loadDataset.ts:126-149, nearestNeighbors.wgsl:12-44):
 * for each of `count` points (positions: count x 3 floats on the device) the
 * mean squared distance to its 3 nearest other points, into distance2 (count
 * floats). Bit-identical to the reference's brute-force scan with the distance
 * fma(dz, dz, fma(dy, dy, dx * dx)); non-finite points get +inf. */
int dio_mean_neighbor_distance2(dio_ctx* ctx, const dio_buf* positions, uint32_t count, dio_buf* distance2);
/* initialGaussians (src/loadDataset.ts:87-110): raw params (count x 16) and zero
 * SH (count x 45) for structure-from-motion points (positions count x 3 floats,
 * colors count x 3 bytes): the point's position and colour, opacity 0.1, no
 * rotation, log scale 0.5 log(max(mean 3-nearest squared distance, 1e-7)). */
int dio_initial_gaussians(dio_ctx* ctx, const dio_buf* positions, const dio_buf* colors, uint32_t count,
                          dio_buf* params, dio_buf* sh);

/* ----------------------------------------------------------- .ply files --
 * The 3DGS .ply format of trained Gaussians; the body is packed and unpacked on
 * the device.
 *
 * writeGaussianPly (src/plyWriter.ts:14-43): the file for `count` raw Gaussians
 * (params: count x 16 floats as the trainer keeps them; sh: count x 45,
 * coefficient-major) written into `out` (host memory of `capacity` bytes): the
 * text header, then count x 62 little-endian floats, packed on the stream of
 * params' context. *length receives the file's size; with out null nothing else
 * happens, so a caller can size `out` first. */
int dio_ply_write(const dio_buf* params, const dio_buf* sh, uint32_t count, void* out, size_t capacity, size_t* length);
/* parseGaussianPly (src/plyLoader.ts:21-98): a binary little-endian .ply of float
 * vertex properties in any order, SH degree 0-3, into two new device buffers the
 * caller owns (dio_buf_destroy, or hand them to dio_renderer_adopt_gaussians):
 * *gaussians count x 16 activated floats (exp scales, sigmoid opacity, 0.5 +
 * SH_C0 dc), *sh_rest count x 45 coefficient-major, zero past the file's degree.
 * Malformed files fail with DIO_ERR_INVALID and the loader's messages. */
int dio_ply_parse(dio_ctx* ctx, const void* file, size_t bytes, dio_buf** gaussians, dio_buf** sh_rest,
                  uint32_t* count, uint32_t* sh_degree);

#ifdef __cplusplus
}
#endif
#endif
