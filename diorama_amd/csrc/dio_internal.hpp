// Internal declarations shared by the HIP kernels and the C-ABI runtime.
// Everything here is gfx950 (CDNA4, wave64) code; nothing is portable by design.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <utility>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace dio {

constexpr int kTile = 16;              // pixels per tile side (common.wgsl:29)
constexpr int kWave = 64;              // CDNA wavefront width
constexpr uint32_t kFloatsPerGaussian = 16;
constexpr uint32_t kFloatsPerSplat = 12;
constexpr uint32_t kShRest = 45;
constexpr uint32_t kMomentSlots = kFloatsPerGaussian + kShRest;  // 61 Adam slots

// Thrown inside the runtime and turned into a status code at the C boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what);
#define DIO_HIP(call) ::dio::check_hip((call), #call)
// Checks the launch itself (bad configuration) without waiting for the kernel.
#define DIO_LAUNCHED(name) ::dio::check_hip(hipGetLastError(), name)

__host__ __device__ inline uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

// ------------------------------------------------------------------ scan --
// Exclusive prefix sum of u32 values (decoupled look-back, one pass over HBM).
// `count` (device, element 0) caps how many of `capacity` values are scanned;
// null scans all of them. `total` (device) receives the grand total.
// Workspaces of the scan and the radix sort start zeroed (allocate them cleared)
// and every call leaves them zeroed again: the last workgroup to finish clears what
// the call used, so no clearing launch precedes a call.
//
// wait_memory() waits until this wave's vector memory operations, stores and
// atomics included, have completed (vmcnt counts all of them on this GPU family).
// A workgroup calls it before the barrier that precedes counting itself finished,
// so the last workgroup sees every other workgroup's device-coherent writes; an
// agent-scope fence would also order them but writes back the whole L2 in every
// workgroup.
__device__ inline void wait_memory() {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#endif
}
struct ScanWorkspace {
  uint64_t* tile_state = nullptr;  // one (flag, value) word per tile
  uint32_t* ticket = nullptr;      // dynamic tile counter
  uint32_t* done = nullptr;        // workgroups finished
  uint32_t tiles = 0;
};
constexpr uint32_t kScanTile = 4096;  // 256 threads x 16 values
size_t scan_workspace_bytes(uint32_t capacity);
ScanWorkspace scan_workspace(void* base, uint32_t capacity);
void scan_exclusive(hipStream_t s, const uint32_t* in, uint32_t* out, const uint32_t* count, uint32_t capacity,
                    uint32_t* total, const ScanWorkspace& ws);
// The allocation-free scan's workspace on `device`, laid out for `capacity` values
// inside a static device array (at most kStaticScanMaxValues; 128 KiB per GPU). Its
// contents are whatever the last call left: clear the words a call uses first.
constexpr uint32_t kStaticScanMaxValues = 1u << 26;
ScanWorkspace static_scan_workspace(int device, uint32_t capacity);

// ------------------------------------------------------------------ sort --
// Stable LSD onesweep radix sort, 8- or 4-bit digits, of u32 keys or u64 keys stored as
// (low word, high word), each carrying a u32 value. Sorts key bits
// [begin_bit, end_bit); the passes ping-pong between (keys, values) and the
// alternate buffers. Returns true when the result ended in the alternate buffers.
constexpr uint32_t kSortTile = 4096;  // 256 threads x 16 keys
constexpr uint32_t kSortGroup = 16;  // tiles per look-back group
struct SortWorkspace {
  uint32_t* hist = nullptr;                 // per XCD copy, per pass: 256 digit counts
  uint32_t* offsets = nullptr;              // per pass: where each digit's run starts
  uint32_t* state[2] = {nullptr, nullptr};  // per tile x digit look-back words, alternating passes
  uint32_t* group_sum[2] = {nullptr, nullptr};     // per group x digit: tiles done << 24 | their count
  uint32_t* group_prefix[2] = {nullptr, nullptr};  // per group x digit: inclusive prefix once known
  uint32_t* tickets = nullptr;              // per pass tile counter; the histogram's finish counter
  uint32_t tiles = 0, groups = 0;
};
size_t sort_workspace_bytes(uint32_t capacity);
SortWorkspace sort_workspace(void* base, uint32_t capacity);
// digit_bits: 8 or 4 (more passes of fewer buckets); the result is the same.
// With gather_src and gather_out, the last pass also writes gather_out[i] =
// gather_src[sorted value i] (e.g. each depth-sorted splat's tile count).
// count_hint (0: none): the key count the host expects for a device-side `count`,
// which sizes the work the idle part of the capacity-sized grids does; any value
// gives the same result.
bool radix_sort(hipStream_t s, bool wide, void* keys, uint32_t* values, void* alt_keys, uint32_t* alt_values,
                const uint32_t* count, uint32_t capacity, uint32_t begin_bit, uint32_t end_bit,
                const SortWorkspace& ws, uint32_t digit_bits = 8, const uint32_t* gather_src = nullptr,
                uint32_t* gather_out = nullptr, uint32_t count_hint = 0);

// ------------------------------------------------------------ renderer --
struct CameraArgs {
  float view[16];
  float proj[16];
  float viewport[2];
  uint32_t bin_shift;  // pair bins of (16 << bin_shift) px: 0 the reference's 16-px tiles, 1 32-px bins
};

// The tiles a splat's square touches: [x0, x1) x [y0, y1) (common.wgsl:37-42).
__device__ inline void tile_rect(float mx, float my, float radius, uint32_t tiles_x, uint32_t tiles_y,
                                 uint32_t rect[4]) {
  const float size = static_cast<float>(kTile);
  const float tx = static_cast<float>(tiles_x), ty = static_cast<float>(tiles_y);
  rect[0] = static_cast<uint32_t>(fminf(fmaxf(floorf((mx - radius) / size), 0.f), tx));
  rect[1] = static_cast<uint32_t>(fminf(fmaxf(floorf((my - radius) / size), 0.f), ty));
  rect[2] = static_cast<uint32_t>(fminf(fmaxf(floorf((mx + radius) / size) + 1.f, 0.f), tx));
  rect[3] = static_cast<uint32_t>(fminf(fmaxf(floorf((my + radius) / size) + 1.f, 0.f), ty));
}

// The bins a splat's square touches: the 16-px tile rectangle merged into bins of
// (1 << shift) x (1 << shift) tiles ([x0, x1) x [y0, y1) in bin units); an empty
// tile rectangle stays empty.
__device__ inline void bin_rect(float mx, float my, float radius, uint32_t tiles_x, uint32_t tiles_y, uint32_t shift,
                                uint32_t rect[4]) {
  tile_rect(mx, my, radius, tiles_x, tiles_y, rect);
  if (shift == 0) return;
  if (rect[2] <= rect[0] || rect[3] <= rect[1]) {
    rect[0] = rect[1] = rect[2] = rect[3] = 0;
    return;
  }
  const uint32_t round = (1u << shift) - 1u;
  rect[0] >>= shift;
  rect[1] >>= shift;
  rect[2] = (rect[2] + round) >> shift;
  rect[3] = (rect[3] + round) >> shift;
}

// A bin rectangle [x0, x1) x [y0, y1) packed two 16-bit coordinates to a word (the
// frames that read it have at most 8192 bins).
__host__ __device__ inline uint2 pack_rect(const uint32_t r[4]) {
  return make_uint2(r[0] | (r[1] << 16), r[2] | (r[3] << 16));
}
__host__ __device__ inline void unpack_rect(uint2 p, uint32_t r[4]) {
  r[0] = p.x & 0xffffu;
  r[1] = p.x >> 16;
  r[2] = p.y & 0xffffu;
  r[3] = p.y >> 16;
}


// `vis` (nullable): each Gaussian's (depth, radius) as two floats, radius 0 when culled.
// `rects` (nullable): each splat's touched tiles (bins) as pack_rect, zero when culled.
void preprocess(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, uint32_t n,
                const CameraArgs& cam, float* splats, uint32_t* tile_counts, uint32_t* depth_keys,
                uint32_t* depth_values, float* vis = nullptr, uint32_t* rects = nullptr);

// The tile stage from existing splats: counts (+ depth keys) only.
// Tile counts of splats in bins of (16 << bin_shift) px, their depth keys and indices.
void count_tiles(hipStream_t s, const float* splats, uint32_t n, uint32_t width, uint32_t height,
                 uint32_t* tile_counts, uint32_t* depth_keys, uint32_t* depth_values, uint32_t bin_shift = 0,
                 uint32_t* rects = nullptr);

// pair_hint (0: none), here and in find_ranges: the pair count the host expects,
// which sizes the grid; the kernels loop over whatever *pair_total holds.
// order (nullable): the splats in the order their pairs are written, offsets
// following that order; null writes them in splat order.
void duplicate(hipStream_t s, const float* splats, const uint32_t* order, const uint32_t* offsets,
               const uint32_t* tile_counts, uint32_t n, uint32_t width, uint32_t height, uint32_t capacity,
               const uint32_t* pair_total, uint32_t* pair_tiles, uint32_t* pair_splats, uint32_t pair_hint = 0,
               uint32_t bin_shift = 0);

// Each tile's [start, end) run of the sorted pairs; with `order` (nullable, one
// word per tile) also the rasterizer's launch order, heaviest tiles first.
// clear (nullable): 1 + kPartClasses words set to zero before the rasterizer runs
// (the segmented backward's item counts). sparse: most tiles hold no pair (a
// batch's tile space, padded per view), so every range is first cleared by a fill
// launch instead of by the thread that meets the gap, which would write a long gap
// alone.
void find_ranges(hipStream_t s, const uint32_t* pair_tiles, const uint32_t* pair_total, uint32_t capacity,
                 uint32_t tiles, uint32_t* ranges, uint32_t* order = nullptr, uint32_t pair_hint = 0,
                 uint32_t* clear = nullptr, bool sparse = false);

// Sorts each tile's run of `list` (pairs binned by tile) by (depth_keys[splat],
// splat), in place: the run of pairs a depth-ordered duplicate would have
// produced. `ordered`: every run arrives in splat order (pairs written in splat
// order, then sorted stably by tile), which saves passes over the splat indices on
// the runs that take radix passes; otherwise runs may arrive in any order.
// tile_order (nullable) is the launch order of the tiles; scratch is a pair-sized
// buffer the runs longer than 2,048 pairs pass through at their own positions.
// keyed (nullable): the runs as (depth key, splat) pairs (two words a pair) at the
// same positions, read instead of `list` and depth_keys[splat] and sorted into
// `list`; scratch then holds two words a pair. keyed is overwritten. queue: 2 tiles + 2
// words, zero before the first call: two queues of the runs that need radix passes,
// used by turns (`parity` counts the calls on this queue).
void depth_order(hipStream_t s, const uint32_t* depth_keys, const uint32_t* ranges, const uint32_t* tile_order,
                 uint32_t tiles, uint32_t* list, void* scratch, bool ordered, uint32_t* queue, uint32_t parity,
                 uint32_t* keyed = nullptr);

// The rasterizer's launch order: tiles by falling pair count, in buckets of
// kOrderStep pairs (ties in any order).
constexpr uint32_t kOrderBuckets = 1024, kOrderStep = 8;

// Pairs binned by tile from counts, without a sort (at most kBinMaxTiles tiles or
// bins). The splats are cut into chunks (bin_layout); each chunk counts its pairs
// per tile, the counts are scanned per tile over the chunks, and each chunk writes
// its pairs into its slots of the tiles' runs (in no particular order inside a
// run: depth_order with ordered = false then orders them). A frame past the pair
// capacity keeps the same pairs as pair_cut (and rewrites tile_counts to them).
constexpr uint32_t kBinMaxTiles = 8192;
struct BinLayout {
  uint32_t groups = 0;  // chunks, one workgroup each (at most 512)
  uint32_t chunk = 0;   // splats per chunk
};
BinLayout bin_layout(uint32_t n);
struct BinWorkspace {
  uint32_t* matrix = nullptr;   // groups x tiles counts, then their prefixes per tile
  uint32_t* partial = nullptr;  // per chunk: its pairs
  uint32_t* totals = nullptr;   // per tile: its pairs
  uint32_t groups = 0, tiles = 0;
};
// Counts, the overflow cut (pair_total[0] gets the frame's uncut pair count) and the
// per-tile scan.
// `rects`: each splat's pack_rect (preprocess, count_tiles).
void bin_count(hipStream_t s, const uint32_t* rects, const uint32_t* depth_keys, uint32_t* tile_counts, uint32_t n,
               uint32_t width, uint32_t height, uint32_t shift, uint32_t capacity, uint32_t* pair_total,
               const BinWorkspace& ws);
// The pairs as (depth key, splat) into `keyed` (two words a pair), each tile's
// [start, end) into `ranges`, the launch order into `order` (nullable) and zeros
// into clear[0 .. kPartClasses] (nullable).
void bin_scatter(hipStream_t s, const uint32_t* rects, const uint32_t* tile_counts, const uint32_t* depth_keys,
                 uint32_t n, uint32_t width, uint32_t height, uint32_t shift, const BinWorkspace& ws, uint32_t* keyed,
                 uint32_t* ranges, uint32_t* order, uint32_t* clear);
// For pairs written in splat order (offsets the exclusive scan of tile_counts in
// splat order, *pair_total their sum): when the pairs overflow `capacity`, rewrites
// tile_counts to the pairs each splat keeps under the depth-order slot rule and
// offsets to their scan, so exactly the capacity's nearest pairs are written.
void pair_cut(hipStream_t s, const uint32_t* depth_keys, uint32_t* tile_counts, uint32_t* offsets, uint32_t n,
              const uint32_t* pair_total, uint32_t capacity);

// The segmented backward (16-px tiles). Segment boundaries are the absolute pair
// positions that are multiples of 2^shift: the forward pass keeps each pixel's
// transmittance and blended colour (T, r, g, b) in front of every boundary inside
// a tile's list that the pixel reaches, at ckpt[(b >> shift) * 256 + quadrant * 64
// + lane]. Its workgroups then list the backward's work items, the parts of their
// tile's list between boundaries up to the tile's last blended splat, as (tile,
// b >> shift), by length class: whole segments at items[0 ...], shorter parts in
// kPartClasses classes of falling length at items[whole_room + (c - 1) part_room
// ...], counted in counts[c] (zeroed by find_ranges). The backward takes them in
// that order, longest first; each pixel's replay of a part starts from the
// checkpoint at its end, so a long list is several workgroups' work, not one's.
constexpr uint32_t kPartClasses = 8;
struct Segments {
  uint32_t shift = 0;             // log2 of the segment length: 8, 9 or 10 (0: off)
  uint32_t whole_room = 0;        // (pair capacity >> shift) + 1
  uint32_t part_room = 0;         // per class: 2 per tile
  uint32_t grid = 0;              // backward workgroups; each loops over the items
  float* ckpt = nullptr;          // whole_room entries of 256 float4
  uint32_t* items = nullptr;      // uint2 per item, whole_room + kPartClasses part_room of them
  uint32_t* counts = nullptr;     // 1 + kPartClasses words
};
// Segment lengths: 2^8 to 2^10 pairs (a multiple of every splat batch).
constexpr uint32_t kSegmentShiftMin = 8, kSegmentShiftMax = 10;

// splat_batch: pairs a wave stages per round (64, 128 or 256); `order` (nullable)
// maps workgroups to tiles. Results depend on neither.
// tile_size: 16 (a workgroup per 16-px tile), 8 (a wave per 8x8 quadrant, same
// 16-px lists) or 32 (a workgroup per 32-px bin: `ranges` and `order` are the bins').
// work (nullable, 16-px tiles): two zeroed words; persistent workgroups take tiles
// from them and leave them zeroed.
// seg (nullable, 16-px tiles, not with work): also writes the segmented backward's
// checkpoints and items.
void rasterize(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
               const uint32_t* order, uint32_t width, uint32_t height, const float* background, uint32_t* target,
               float* pixel_color, uint32_t* pixel_last, uint32_t splat_batch = 64, uint32_t tile_size = 16,
               uint32_t* work = nullptr, const Segments* seg = nullptr);

// footprint (nullable; tile size 16): per splat, the sums of alpha T x the pixel's
// loss share (`pixel_loss`) and of alpha T over its blended pixels are added into
// footprint[2 i], footprint[2 i + 1].
// seg (nullable): the segmented backward over the checkpoints and items the
// forward pass wrote with the same `seg`.
void rasterize_backward(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
                        const uint32_t* order, uint32_t width, uint32_t height, const float* background,
                        const float* pixel_color, const uint32_t* pixel_last, const float* pixel_grads,
                        float* splat_grads, uint32_t splat_batch = 64, uint32_t tile_size = 16,
                        const float* pixel_loss = nullptr, float* footprint = nullptr, uint32_t* work = nullptr,
                        const Segments* seg = nullptr);

void preprocess_backward(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, uint32_t n,
                         const CameraArgs& cam, const float* splats, const float* splat_grads,
                         float* gaussian_grads, float* sh_grads);

// ------------------------------------------------------ multi-view batch --
// A batch renders several views of one Gaussian set in the same launches: view v's
// splats sit at [v N, (v + 1) N), its 16x16 tiles at [tile_base, tile_base +
// tiles_x tiles_y) of the batch's tile space and its pixels at pixel_base of the
// batch's pixel buffers. The header and the views live in device memory, so a
// captured graph replays with other views by rewriting them.
constexpr uint32_t kMaxBatchViews = 64;

struct ViewDesc {
  CameraArgs cam;
  uint32_t width, height, tiles_x, tiles_y;
  uint32_t tile_base, pixel_base;
  float loss_scale;  // 1 / (3 width height)
  uint32_t pad;
  const uint32_t* target;  // the view's rgba8 photo
};

struct BatchHeader {
  uint32_t views, tiles, pixels, pad;
};

struct BatchArgs {
  const BatchHeader* header;
  const ViewDesc* views;
  uint32_t view_count;  // the batch's views (fixed while a graph is replayed)
  uint32_t per_view;    // N: Gaussians per view
  uint32_t max_tiles;   // the tile space the launches cover
};

void preprocess_batch(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, const BatchArgs& b,
                      float* splats, uint32_t* tile_counts, uint32_t* depth_keys, uint32_t* depth_values);
void duplicate_batch(hipStream_t s, const float* splats, const uint32_t* order, const uint32_t* offsets,
                     const uint32_t* tile_counts, const BatchArgs& b, uint32_t capacity, const uint32_t* pair_total,
                     uint32_t* pair_tiles, uint32_t* pair_splats);
void rasterize_batch(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
                     const BatchArgs& b, const float* background, float* pixel_color, uint32_t* pixel_last,
                     uint32_t splat_batch = 64);
void training_loss_batch(hipStream_t s, const BatchArgs& b, const float* rendered, float* partials,
                         float* pixel_grads, float* pixel_loss);
void rasterize_backward_batch(hipStream_t s, const float* splats, const uint32_t* pair_splats,
                              const uint32_t* ranges, const BatchArgs& b, const float* background,
                              const float* pixel_color, const uint32_t* pixel_last, const float* pixel_grads,
                              float* splat_grads, uint32_t splat_batch = 64);
// Gradients summed over the batch's views, and the densification statistics of
// every view (adam.wgsl:125-134) in the same pass.
void preprocess_backward_batch(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree,
                               const BatchArgs& b, const float* splats, const float* splat_grads,
                               float* gaussian_grads, float* sh_grads, float* stats, float* extra = nullptr);

// ---------------------------------------------------------------- loss --
// `clear` (optional, 16-byte aligned): `clear_words` words that the partials kernel
// zeroes on the side (the trainer's splat gradients, which the rasterizer's backward
// adds into next); the loss kernels are compute-bound and leave HBM time to spare.
void training_loss(hipStream_t s, uint32_t width, uint32_t height, const float* rendered, const uint32_t* target,
                   float* partials, float* pixel_grads, float* pixel_loss, uint32_t* clear = nullptr,
                   size_t clear_words = 0);

// ----------------------------------------------------------- optimizer --
struct AdamArgs {
  float bias_correction1, bias_correction2;
  float inv_bias_correction1, inv_bias_correction2;
  float position_lr, opacity_lr, scale_lr, rotation_lr, color_lr, sh_lr;
  float viewport[2];
};
// `extra` (nullable): per Gaussian (largest gradient length, summed depth,
// summed gradient x, summed gradient y), for the learned densification.
void accumulate_stats(hipStream_t s, uint32_t n, const float* splats, const float* splat_grads,
                      const AdamArgs& a, float* stats, float* extra = nullptr);
// The plain trainer's preprocess backward, statistics and Adam step in one pass
// (the per-Gaussian gradients stay on chip): reads the activated Gaussians and SH
// before the step, each splat's (depth, radius) from the preprocess's `vis` and the
// splat gradients; updates params, SH, moments, the activated copy and the
// statistics (and `extra`, optional).
// With `prep`, the kernel also runs the next step's preprocess (preprocess with
// that step's camera and SH degree) on the updated Gaussians it still holds on chip,
// writing that step's splats, tile counts, depth keys and values and (depth, radius)
// where `prep` says: the next step then skips its own preprocess and the re-read of
// the activated Gaussians and SH.
struct PrepArgs {
  CameraArgs cam;
  uint32_t sh_degree = 0;
  float* splats = nullptr;
  uint32_t* tile_counts = nullptr;
  uint32_t* depth_keys = nullptr;
  uint32_t* depth_values = nullptr;
  float* vis = nullptr;
  uint32_t* rects = nullptr;  // nullable
};
void backward_adam(hipStream_t s, uint32_t n, float* gaussians, float* sh, uint32_t sh_degree, const CameraArgs& cam,
                   const float* vis, const float* splat_grads, float* params, float* moments, const AdamArgs& a,
                   float* stats, float* extra, const PrepArgs* prep = nullptr);
void adam_step(hipStream_t s, uint32_t n, float* params, float* sh, float* moments, const float* gaussian_grads,
               const float* sh_grads, const AdamArgs& a, float* gaussians);
// The same step with its arguments read from device memory (for graph replays).
void adam_step_device_args(hipStream_t s, uint32_t n, float* params, float* sh, float* moments,
                           const float* gaussian_grads, const float* sh_grads, const AdamArgs* a, float* gaussians);
void activate(hipStream_t s, uint32_t n, const float* params, float* gaussians);
void reset_opacity(hipStream_t s, uint32_t n, float* params, float* moments, float* gaussians);

struct DensifyArgs {
  float grad_threshold, clone_max_size, min_opacity, max_size;
  uint32_t seed;
};
// Action codes: 0 prune, 1 keep, 2 clone, 3 split. `forced` (nullable) gives
// every Gaussian's action instead of the rule.
void densify_count(hipStream_t s, uint32_t n, const float* params, const float* stats, const DensifyArgs& a,
                   const uint32_t* forced, uint32_t* outputs, uint32_t* counters);
void densify_scatter(hipStream_t s, uint32_t n, const float* params, const float* stats, const uint32_t* outputs,
                     const float* sh, const float* moments, const DensifyArgs& a, const uint32_t* forced,
                     float* new_params, float* new_sh, float* new_moments);
// The rule's codes and each Gaussian's mean view-space gradient as its growth priority.
void densify_rule_codes(hipStream_t s, uint32_t n, const float* params, const float* stats, const DensifyArgs& a,
                        uint32_t* codes, float* priority);
// Budget: sort keys ranking clone/split candidates by priority, then demotion of
// the candidates past the first `allowed`.
void densify_growth_keys(hipStream_t s, uint32_t n, const uint32_t* codes, const float* priority, uint32_t* keys,
                         uint32_t* values);
void densify_demote(hipStream_t s, uint32_t n, const uint32_t* sorted_keys, const uint32_t* sorted_values,
                    uint32_t allowed, const float* params, const DensifyArgs& a, uint32_t* codes);

// ------------------------------------------------ learned densification --
constexpr uint32_t kMlpInputs = 12;
constexpr uint32_t kMlpParams = 12 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4;  // 5,252
// `views_per_step` turns ages in steps into ages in views seen.
void mlp_features(hipStream_t s, uint32_t n, const float* params, const float* stats, const float* extra,
                  float extent, uint32_t iteration, uint32_t views_per_step, float* features);
// `norm`: the input standardisation, mean[12] then scale[12].
void mlp_forward(hipStream_t s, uint32_t n, const float* features, const uint16_t* weights, const float* norm,
                 float* logits);
// Mean class-weighted softmax cross-entropy gradients over `batch` samples added
// into `grads` (cleared by the caller), the loss sum into loss[0]. seed 0: samples
// 0..batch-1; else `batch` hashed draws from `count` samples.
void mlp_gradients(hipStream_t s, uint32_t batch, uint32_t count, uint32_t seed, const float* features,
                   const uint32_t* labels, const uint16_t* weights, const float* norm, const float class_weights[4],
                   float* grads, float* loss);
void mlp_adam(hipStream_t s, float* master, float* moments, float* grads, uint16_t* weights, float lr,
              float inv_bias1, float inv_bias2);
void mlp_round(hipStream_t s, const float* master, uint16_t* weights);
// The MLP's action where its top probability reaches `confidence` (codes holds
// the rule's codes on entry), and P(clone) + P(split) as the growth priority.
void mlp_learned_codes(hipStream_t s, uint32_t n, const float* logits, float confidence, uint32_t* codes,
                       float* priority);
// Sample logging: Gaussians drawn with probability (threshold + 1) / 2^32 become
// pending samples (16 words each); at the next densification the ones whose
// outcome held join the training set.
constexpr uint32_t kMlpPendingWords = 16;
// Logs a hashed subset of the Gaussians as pending samples: inputs, the rule's (or
// forced) class, the first output index and the loss over the footprint since the
// last densification (`footprint`: per Gaussian the sums of w x loss share and of w).
void mlp_log_pending(hipStream_t s, uint32_t n, const float* features, const float* footprint, const uint32_t* codes,
                     const uint32_t* outputs, uint32_t step, uint32_t threshold, uint32_t capacity, float* pending,
                     uint32_t* pending_count);
// Moves the pending samples whose outcome held into the training set: pruned, or
// the first descendant's footprint loss since (`footprint`, n Gaussians) below the
// Gaussian's before.
void mlp_resolve(hipStream_t s, const uint32_t* pending_count, uint32_t pending_capacity, const float* pending,
                 const float* footprint, uint32_t n, uint32_t store_capacity, float* store_features,
                 uint32_t* store_labels, uint32_t* store_count);

// ------------------------------------------------------------------- ply --
// The 3DGS .ply body: 62 float columns per vertex (ply.hip).
constexpr uint32_t kPlyColumns = 62;
constexpr uint32_t kPlyLdsFloats = 16384;  // rows staged per unpack workgroup x stride stay within 64 KiB
// Where each field sits in a file's rows of `stride` floats; `per_channel`
// higher-order coefficients per colour channel (0, 3, 8 or 15).
struct PlyColumns {
  uint32_t stride;
  uint32_t per_channel;
  uint32_t position[3];
  uint32_t dc[3];
  uint32_t opacity;
  uint32_t scale[3];
  uint32_t rotation[4];
  uint32_t rest[kShRest];
};
// Raw params (16 floats) + SH (45, coefficient-major) -> count x 62 body floats.
void ply_pack(hipStream_t s, const float* params, const float* sh, uint32_t count, float* body);
// count rows of cols.stride floats -> activated Gaussians (16 floats) + SH (45).
void ply_unpack(hipStream_t s, const float* body, uint32_t count, const PlyColumns& cols, float* gaussians, float* sh);
uint32_t ply_unpack_rows(uint32_t stride);

// ------------------------------------------------------------- neighbors --
// Mean squared distance to the 3 nearest other points of n points (n x 3
// floats), bit-identical to the brute-force scan (knn.hip).
size_t neighbors_workspace_bytes(uint32_t n);
void mean_neighbor_distance2(hipStream_t s, const float* pts, uint32_t n, float* out, void* workspace);
// Starting raw parameters (n x 16) and zero SH (n x 45) from points, rgb bytes and those distances.
void initial_gaussians(hipStream_t s, const float* pts, const uint8_t* colors, const float* d2, uint32_t n,
                       float* params, float* sh);

// --------------------------------------------------------------- metrics --
// Per 16x16 tile of a render against an rgba8 photo (metrics.hip): the exact sum
// of squared 8-bit errors over rgb and, with_ssim, the sum of per-pixel SSIM.
uint32_t evaluate_tiles(uint32_t width, uint32_t height);
void evaluate_image(hipStream_t s, const float4* rendered, const uint32_t* target, uint32_t width, uint32_t height,
                    bool with_ssim, uint32_t* sse, double* ssim);

void fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t count);

}  // namespace dio
