// The C ABI (include/diorama.h): the reference's operator classes as opaque
// handles over one HIP stream. The host logic follows src/prefixSum.ts,
// src/radixSort.ts, src/tileRasterizer.ts, src/gaussianRenderer.ts,
// src/trainingLoss.ts and src/trainer.ts; the arithmetic is in the HIP kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace dio {
void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(DIO_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace dio

using dio::Error;
using dio::alloc;
using dio::dev;
using dio::guard;
using dio::guard_read;
using dio::invalid;
using dio::need;
using dio::release;
using dio::sync;
using dio::within;

namespace dio {
std::string& last_error() {
  thread_local std::string message;
  return message;
}
}  // namespace dio

namespace {

// `bin_shift`: the tile counts the preprocess writes are per (16 << bin_shift)-px bin.
void to_camera(const dio_camera& c, dio::CameraArgs& out, uint32_t bin_shift = 0) {
  std::memcpy(out.view, c.view, sizeof(out.view));
  std::memcpy(out.proj, c.proj, sizeof(out.proj));
  out.viewport[0] = c.viewport[0];
  out.viewport[1] = c.viewport[1];
  out.bin_shift = bin_shift;
}

// One side of a viewport as a pixel count. Cameras carry it as a float; one that is
// negative, not finite or past 2^24 is refused before any cast or size product.
uint32_t viewport_dim(float v) {
  if (!(v >= 0.f && v <= 16777216.f)) invalid("a viewport side must be between 0 and 2^24 pixels");
  return static_cast<uint32_t>(v);
}

// 16-pixel tiles of a w x h viewport, in 64 bits (a 2^20 x 2^20 viewport has 2^32).
uint64_t viewport_tiles(uint32_t w, uint32_t h) {
  return static_cast<uint64_t>(dio::ceil_div(w, dio::kTile)) * dio::ceil_div(h, dio::kTile);
}

// The tile sort tells at most 2^16 tiles apart (src/tileRasterizer.ts:10,136-138).
void check_tile_limit(uint32_t w, uint32_t h) {
  if (viewport_tiles(w, h) > 65536u) {
    invalid("A " + std::to_string(w) + "×" + std::to_string(h) + " viewport has more tiles than the sort can tell apart.");
  }
}

// Exchanges the device memory behind two buffers of one context; the handles stay.
void swap_storage(dio_buf* a, dio_buf* b) {
  std::swap(a->ptr, b->ptr);
  std::swap(a->bytes, b->bytes);
}

uint32_t next_pow2_capacity(uint32_t count) {
  double want = std::ceil(std::log2(count * 1.25));
  return static_cast<uint32_t>(std::min(std::pow(2.0, want), 4294967295.0));
}

}  // namespace

// =================================================================== runtime

extern "C" const char* dio_last_error(void) { return dio::last_error().c_str(); }
extern "C" const char* dio_version(void) { return "diorama-mi355x 0.1 gfx950"; }

extern "C" int dio_ctx_create(int device, dio_ctx** out) {
  return guard([&] {
    need(out, "out");
    DIO_HIP(hipSetDevice(device));
    auto* c = new dio_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      dio::check_hip(e, "hipStreamCreate");
    }
    *out = c;
  });
}

extern "C" int dio_ctx_synchronize(dio_ctx* ctx) {
  return guard_read([&] {
    need(ctx, "ctx");
    sync(ctx);
  });
}

extern "C" void dio_ctx_destroy(dio_ctx* ctx) {
  if (!ctx || ctx->closed.exchange(true)) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (hipEvent_t& e : ctx->stamps)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  (void)hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;  // later syncs of buffers or objects that outlive it touch no stream
  dio::drop(ctx);              // the open context's own reference
}

extern "C" int dio_ctx_timestamp(dio_ctx* ctx, uint32_t slot) {
  return guard([&] {
    need(ctx, "ctx");
    if (slot >= DIO_TIMESTAMP_SLOTS) invalid("no such timestamp slot");
    if (!ctx->stamps[slot]) DIO_HIP(hipEventCreate(&ctx->stamps[slot]));
    DIO_HIP(hipEventRecord(ctx->stamps[slot], ctx->stream));
    ctx->stamped[slot] = true;
  });
}

extern "C" int dio_ctx_elapsed(dio_ctx* ctx, uint32_t from_slot, uint32_t to_slot, double* ms) {
  return guard([&] {
    need(ctx, "ctx");
    need(ms, "ms");
    if (from_slot >= DIO_TIMESTAMP_SLOTS || to_slot >= DIO_TIMESTAMP_SLOTS) invalid("no such timestamp slot");
    if (!ctx->stamped[from_slot] || !ctx->stamped[to_slot]) invalid("timestamp slot not recorded");
    DIO_HIP(hipEventSynchronize(ctx->stamps[to_slot]));
    float elapsed = 0.f;
    DIO_HIP(hipEventElapsedTime(&elapsed, ctx->stamps[from_slot], ctx->stamps[to_slot]));
    *ms = elapsed;
  });
}

extern "C" int dio_buf_create(dio_ctx* ctx, size_t bytes, dio_buf** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    *out = alloc(ctx, bytes, true);
  });
}

extern "C" int dio_buf_upload(dio_buf* buf, size_t offset, const void* src, size_t bytes) {
  return guard([&] {
    need(buf, "buffer");
    if (!within(buf, offset, bytes)) invalid("upload past the end of the buffer");
    if (bytes == 0) return;
    need(src, "source");
    DIO_HIP(hipMemcpyAsync(static_cast<char*>(buf->ptr) + offset, src, bytes, hipMemcpyHostToDevice, buf->ctx->stream));
    sync(buf->ctx);  // the caller may reuse its host memory at once, as with queue.writeBuffer
  });
}

extern "C" int dio_buf_download(const dio_buf* buf, size_t offset, void* dst, size_t bytes) {
  return guard_read([&] {
    need(buf, "buffer");
    if (!within(buf, offset, bytes)) invalid("download past the end of the buffer");
    if (bytes == 0) return;
    need(dst, "destination");
    DIO_HIP(hipMemcpyAsync(dst, static_cast<const char*>(buf->ptr) + offset, bytes, hipMemcpyDeviceToHost,
                           buf->ctx->stream));
    sync(buf->ctx);
  });
}

extern "C" int dio_buf_clear(dio_buf* buf) {
  return guard([&] {
    need(buf, "buffer");
    dio::fill_u32(buf->ctx->stream, static_cast<uint32_t*>(buf->ptr), 0u, (buf->bytes + 3) / 4);
  });
}

extern "C" int dio_buf_copy(dio_buf* dst, size_t dst_offset, const dio_buf* src, size_t src_offset, size_t bytes) {
  return guard([&] {
    need(dst, "destination");
    need(src, "source");
    if (!within(dst, dst_offset, bytes) || !within(src, src_offset, bytes)) invalid("copy past the end of a buffer");
    if (bytes == 0) return;
    DIO_HIP(hipMemcpyAsync(static_cast<char*>(dst->ptr) + dst_offset, static_cast<const char*>(src->ptr) + src_offset,
                           bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
  });
}

extern "C" size_t dio_buf_size(const dio_buf* buf) { return buf ? buf->bytes : 0; }

extern "C" void dio_buf_destroy(dio_buf* buf) {
  if (!buf) return;
  if (!buf->ctx->closed) (void)hipStreamSynchronize(buf->ctx->stream);
  release(buf);
}

// ================================================================= PrefixSum

struct dio_prefix_sum {
  dio_ctx* ctx = nullptr;
  dio_buf* data = nullptr;
  dio_buf* count = nullptr;
  uint32_t capacity = 0;
  dio_buf* total = nullptr;
  dio_buf* ws = nullptr;
};

extern "C" int dio_prefix_sum_create(dio_ctx* ctx, dio_buf* data, uint32_t capacity, dio_buf* count,
                                     dio_prefix_sum** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(data, "data");
    need(out, "out");
    if (static_cast<size_t>(capacity) * 4 > data->bytes) invalid("the data buffer is smaller than the capacity");
    auto* s = new dio_prefix_sum;
    s->ctx = ctx;
    dio::hold(s->ctx);
    s->data = data;
    s->count = count;
    s->capacity = capacity;
    try {
      s->total = alloc(ctx, 16, true);
      s->ws = alloc(ctx, dio::scan_workspace_bytes(capacity), true);
    } catch (...) {
      release(s->total);
      delete s;
      throw;
    }
    *out = s;
  });
}

extern "C" dio_buf* dio_prefix_sum_total(dio_prefix_sum* scan) { return scan ? scan->total : nullptr; }

extern "C" int dio_prefix_sum_encode(dio_prefix_sum* s) {
  return guard([&] {
    need(s, "prefix sum");
    dio::scan_exclusive(s->ctx->stream, dev<uint32_t>(s->data), dev<uint32_t>(s->data), dev<uint32_t>(s->count),
                        s->capacity, dev<uint32_t>(s->total), dio::scan_workspace(s->ws->ptr, s->capacity));
  });
}

extern "C" int dio_prefix_sum_encode_static(dio_ctx* ctx, dio_buf* data, uint32_t capacity, const dio_buf* count,
                                            dio_buf* total) {
  return guard([&] {
    need(ctx, "ctx");
    need(data, "data");
    need(total, "total");
    if (capacity > dio::kStaticScanMaxValues) {
      invalid("the allocation-free prefix sum takes at most " + std::to_string(dio::kStaticScanMaxValues) + " values");
    }
    if (static_cast<size_t>(capacity) * 4 > data->bytes) invalid("the data buffer is smaller than the capacity");
    if (total->bytes < 4) invalid("the total buffer is smaller than 4 bytes");
    if (count && count->bytes < 4) invalid("the count buffer is smaller than 4 bytes");
    const dio::ScanWorkspace ws = dio::static_scan_workspace(ctx->device, capacity);
    // The tile states, the ticket and the finish counter are one run of words; a
    // fill kernel (no memset node) clears them, whatever an earlier call left.
    dio::fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(ws.tile_state), 0u, static_cast<size_t>(ws.tiles) * 2 + 2);
    dio::scan_exclusive(ctx->stream, dev<uint32_t>(data), dev<uint32_t>(data), count ? dev<uint32_t>(const_cast<dio_buf*>(count)) : nullptr,
                        capacity, dev<uint32_t>(total), ws);
  });
}

extern "C" void dio_prefix_sum_destroy(dio_prefix_sum* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->ctx->stream);
  release(s->total);
  release(s->ws);
  dio_ctx* owner = s->ctx;
  delete s;
  dio::drop(owner);
}

// ================================================================= RadixSort

struct dio_radix_sort {
  dio_ctx* ctx = nullptr;
  uint32_t capacity = 0, high_bits = 0;
  bool wide = false;
  dio_buf* count = nullptr;
  dio_buf *keys = nullptr, *values = nullptr, *alt_keys = nullptr, *alt_values = nullptr, *ws = nullptr;
};

extern "C" int dio_radix_sort_create(dio_ctx* ctx, uint32_t capacity, uint32_t high_bits, dio_buf* count,
                                     dio_radix_sort** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    if (high_bits > 32) invalid("highBits must be at most 32");
    std::unique_ptr<dio_radix_sort> s(new dio_radix_sort);
    s->ctx = ctx;
    dio::hold(s->ctx);
    s->capacity = capacity;
    s->high_bits = high_bits;
    s->wide = high_bits > 0;
    s->count = count;
    const size_t kb = s->wide ? 8 : 4;
    try {
      s->keys = alloc(ctx, static_cast<size_t>(capacity) * kb, true);
      s->values = alloc(ctx, static_cast<size_t>(capacity) * 4, true);
      s->alt_keys = alloc(ctx, static_cast<size_t>(capacity) * kb);
      s->alt_values = alloc(ctx, static_cast<size_t>(capacity) * 4);
      s->ws = alloc(ctx, dio::sort_workspace_bytes(capacity), true);
    } catch (...) {
      release(s->keys); release(s->values); release(s->alt_keys); release(s->alt_values); release(s->ws);
      throw;
    }
    *out = s.release();
  });
}

extern "C" dio_buf* dio_radix_sort_keys(dio_radix_sort* s) { return s ? s->keys : nullptr; }
extern "C" dio_buf* dio_radix_sort_values(dio_radix_sort* s) { return s ? s->values : nullptr; }
extern "C" uint32_t dio_radix_sort_capacity(const dio_radix_sort* s) { return s ? s->capacity : 0; }

extern "C" int dio_radix_sort_encode(dio_radix_sort* s) {
  return guard([&] {
    need(s, "radix sort");
    // Four passes over the low word, then two per 16 bits of the high word
    // (src/radixSort.ts:103-104): an even count, so the result lands back in place.
    const uint32_t end = s->wide ? 32u + 16u * ((s->high_bits + 15) / 16) : 32u;
    const bool flipped = dio::radix_sort(s->ctx->stream, s->wide, s->keys->ptr, dev<uint32_t>(s->values),
                                         s->alt_keys->ptr, dev<uint32_t>(s->alt_values), dev<uint32_t>(s->count),
                                         s->capacity, 0, end, dio::sort_workspace(s->ws->ptr, s->capacity));
    if (flipped) throw Error(DIO_ERR_INTERNAL, "radix sort ended in its scratch buffers");
  });
}

extern "C" void dio_radix_sort_destroy(dio_radix_sort* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->ctx->stream);
  release(s->keys); release(s->values); release(s->alt_keys); release(s->alt_values); release(s->ws);
  dio_ctx* owner = s->ctx;
  delete s;
  dio::drop(owner);
}

// ============================================================ TileRasterizer

namespace {

// Stage boundaries of a plain training step, in launch order: event k closes stage
// k - 1 (kStageNames) and opens stage k.
enum StageMark : uint32_t {
  kMarkStart, kMarkPreprocess, kMarkDepthSort, kMarkScan, kMarkDuplicate, kMarkTileSort, kMarkRanges,
  kMarkListOrder, kMarkRasterize, kMarkLoss, kMarkClearGrads, kMarkBackward, kMarkBackwardAdam, kMarks
};
constexpr const char* kStageNames[kMarks - 1] = {
    "preprocess", "depth_sort", "scan", "duplicate", "tile_sort", "ranges", "list_order",
    "rasterize", "loss", "clear_grads", "rasterize_backward", "backward_adam"};

// HIP events between the stages of clocked steps, summed per stage, and each step's
// pair count. A step's records are read once they are old (kKeep steps behind) or
// when the totals are asked for, so the clock never waits on the step being enqueued.
struct StageClock {
  struct Step {
    hipEvent_t marks[kMarks] = {};
    hipEvent_t read = nullptr;   // after the pair-count copy
    uint32_t* pairs = nullptr;   // pinned host word
    void init() {
      // The stage marks only time: without the system-scope fence a mark does not
      // write back and invalidate the caches under the next stage.
      for (auto& e : marks) DIO_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      DIO_HIP(hipEventCreate(&read));
      DIO_HIP(hipHostMalloc(reinterpret_cast<void**>(&pairs), 4, hipHostMallocDefault));
    }
    ~Step() {
      for (auto e : marks)
        if (e) (void)hipEventDestroy(e);
      if (read) (void)hipEventDestroy(read);
      if (pairs) (void)hipHostFree(pairs);
    }
  };
  static constexpr size_t kKeep = 16;
  bool enabled = false;
  std::deque<std::unique_ptr<Step>> pending;
  std::vector<std::unique_ptr<Step>> spare;
  std::unique_ptr<Step> current;
  double totals_ms[kMarks - 1] = {};
  uint64_t pairs = 0;
  uint32_t steps = 0;

  void mark(hipStream_t s, StageMark k) {
    if (current) DIO_HIP(hipEventRecord(current->marks[k], s));
  }
  void begin(hipStream_t s) {
    drop();  // a step that an error cut short
    while (pending.size() >= kKeep) harvest_oldest();
    if (!spare.empty()) {
      current = std::move(spare.back());
      spare.pop_back();
    } else {
      current = std::make_unique<Step>();
      current->init();
    }
    mark(s, kMarkStart);
  }
  // Closes a step that ran every stage: its pair count is copied after the last mark,
  // outside every timed stage.
  void finish(hipStream_t s, const void* pair_total) {
    if (!current) return;
    DIO_HIP(hipMemcpyAsync(current->pairs, pair_total, 4, hipMemcpyDeviceToHost, s));
    DIO_HIP(hipEventRecord(current->read, s));
    pending.push_back(std::move(current));
  }
  // Drops a step that left the plain stage order (a densification step) unread.
  void drop() {
    if (current) spare.push_back(std::move(current));
  }
  void harvest_oldest() {
    Step& m = *pending.front();
    DIO_HIP(hipEventSynchronize(m.read));
    for (uint32_t k = 0; k + 1 < kMarks; k++) {
      float ms = 0.f;
      DIO_HIP(hipEventElapsedTime(&ms, m.marks[k], m.marks[k + 1]));
      totals_ms[k] += ms;
    }
    pairs += *m.pairs;
    steps++;
    spare.push_back(std::move(pending.front()));
    pending.pop_front();
  }
  void harvest_all() {
    while (!pending.empty()) harvest_oldest();
  }
  void reset() {
    for (double& v : totals_ms) v = 0;
    pairs = 0;
    steps = 0;
  }
  void destroy() {
    pending.clear();
    spare.clear();
    current.reset();
  }
};

}  // namespace

struct dio_tile_rasterizer {
  dio_ctx* ctx = nullptr;
  dio_buf* splats = nullptr;  // not owned
  uint32_t count = 0;
  uint32_t capacity = 0;
  // Per splat: tile counts, depth sort keys/values (+ scratch), the tile counts in
  // depth order (written by the depth sort's last pass), pair offsets, and the tile
  // (or bin) rectangle, packed (dio::pack_rect), for binning by counts.
  dio_buf* rects = nullptr;
  dio_buf *tile_counts = nullptr, *depth_keys = nullptr, *depth_values = nullptr, *depth_alt_keys = nullptr,
          *depth_alt_values = nullptr, *sorted_counts = nullptr, *offsets = nullptr, *splat_scan_ws = nullptr,
          *splat_sort_ws = nullptr;
  dio_buf* pair_total = nullptr;  // element 0: pairs this frame needs
  // Pairs: tile index and splat, ping-pong buffers for the tile sort.
  dio_buf *pair_tiles = nullptr, *pair_splats = nullptr, *alt_tiles = nullptr, *alt_splats = nullptr,
          *pair_sort_ws = nullptr;
  // Binning by counts: the pairs as (depth key, splat), then a scratch of the same
  // size for the long lists' depth order (allocated with its first frame).
  dio_buf* pair_keyed = nullptr;
  bool sorted_in_alt = false;
  dio_buf* ranges = nullptr;
  dio_buf* tile_order = nullptr;  // the rasterizer's launch order, heaviest tiles first
  dio_buf* order_queue = nullptr;  // per-tile depth order: the lists left for radix passes (2 tiles + 2 words)
  uint32_t order_parity = 0;       // which of its two queues the next depth order fills
  uint32_t ranges_tiles = 0;
  dio_buf *pixel_color = nullptr, *pixel_last = nullptr;
  uint32_t pixel_count = 0;
  uint32_t last_pair_count = 0;
  bool readback_queued = false;
  uint32_t* readback_host = nullptr;  // pinned
  hipEvent_t readback_done = nullptr;
  bool framed = false;
  uint32_t frame_w = 0, frame_h = 0;
  float frame_bg[3] = {0.f, 0.f, 0.f};
  uint32_t splat_batch = 64, digit_bits = 8;  // tuning; results do not depend on them
  uint32_t tile_size = 16;  // tuning: 8, 16 or 32 (32: pairs binned by 32-px bins)
  bool persistent = false;  // tuning: persistent rasterizer workgroups (16-px tiles)
  // Tuning: 0 sorts the splats by depth before writing their pairs; 1 writes the
  // pairs in splat order, sorts them stably by tile and then each tile's list by
  // depth; 2 bins the pairs by tile from per-tile counts (no sort over the pairs)
  // and then sorts each tile's list by (depth, splat). The lists are the same.
  uint32_t per_tile_depth = 0;
  // Binning by counts (per_tile_depth 2): the chunk x tile count matrix, the chunk
  // totals and the tile totals.
  dio_buf *bin_matrix = nullptr, *bin_partial = nullptr, *bin_totals = nullptr;
  uint32_t bin_groups = 0, bin_tiles = 0;
  dio_buf* work = nullptr;  // their two tile counters, zeroed between launches
  // Tuning: the segmented backward's segment length 2^segment_shift (0: a workgroup
  // per tile); framed_shift is the one the last forward pass wrote checkpoints for.
  uint32_t segment_shift = 0, framed_shift = 0;
  dio_buf *ckpt = nullptr, *items = nullptr;
  size_t ckpt_entries = 0, item_room = 0;
  uint32_t last_item_count = 0;  // items of an earlier frame, from the readback
  uint32_t bin_shift() const { return tile_size == 32 ? 1u : 0u; }
  StageClock* clock = nullptr;  // set by a trainer only while it runs a clocked step
};

namespace {

void free_pairs(dio_tile_rasterizer* t) {
  release(t->pair_tiles); release(t->pair_splats); release(t->alt_tiles); release(t->alt_splats);
  release(t->pair_sort_ws); release(t->pair_keyed);
}

void create_pairs(dio_tile_rasterizer* t, uint32_t capacity) {
  t->capacity = capacity;
  t->pair_tiles = alloc(t->ctx, static_cast<size_t>(capacity) * 4);
  t->pair_splats = alloc(t->ctx, static_cast<size_t>(capacity) * 4, true);
  t->alt_tiles = alloc(t->ctx, static_cast<size_t>(capacity) * 4);
  t->alt_splats = alloc(t->ctx, static_cast<size_t>(capacity) * 4, true);
  t->pair_sort_ws = alloc(t->ctx, dio::sort_workspace_bytes(capacity), true);
}

void destroy_tiles(dio_tile_rasterizer* t) {
  if (!t) return;
  (void)hipStreamSynchronize(t->ctx->stream);
  release(t->rects);
  release(t->tile_counts); release(t->depth_keys); release(t->depth_values); release(t->depth_alt_keys);
  release(t->depth_alt_values); release(t->sorted_counts); release(t->offsets); release(t->splat_scan_ws);
  release(t->splat_sort_ws);
  release(t->pair_total);
  free_pairs(t);
  release(t->ranges);
  release(t->tile_order);
  release(t->order_queue);
  release(t->bin_matrix);
  release(t->bin_partial);
  release(t->bin_totals);
  release(t->work);
  release(t->ckpt);
  release(t->items);
  release(t->pixel_color);
  release(t->pixel_last);
  if (t->readback_host) (void)hipHostFree(t->readback_host);
  if (t->readback_done) (void)hipEventDestroy(t->readback_done);
  dio_ctx* owner = t->ctx;
  delete t;
  dio::drop(owner);
}

dio_tile_rasterizer* create_tiles(dio_ctx* ctx, dio_buf* splats, uint32_t count, uint32_t initial) {
  auto* t = new dio_tile_rasterizer;
  t->ctx = ctx;
  dio::hold(t->ctx);
  t->splats = splats;
  t->count = count;
  try {
    const size_t n = std::max<uint32_t>(count, 1);
    t->tile_counts = alloc(ctx, n * 4, true);
    t->rects = alloc(ctx, n * 8, true);
    t->depth_keys = alloc(ctx, n * 4);
    t->depth_values = alloc(ctx, n * 4);
    t->depth_alt_keys = alloc(ctx, n * 4);
    t->depth_alt_values = alloc(ctx, n * 4);
    t->sorted_counts = alloc(ctx, n * 4);
    t->offsets = alloc(ctx, n * 4);
    t->splat_scan_ws = alloc(ctx, dio::scan_workspace_bytes(count), true);
    t->splat_sort_ws = alloc(ctx, dio::sort_workspace_bytes(count), true);
    t->pair_total = alloc(ctx, 64, true);  // the pair count, then the segmented backward's item counts
    t->work = alloc(ctx, 16, true);
    create_pairs(t, initial ? initial : std::max<uint32_t>(1u << 16, count * 8u));
    DIO_HIP(hipHostMalloc(reinterpret_cast<void**>(&t->readback_host), 64, hipHostMallocDefault));
    DIO_HIP(hipEventCreateWithFlags(&t->readback_done, hipEventDisableTiming));
  } catch (...) {
    destroy_tiles(t);
    throw;
  }
  return t;
}

void ensure_frame_buffers(dio_tile_rasterizer* t, uint32_t tiles, uint32_t pixels) {
  if (!t->ranges || t->ranges_tiles < tiles) {
    release(t->ranges);
    release(t->tile_order);
    release(t->order_queue);
    t->ranges = alloc(t->ctx, static_cast<size_t>(tiles) * 8, true);
    t->tile_order = alloc(t->ctx, static_cast<size_t>(tiles) * 4, true);
    t->order_queue = alloc(t->ctx, (2 * static_cast<size_t>(tiles) + 2) * 4, true);
    t->order_parity = 0;
    t->ranges_tiles = tiles;
  }
  if (!t->pixel_color || t->pixel_count < pixels) {
    release(t->pixel_color);
    release(t->pixel_last);
    t->pixel_color = alloc(t->ctx, static_cast<size_t>(pixels) * 16, true);
    t->pixel_last = alloc(t->ctx, static_cast<size_t>(pixels) * 4, true);
    t->pixel_count = pixels;
  }
}

// The binning-by-counts workspace for `tiles` tiles (at most dio::kBinMaxTiles).
dio::BinWorkspace bin_workspace(dio_tile_rasterizer* t, uint32_t tiles) {
  const uint32_t groups = dio::bin_layout(t->count).groups;
  if (!t->bin_matrix || t->bin_groups < groups || t->bin_tiles < tiles) {
    release(t->bin_matrix);
    release(t->bin_partial);
    release(t->bin_totals);
    t->bin_matrix = alloc(t->ctx, static_cast<size_t>(groups) * tiles * 4);
    t->bin_partial = alloc(t->ctx, static_cast<size_t>(groups) * 4);
    t->bin_totals = alloc(t->ctx, static_cast<size_t>(tiles) * 4);
    t->bin_groups = groups;
    t->bin_tiles = tiles;
  }
  dio::BinWorkspace ws;
  ws.matrix = dev<uint32_t>(t->bin_matrix);
  ws.partial = dev<uint32_t>(t->bin_partial);
  ws.totals = dev<uint32_t>(t->bin_totals);
  ws.groups = t->bin_groups;
  ws.tiles = t->bin_tiles;
  return ws;
}

// Whether a frame of `tiles` tiles (or bins) bins its pairs by counts.
bool bins_by_counts(const dio_tile_rasterizer* t, uint32_t tiles) {
  return t->per_tile_depth == 2 && tiles <= dio::kBinMaxTiles;
}

// The keyed pairs of binning by counts (two words a pair), then their scratch.
uint32_t* keyed_pairs(dio_tile_rasterizer* t) {
  if (!t->pair_keyed) t->pair_keyed = alloc(t->ctx, static_cast<size_t>(t->capacity) * 16);
  return dev<uint32_t>(t->pair_keyed);
}

// The segmented backward's buffers for `tiles` tiles and the pair capacity, or
// nullptr when the segment length is 0. The forward pass (re)allocates them; the
// backward (allocate false) takes the ones its forward pass wrote, or nullptr if
// they no longer fit, and then replays whole lists. The item counts are words 1
// to 1 + kPartClasses of pair_total, read back with the pair count.
const dio::Segments* segments_for(dio_tile_rasterizer* t, uint32_t tiles, dio::Segments& seg, bool allocate) {
  const uint32_t shift = t->segment_shift;
  if (shift == 0) return nullptr;
  const size_t whole = (static_cast<size_t>(t->capacity) >> shift) + 1;
  const size_t room = whole + dio::kPartClasses * 2 * static_cast<size_t>(tiles);
  const bool fits = t->ckpt && t->ckpt_entries >= whole && t->items && t->item_room >= room;
  if (!allocate && !fits) return nullptr;
  if (!t->ckpt || t->ckpt_entries < whole) {
    release(t->ckpt);
    t->ckpt = alloc(t->ctx, whole * 4 * dio::kWave * 16);
    t->ckpt_entries = whole;
  }
  if (!t->items || t->item_room < room) {
    release(t->items);
    t->items = alloc(t->ctx, room * 8);
    t->item_room = room;
  }
  seg.shift = shift;
  seg.whole_room = static_cast<uint32_t>(whole);
  seg.part_room = 2 * tiles;
  seg.ckpt = dev<float>(t->ckpt);
  seg.items = dev<uint32_t>(t->items);
  seg.counts = dev<uint32_t>(t->pair_total) + 1;
  // Workgroups for about the last counted items; each loops over whatever there are.
  const uint32_t hint = t->last_item_count;
  seg.grid = hint ? hint + hint / 8 + 64 : std::max(2 * tiles, (t->last_pair_count >> shift) + 2 * tiles);
  return &seg;
}

// Bits of the tile index the tile sort must look at, in whole digits.
uint32_t tile_sort_bits(uint32_t tiles, uint32_t digit_bits = 8) {
  uint32_t bits = 0;
  while ((1u << bits) < tiles && bits < 32) bits++;
  return std::max(digit_bits, (bits + digit_bits - 1) / digit_bits * digit_bits);
}

// The tile stage of TileRasterizer.encode (src/tileRasterizer.ts:127-192) after
// the tile counts and depth keys exist.
void encode_tiles(dio_tile_rasterizer* t, dio_buf* target, uint32_t w, uint32_t h, const float bg[3],
                  bool counts_ready) {
  check_tile_limit(w, h);
  // Pair lists per 16-px tile, or per 32-px bin (tile size 32).
  const uint32_t shift = t->bin_shift();
  const uint32_t tiles_x = dio::ceil_div(dio::ceil_div(w, dio::kTile), 1u << shift);
  const uint32_t tiles_y = dio::ceil_div(dio::ceil_div(h, dio::kTile), 1u << shift);
  const uint32_t tiles = tiles_x * tiles_y;
  if (target && target->bytes < static_cast<size_t>(w) * h * 4) invalid("the target is smaller than the viewport");
  ensure_frame_buffers(t, tiles, w * h);
  hipStream_t s = t->ctx->stream;
  const float* splats = dev<float>(t->splats);
  if (!counts_ready)
    dio::count_tiles(s, splats, t->count, w, h, dev<uint32_t>(t->tile_counts), dev<uint32_t>(t->depth_keys),
                     dev<uint32_t>(t->depth_values), shift, dev<uint32_t>(t->rects));
  auto mark = [&](StageMark k) {
    if (t->clock) t->clock->mark(s, k);
  };
  dio::Segments seg;
  const dio::Segments* segments =
      t->tile_size == 16 && !t->persistent ? segments_for(t, tiles, seg, true) : nullptr;
  dio_buf* sorted_splats = nullptr;
  if (bins_by_counts(t, tiles)) {
    // Pairs binned by tile from counts, then each tile's list by (depth, splat).
    const dio::BinWorkspace ws = bin_workspace(t, tiles);
    mark(kMarkDepthSort);
    dio::bin_count(s, dev<uint32_t>(t->rects), dev<uint32_t>(t->depth_keys), dev<uint32_t>(t->tile_counts), t->count,
                   w, h, shift, t->capacity, dev<uint32_t>(t->pair_total), ws);
    mark(kMarkScan);
    uint32_t* keyed = keyed_pairs(t);
    dio::bin_scatter(s, dev<uint32_t>(t->rects), dev<uint32_t>(t->tile_counts), dev<uint32_t>(t->depth_keys), t->count,
                     w, h, shift, ws, keyed, dev<uint32_t>(t->ranges), dev<uint32_t>(t->tile_order),
                     segments ? segments->counts : nullptr);
    mark(kMarkDuplicate);
    mark(kMarkTileSort);
    mark(kMarkRanges);
    t->sorted_in_alt = false;
    sorted_splats = t->pair_splats;
    dio::depth_order(s, dev<uint32_t>(t->depth_keys), dev<uint32_t>(t->ranges), dev<uint32_t>(t->tile_order), tiles,
                     dev<uint32_t>(t->pair_splats), keyed + 2 * static_cast<size_t>(t->capacity), false,
                     dev<uint32_t>(t->order_queue), t->order_parity++, keyed);
    mark(kMarkListOrder);
  } else {
    // Depth order of the splats (culled ones last), with their tile counts in that
    // order, then each splat's first pair slot. Per-tile depth order skips the sort:
    // pairs are written in splat order and each tile's list is sorted after binning.
    const bool per_tile = t->per_tile_depth != 0;
    if (!per_tile &&
        dio::radix_sort(s, false, t->depth_keys->ptr, dev<uint32_t>(t->depth_values), t->depth_alt_keys->ptr,
                        dev<uint32_t>(t->depth_alt_values), nullptr, t->count, 0, 32,
                        dio::sort_workspace(t->splat_sort_ws->ptr, t->count), t->digit_bits,
                        dev<uint32_t>(t->tile_counts), dev<uint32_t>(t->sorted_counts)))
      throw Error(DIO_ERR_INTERNAL, "depth sort ended in its scratch buffers");
    mark(kMarkDepthSort);
    dio::scan_exclusive(s, dev<uint32_t>(per_tile ? t->tile_counts : t->sorted_counts), dev<uint32_t>(t->offsets),
                        nullptr, t->count, dev<uint32_t>(t->pair_total),
                        dio::scan_workspace(t->splat_scan_ws->ptr, t->count));
    // Pairs in splat order: a frame past the capacity keeps the nearest pair slots, as
    // the depth-ordered duplicate does.
    if (per_tile)
      dio::pair_cut(s, dev<uint32_t>(t->depth_keys), dev<uint32_t>(t->tile_counts), dev<uint32_t>(t->offsets),
                    t->count, dev<uint32_t>(t->pair_total), t->capacity);
    mark(kMarkScan);
    // The pair count is known on the device only; the last frame's (0 before the first
    // readback) sizes the grids of the pair kernels, which loop over whatever it is.
    const uint32_t pair_hint = t->last_pair_count;
    dio::duplicate(s, splats, per_tile ? nullptr : dev<uint32_t>(t->depth_values), dev<uint32_t>(t->offsets),
                   dev<uint32_t>(t->tile_counts), t->count, w, h, t->capacity, dev<uint32_t>(t->pair_total),
                   dev<uint32_t>(t->pair_tiles), dev<uint32_t>(t->pair_splats), pair_hint, shift);
    mark(kMarkDuplicate);
    t->sorted_in_alt = dio::radix_sort(s, false, t->pair_tiles->ptr, dev<uint32_t>(t->pair_splats),
                                       t->alt_tiles->ptr, dev<uint32_t>(t->alt_splats), dev<uint32_t>(t->pair_total),
                                       t->capacity, 0, tile_sort_bits(tiles, t->digit_bits),
                                       dio::sort_workspace(t->pair_sort_ws->ptr, t->capacity), t->digit_bits, nullptr,
                                       nullptr, pair_hint ? pair_hint + pair_hint / 8 + 4096 : 0);
    mark(kMarkTileSort);
    dio_buf* sorted_tiles = t->sorted_in_alt ? t->alt_tiles : t->pair_tiles;
    sorted_splats = t->sorted_in_alt ? t->alt_splats : t->pair_splats;
    dio::find_ranges(s, dev<uint32_t>(sorted_tiles), dev<uint32_t>(t->pair_total), t->capacity, tiles,
                     dev<uint32_t>(t->ranges), dev<uint32_t>(t->tile_order), pair_hint,
                     segments ? segments->counts : nullptr);
    mark(kMarkRanges);
    if (per_tile)
      dio::depth_order(s, dev<uint32_t>(t->depth_keys), dev<uint32_t>(t->ranges), dev<uint32_t>(t->tile_order),
                       tiles, dev<uint32_t>(sorted_splats),
                       dev<uint32_t>(t->sorted_in_alt ? t->pair_splats : t->alt_splats), true,
                       dev<uint32_t>(t->order_queue), t->order_parity++);
    mark(kMarkListOrder);
  }
  dio::rasterize(s, splats, dev<uint32_t>(sorted_splats), dev<uint32_t>(t->ranges), dev<uint32_t>(t->tile_order), w,
                 h, bg, dev<uint32_t>(target), dev<float>(t->pixel_color), dev<uint32_t>(t->pixel_last),
                 t->splat_batch, t->tile_size, t->persistent ? dev<uint32_t>(t->work) : nullptr, segments);
  t->framed_shift = segments ? segments->shift : 0;
  mark(kMarkRasterize);
  t->framed = true;
  t->frame_w = w;
  t->frame_h = h;
  for (int k = 0; k < 3; k++) t->frame_bg[k] = bg[k];
  if (!t->readback_queued) {
    // The pair count, and the segmented backward's item counts.
    DIO_HIP(hipMemcpyAsync(t->readback_host, t->pair_total->ptr, 4 * (2 + dio::kPartClasses), hipMemcpyDeviceToHost,
                           s));
    DIO_HIP(hipEventRecord(t->readback_done, s));
    t->readback_queued = true;
  }
}

// afterSubmit: `wait` blocks for the readback; otherwise only a finished one is taken.
void tiles_after_submit(dio_tile_rasterizer* t, bool wait) {
  if (!t->readback_queued) return;
  if (wait) {
    DIO_HIP(hipEventSynchronize(t->readback_done));
  } else {
    hipError_t e = hipEventQuery(t->readback_done);
    if (e == hipErrorNotReady) return;
    dio::check_hip(e, "hipEventQuery");
  }
  t->readback_queued = false;
  t->last_pair_count = t->readback_host[0];
  uint32_t items = 0;
  for (uint32_t c = 0; c <= dio::kPartClasses; c++) items += t->readback_host[1 + c];
  if (items) t->last_item_count = items;
  if (t->last_pair_count > t->capacity) {
    sync(t->ctx);
    free_pairs(t);
    create_pairs(t, next_pow2_capacity(t->last_pair_count));
  }
}

// encodeBackward (src/tileRasterizer.ts:200-222) of the frame encoded last. The pass
// adds into splat_grads, which is cleared first unless the caller already did.
void backward_tiles(dio_tile_rasterizer* t, uint32_t w, uint32_t h, dio_buf* pixel_grads, dio_buf* splat_grads,
                    bool cleared = false, dio_buf* pixel_loss = nullptr, dio_buf* footprint = nullptr) {
  hipStream_t s = t->ctx->stream;
  if (!cleared) dio::fill_u32(s, dev<uint32_t>(splat_grads), 0u, static_cast<size_t>(t->count) * 9);
  if (t->clock) t->clock->mark(s, kMarkClearGrads);
  dio_buf* sorted_splats = t->sorted_in_alt ? t->alt_splats : t->pair_splats;
  // Segmented only over the checkpoints of this frame's forward pass.
  dio::Segments seg;
  const dio::Segments* segments =
      t->framed_shift && t->framed_shift == t->segment_shift
          ? segments_for(t, dio::ceil_div(w, dio::kTile) * dio::ceil_div(h, dio::kTile), seg, false)
          : nullptr;
  dio::rasterize_backward(s, dev<float>(t->splats), dev<uint32_t>(sorted_splats), dev<uint32_t>(t->ranges),
                          dev<uint32_t>(t->tile_order), w, h, t->frame_bg, dev<float>(t->pixel_color),
                          dev<uint32_t>(t->pixel_last), dev<float>(pixel_grads), dev<float>(splat_grads),
                          t->splat_batch, t->tile_size, dev<float>(pixel_loss), dev<float>(footprint),
                          t->persistent ? dev<uint32_t>(t->work) : nullptr, segments);
  if (t->clock) t->clock->mark(s, kMarkBackward);
}

}  // namespace

extern "C" int dio_tile_rasterizer_create(dio_ctx* ctx, dio_buf* splats, uint32_t count,
                                          uint32_t initial_pair_capacity, dio_tile_rasterizer** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(splats, "splats");
    need(out, "out");
    if (splats->bytes < static_cast<size_t>(count) * dio::kFloatsPerSplat * 4)
      invalid("the splat buffer is smaller than count splats");
    *out = create_tiles(ctx, splats, count, initial_pair_capacity);
  });
}

extern "C" int dio_tile_rasterizer_encode(dio_tile_rasterizer* t, dio_buf* target, uint32_t width, uint32_t height,
                                          const float background[3]) {
  return guard([&] {
    need(t, "tile rasterizer");
    need(background, "background");
    encode_tiles(t, target, width, height, background, false);
  });
}

extern "C" int dio_tile_rasterizer_encode_backward(dio_tile_rasterizer* t, uint32_t width, uint32_t height,
                                                   dio_buf* pixel_grads, dio_buf* splat_grads) {
  return guard([&] {
    need(t, "tile rasterizer");
    if (!t->framed) invalid("Encode a frame before its backward pass.");
    need(pixel_grads, "pixel gradients");
    need(splat_grads, "splat gradients");
    if (splat_grads->bytes < static_cast<size_t>(t->count) * 36) invalid("the splat gradient buffer is too small");
    if (pixel_grads->bytes < static_cast<size_t>(width) * height * 16) invalid("the pixel gradient buffer is too small");
    backward_tiles(t, width, height, pixel_grads, splat_grads);
  });
}

namespace {

// The launch settings any tile rasterizer takes (dio_tuning); trainers add their own.
void check_launch_tuning(const dio_tuning& k) {
  if (k.tile_size != 8 && k.tile_size != 16 && k.tile_size != 32) invalid("the tile size must be 8, 16 or 32 pixels");
  if (k.splat_batch != 64 && k.splat_batch != 128 && k.splat_batch != 256) invalid("splat batch must be 64, 128 or 256");
  if (k.tile_size == 32 && k.splat_batch == 256) invalid("32-pixel tiles take a splat batch of 64 or 128");
  if (k.digit_bits != 4 && k.digit_bits != 8) invalid("digit width must be 4 or 8 bits");
  if (k.persistent > 1) invalid("persistent is 0 or 1");
  if (k.persistent && k.tile_size != 16) invalid("persistent workgroups rasterize 16-pixel tiles");
  if (k.segment != 0 && k.segment != 256 && k.segment != 512 && k.segment != 1024)
    invalid("segment must be 0, 256, 512 or 1024 pairs");
  if (k.segment && (k.tile_size != 16 || k.persistent))
    invalid("the segmented backward runs on 16-pixel tiles, one workgroup each");
  if (k.per_tile_depth > 2) invalid("per_tile_depth is 0, 1 or 2");
}

}  // namespace

extern "C" int dio_tile_rasterizer_set_tuning(dio_tile_rasterizer* t, const dio_tuning* k) {
  return guard([&] {
    need(t, "tile rasterizer");
    need(k, "tuning");
    check_launch_tuning(*k);
    t->tile_size = k->tile_size;
    t->splat_batch = k->splat_batch;
    t->digit_bits = k->digit_bits;
    t->persistent = k->persistent != 0;
    t->segment_shift = k->segment ? static_cast<uint32_t>(__builtin_ctz(k->segment)) : 0u;
    t->per_tile_depth = k->per_tile_depth;
    t->framed = false;  // the next backward pass needs a frame of these settings
  });
}

extern "C" int dio_tile_rasterizer_after_submit(dio_tile_rasterizer* t) {
  return guard_read([&] {
    need(t, "tile rasterizer");
    tiles_after_submit(t, true);
  });
}

extern "C" uint32_t dio_tile_rasterizer_last_pair_count(const dio_tile_rasterizer* t) {
  return t ? t->last_pair_count : 0;
}
extern "C" uint32_t dio_tile_rasterizer_pair_capacity(const dio_tile_rasterizer* t) { return t ? t->capacity : 0; }
extern "C" dio_buf* dio_tile_rasterizer_pair_splats(dio_tile_rasterizer* t) {
  return t ? (t->sorted_in_alt ? t->alt_splats : t->pair_splats) : nullptr;
}
extern "C" dio_buf* dio_tile_rasterizer_ranges(dio_tile_rasterizer* t) { return t ? t->ranges : nullptr; }
extern "C" dio_buf* dio_tile_rasterizer_pixel_color(dio_tile_rasterizer* t) { return t ? t->pixel_color : nullptr; }
extern "C" dio_buf* dio_tile_rasterizer_pixel_last(dio_tile_rasterizer* t) { return t ? t->pixel_last : nullptr; }
extern "C" void dio_tile_rasterizer_destroy(dio_tile_rasterizer* t) { destroy_tiles(t); }

// ========================================================== GaussianRenderer

struct dio_renderer {
  dio_ctx* ctx = nullptr;
  uint32_t initial_pair_capacity = 0;
  struct Scene {
    uint32_t count = 0;
    dio_buf *gaussians = nullptr, *splats = nullptr, *sh = nullptr;
    uint32_t degree = 0;
    dio_tile_rasterizer* rasterizer = nullptr;
    dio_buf *grad_splats = nullptr, *grad_gaussians = nullptr, *grad_sh = nullptr;
  };
  Scene* scene = nullptr;
  dio_tile_rasterizer* draw_rasterizer = nullptr;
  dio::CameraArgs camera{};
  bool preprocessed = false;
};

namespace {

void destroy_scene(dio_renderer* r) {
  auto* sc = r->scene;
  if (!sc) return;
  (void)hipStreamSynchronize(r->ctx->stream);
  release(sc->gaussians); release(sc->splats); release(sc->sh);
  release(sc->grad_splats); release(sc->grad_gaussians); release(sc->grad_sh);
  destroy_tiles(sc->rasterizer);
  delete sc;
  r->scene = nullptr;
}

// Takes ownership of `gaussians` and `sh` (createScene, src/gaussianRenderer.ts:246-270).
void create_scene(dio_renderer* r, uint32_t count, dio_buf* gaussians, dio_buf* sh, uint32_t degree,
                  uint32_t pair_capacity) {
  auto* sc = new dio_renderer::Scene;
  sc->count = count;
  sc->gaussians = gaussians;
  sc->sh = sh;
  sc->degree = degree;
  r->scene = sc;
  sc->splats = alloc(r->ctx, static_cast<size_t>(count) * dio::kFloatsPerSplat * 4, true);
  sc->rasterizer = create_tiles(r->ctx, sc->splats, count, pair_capacity);
}

void fill_background(dio_ctx* ctx, dio_buf* target, uint32_t w, uint32_t h, const float bg[3]) {
  if (!target) return;
  auto u8 = [](float v) { return static_cast<uint32_t>(std::min(std::max(v, 0.f), 1.f) * 255.f + 0.5f); };
  dio::fill_u32(ctx->stream, dev<uint32_t>(target), u8(bg[0]) | (u8(bg[1]) << 8) | (u8(bg[2]) << 16) | (255u << 24),
                std::min<size_t>(static_cast<size_t>(w) * h, target->bytes / 4));
}

// `vis` (nullable): each Gaussian's (depth, radius) for the trainer's fused backward.
void renderer_preprocess(dio_renderer* r, const dio_camera& camera, dio_buf* vis = nullptr) {
  auto* sc = r->scene;
  to_camera(camera, r->camera, sc ? sc->rasterizer->bin_shift() : 0u);
  if (!sc) return;
  auto* t = sc->rasterizer;
  dio::preprocess(r->ctx->stream, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree, sc->count, r->camera,
                  dev<float>(sc->splats), dev<uint32_t>(t->tile_counts), dev<uint32_t>(t->depth_keys),
                  dev<uint32_t>(t->depth_values), vis ? dev<float>(vis) : nullptr, dev<uint32_t>(t->rects));
  if (t->clock) t->clock->mark(r->ctx->stream, kMarkPreprocess);
  r->preprocessed = true;
}

void renderer_tiles(dio_renderer* r, dio_buf* target, const float bg[3]) {
  const uint32_t w = viewport_dim(r->camera.viewport[0]), h = viewport_dim(r->camera.viewport[1]);
  if (!r->scene) {
    fill_background(r->ctx, target, w, h, bg);
    return;
  }
  if (!r->preprocessed) invalid("encode_preprocess must come before encode_tiles");
  r->preprocessed = false;
  encode_tiles(r->scene->rasterizer, target, w, h, bg, true);
}

void ensure_grad_buffers(dio_renderer* r) {
  auto* sc = r->scene;
  if (!sc->grad_splats) {
    sc->grad_splats = alloc(r->ctx, static_cast<size_t>(sc->count) * 9 * 4, true);
    sc->grad_gaussians = alloc(r->ctx, static_cast<size_t>(sc->count) * dio::kFloatsPerGaussian * 4, true);
    sc->grad_sh = alloc(r->ctx, static_cast<size_t>(sc->count) * dio::kShRest * 4, true);
  }
}

// The rasterizer's half of encodeBackward: per-splat gradients into grad_splats
// (already zeroed when `cleared`).
// pixel_loss and footprint (both or neither): also gather the footprint terms.
void renderer_backward_splats(dio_renderer* r, dio_buf* pixel_grads, bool cleared = false,
                              dio_buf* pixel_loss = nullptr, dio_buf* footprint = nullptr) {
  auto* sc = r->scene;
  ensure_grad_buffers(r);
  auto* t = sc->rasterizer;
  if (!t->framed) invalid("Encode a frame before its backward pass.");
  backward_tiles(t, t->frame_w, t->frame_h, pixel_grads, sc->grad_splats, cleared, pixel_loss, footprint);
}

void renderer_backward(dio_renderer* r, dio_buf* pixel_grads, bool cleared = false, dio_buf* pixel_loss = nullptr,
                       dio_buf* footprint = nullptr) {
  auto* sc = r->scene;
  if (!sc) return;
  renderer_backward_splats(r, pixel_grads, cleared, pixel_loss, footprint);
  dio::preprocess_backward(r->ctx->stream, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree, sc->count,
                           r->camera, dev<float>(sc->splats), dev<float>(sc->grad_splats),
                           dev<float>(sc->grad_gaussians), dev<float>(sc->grad_sh));
}

}  // namespace

extern "C" int dio_renderer_create(dio_ctx* ctx, uint32_t initial_pair_capacity, dio_renderer** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    auto* r = new dio_renderer;
    r->ctx = ctx;
    dio::hold(r->ctx);
    r->initial_pair_capacity = initial_pair_capacity;
    *out = r;
  });
}

extern "C" int dio_renderer_set_gaussians(dio_renderer* r, const float* gaussians, uint32_t count, const float* sh_rest,
                                          uint32_t sh_degree) {
  return guard([&] {
    need(r, "renderer");
    if (count == 0) {
      destroy_scene(r);
      return;
    }
    need(gaussians, "gaussians");
    dio_buf* g = alloc(r->ctx, static_cast<size_t>(count) * dio::kFloatsPerGaussian * 4);
    dio_buf* s = nullptr;
    try {
      s = alloc(r->ctx, sh_rest ? static_cast<size_t>(count) * dio::kShRest * 4 : 16, true);
      DIO_HIP(hipMemcpyAsync(g->ptr, gaussians, g->bytes, hipMemcpyHostToDevice, r->ctx->stream));
      if (sh_rest) DIO_HIP(hipMemcpyAsync(s->ptr, sh_rest, s->bytes, hipMemcpyHostToDevice, r->ctx->stream));
      sync(r->ctx);
    } catch (...) {
      release(g);
      release(s);
      throw;
    }
    destroy_scene(r);
    // Without coefficients the shader never reads them; the degree is then 0.
    create_scene(r, count, g, s, sh_rest ? sh_degree : 0, r->initial_pair_capacity);
    if (!sh_rest) r->scene->degree = 0;
  });
}

extern "C" int dio_renderer_adopt_gaussians(dio_renderer* r, uint32_t count, dio_buf* gaussians, dio_buf* sh,
                                            uint32_t sh_degree) {
  return guard([&] {
    need(r, "renderer");
    need(gaussians, "gaussians");
    need(sh, "sh");
    const uint32_t capacity = r->scene ? r->scene->rasterizer->capacity : r->initial_pair_capacity;
    destroy_scene(r);
    create_scene(r, count, gaussians, sh, sh_degree, capacity);
  });
}

extern "C" int dio_renderer_set_sh_degree(dio_renderer* r, uint32_t degree) {
  return guard([&] {
    need(r, "renderer");
    if (degree > 3) invalid("the SH degree is at most 3");
    if (r->scene) r->scene->degree = degree;
  });
}

extern "C" int dio_renderer_encode(dio_renderer* r, dio_buf* target, const dio_camera* camera,
                                   const float background[3]) {
  return guard([&] {
    need(r, "renderer");
    need(camera, "camera");
    need(background, "background");
    renderer_preprocess(r, *camera);
    renderer_tiles(r, target, background);
  });
}

extern "C" int dio_renderer_encode_preprocess(dio_renderer* r, const dio_camera* camera) {
  return guard([&] {
    need(r, "renderer");
    need(camera, "camera");
    renderer_preprocess(r, *camera);
  });
}

extern "C" int dio_renderer_encode_tiles(dio_renderer* r, dio_buf* target, const float background[3]) {
  return guard([&] {
    need(r, "renderer");
    need(background, "background");
    renderer_tiles(r, target, background);
  });
}

extern "C" int dio_renderer_encode_draw(dio_renderer* r, dio_buf* target, dio_buf* splats, uint32_t count,
                                        uint32_t width, uint32_t height, const float background[3]) {
  return guard([&] {
    need(r, "renderer");
    need(splats, "splats");
    need(background, "background");
    if (splats->bytes < static_cast<size_t>(count) * dio::kFloatsPerSplat * 4)
      invalid("the splat buffer is smaller than count splats");
    destroy_tiles(r->draw_rasterizer);
    r->draw_rasterizer = nullptr;
    r->draw_rasterizer = create_tiles(r->ctx, splats, count, r->initial_pair_capacity);
    encode_tiles(r->draw_rasterizer, target, width, height, background, false);
  });
}

extern "C" int dio_renderer_encode_backward(dio_renderer* r, dio_buf* pixel_grads) {
  return guard([&] {
    need(r, "renderer");
    need(pixel_grads, "pixel gradients");
    renderer_backward(r, pixel_grads);
  });
}

extern "C" int dio_renderer_after_submit(dio_renderer* r) {
  return guard_read([&] {
    need(r, "renderer");
    if (r->scene) tiles_after_submit(r->scene->rasterizer, true);
  });
}

extern "C" uint32_t dio_renderer_count(const dio_renderer* r) { return r && r->scene ? r->scene->count : 0; }
extern "C" dio_buf* dio_renderer_splat_buffer(dio_renderer* r) { return r && r->scene ? r->scene->splats : nullptr; }
extern "C" dio_buf* dio_renderer_gaussian_buffer(dio_renderer* r) {
  return r && r->scene ? r->scene->gaussians : nullptr;
}
extern "C" dio_buf* dio_renderer_sh_buffer(dio_renderer* r) { return r && r->scene ? r->scene->sh : nullptr; }
extern "C" dio_tile_rasterizer* dio_renderer_tiles(dio_renderer* r) {
  if (!r) return nullptr;
  if (r->draw_rasterizer) return r->draw_rasterizer;
  return r->scene ? r->scene->rasterizer : nullptr;
}
extern "C" dio_buf* dio_renderer_pixel_buffer(dio_renderer* r) {
  return r && r->scene ? r->scene->rasterizer->pixel_color : nullptr;
}
extern "C" dio_buf* dio_renderer_pixel_last_buffer(dio_renderer* r) {
  return r && r->scene ? r->scene->rasterizer->pixel_last : nullptr;
}
extern "C" dio_buf* dio_renderer_gradient_gaussians(dio_renderer* r) {
  return r && r->scene ? r->scene->grad_gaussians : nullptr;
}
extern "C" dio_buf* dio_renderer_gradient_sh(dio_renderer* r) { return r && r->scene ? r->scene->grad_sh : nullptr; }
extern "C" dio_buf* dio_renderer_gradient_splats(dio_renderer* r) {
  return r && r->scene ? r->scene->grad_splats : nullptr;
}

extern "C" void dio_renderer_destroy(dio_renderer* r) {
  if (!r) return;
  destroy_scene(r);
  destroy_tiles(r->draw_rasterizer);
  dio_ctx* owner = r->ctx;
  delete r;
  dio::drop(owner);
}

// ============================================================== TrainingLoss

struct dio_training_loss {
  dio_ctx* ctx = nullptr;
  uint32_t width = 0, height = 0;
  dio_buf *partials = nullptr, *pixel_grads = nullptr, *pixel_loss = nullptr;
};

namespace {
void destroy_loss(dio_training_loss* l) {
  if (!l) return;
  (void)hipStreamSynchronize(l->ctx->stream);
  release(l->partials); release(l->pixel_grads); release(l->pixel_loss);
  dio_ctx* owner = l->ctx;
  delete l;
  dio::drop(owner);
}

double loss_read(dio_training_loss* l) {
  std::vector<float> shares(static_cast<size_t>(l->width) * l->height);
  DIO_HIP(hipMemcpyAsync(shares.data(), l->pixel_loss->ptr, shares.size() * 4, hipMemcpyDeviceToHost, l->ctx->stream));
  sync(l->ctx);
  double sum = 0;
  for (float v : shares) sum += v;
  return sum + 0.2;
}
}  // namespace

extern "C" int dio_training_loss_create(dio_ctx* ctx, uint32_t width, uint32_t height, dio_training_loss** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    auto* l = new dio_training_loss;
    l->ctx = ctx;
    dio::hold(l->ctx);
    l->width = width;
    l->height = height;
    const size_t px = std::max<size_t>(static_cast<size_t>(width) * height, 1);
    try {
      l->partials = alloc(ctx, px * 48);
      l->pixel_grads = alloc(ctx, px * 16, true);
      l->pixel_loss = alloc(ctx, px * 4, true);
    } catch (...) {
      destroy_loss(l);
      throw;
    }
    *out = l;
  });
}

extern "C" int dio_training_loss_encode(dio_training_loss* l, dio_buf* rendered, dio_buf* target) {
  return guard([&] {
    need(l, "loss");
    need(rendered, "rendered");
    need(target, "target");
    const size_t px = static_cast<size_t>(l->width) * l->height;
    if (rendered->bytes < px * 16 || target->bytes < px * 4) invalid("the loss inputs are smaller than the image");
    dio::training_loss(l->ctx->stream, l->width, l->height, dev<float>(rendered), dev<uint32_t>(target),
                       dev<float>(l->partials), dev<float>(l->pixel_grads), dev<float>(l->pixel_loss));
  });
}

extern "C" int dio_training_loss_read(dio_training_loss* l, double* out) {
  return guard_read([&] {
    need(l, "loss");
    need(out, "out");
    *out = loss_read(l);
  });
}

extern "C" dio_buf* dio_training_loss_pixel_grads(dio_training_loss* l) { return l ? l->pixel_grads : nullptr; }
extern "C" dio_buf* dio_training_loss_pixel_loss(dio_training_loss* l) { return l ? l->pixel_loss : nullptr; }
extern "C" void dio_training_loss_destroy(dio_training_loss* l) { destroy_loss(l); }

// =================================================================== Trainer

namespace {

// mulberry32 (src/random.ts:3-12).
struct Mulberry32 {
  uint32_t state;
  explicit Mulberry32(uint32_t seed) : state(seed) {}
  double next() {
    state += 0x6d2b79f5u;
    uint32_t t = state;
    t = (t ^ (t >> 15)) * (t | 1u);
    t ^= t + (t ^ (t >> 7)) * (t | 61u);
    return static_cast<double>(t ^ (t >> 14)) / 4294967296.0;
  }
};


}  // namespace

// The device state of a multi-view batch trainer (no counterpart in the
// reference): each step takes `views` training views of any sizes, renders,
// scores and back-propagates them in the same launches (batch.hip) and sums
// their gradients into one Adam step. The whole step is one captured HIP graph;
// the views, their cameras and the Adam arguments sit in device memory and are
// rewritten before each replay.
struct BatchState {
  uint32_t views = 0;                   // views per step
  uint32_t n = 0;                       // Gaussians per view in the frame layout (the trainer's count)
  uint32_t n_cap = 0;                   // Gaussians per view the frame buffers hold (n_cap >= n)
  uint32_t max_tiles = 0, max_pixels = 0;
  uint32_t capacity = 0;                // (splat, tile) pairs
  dio_buf *desc = nullptr, *adam = nullptr;
  dio_buf *splats = nullptr, *tile_counts = nullptr, *depth_keys = nullptr, *depth_values = nullptr,
          *depth_alt_keys = nullptr, *depth_alt_values = nullptr, *sorted_counts = nullptr, *offsets = nullptr,
          *scan_ws = nullptr,
          *sort_ws = nullptr;
  dio_buf *pair_total = nullptr, *pair_tiles = nullptr, *pair_splats = nullptr, *alt_tiles = nullptr,
          *alt_splats = nullptr, *pair_sort_ws = nullptr, *ranges = nullptr;
  dio_buf *pixel_color = nullptr, *pixel_last = nullptr, *partials = nullptr, *pixel_grads = nullptr,
          *pixel_loss = nullptr;
  dio_buf *splat_grads = nullptr, *grad_gaussians = nullptr, *grad_sh = nullptr;
  char* host = nullptr;            // pinned staging: header, views, Adam arguments
  uint32_t* readback = nullptr;    // pinned: the last step's pair count
  hipEvent_t uploaded = nullptr, read = nullptr;
  bool upload_pending = false, read_pending = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_degree = ~0u, graph_generation = ~0u;
  bool sorted_in_alt = false;
  std::vector<uint32_t> step_views;  // training views of the last step
  uint32_t step_pixels = 0;
  uint32_t last_pairs = 0;
};

constexpr size_t kBatchViewsOffset = 256;
constexpr size_t kBatchDescBytes = kBatchViewsOffset + dio::kMaxBatchViews * sizeof(dio::ViewDesc);
constexpr size_t kBatchHostBytes = kBatchDescBytes + sizeof(dio::AdamArgs);


// The learned densification proposal (no counterpart in the reference): the MLP's
// BF16 weights with their FP32 master copy and Adam state (mlp.hip).
struct dio_densify_mlp {
  dio_ctx* ctx = nullptr;
  dio_buf *weights = nullptr, *master = nullptr, *moments = nullptr, *grads = nullptr, *loss = nullptr;
  dio_buf* norm = nullptr;  // input standardisation: mean[12], scale[12]
  uint32_t adam_steps = 0;
  float class_weights[4] = {1.f, 1.f, 1.f, 1.f};
};

// A trainer's learned-densification state. mode 0: the plain rule; 1: the plain
// rule, logging (inputs, action) samples whose outcome held; 2: the MLP's action
// where confident, else the rule's. A budget caps the count after densifying.
struct LearnedState {
  uint32_t mode = 0;
  dio_densify_mlp* mlp = nullptr;  // not owned
  uint32_t budget = 0;
  float confidence = 0.5f;
  uint32_t samples_per_densify = 80000;
  dio_buf* extra = nullptr;        // per Gaussian: largest gradient, summed depth, summed gradient
  // mode 1: per Gaussian, since the last densification, the sums over its blended
  // pixels of blend weight x loss share and of blend weight (the footprint loss)
  dio_buf* footprint = nullptr;
  dio_buf *features = nullptr, *logits = nullptr;
  uint32_t features_count = 0;
  dio_buf *pending = nullptr, *pending_count = nullptr;
  uint32_t pending_capacity = 0;
  bool pending_valid = false;
  dio_buf *store_features = nullptr, *store_labels = nullptr, *store_count = nullptr;
  uint32_t store_capacity = 0;
};

struct dio_trainer {
  dio_ctx* ctx = nullptr;
  dio_training_settings settings{};
  uint32_t iteration = 0, sh_degree = 0, adam_steps = 0, count = 0;
  uint32_t width = 0, height = 0;
  dio_renderer* renderer = nullptr;
  dio_training_loss* loss = nullptr;
  dio_buf *params = nullptr, *moments = nullptr, *stats = nullptr;
  struct View {
    dio_camera camera;
    dio_buf* image;
  };
  std::vector<View> views;
  std::vector<uint32_t> queue;
  Mulberry32 random{1};
  std::vector<dio_densification> densifications;
  bool due_densify = false, due_reset = false;
  dio::AdamArgs last_adam{};
  dio_buf* vis = nullptr;  // each Gaussian's (depth, radius) from the step's preprocess
  uint32_t last_view = 0;
  BatchState* batch = nullptr;  // set for a multi-view batch trainer
  uint32_t generation = 0;      // bumped whenever densification replaces the buffers
  LearnedState learned;
  // The default launch settings: 512-pair backward segments and pairs binned by
  // counts for a single-view trainer (measured fastest at 1.1M Gaussians, 980x545);
  // a batch trainer replays whole lists.
  dio_tuning tuning{16, 64, 8, 0, 512, 2};
  StageClock clock;
  // The next step's frame, prepared at the end of a plain step by the fused
  // backward (its preprocess for the next view), and what it was prepared for. The
  // next step swaps these buffers in instead of running its preprocess while they
  // are still valid: same iteration, view and SH degree, the same tile rasterizer,
  // and no mutating call through the C ABI since (mutation_epoch).
  struct Prepared {
    dio_buf *splats = nullptr, *tile_counts = nullptr, *depth_keys = nullptr, *depth_values = nullptr,
            *vis = nullptr, *rects = nullptr;
    uint32_t count = 0;  // Gaussians the buffers are sized for
    bool ready = false;
    uint32_t iteration = 0, view = 0, sh_degree = 0, bin_shift = 0;
    uint64_t epoch = 0;
    const dio_tile_rasterizer* tiles = nullptr;
  } prepared;
  // A view drawn ahead for the next step (the draw order is the reference's either way).
  bool has_pending_view = false;
  uint32_t pending_view = 0;
};

namespace {

void batch_free_frames(BatchState* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (dio_buf** p : {&b->splats, &b->tile_counts, &b->depth_keys, &b->depth_values, &b->depth_alt_keys,
                      &b->depth_alt_values, &b->sorted_counts, &b->offsets, &b->scan_ws, &b->sort_ws,
                      &b->pair_total, &b->pair_tiles, &b->pair_splats, &b->alt_tiles, &b->alt_splats,
                      &b->pair_sort_ws, &b->splat_grads, &b->grad_gaussians, &b->grad_sh})
    release(*p);
  b->n = 0;
  b->n_cap = 0;
  b->capacity = 0;
}

void batch_destroy(BatchState* b) {
  if (!b) return;
  batch_free_frames(b);
  for (dio_buf** p : {&b->desc, &b->adam, &b->ranges, &b->pixel_color, &b->pixel_last, &b->partials,
                      &b->pixel_grads, &b->pixel_loss})
    release(*p);
  if (b->host) (void)hipHostFree(b->host);
  if (b->readback) (void)hipHostFree(b->readback);
  if (b->uploaded) (void)hipEventDestroy(b->uploaded);
  if (b->read) (void)hipEventDestroy(b->read);
  delete b;
}

void release_prepared(dio_trainer::Prepared& p) {
  release(p.splats); release(p.tile_counts); release(p.depth_keys); release(p.depth_values); release(p.vis);
  release(p.rects);
  p.count = 0;
  p.ready = false;
}

void destroy_trainer(dio_trainer* t) {
  if (!t) return;
  (void)hipStreamSynchronize(t->ctx->stream);
  t->clock.destroy();
  batch_destroy(t->batch);
  for (dio_buf** p : {&t->learned.extra, &t->learned.footprint, &t->learned.features, &t->learned.logits,
                      &t->learned.pending,
                      &t->learned.pending_count, &t->learned.store_features, &t->learned.store_labels,
                      &t->learned.store_count})
    release(*p);
  dio_renderer_destroy(t->renderer);
  destroy_loss(t->loss);
  release(t->params); release(t->moments); release(t->stats); release(t->vis);
  release_prepared(t->prepared);
  for (auto& v : t->views) release(v.image);
  dio_ctx* owner = t->ctx;
  delete t;
  dio::drop(owner);
}

// The next view of the shuffled queue (src/trainer.ts:188-199).
uint32_t draw_view(dio_trainer* t) {
  if (t->queue.empty()) {
    for (uint32_t i = 0; i < t->views.size(); i++) t->queue.push_back(i);
    for (size_t i = t->queue.size() - 1; i > 0; i--) {
      size_t j = static_cast<size_t>(std::floor(t->random.next() * static_cast<double>(i + 1)));
      std::swap(t->queue[i], t->queue[j]);
    }
  }
  const uint32_t v = t->queue.back();
  t->queue.pop_back();
  return v;
}

// This step's view: the one drawn ahead for it, if any, else the next of the queue.
const dio_trainer::View& next_view(dio_trainer* t) {
  t->last_view = t->has_pending_view ? t->pending_view : draw_view(t);
  t->has_pending_view = false;
  return t->views[t->last_view];
}

void batch_optimize(dio_trainer* t);

void trainer_optimize(dio_trainer* t) {
  if (t->batch) {
    batch_optimize(t);
    return;
  }
  const uint32_t step = ++t->iteration;
  const auto& st = t->settings;
  if (st.sh_degree_interval && step % st.sh_degree_interval == 0 && t->sh_degree < 3) {
    t->sh_degree++;
    if (t->renderer->scene) t->renderer->scene->degree = t->sh_degree;
  }
  const bool densifying = step < st.densify_until;
  const bool densify_now = densifying && step > st.densify_from && st.densify_interval &&
                           step % st.densify_interval == 0;
  const auto& view = next_view(t);
  dio_tile_rasterizer* tiles = t->renderer->scene->rasterizer;
  tiles->splat_batch = t->tuning.splat_batch;
  tiles->digit_bits = t->tuning.digit_bits;
  tiles->tile_size = t->tuning.tile_size;
  tiles->persistent = t->tuning.persistent != 0;
  tiles->per_tile_depth = t->tuning.per_tile_depth;
  tiles->segment_shift = t->tuning.segment ? static_cast<uint32_t>(__builtin_ctz(t->tuning.segment)) : 0u;
  if (t->clock.enabled) {
    t->clock.begin(t->ctx->stream);
    tiles->clock = &t->clock;
  }
  auto& prep = t->prepared;
  const bool use_prepared = prep.ready && prep.iteration == step && prep.view == t->last_view &&
                            prep.sh_degree == t->sh_degree && prep.tiles == tiles && prep.count == t->count &&
                            prep.bin_shift == tiles->bin_shift() &&
                            prep.epoch == dio::mutation_epoch().load(std::memory_order_relaxed);
  prep.ready = false;
  if (use_prepared) {
    // The previous step's fused backward already ran this preprocess: swap its
    // outputs in (the previous frame's buffers take their place for the next one).
    // The storage moves, not the handles, so a handle a caller took from
    // dio_renderer_splat_buffer keeps naming the renderer's current splats.
    dio_renderer::Scene* sc = t->renderer->scene;
    swap_storage(sc->splats, prep.splats);
    swap_storage(tiles->tile_counts, prep.tile_counts);
    swap_storage(tiles->depth_keys, prep.depth_keys);
    swap_storage(tiles->depth_values, prep.depth_values);
    swap_storage(tiles->rects, prep.rects);
    std::swap(t->vis, prep.vis);
    to_camera(view.camera, t->renderer->camera, tiles->bin_shift());
    t->renderer->preprocessed = true;
    if (tiles->clock) tiles->clock->mark(t->ctx->stream, kMarkPreprocess);
  } else {
    if (!t->vis || t->vis->bytes < static_cast<size_t>(t->count) * 8) {
      release(t->vis);
      t->vis = alloc(t->ctx, static_cast<size_t>(t->count) * 8);
    }
    renderer_preprocess(t->renderer, view.camera, t->vis);
  }
  renderer_tiles(t->renderer, nullptr, st.background);
  // The loss's partials kernel also clears the splat gradients the backward adds into.
  ensure_grad_buffers(t->renderer);
  dio::training_loss(t->ctx->stream, t->width, t->height, dev<float>(tiles->pixel_color),
                     dev<uint32_t>(view.image), dev<float>(t->loss->partials), dev<float>(t->loss->pixel_grads),
                     dev<float>(t->loss->pixel_loss), dev<uint32_t>(t->renderer->scene->grad_splats),
                     static_cast<size_t>(t->renderer->scene->count) * 9);
  if (tiles->clock) tiles->clock->mark(t->ctx->stream, kMarkLoss);
  if (!densify_now) t->adam_steps++;
  const double progress = std::min(1.0, static_cast<double>(step) / st.position_lr_steps);
  const double position_lr =
      std::exp(std::log(static_cast<double>(st.position_lr[0])) * (1 - progress) +
               std::log(static_cast<double>(st.position_lr[1])) * progress) * st.extent;
  dio::AdamArgs a;
  a.bias_correction1 = static_cast<float>(1 - std::pow(0.9, t->adam_steps));
  a.bias_correction2 = static_cast<float>(1 - std::pow(0.999, t->adam_steps));
  a.inv_bias_correction1 = static_cast<float>(1 / (1 - std::pow(0.9, t->adam_steps)));
  a.inv_bias_correction2 = static_cast<float>(1 / (1 - std::pow(0.999, t->adam_steps)));
  a.position_lr = static_cast<float>(position_lr);
  a.opacity_lr = st.opacity_lr;
  a.scale_lr = st.scale_lr;
  a.rotation_lr = st.rotation_lr;
  a.color_lr = st.color_lr;
  a.sh_lr = st.sh_lr;
  a.viewport[0] = view.camera.viewport[0];
  a.viewport[1] = view.camera.viewport[1];
  t->last_adam = a;
  auto* sc = t->renderer->scene;
  dio_buf* const foot = t->learned.mode == 1 ? t->learned.footprint : nullptr;
  dio_buf* const foot_loss = foot ? t->loss->pixel_loss : nullptr;
  if (densify_now) {
    // No Adam step on a densification step: the gradients feed only the statistics.
    renderer_backward(t->renderer, t->loss->pixel_grads, true, foot_loss, foot);
    dio::accumulate_stats(t->ctx->stream, t->count, dev<float>(sc->splats), dev<float>(sc->grad_splats), a,
                          dev<float>(t->stats), dev<float>(t->learned.extra));
  } else {
    // The rasterizer's backward, then preprocess backward + statistics + Adam in
    // one pass; the renderer's per-Gaussian gradient buffers are not written. Unless
    // this step's maintenance resets the opacities, the same pass also runs the next
    // step's preprocess (next view drawn now, SH degree as that step will have it).
    renderer_backward_splats(t->renderer, t->loss->pixel_grads, true, foot_loss, foot);
    const bool reset_next = densifying && st.opacity_reset_interval && step % st.opacity_reset_interval == 0;
    dio::PrepArgs next;
    const bool prepare = !reset_next;
    if (prepare) {
      if (!t->has_pending_view) {
        t->pending_view = draw_view(t);
        t->has_pending_view = true;
      }
      uint32_t next_degree = t->sh_degree;
      if (st.sh_degree_interval && (step + 1) % st.sh_degree_interval == 0 && next_degree < 3) next_degree++;
      if (prep.count != t->count || !prep.splats) {
        release_prepared(prep);
        const size_t n = t->count;
        prep.splats = alloc(t->ctx, n * dio::kFloatsPerSplat * 4);
        prep.tile_counts = alloc(t->ctx, n * 4);
        prep.depth_keys = alloc(t->ctx, n * 4);
        prep.depth_values = alloc(t->ctx, n * 4);
        prep.vis = alloc(t->ctx, n * 8);
        prep.rects = alloc(t->ctx, n * 8);
        prep.count = t->count;
      }
      to_camera(t->views[t->pending_view].camera, next.cam, tiles->bin_shift());
      next.sh_degree = next_degree;
      next.splats = dev<float>(prep.splats);
      next.tile_counts = dev<uint32_t>(prep.tile_counts);
      next.depth_keys = dev<uint32_t>(prep.depth_keys);
      next.depth_values = dev<uint32_t>(prep.depth_values);
      next.vis = dev<float>(prep.vis);
      next.rects = dev<uint32_t>(prep.rects);
      prep.iteration = step + 1;
      prep.view = t->pending_view;
      prep.sh_degree = next_degree;
      prep.tiles = tiles;
      prep.bin_shift = tiles->bin_shift();
    }
    dio::backward_adam(t->ctx->stream, t->count, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree,
                       t->renderer->camera, dev<float>(t->vis), dev<float>(sc->grad_splats),
                       dev<float>(t->params), dev<float>(t->moments), a, dev<float>(t->stats),
                       dev<float>(t->learned.extra), prepare ? &next : nullptr);
    if (prepare) {
      prep.epoch = dio::mutation_epoch().load(std::memory_order_relaxed);
      prep.ready = true;
    }
    if (tiles->clock) tiles->clock->mark(t->ctx->stream, kMarkBackwardAdam);
  }
  if (tiles->clock) {
    if (densify_now)
      t->clock.drop();
    else
      t->clock.finish(t->ctx->stream, tiles->pair_total->ptr);
    tiles->clock = nullptr;
  }
  tiles_after_submit(sc->rasterizer, false);
  t->due_densify = densify_now;
  t->due_reset = densifying && st.opacity_reset_interval && step % st.opacity_reset_interval == 0;
}

void trainer_reset_opacity(dio_trainer* t) {
  dio::reset_opacity(t->ctx->stream, t->count, dev<float>(t->params), dev<float>(t->moments),
                     dev<float>(t->renderer->scene->gaussians));
}

// The per-Gaussian buffers of the learned densification, sized for the count.
void learned_resize(dio_trainer* t) {
  LearnedState& L = t->learned;
  if (L.features_count == t->count && L.features) return;
  release(L.features);
  release(L.logits);
  L.features = alloc(t->ctx, static_cast<size_t>(t->count) * dio::kMlpInputs * 4, true);
  L.logits = alloc(t->ctx, static_cast<size_t>(t->count) * 4 * 4, true);
  L.features_count = t->count;
}

// The plain rule's actions, or the MLP's (mode 2), and the growth budget: fills
// `codes` and returns true when the densification must follow them rather than
// the rule (mode 0 without a budget returns false and leaves codes alone).
bool densify_codes(dio_trainer* t, const dio::DensifyArgs& a, dio_buf* codes, dio_buf* priority,
                   uint32_t* counters) {
  LearnedState& L = t->learned;
  const uint32_t n = t->count;
  hipStream_t s = t->ctx->stream;
  if (L.mode == 0 && L.budget == 0) return false;
  dio::densify_rule_codes(s, n, dev<float>(t->params), dev<float>(t->stats), a, dev<uint32_t>(codes),
                          dev<float>(priority));
  if (L.mode != 0) {
    learned_resize(t);
    dio::mlp_features(s, n, dev<float>(t->params), dev<float>(t->stats), dev<float>(L.extra), t->settings.extent,
                      t->iteration, t->batch ? t->batch->views : 1u, dev<float>(L.features));
    if (L.mode == 2) {
      dio::mlp_forward(s, n, dev<float>(L.features), dev<uint16_t>(L.mlp->weights), dev<float>(L.mlp->norm),
                       dev<float>(L.logits));
      dio::mlp_learned_codes(s, n, dev<float>(L.logits), L.confidence, dev<uint32_t>(codes), dev<float>(priority));
    }
    if (L.mode == 1 && L.pending_valid)  // the last logging step's samples, judged on these statistics
      dio::mlp_resolve(s, dev<uint32_t>(L.pending_count), L.pending_capacity, dev<float>(L.pending),
                       dev<float>(L.footprint), n, L.store_capacity, dev<float>(L.store_features),
                       dev<uint32_t>(L.store_labels), dev<uint32_t>(L.store_count));
  }
  if (L.budget) {
    dio_buf *scratch = nullptr, *keys = nullptr, *values = nullptr, *alt_keys = nullptr, *alt_values = nullptr,
            *ws = nullptr;
    try {
      scratch = alloc(t->ctx, static_cast<size_t>(n) * 4);
      dio::densify_count(s, n, dev<float>(t->params), dev<float>(t->stats), a, dev<uint32_t>(codes),
                         dev<uint32_t>(scratch), counters);
      uint32_t tally[4];
      DIO_HIP(hipMemcpyAsync(tally, counters, sizeof(tally), hipMemcpyDeviceToHost, s));
      sync(t->ctx);
      const uint64_t kept = static_cast<uint64_t>(n) - tally[0];
      const uint64_t grown = kept + tally[2] + tally[3];
      if (grown > L.budget) {
        const uint32_t allowed = static_cast<uint32_t>(L.budget > kept ? L.budget - kept : 0);
        keys = alloc(t->ctx, static_cast<size_t>(n) * 4);
        values = alloc(t->ctx, static_cast<size_t>(n) * 4);
        alt_keys = alloc(t->ctx, static_cast<size_t>(n) * 4);
        alt_values = alloc(t->ctx, static_cast<size_t>(n) * 4);
        ws = alloc(t->ctx, dio::sort_workspace_bytes(n), true);
        dio::densify_growth_keys(s, n, dev<uint32_t>(codes), dev<float>(priority), dev<uint32_t>(keys),
                                 dev<uint32_t>(values));
        if (dio::radix_sort(s, false, keys->ptr, dev<uint32_t>(values), alt_keys->ptr, dev<uint32_t>(alt_values),
                            nullptr, n, 0, 32, dio::sort_workspace(ws->ptr, n)))
          throw Error(DIO_ERR_INTERNAL, "budget sort ended in its scratch buffers");
        dio::densify_demote(s, n, dev<uint32_t>(keys), dev<uint32_t>(values), allowed, dev<float>(t->params), a,
                            dev<uint32_t>(codes));
      }
      sync(t->ctx);
    } catch (...) {
      release(scratch); release(keys); release(values); release(alt_keys); release(alt_values); release(ws);
      throw;
    }
    release(scratch); release(keys); release(values); release(alt_keys); release(alt_values); release(ws);
  }
  return true;
}

void trainer_densify(dio_trainer* t, bool prune_large) {
  const uint32_t n = t->count;
  const auto& st = t->settings;
  LearnedState& L = t->learned;
  dio::DensifyArgs a;
  a.grad_threshold = st.densify_grad_threshold;
  a.clone_max_size = st.percent_dense * st.extent;
  a.min_opacity = st.min_opacity;
  a.max_size = prune_large ? 0.1f * st.extent : 0.f;
  a.seed = t->iteration;
  hipStream_t s = t->ctx->stream;
  dio_buf* outputs = alloc(t->ctx, static_cast<size_t>(n) * 4);
  dio_buf *scratch = nullptr, *codes = nullptr, *priority = nullptr, *extra = nullptr, *foot = nullptr;
  dio_buf *params = nullptr, *sh = nullptr, *moments = nullptr, *stats = nullptr, *gaussians = nullptr;
  try {
    scratch = alloc(t->ctx, 64 + dio::scan_workspace_bytes(n), true);
    uint32_t* counters = dev<uint32_t>(scratch);  // [0..3] tally, [4] total
    codes = alloc(t->ctx, static_cast<size_t>(n) * 4);
    priority = alloc(t->ctx, static_cast<size_t>(n) * 4);
    const bool forced = densify_codes(t, a, codes, priority, counters);
    const uint32_t* forced_codes = forced ? dev<uint32_t>(codes) : nullptr;
    dio::densify_count(s, n, dev<float>(t->params), dev<float>(t->stats), a, forced_codes, dev<uint32_t>(outputs),
                       counters);
    dio::scan_exclusive(s, dev<uint32_t>(outputs), dev<uint32_t>(outputs), nullptr, n, counters + 4,
                        dio::scan_workspace(static_cast<char*>(scratch->ptr) + 64, n));
    if (L.mode == 1) {  // log this step's samples; their outcome is read at the next densification
      const double fraction = std::min(1.0, static_cast<double>(L.samples_per_densify) / std::max(n, 1u));
      const uint32_t threshold = static_cast<uint32_t>(std::min(4294967295.0, fraction * 4294967296.0));
      dio::fill_u32(s, dev<uint32_t>(L.pending_count), 0u, 1);
      dio::mlp_log_pending(s, n, dev<float>(L.features), dev<float>(L.footprint), forced_codes,
                           dev<uint32_t>(outputs), t->iteration, threshold, L.pending_capacity, dev<float>(L.pending),
                           dev<uint32_t>(L.pending_count));
      L.pending_valid = true;
    }
    uint32_t host[5];
    DIO_HIP(hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, s));
    sync(t->ctx);
    const uint32_t new_count = host[4];
    if (new_count == 0) invalid("Densification removed every Gaussian.");
    t->densifications.push_back({t->iteration, n, new_count, host[1], host[2], host[3], host[0]});
    params = alloc(t->ctx, static_cast<size_t>(new_count) * dio::kFloatsPerGaussian * 4, true);
    sh = alloc(t->ctx, static_cast<size_t>(new_count) * dio::kShRest * 4, true);
    moments = alloc(t->ctx, static_cast<size_t>(new_count) * dio::kMomentSlots * 8, true);
    stats = alloc(t->ctx, static_cast<size_t>(new_count) * 16, true);
    gaussians = alloc(t->ctx, static_cast<size_t>(new_count) * dio::kFloatsPerGaussian * 4, true);
    if (L.extra) extra = alloc(t->ctx, static_cast<size_t>(new_count) * 16, true);
    if (L.footprint) foot = alloc(t->ctx, static_cast<size_t>(new_count) * 8, true);
    dio::densify_scatter(s, n, dev<float>(t->params), dev<float>(t->stats), dev<uint32_t>(outputs),
                         dev<float>(t->renderer->scene->sh), dev<float>(t->moments), a, forced_codes,
                         dev<float>(params), dev<float>(sh), dev<float>(moments));
    dio::activate(s, new_count, dev<float>(params), dev<float>(gaussians));
  } catch (...) {
    release(outputs); release(scratch); release(codes); release(priority); release(params); release(sh);
    release(moments); release(stats); release(gaussians); release(extra); release(foot);
    throw;
  }
  sync(t->ctx);
  release(outputs);
  release(scratch);
  release(codes);
  release(priority);
  const uint32_t new_count = t->densifications.back().after;
  dio_renderer* r = t->renderer;
  const uint32_t capacity = r->scene->rasterizer->capacity;
  destroy_scene(r);
  create_scene(r, new_count, gaussians, sh, t->sh_degree, capacity);
  release(t->params); release(t->moments); release(t->stats);
  t->params = params;
  t->moments = moments;
  t->stats = stats;
  t->count = new_count;
  if (L.extra) {
    release(L.extra);
    L.extra = extra;
  }
  if (L.footprint) {
    release(L.footprint);
    L.footprint = foot;
  }
  t->generation++;
}

// ------------------------------------------------------ multi-view batches --

// The tile space and pixel buffers any `views` of the training views fit in:
// the sums of the largest tile and pixel counts.
void batch_limits(dio_trainer* t, BatchState* b) {
  std::vector<uint32_t> tiles, pixels;
  for (const auto& v : t->views) {
    const uint32_t w = viewport_dim(v.camera.viewport[0]), h = viewport_dim(v.camera.viewport[1]);
    tiles.push_back(dio::ceil_div(w, dio::kTile) * dio::ceil_div(h, dio::kTile));
    pixels.push_back(w * h);
  }
  std::sort(tiles.rbegin(), tiles.rend());
  std::sort(pixels.rbegin(), pixels.rend());
  uint64_t tile_sum = 0, pixel_sum = 0;
  for (uint32_t k = 0; k < b->views; k++) {
    tile_sum += tiles[k];
    pixel_sum += pixels[k];
  }
  if (pixel_sum >= (1ull << 31)) invalid("the batch's views hold too many pixels");
  b->max_tiles = static_cast<uint32_t>(tile_sum);
  b->max_pixels = static_cast<uint32_t>(pixel_sum);
}

void batch_alloc_frames(dio_trainer* t, uint32_t capacity) {
  BatchState* b = t->batch;
  dio_ctx* ctx = t->ctx;
  // Room for a quarter more Gaussians than now: densification grows the count a
  // little at a time, and each reallocation of these buffers costs milliseconds.
  if (static_cast<uint64_t>(t->count) * b->views >= (1ull << 30))
    invalid("a batch holds at most 2^30 splats (views x Gaussians)");
  const uint32_t n = static_cast<uint32_t>(
      std::min<uint64_t>(static_cast<uint64_t>(t->count) + t->count / 4, ((1ull << 30) - 1) / b->views));
  const uint64_t total = static_cast<uint64_t>(n) * b->views;
  b->splats = alloc(ctx, total * dio::kFloatsPerSplat * 4, true);
  b->tile_counts = alloc(ctx, total * 4, true);
  b->depth_keys = alloc(ctx, total * 4);
  b->depth_values = alloc(ctx, total * 4);
  b->depth_alt_keys = alloc(ctx, total * 4);
  b->depth_alt_values = alloc(ctx, total * 4);
  b->sorted_counts = alloc(ctx, total * 4);
  b->offsets = alloc(ctx, total * 4);
  b->scan_ws = alloc(ctx, dio::scan_workspace_bytes(static_cast<uint32_t>(total)), true);
  b->sort_ws = alloc(ctx, dio::sort_workspace_bytes(static_cast<uint32_t>(total)), true);
  b->pair_total = alloc(ctx, 16, true);
  b->pair_tiles = alloc(ctx, static_cast<size_t>(capacity) * 4);
  b->pair_splats = alloc(ctx, static_cast<size_t>(capacity) * 4, true);
  b->alt_tiles = alloc(ctx, static_cast<size_t>(capacity) * 4);
  b->alt_splats = alloc(ctx, static_cast<size_t>(capacity) * 4, true);
  b->pair_sort_ws = alloc(ctx, dio::sort_workspace_bytes(capacity), true);
  b->splat_grads = alloc(ctx, total * 9 * 4, true);
  b->grad_gaussians = alloc(ctx, static_cast<size_t>(n) * dio::kFloatsPerGaussian * 4, true);
  b->grad_sh = alloc(ctx, static_cast<size_t>(n) * dio::kShRest * 4, true);
  b->n = t->count;
  b->n_cap = n;
  b->capacity = capacity;
}

dio::BatchArgs batch_args(const BatchState* b) {
  dio::BatchArgs a;
  a.header = static_cast<const dio::BatchHeader*>(b->desc->ptr);
  a.views = reinterpret_cast<const dio::ViewDesc*>(static_cast<const char*>(b->desc->ptr) + kBatchViewsOffset);
  a.view_count = b->views;
  a.per_view = b->n;
  a.max_tiles = b->max_tiles;
  return a;
}

// One batch step's launches in stream order: the body of the captured graph.
void batch_launch(dio_trainer* t, bool with_adam) {
  BatchState* b = t->batch;
  hipStream_t s = t->ctx->stream;
  auto* sc = t->renderer->scene;
  const dio::BatchArgs a = batch_args(b);
  const uint32_t total = b->views * b->n;
  // The workspaces keep the layout of the buffers' allocation (n_cap Gaussians a
  // view): their counters sit after the per-tile states, at an offset that a layout
  // by the current count would move onto words another layout left non-zero.
  const uint32_t held = b->views * b->n_cap;
  const float* bg = t->settings.background;
  dio::preprocess_batch(s, dev<float>(sc->gaussians), dev<float>(sc->sh), t->sh_degree, a, dev<float>(b->splats),
                        dev<uint32_t>(b->tile_counts), dev<uint32_t>(b->depth_keys), dev<uint32_t>(b->depth_values));
  // One depth order over every view's splats: the stable tile sort below keeps
  // each tile's pairs front to back, and tiles of different views never mix.
  const dio_tuning& tune = t->tuning;
  if (dio::radix_sort(s, false, b->depth_keys->ptr, dev<uint32_t>(b->depth_values), b->depth_alt_keys->ptr,
                      dev<uint32_t>(b->depth_alt_values), nullptr, total, 0, 32,
                      dio::sort_workspace(b->sort_ws->ptr, held), tune.digit_bits, dev<uint32_t>(b->tile_counts),
                      dev<uint32_t>(b->sorted_counts)))
    throw Error(DIO_ERR_INTERNAL, "depth sort ended in its scratch buffers");
  dio::scan_exclusive(s, dev<uint32_t>(b->sorted_counts), dev<uint32_t>(b->offsets), nullptr, total,
                      dev<uint32_t>(b->pair_total), dio::scan_workspace(b->scan_ws->ptr, held));
  dio::duplicate_batch(s, dev<float>(b->splats), dev<uint32_t>(b->depth_values), dev<uint32_t>(b->offsets),
                       dev<uint32_t>(b->tile_counts), a, b->capacity, dev<uint32_t>(b->pair_total),
                       dev<uint32_t>(b->pair_tiles), dev<uint32_t>(b->pair_splats));
  b->sorted_in_alt = dio::radix_sort(s, false, b->pair_tiles->ptr, dev<uint32_t>(b->pair_splats), b->alt_tiles->ptr,
                                     dev<uint32_t>(b->alt_splats), dev<uint32_t>(b->pair_total), b->capacity, 0,
                                     tile_sort_bits(b->max_tiles, tune.digit_bits),
                                     dio::sort_workspace(b->pair_sort_ws->ptr, b->capacity), tune.digit_bits);
  dio_buf* sorted_tiles = b->sorted_in_alt ? b->alt_tiles : b->pair_tiles;
  dio_buf* sorted_splats = b->sorted_in_alt ? b->alt_splats : b->pair_splats;
  dio::find_ranges(s, dev<uint32_t>(sorted_tiles), dev<uint32_t>(b->pair_total), b->capacity, b->max_tiles,
                   dev<uint32_t>(b->ranges), nullptr, 0, nullptr, true);
  dio::rasterize_batch(s, dev<float>(b->splats), dev<uint32_t>(sorted_splats), dev<uint32_t>(b->ranges), a, bg,
                       dev<float>(b->pixel_color), dev<uint32_t>(b->pixel_last), tune.splat_batch);
  dio::training_loss_batch(s, a, dev<float>(b->pixel_color), dev<float>(b->partials), dev<float>(b->pixel_grads),
                           dev<float>(b->pixel_loss));
  dio::fill_u32(s, dev<uint32_t>(b->splat_grads), 0u, static_cast<size_t>(total) * 9);  // the backward adds into it
  dio::rasterize_backward_batch(s, dev<float>(b->splats), dev<uint32_t>(sorted_splats), dev<uint32_t>(b->ranges), a,
                                bg, dev<float>(b->pixel_color), dev<uint32_t>(b->pixel_last),
                                dev<float>(b->pixel_grads), dev<float>(b->splat_grads), tune.splat_batch);
  dio::preprocess_backward_batch(s, dev<float>(sc->gaussians), dev<float>(sc->sh), t->sh_degree, a,
                                 dev<float>(b->splats), dev<float>(b->splat_grads), dev<float>(b->grad_gaussians),
                                 dev<float>(b->grad_sh), dev<float>(t->stats), dev<float>(t->learned.extra));
  if (with_adam)
    dio::adam_step_device_args(s, b->n, dev<float>(t->params), dev<float>(sc->sh), dev<float>(t->moments),
                               dev<float>(b->grad_gaussians), dev<float>(b->grad_sh),
                               static_cast<const dio::AdamArgs*>(b->adam->ptr), dev<float>(sc->gaussians));
}

void batch_capture(dio_trainer* t) {
  BatchState* b = t->batch;
  hipStream_t s = t->ctx->stream;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  DIO_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    batch_launch(t, true);
  } catch (...) {
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(s, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  DIO_HIP(hipStreamEndCapture(s, &b->graph));
  DIO_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_degree = t->sh_degree;
  b->graph_generation = t->generation;
}

// One optimize() step of a batch trainer: the plain schedule (src/trainer.ts:192-244)
// with `views` training views per step and their gradients summed.
void batch_optimize(dio_trainer* t) {
  BatchState* b = t->batch;
  hipStream_t s = t->ctx->stream;
  const uint32_t step = ++t->iteration;
  const auto& st = t->settings;
  if (st.sh_degree_interval && step % st.sh_degree_interval == 0 && t->sh_degree < 3) {
    t->sh_degree++;
    if (t->renderer->scene) t->renderer->scene->degree = t->sh_degree;
  }
  const bool densifying = step < st.densify_until;
  const bool densify_now = densifying && step > st.densify_from && st.densify_interval &&
                           step % st.densify_interval == 0;
  // The last step's pair count: grow the pair buffers after an overflow.
  uint32_t capacity = b->capacity;
  if (b->read_pending) {
    DIO_HIP(hipEventSynchronize(b->read));
    b->read_pending = false;
    b->last_pairs = b->readback[0];
    if (b->last_pairs > b->capacity) capacity = next_pow2_capacity(b->last_pairs);
  }
  if (t->count > b->n_cap || capacity != b->capacity) {
    sync(t->ctx);
    batch_free_frames(b);
    batch_alloc_frames(t, capacity);
  }
  b->n = t->count;  // the frame layout's stride; a changed count also changes the generation (recapture)
  if (b->upload_pending) DIO_HIP(hipEventSynchronize(b->uploaded));  // the staging memory is free again
  auto* header = reinterpret_cast<dio::BatchHeader*>(b->host);
  auto* views = reinterpret_cast<dio::ViewDesc*>(b->host + kBatchViewsOffset);
  uint32_t tile_base = 0, pixel_base = 0;
  b->step_views.clear();
  for (uint32_t k = 0; k < b->views; k++) {
    const auto& view = next_view(t);
    b->step_views.push_back(t->last_view);
    dio::ViewDesc& d = views[k];
    d = dio::ViewDesc{};
    std::memcpy(d.cam.view, view.camera.view, sizeof(d.cam.view));
    std::memcpy(d.cam.proj, view.camera.proj, sizeof(d.cam.proj));
    d.cam.viewport[0] = view.camera.viewport[0];
    d.cam.viewport[1] = view.camera.viewport[1];
    d.width = viewport_dim(view.camera.viewport[0]);
    d.height = viewport_dim(view.camera.viewport[1]);
    d.tiles_x = dio::ceil_div(d.width, dio::kTile);
    d.tiles_y = dio::ceil_div(d.height, dio::kTile);
    d.tile_base = tile_base;
    d.pixel_base = pixel_base;
    d.loss_scale = static_cast<float>(1.0 / (3.0 * static_cast<double>(d.width) * d.height));
    d.target = dev<uint32_t>(view.image);
    tile_base += d.tiles_x * d.tiles_y;
    pixel_base += d.width * d.height;
  }
  header->views = b->views;
  header->tiles = tile_base;
  header->pixels = pixel_base;
  header->pad = 0;
  if (!densify_now) t->adam_steps++;
  const double progress = std::min(1.0, static_cast<double>(step) / st.position_lr_steps);
  const double position_lr =
      std::exp(std::log(static_cast<double>(st.position_lr[0])) * (1 - progress) +
               std::log(static_cast<double>(st.position_lr[1])) * progress) * st.extent;
  dio::AdamArgs a;
  a.bias_correction1 = static_cast<float>(1 - std::pow(0.9, t->adam_steps));
  a.bias_correction2 = static_cast<float>(1 - std::pow(0.999, t->adam_steps));
  a.inv_bias_correction1 = static_cast<float>(1 / (1 - std::pow(0.9, t->adam_steps)));
  a.inv_bias_correction2 = static_cast<float>(1 / (1 - std::pow(0.999, t->adam_steps)));
  a.position_lr = static_cast<float>(position_lr);
  a.opacity_lr = st.opacity_lr;
  a.scale_lr = st.scale_lr;
  a.rotation_lr = st.rotation_lr;
  a.color_lr = st.color_lr;
  a.sh_lr = st.sh_lr;
  a.viewport[0] = a.viewport[1] = 0.f;  // statistics use each view's own viewport
  std::memcpy(b->host + kBatchDescBytes, &a, sizeof(a));
  t->last_adam = a;
  DIO_HIP(hipMemcpyAsync(b->desc->ptr, b->host, kBatchDescBytes, hipMemcpyHostToDevice, s));
  DIO_HIP(hipMemcpyAsync(b->adam->ptr, b->host + kBatchDescBytes, sizeof(a), hipMemcpyHostToDevice, s));
  DIO_HIP(hipEventRecord(b->uploaded, s));
  b->upload_pending = true;
  if (densify_now) {
    batch_launch(t, false);  // densification steps skip Adam (src/trainer.ts:236-239)
  } else {
    if (!b->exec || b->graph_degree != t->sh_degree || b->graph_generation != t->generation) batch_capture(t);
    DIO_HIP(hipGraphLaunch(b->exec, s));
  }
  DIO_HIP(hipMemcpyAsync(b->readback, b->pair_total->ptr, 4, hipMemcpyDeviceToHost, s));
  DIO_HIP(hipEventRecord(b->read, s));
  b->read_pending = true;
  b->step_pixels = pixel_base;
  t->due_densify = densify_now;
  t->due_reset = densifying && st.opacity_reset_interval && step % st.opacity_reset_interval == 0;
}

void trainer_maintain(dio_trainer* t) {
  const bool densify = t->due_densify, reset = t->due_reset;
  t->due_densify = t->due_reset = false;
  if (densify || reset) t->prepared.ready = false;  // the Gaussians change under the prepared frame
  if (densify) trainer_densify(t, t->iteration > t->settings.opacity_reset_interval);
  if (reset) trainer_reset_opacity(t);
}

}  // namespace

extern "C" void dio_default_settings(float extent, dio_training_settings* out) {
  if (!out) return;
  dio_training_settings s{};
  s.extent = extent;
  s.position_lr_steps = 30000;
  s.position_lr[0] = 0.00016f;
  s.position_lr[1] = 0.0000016f;
  s.opacity_lr = 0.025f;
  s.scale_lr = 0.005f;
  s.rotation_lr = 0.001f;
  s.color_lr = 0.0025f;
  s.sh_lr = static_cast<float>(0.0025 / 20.0);  // formed in double, as src/trainer.ts:72
  s.sh_degree_interval = 1000;
  s.densify_from = 500;
  s.densify_until = 15000;
  s.densify_interval = 100;
  s.densify_grad_threshold = 0.0002f;
  s.percent_dense = 0.01f;
  s.min_opacity = 0.005f;
  s.opacity_reset_interval = 3000;
  s.background[0] = s.background[1] = s.background[2] = 0.f;
  *out = s;
}

namespace {

// The trainer both creation calls share. `batch` 0: the plain trainer, whose views
// must share one size (src/trainer.ts:133-135); otherwise a batch trainer taking
// `batch` views of any sizes per step.
void trainer_create(dio_ctx* ctx, const float* params, const float* sh, uint32_t count, const dio_camera* cameras,
                    const uint32_t* const* images, uint32_t view_count, uint32_t batch,
                    const dio_training_settings* settings, dio_trainer** out) {
  {
    need(ctx, "ctx");
    need(params, "params");
    need(sh, "sh");
    need(cameras, "cameras");
    need(images, "images");
    need(settings, "settings");
    need(out, "out");
    if (view_count == 0) invalid("Training needs at least one view.");
    if (count == 0) invalid("Training needs at least one Gaussian.");
    for (uint32_t v = 0; v < view_count; v++)
      check_tile_limit(viewport_dim(cameras[v].viewport[0]), viewport_dim(cameras[v].viewport[1]));
    const float w = cameras[0].viewport[0], h = cameras[0].viewport[1];
    if (batch == 0) {
      for (uint32_t v = 0; v < view_count; v++)
        if (cameras[v].viewport[0] != w || cameras[v].viewport[1] != h)
          invalid("All training views must be the same size.");
    } else if (batch > view_count || batch > dio::kMaxBatchViews) {
      invalid("a batch takes between 1 and min(views, " + std::to_string(dio::kMaxBatchViews) + ") views");
    }
    auto* t = new dio_trainer;
    try {
      t->ctx = ctx;
      dio::hold(t->ctx);
      t->settings = *settings;
      t->count = count;
      t->width = static_cast<uint32_t>(w);
      t->height = static_cast<uint32_t>(h);
      for (uint32_t v = 0; v < view_count; v++) {
        need(images[v], "image");
        const size_t px = static_cast<size_t>(viewport_dim(cameras[v].viewport[0])) * viewport_dim(cameras[v].viewport[1]);
        dio_buf* img = alloc(ctx, px * 4);
        t->views.push_back({cameras[v], img});
        DIO_HIP(hipMemcpyAsync(img->ptr, images[v], px * 4, hipMemcpyHostToDevice, ctx->stream));
      }
      t->params = alloc(ctx, static_cast<size_t>(count) * dio::kFloatsPerGaussian * 4);
      DIO_HIP(hipMemcpyAsync(t->params->ptr, params, t->params->bytes, hipMemcpyHostToDevice, ctx->stream));
      t->moments = alloc(ctx, static_cast<size_t>(count) * dio::kMomentSlots * 8, true);
      t->stats = alloc(ctx, static_cast<size_t>(count) * 16, true);
      t->renderer = new dio_renderer;
      t->renderer->ctx = ctx;
      dio::hold(t->renderer->ctx);
      // The renderer gets the activated copy; from here on Adam keeps it in step.
      dio_buf* g = alloc(ctx, static_cast<size_t>(count) * dio::kFloatsPerGaussian * 4, true);
      dio_buf* s = nullptr;
      try {
        s = alloc(ctx, static_cast<size_t>(count) * dio::kShRest * 4);
        DIO_HIP(hipMemcpyAsync(s->ptr, sh, s->bytes, hipMemcpyHostToDevice, ctx->stream));
        dio::activate(ctx->stream, count, dev<float>(t->params), dev<float>(g));
      } catch (...) {
        release(g);
        release(s);
        throw;
      }
      create_scene(t->renderer, count, g, s, 0, 0);
      dio_training_loss* l = nullptr;
      if (dio_training_loss_create(ctx, t->width, t->height, &l) != DIO_OK) throw Error(DIO_ERR_MEMORY, dio::last_error());
      t->loss = l;
      if (batch) {
        t->tuning.segment = 0;
        t->tuning.per_tile_depth = 0;
        t->batch = new BatchState;
        BatchState* b = t->batch;
        b->views = batch;
        batch_limits(t, b);
        b->desc = alloc(ctx, kBatchDescBytes, true);
        b->adam = alloc(ctx, sizeof(dio::AdamArgs), true);
        b->ranges = alloc(ctx, static_cast<size_t>(b->max_tiles) * 8, true);
        b->pixel_color = alloc(ctx, static_cast<size_t>(b->max_pixels) * 16, true);
        b->pixel_last = alloc(ctx, static_cast<size_t>(b->max_pixels) * 4, true);
        b->partials = alloc(ctx, static_cast<size_t>(b->max_pixels) * 48);
        b->pixel_grads = alloc(ctx, static_cast<size_t>(b->max_pixels) * 16, true);
        b->pixel_loss = alloc(ctx, static_cast<size_t>(b->max_pixels) * 4, true);
        DIO_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->host), kBatchHostBytes, hipHostMallocDefault));
        std::memset(b->host, 0, kBatchHostBytes);
        DIO_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->readback), 16, hipHostMallocDefault));
        DIO_HIP(hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming));
        DIO_HIP(hipEventCreateWithFlags(&b->read, hipEventDisableTiming));
        // Room for two pairs per splat to start with; an overflowing step drops
        // its farthest pairs and the buffers grow for the next one.
        const uint64_t guess = std::max<uint64_t>(1u << 20, 2ull * count * batch);
        batch_alloc_frames(t, next_pow2_capacity(static_cast<uint32_t>(std::min<uint64_t>(guess, 1u << 30))));
      }
      sync(ctx);
    } catch (...) {
      destroy_trainer(t);
      throw;
    }
    *out = t;
  }
}

}  // namespace

extern "C" int dio_trainer_create(dio_ctx* ctx, const float* params, const float* sh, uint32_t count,
                                  const dio_camera* cameras, const uint32_t* const* images, uint32_t view_count,
                                  const dio_training_settings* settings, dio_trainer** out) {
  return guard([&] { trainer_create(ctx, params, sh, count, cameras, images, view_count, 0, settings, out); });
}

extern "C" int dio_trainer_create_batched(dio_ctx* ctx, const float* params, const float* sh, uint32_t count,
                                          const dio_camera* cameras, const uint32_t* const* images,
                                          uint32_t view_count, uint32_t batch, const dio_training_settings* settings,
                                          dio_trainer** out) {
  return guard([&] {
    if (batch == 0) invalid("a batch takes at least one view");
    trainer_create(ctx, params, sh, count, cameras, images, view_count, batch, settings, out);
  });
}

extern "C" uint32_t dio_trainer_batch_size(const dio_trainer* t) { return t && t->batch ? t->batch->views : 0; }

extern "C" int dio_trainer_step(dio_trainer* t) {
  return guard_read([&] {
    need(t, "trainer");
    trainer_optimize(t);
    trainer_maintain(t);
  });
}

extern "C" int dio_trainer_optimize(dio_trainer* t) {
  return guard_read([&] {
    need(t, "trainer");
    trainer_optimize(t);
  });
}

extern "C" int dio_trainer_maintain(dio_trainer* t) {
  return guard_read([&] {
    need(t, "trainer");
    trainer_maintain(t);
  });
}

extern "C" int dio_trainer_densify(dio_trainer* t, int prune_large) {
  return guard([&] {
    need(t, "trainer");
    trainer_densify(t, prune_large != 0);
  });
}

extern "C" int dio_trainer_reset_opacity(dio_trainer* t) {
  return guard([&] {
    need(t, "trainer");
    trainer_reset_opacity(t);
  });
}

extern "C" int dio_trainer_export_raw(dio_trainer* t, float* params, float* sh) {
  return guard_read([&] {
    need(t, "trainer");
    need(params, "params");
    need(sh, "sh");
    DIO_HIP(hipMemcpyAsync(params, t->params->ptr, static_cast<size_t>(t->count) * dio::kFloatsPerGaussian * 4,
                           hipMemcpyDeviceToHost, t->ctx->stream));
    DIO_HIP(hipMemcpyAsync(sh, t->renderer->scene->sh->ptr, static_cast<size_t>(t->count) * dio::kShRest * 4,
                           hipMemcpyDeviceToHost, t->ctx->stream));
    sync(t->ctx);
  });
}

extern "C" int dio_trainer_last_loss(dio_trainer* t, double* out) {
  return guard_read([&] {
    need(t, "trainer");
    need(out, "out");
    if (!t->batch) {
      *out = loss_read(t->loss);
      return;
    }
    // The mean of the batch's per-view losses, each read() of its own view.
    BatchState* b = t->batch;
    std::vector<float> shares(b->step_pixels);
    if (!shares.empty())
      DIO_HIP(hipMemcpyAsync(shares.data(), b->pixel_loss->ptr, shares.size() * 4, hipMemcpyDeviceToHost,
                             t->ctx->stream));
    sync(t->ctx);
    double sum = 0;
    for (float v : shares) sum += v;
    *out = sum / b->views + 0.2;
  });
}

namespace {

// Renders each view and scores it on the device (metrics.hip); one download of
// every tile's sums at the end. Photos come from the host (uploaded through one
// staging buffer) or are already device buffers.
void trainer_score(dio_trainer* t, const dio_camera* cameras, const uint32_t* const* host_images,
                   const dio_buf* const* device_images, uint32_t view_count, bool with_ssim, double* psnr,
                   double* ssim) {
  need(cameras, "cameras");
  if (!host_images && !device_images) invalid("images is null");
  if (view_count == 0) invalid("no views to evaluate");
  std::vector<uint32_t> first(view_count + 1, 0);
  size_t largest = 0;
  for (uint32_t v = 0; v < view_count; v++) {
    const uint32_t w = viewport_dim(cameras[v].viewport[0]), h = viewport_dim(cameras[v].viewport[1]);
    if (w == 0 || h == 0) invalid("a view has an empty viewport");
    check_tile_limit(w, h);
    first[v + 1] = first[v] + dio::evaluate_tiles(w, h);
    largest = std::max<size_t>(largest, static_cast<size_t>(w) * h);
    if (device_images) {
      need(device_images[v], "image");
      if (device_images[v]->bytes < static_cast<size_t>(w) * h * 4) invalid("an image is smaller than its viewport");
    } else {
      need(host_images[v], "image");
    }
  }
  const uint32_t tiles = first[view_count];
  dio_buf* staging = nullptr;
  dio_buf* sse = nullptr;
  dio_buf* structure = nullptr;
  std::vector<uint32_t> squared(tiles);
  std::vector<double> sums(with_ssim ? tiles : 0);
  try {
    if (!device_images) staging = alloc(t->ctx, largest * 4);
    sse = alloc(t->ctx, static_cast<size_t>(tiles) * 4);
    structure = alloc(t->ctx, with_ssim ? static_cast<size_t>(tiles) * 8 : 0);
    for (uint32_t v = 0; v < view_count; v++) {
      const uint32_t w = viewport_dim(cameras[v].viewport[0]),
                     h = viewport_dim(cameras[v].viewport[1]);
      const uint32_t* photo = nullptr;
      if (device_images) {
        photo = static_cast<const uint32_t*>(device_images[v]->ptr);
      } else {
        DIO_HIP(hipMemcpyAsync(staging->ptr, host_images[v], static_cast<size_t>(w) * h * 4, hipMemcpyHostToDevice,
                               t->ctx->stream));
        photo = static_cast<const uint32_t*>(staging->ptr);
      }
      renderer_preprocess(t->renderer, cameras[v]);
      renderer_tiles(t->renderer, nullptr, t->settings.background);
      dio::evaluate_image(t->ctx->stream, static_cast<const float4*>(t->renderer->scene->rasterizer->pixel_color->ptr),
                          photo, w, h, with_ssim, dev<uint32_t>(sse) + first[v],
                          with_ssim ? dev<double>(structure) + first[v] : nullptr);
    }
    DIO_HIP(hipMemcpyAsync(squared.data(), sse->ptr, squared.size() * 4, hipMemcpyDeviceToHost, t->ctx->stream));
    if (with_ssim)
      DIO_HIP(hipMemcpyAsync(sums.data(), structure->ptr, sums.size() * 8, hipMemcpyDeviceToHost, t->ctx->stream));
    sync(t->ctx);
  } catch (...) {
    release(staging);
    release(sse);
    release(structure);
    throw;
  }
  release(staging);
  release(sse);
  release(structure);
  // src/metrics.ts:5-15 per view: 10 log10(1 / (SSE / 3 pixels)) with SSE over
  // 0-1 values, here the exact 8-bit sum / 255^2; the views' mean.
  double psnr_total = 0, ssim_total = 0;
  for (uint32_t v = 0; v < view_count; v++) {
    const double pixels = cameras[v].viewport[0] * static_cast<double>(cameras[v].viewport[1]);
    uint64_t se = 0;
    double st = 0;
    for (uint32_t k = first[v]; k < first[v + 1]; k++) {
      se += squared[k];
      if (with_ssim) st += sums[k];
    }
    psnr_total += 10.0 * std::log10(3.0 * pixels * 65025.0 / static_cast<double>(se));
    ssim_total += st / (3.0 * pixels);
  }
  if (psnr) *psnr = psnr_total / view_count;
  if (ssim) *ssim = with_ssim ? ssim_total / view_count : NAN;
}

}  // namespace

extern "C" int dio_trainer_evaluate(dio_trainer* t, const dio_camera* cameras, const uint32_t* const* images,
                                    uint32_t view_count, int with_ssim, double* psnr, double* ssim) {
  return guard([&] {
    need(t, "trainer");
    need(images, "images");
    trainer_score(t, cameras, images, nullptr, view_count, with_ssim != 0, psnr, ssim);
  });
}

extern "C" int dio_trainer_measure_psnr(dio_trainer* t, const dio_camera* cameras, const dio_buf* const* images,
                                        uint32_t view_count, double* psnr) {
  return guard([&] {
    need(t, "trainer");
    need(images, "images");
    need(psnr, "psnr");
    trainer_score(t, cameras, nullptr, images, view_count, false, psnr, nullptr);
  });
}

extern "C" uint32_t dio_trainer_count(const dio_trainer* t) { return t ? t->count : 0; }
extern "C" uint32_t dio_trainer_iteration(const dio_trainer* t) { return t ? t->iteration : 0; }
extern "C" void dio_trainer_set_iteration(dio_trainer* t, uint32_t iteration) {
  if (t) t->iteration = iteration;
}
extern "C" uint32_t dio_trainer_sh_degree(const dio_trainer* t) { return t ? t->sh_degree : 0; }
extern "C" uint32_t dio_trainer_last_view(const dio_trainer* t) { return t ? t->last_view : 0; }
extern "C" int dio_trainer_set_sh_degree(dio_trainer* t, uint32_t degree) {
  return guard([&] {
    need(t, "trainer");
    if (degree > 3) invalid("the SH degree is at most 3");
    t->sh_degree = degree;
    if (t->renderer->scene) t->renderer->scene->degree = degree;
  });
}
extern "C" uint32_t dio_trainer_densification_count(const dio_trainer* t) {
  return t ? static_cast<uint32_t>(t->densifications.size()) : 0;
}
extern "C" int dio_trainer_densification(const dio_trainer* t, uint32_t index, dio_densification* out) {
  return guard_read([&] {
    need(t, "trainer");
    need(out, "out");
    if (index >= t->densifications.size()) invalid("no such densification");
    *out = t->densifications[index];
  });
}
extern "C" dio_buf* dio_trainer_params(dio_trainer* t) { return t ? t->params : nullptr; }
extern "C" dio_buf* dio_trainer_moments(dio_trainer* t) { return t ? t->moments : nullptr; }
extern "C" dio_buf* dio_trainer_stats(dio_trainer* t) { return t ? t->stats : nullptr; }
extern "C" dio_renderer* dio_trainer_renderer(dio_trainer* t) { return t ? t->renderer : nullptr; }
extern "C" void dio_trainer_destroy(dio_trainer* t) { destroy_trainer(t); }

// =============================================================== measurement

extern "C" int dio_trainer_time_stage(dio_trainer* t, const char* stage, uint32_t repeats, double* ms) {
  return guard([&] {
    need(t, "trainer");
    need(stage, "stage");
    need(ms, "ms");
    if (repeats == 0) invalid("repeats must be positive");
    const std::string name(stage);
    if (t->batch) {
      // A batch trainer times its whole captured step ("batch_step").
      BatchState* b = t->batch;
      if (name != "batch_step") invalid("a batch trainer times only the stage batch_step");
      if (!b->exec) invalid("run an optimize() step before timing a stage");
      hipStream_t s = t->ctx->stream;
      hipEvent_t e0, e1;
      DIO_HIP(hipEventCreate(&e0));
      DIO_HIP(hipEventCreate(&e1));
      DIO_HIP(hipGraphLaunch(b->exec, s));  // warm
      DIO_HIP(hipEventRecord(e0, s));
      for (uint32_t k = 0; k < repeats; k++) DIO_HIP(hipGraphLaunch(b->exec, s));
      DIO_HIP(hipEventRecord(e1, s));
      DIO_HIP(hipEventSynchronize(e1));
      float elapsed = 0.f;
      DIO_HIP(hipEventElapsedTime(&elapsed, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      *ms = elapsed / repeats;
      return;
    }
    auto* r = t->renderer;
    auto* sc = r->scene;
    auto* tr = sc->rasterizer;
    if (!tr->framed || !sc->grad_splats) invalid("run an optimize() step before timing a stage");
    hipStream_t s = t->ctx->stream;
    const auto& view = t->views[t->last_view];
    dio_buf* sorted_splats = tr->sorted_in_alt ? tr->alt_splats : tr->pair_splats;
    dio::Segments seg;
    const dio::Segments* framed_segments =
        tr->framed_shift && tr->framed_shift == tr->segment_shift
            ? segments_for(tr, dio::ceil_div(tr->frame_w, dio::kTile) * dio::ceil_div(tr->frame_h, dio::kTile), seg,
                           false)
            : nullptr;
    std::function<void()> run;
    if (name == "preprocess") {
      run = [&] {
        dio::preprocess(s, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree, sc->count, r->camera,
                        dev<float>(sc->splats), dev<uint32_t>(tr->tile_counts), dev<uint32_t>(tr->depth_keys),
                        dev<uint32_t>(tr->depth_values), nullptr, dev<uint32_t>(tr->rects));
      };
    } else if (name == "scan") {
      run = [&] {
        dio::scan_exclusive(s, dev<uint32_t>(tr->per_tile_depth ? tr->tile_counts : tr->sorted_counts),
                            dev<uint32_t>(tr->offsets), nullptr, tr->count,
                            dev<uint32_t>(tr->pair_total), dio::scan_workspace(tr->splat_scan_ws->ptr, tr->count));
      };
    } else if (name == "sort") {
      // The whole binning sort (Gaussians by depth, then pairs by tile; or pairs by
      // tile, then each tile's list by depth) on the data a step really sorts: the
      // chain preprocess -> depth sort -> scan -> duplicate -> tile sort (-> ranges ->
      // list sort), less the same chain without the sorts (timed apart below).
      const uint32_t shift = tr->bin_shift();
      const uint32_t tiles = dio::ceil_div(dio::ceil_div(tr->frame_w, dio::kTile), 1u << shift) *
                             dio::ceil_div(dio::ceil_div(tr->frame_h, dio::kTile), 1u << shift);
      const bool per_tile = tr->per_tile_depth != 0;
      const bool binned = bins_by_counts(tr, tiles);
      const bool in_alt = (tile_sort_bits(tiles, tr->digit_bits) / tr->digit_bits) % 2 == 1;
      auto chain = [&](bool sorts) {
        dio::preprocess(s, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree, sc->count, r->camera,
                        dev<float>(sc->splats), dev<uint32_t>(tr->tile_counts), dev<uint32_t>(tr->depth_keys),
                        dev<uint32_t>(tr->depth_values), nullptr, dev<uint32_t>(tr->rects));
        if (binned) {
          // Everything from the counts to the ordered lists is the sort here.
          if (!sorts) return;
          const dio::BinWorkspace ws = bin_workspace(tr, tiles);
          dio::bin_count(s, dev<uint32_t>(tr->rects), dev<uint32_t>(tr->depth_keys), dev<uint32_t>(tr->tile_counts),
                         tr->count, tr->frame_w, tr->frame_h, shift, tr->capacity, dev<uint32_t>(tr->pair_total), ws);
          uint32_t* keyed = keyed_pairs(tr);
          dio::bin_scatter(s, dev<uint32_t>(tr->rects), dev<uint32_t>(tr->tile_counts), dev<uint32_t>(tr->depth_keys),
                           tr->count, tr->frame_w, tr->frame_h, shift, ws, keyed, dev<uint32_t>(tr->ranges),
                           dev<uint32_t>(tr->tile_order), nullptr);
          dio::depth_order(s, dev<uint32_t>(tr->depth_keys), dev<uint32_t>(tr->ranges), dev<uint32_t>(tr->tile_order),
                           tiles, dev<uint32_t>(tr->pair_splats), keyed + 2 * static_cast<size_t>(tr->capacity), false,
                           dev<uint32_t>(tr->order_queue), tr->order_parity++, keyed);
          return;
        }
        if (sorts && !per_tile)
          dio::radix_sort(s, false, tr->depth_keys->ptr, dev<uint32_t>(tr->depth_values), tr->depth_alt_keys->ptr,
                          dev<uint32_t>(tr->depth_alt_values), nullptr, tr->count, 0, 32,
                          dio::sort_workspace(tr->splat_sort_ws->ptr, tr->count), tr->digit_bits,
                          dev<uint32_t>(tr->tile_counts), dev<uint32_t>(tr->sorted_counts));
        dio::scan_exclusive(s, dev<uint32_t>(per_tile ? tr->tile_counts : tr->sorted_counts), dev<uint32_t>(tr->offsets),
                            nullptr, tr->count, dev<uint32_t>(tr->pair_total),
                            dio::scan_workspace(tr->splat_scan_ws->ptr, tr->count));
        if (per_tile)
          dio::pair_cut(s, dev<uint32_t>(tr->depth_keys), dev<uint32_t>(tr->tile_counts), dev<uint32_t>(tr->offsets),
                        tr->count, dev<uint32_t>(tr->pair_total), tr->capacity);
        dio::duplicate(s, dev<float>(sc->splats), per_tile ? nullptr : dev<uint32_t>(tr->depth_values),
                       dev<uint32_t>(tr->offsets), dev<uint32_t>(tr->tile_counts), tr->count, tr->frame_w, tr->frame_h,
                       tr->capacity, dev<uint32_t>(tr->pair_total), dev<uint32_t>(tr->pair_tiles),
                       dev<uint32_t>(tr->pair_splats), 0, shift);
        if (sorts)
          dio::radix_sort(s, false, tr->pair_tiles->ptr, dev<uint32_t>(tr->pair_splats), tr->alt_tiles->ptr,
                          dev<uint32_t>(tr->alt_splats), dev<uint32_t>(tr->pair_total), tr->capacity, 0,
                          tile_sort_bits(tiles, tr->digit_bits), dio::sort_workspace(tr->pair_sort_ws->ptr, tr->capacity),
                          tr->digit_bits);
        if (sorts && per_tile) {
          dio::find_ranges(s, dev<uint32_t>(in_alt ? tr->alt_tiles : tr->pair_tiles), dev<uint32_t>(tr->pair_total),
                           tr->capacity, tiles, dev<uint32_t>(tr->ranges), dev<uint32_t>(tr->tile_order));
          dio::depth_order(s, dev<uint32_t>(tr->depth_keys), dev<uint32_t>(tr->ranges), dev<uint32_t>(tr->tile_order),
                           tiles, dev<uint32_t>(in_alt ? tr->alt_splats : tr->pair_splats),
                           dev<uint32_t>(in_alt ? tr->pair_splats : tr->alt_splats), true,
                           dev<uint32_t>(tr->order_queue), tr->order_parity++);
        }
      };
      hipEvent_t e0, e1, e2;
      DIO_HIP(hipEventCreate(&e0));
      DIO_HIP(hipEventCreate(&e1));
      DIO_HIP(hipEventCreate(&e2));
      chain(true);  // warm
      DIO_HIP(hipEventRecord(e0, s));
      for (uint32_t k = 0; k < repeats; k++) chain(true);
      DIO_HIP(hipEventRecord(e1, s));
      for (uint32_t k = 0; k < repeats; k++) chain(false);
      DIO_HIP(hipEventRecord(e2, s));
      DIO_HIP(hipEventSynchronize(e2));
      float with_sorts = 0.f, without = 0.f;
      DIO_HIP(hipEventElapsedTime(&with_sorts, e0, e1));
      DIO_HIP(hipEventElapsedTime(&without, e1, e2));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipEventDestroy(e2);
      // Leave the frame as a step leaves it.
      chain(true);
      tr->sorted_in_alt = !binned && (tile_sort_bits(tiles, tr->digit_bits) / tr->digit_bits) % 2 == 1;
      sync(t->ctx);
      *ms = std::max(0.0, static_cast<double>(with_sorts - without)) / repeats;
      return;
    } else if (name == "rasterize") {
      run = [&] {
        // The forward pass appends the segmented backward's items: start them afresh.
        if (framed_segments) dio::fill_u32(s, framed_segments->counts, 0u, 1 + dio::kPartClasses);
        dio::rasterize(s, dev<float>(sc->splats), dev<uint32_t>(sorted_splats), dev<uint32_t>(tr->ranges),
                       dev<uint32_t>(tr->tile_order), tr->frame_w, tr->frame_h, tr->frame_bg, nullptr,
                       dev<float>(tr->pixel_color), dev<uint32_t>(tr->pixel_last), tr->splat_batch, tr->tile_size,
                       tr->persistent ? dev<uint32_t>(tr->work) : nullptr, framed_segments);
      };
    } else if (name == "rasterize_backward") {
      run = [&] {
        dio::rasterize_backward(s, dev<float>(sc->splats), dev<uint32_t>(sorted_splats), dev<uint32_t>(tr->ranges),
                                dev<uint32_t>(tr->tile_order), tr->frame_w, tr->frame_h, tr->frame_bg,
                                dev<float>(tr->pixel_color),
                                dev<uint32_t>(tr->pixel_last), dev<float>(t->loss->pixel_grads),
                                dev<float>(sc->grad_splats), tr->splat_batch, tr->tile_size, nullptr, nullptr,
                                tr->persistent ? dev<uint32_t>(tr->work) : nullptr, framed_segments);
      };
    } else if (name == "preprocess_backward") {
      run = [&] {
        dio::preprocess_backward(s, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree, sc->count, r->camera,
                                 dev<float>(sc->splats), dev<float>(sc->grad_splats), dev<float>(sc->grad_gaussians),
                                 dev<float>(sc->grad_sh));
      };
    } else if (name == "loss") {
      run = [&] {
        dio::training_loss(s, t->width, t->height, dev<float>(tr->pixel_color), dev<uint32_t>(view.image),
                           dev<float>(t->loss->partials), dev<float>(t->loss->pixel_grads),
                           dev<float>(t->loss->pixel_loss), dev<uint32_t>(sc->grad_splats),
                           static_cast<size_t>(sc->count) * 9);
      };
    } else if (name == "backward_adam") {
      run = [&] {
        dio::backward_adam(s, t->count, dev<float>(sc->gaussians), dev<float>(sc->sh), sc->degree, r->camera,
                           dev<float>(t->vis), dev<float>(sc->grad_splats), dev<float>(t->params),
                           dev<float>(t->moments), t->last_adam, dev<float>(t->stats), dev<float>(t->learned.extra));
      };
    } else if (name == "adam") {
      run = [&] {
        dio::adam_step(s, t->count, dev<float>(t->params), dev<float>(sc->sh), dev<float>(t->moments),
                       dev<float>(sc->grad_gaussians), dev<float>(sc->grad_sh), t->last_adam, dev<float>(sc->gaussians));
      };
    } else {
      invalid("unknown stage " + name);
    }
    hipEvent_t start, stop;
    DIO_HIP(hipEventCreate(&start));
    DIO_HIP(hipEventCreate(&stop));
    run();  // warm
    DIO_HIP(hipEventRecord(start, s));
    for (uint32_t k = 0; k < repeats; k++) run();
    DIO_HIP(hipEventRecord(stop, s));
    DIO_HIP(hipEventSynchronize(stop));
    float elapsed = 0;
    DIO_HIP(hipEventElapsedTime(&elapsed, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    *ms = elapsed / repeats;
  });
}

extern "C" int dio_trainer_stage_clock(dio_trainer* t, int enable) {
  return guard_read([&] {
    need(t, "trainer");
    if (t->batch) invalid("the stage clock times the plain trainer's step; a batch trainer has time_stage(\"batch_step\")");
    t->clock.harvest_all();
    t->clock.reset();
    t->clock.enabled = enable != 0;
  });
}

extern "C" int dio_trainer_stage_totals(dio_trainer* t, double* ms, uint32_t capacity, uint32_t* stages,
                                        uint32_t* steps, uint64_t* pairs) {
  return guard_read([&] {
    need(t, "trainer");
    if (ms && capacity < kMarks - 1) invalid("ms holds " + std::to_string(capacity) + " stages; the step has " +
                                             std::to_string(kMarks - 1));
    t->clock.harvest_all();
    if (ms)
      for (uint32_t k = 0; k + 1 < kMarks; k++) ms[k] = t->clock.totals_ms[k];
    if (stages) *stages = kMarks - 1;
    if (steps) *steps = t->clock.steps;
    if (pairs) *pairs = t->clock.pairs;
  });
}

extern "C" const char* dio_trainer_stage_name(uint32_t stage) {
  return stage + 1 < kMarks ? kStageNames[stage] : nullptr;
}

// ================================================== learned densification

namespace {

void destroy_mlp(dio_densify_mlp* m) {
  if (!m) return;
  (void)hipStreamSynchronize(m->ctx->stream);
  release(m->weights); release(m->master); release(m->moments); release(m->grads); release(m->loss);
  release(m->norm);
  dio_ctx* owner = m->ctx;
  delete m;
  dio::drop(owner);
}

void mlp_upload_master(dio_densify_mlp* m, const float* w) {
  hipStream_t s = m->ctx->stream;
  DIO_HIP(hipMemcpyAsync(m->master->ptr, w, dio::kMlpParams * 4, hipMemcpyHostToDevice, s));
  dio::mlp_round(s, dev<float>(m->master), dev<uint16_t>(m->weights));
  dio::fill_u32(s, dev<uint32_t>(m->moments), 0u, static_cast<size_t>(dio::kMlpParams) * 2);
  dio::fill_u32(s, dev<uint32_t>(m->grads), 0u, dio::kMlpParams);
  m->adam_steps = 0;
  sync(m->ctx);
}

}  // namespace

extern "C" int dio_mlp_create(dio_ctx* ctx, uint32_t seed, dio_densify_mlp** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    auto* m = new dio_densify_mlp;
    m->ctx = ctx;
    dio::hold(m->ctx);
    try {
      m->weights = alloc(ctx, dio::kMlpParams * 2, true);
      m->master = alloc(ctx, dio::kMlpParams * 4, true);
      m->moments = alloc(ctx, dio::kMlpParams * 8, true);
      m->grads = alloc(ctx, dio::kMlpParams * 4, true);
      m->loss = alloc(ctx, 16, true);
      m->norm = alloc(ctx, 2 * dio::kMlpInputs * 4, true);
      std::vector<float> identity(2 * dio::kMlpInputs, 0.f);
      for (uint32_t k = 0; k < dio::kMlpInputs; k++) identity[dio::kMlpInputs + k] = 1.f;
      DIO_HIP(hipMemcpyAsync(m->norm->ptr, identity.data(), identity.size() * 4, hipMemcpyHostToDevice, ctx->stream));
      // He-uniform layers, zero biases, from mulberry32(seed).
      std::vector<float> w(dio::kMlpParams, 0.f);
      Mulberry32 random(seed);
      const int fan_in[3] = {static_cast<int>(dio::kMlpInputs), 64, 64};
      const int offset[3] = {0, 12 * 64 + 64, 12 * 64 + 64 + 64 * 64 + 64};
      const int size[3] = {12 * 64, 64 * 64, 64 * 4};
      for (int l = 0; l < 3; l++) {
        const double bound = std::sqrt(6.0 / fan_in[l]);
        for (int k = 0; k < size[l]; k++) w[offset[l] + k] = static_cast<float>((2 * random.next() - 1) * bound);
      }
      mlp_upload_master(m, w.data());
    } catch (...) {
      destroy_mlp(m);
      throw;
    }
    *out = m;
  });
}

extern "C" int dio_mlp_get_weights(dio_densify_mlp* m, float* out) {
  return guard_read([&] {
    need(m, "mlp");
    need(out, "out");
    std::vector<uint16_t> bf(dio::kMlpParams);
    DIO_HIP(hipMemcpyAsync(bf.data(), m->weights->ptr, bf.size() * 2, hipMemcpyDeviceToHost, m->ctx->stream));
    sync(m->ctx);
    for (uint32_t k = 0; k < dio::kMlpParams; k++) {
      const uint32_t bits = static_cast<uint32_t>(bf[k]) << 16;
      std::memcpy(&out[k], &bits, 4);
    }
  });
}

extern "C" int dio_mlp_set_weights(dio_densify_mlp* m, const float* weights) {
  return guard([&] {
    need(m, "mlp");
    need(weights, "weights");
    mlp_upload_master(m, weights);
  });
}

extern "C" int dio_mlp_forward(dio_densify_mlp* m, dio_buf* features, uint32_t count, dio_buf* logits) {
  return guard([&] {
    need(m, "mlp");
    need(features, "features");
    need(logits, "logits");
    if (features->bytes < static_cast<size_t>(count) * dio::kMlpInputs * 4 || logits->bytes < static_cast<size_t>(count) * 16)
      invalid("the feature or logit buffer is smaller than count samples");
    dio::mlp_forward(m->ctx->stream, count, dev<float>(features), dev<uint16_t>(m->weights), dev<float>(m->norm),
                     dev<float>(logits));
  });
}

extern "C" int dio_mlp_gradients(dio_densify_mlp* m, dio_buf* features, dio_buf* labels, uint32_t count,
                                 float* grads, double* loss) {
  return guard([&] {
    need(m, "mlp");
    need(features, "features");
    need(labels, "labels");
    need(grads, "grads");
    need(loss, "loss");
    if (features->bytes < static_cast<size_t>(count) * dio::kMlpInputs * 4 || labels->bytes < static_cast<size_t>(count) * 4)
      invalid("the feature or label buffer is smaller than count samples");
    hipStream_t s = m->ctx->stream;
    dio::fill_u32(s, dev<uint32_t>(m->loss), 0u, 1);
    dio::mlp_gradients(s, count, count, 0, dev<float>(features), dev<uint32_t>(labels), dev<uint16_t>(m->weights),
                       dev<float>(m->norm), m->class_weights, dev<float>(m->grads), dev<float>(m->loss));
    float total = 0.f;
    DIO_HIP(hipMemcpyAsync(grads, m->grads->ptr, dio::kMlpParams * 4, hipMemcpyDeviceToHost, s));
    DIO_HIP(hipMemcpyAsync(&total, m->loss->ptr, 4, hipMemcpyDeviceToHost, s));
    dio::fill_u32(s, dev<uint32_t>(m->grads), 0u, dio::kMlpParams);
    sync(m->ctx);
    *loss = count ? total / count : 0.0;
  });
}

extern "C" int dio_mlp_train(dio_densify_mlp* m, dio_buf* features, dio_buf* labels, uint32_t count, uint32_t steps,
                             uint32_t batch, float lr, uint32_t seed, double* last_loss) {
  return guard([&] {
    need(m, "mlp");
    need(features, "features");
    need(labels, "labels");
    if (count == 0 || batch == 0) invalid("training needs samples and a batch size");
    if (features->bytes < static_cast<size_t>(count) * dio::kMlpInputs * 4 || labels->bytes < static_cast<size_t>(count) * 4)
      invalid("the feature or label buffer is smaller than count samples");
    hipStream_t s = m->ctx->stream;
    for (uint32_t k = 0; k < steps; k++) {
      dio::fill_u32(s, dev<uint32_t>(m->loss), 0u, 1);
      // a nonzero seed per step: `batch` hashed draws from the samples
      const uint32_t step_seed = ((seed * 2654435761u) ^ (k * 40503u + 1u)) | 1u;
      dio::mlp_gradients(s, batch, count, step_seed, dev<float>(features), dev<uint32_t>(labels),
                         dev<uint16_t>(m->weights), dev<float>(m->norm), m->class_weights, dev<float>(m->grads),
                         dev<float>(m->loss));
      m->adam_steps++;
      const float inv1 = static_cast<float>(1.0 / (1.0 - std::pow(0.9, m->adam_steps)));
      const float inv2 = static_cast<float>(1.0 / (1.0 - std::pow(0.999, m->adam_steps)));
      dio::mlp_adam(s, dev<float>(m->master), dev<float>(m->moments), dev<float>(m->grads), dev<uint16_t>(m->weights),
                    lr, inv1, inv2);
    }
    float total = 0.f;
    DIO_HIP(hipMemcpyAsync(&total, m->loss->ptr, 4, hipMemcpyDeviceToHost, s));
    sync(m->ctx);
    if (last_loss) *last_loss = total / batch;
  });
}

extern "C" int dio_mlp_set_input_norm(dio_densify_mlp* m, const float mean[12], const float scale[12]) {
  return guard([&] {
    need(m, "mlp");
    need(mean, "mean");
    need(scale, "scale");
    float host[2 * dio::kMlpInputs];
    for (uint32_t k = 0; k < dio::kMlpInputs; k++) {
      if (!std::isfinite(mean[k]) || !std::isfinite(scale[k])) invalid("the input standardisation must be finite");
      host[k] = mean[k];
      host[dio::kMlpInputs + k] = scale[k];
    }
    DIO_HIP(hipMemcpyAsync(m->norm->ptr, host, sizeof(host), hipMemcpyHostToDevice, m->ctx->stream));
    sync(m->ctx);
  });
}

extern "C" int dio_mlp_set_class_weights(dio_densify_mlp* m, const float weights[4]) {
  return guard([&] {
    need(m, "mlp");
    need(weights, "weights");
    for (int c = 0; c < 4; c++) {
      if (!(weights[c] >= 0.f)) invalid("class weights must be non-negative");
      m->class_weights[c] = weights[c];
    }
  });
}

extern "C" void dio_mlp_destroy(dio_densify_mlp* m) { destroy_mlp(m); }

extern "C" int dio_trainer_set_densify_mode(dio_trainer* t, uint32_t mode, dio_densify_mlp* mlp, uint32_t budget,
                                            float confidence, uint32_t samples_per_densify) {
  return guard([&] {
    need(t, "trainer");
    if (mode > 2) invalid("densify mode is 0 (rule), 1 (rule, logging samples) or 2 (learned)");
    if (mode == 2) need(mlp, "mlp");
    if (mode == 1 && t->batch) invalid("samples are logged by the plain (one view per step) trainer");
    LearnedState& L = t->learned;
    L.mode = mode;
    L.mlp = mlp;
    L.budget = budget;
    L.confidence = confidence;
    L.samples_per_densify = samples_per_densify ? samples_per_densify : 80000;
    if (mode != 0 && !L.extra) L.extra = alloc(t->ctx, static_cast<size_t>(t->count) * 16, true);
    if (mode == 1 && !L.footprint) L.footprint = alloc(t->ctx, static_cast<size_t>(t->count) * 8, true);
    // The footprint terms come from the 16-px backward.
    if (mode == 1) t->tuning.tile_size = 16;
    if (mode == 1 && !L.pending) {
      L.pending_capacity = 2 * L.samples_per_densify + 1024;
      L.pending = alloc(t->ctx, static_cast<size_t>(L.pending_capacity) * dio::kMlpPendingWords * 4, true);
      L.pending_count = alloc(t->ctx, 16, true);
      L.store_capacity = 8u << 20;
      L.store_features = alloc(t->ctx, static_cast<size_t>(L.store_capacity) * dio::kMlpInputs * 4, true);
      L.store_labels = alloc(t->ctx, static_cast<size_t>(L.store_capacity) * 4, true);
      L.store_count = alloc(t->ctx, 16, true);
    }
    if (mode != 1) L.pending_valid = false;
    learned_resize(t);
    t->generation++;  // a captured batch graph must pick up the statistics buffer
  });
}

extern "C" int dio_trainer_samples(dio_trainer* t, dio_buf** features, dio_buf** labels, uint32_t* count) {
  return guard([&] {
    need(t, "trainer");
    need(features, "features");
    need(labels, "labels");
    need(count, "count");
    LearnedState& L = t->learned;
    uint32_t n = 0;
    if (L.store_count) {
      DIO_HIP(hipMemcpyAsync(&n, L.store_count->ptr, 4, hipMemcpyDeviceToHost, t->ctx->stream));
      sync(t->ctx);
    }
    *features = L.store_features;
    *labels = L.store_labels;
    *count = std::min(n, L.store_capacity);
  });
}

extern "C" int dio_trainer_features(dio_trainer* t, dio_buf** features) {
  return guard([&] {
    need(t, "trainer");
    need(features, "features");
    LearnedState& L = t->learned;
    if (!L.extra) invalid("set a densify mode first: the inputs need its statistics");
    learned_resize(t);
    dio::mlp_features(t->ctx->stream, t->count, dev<float>(t->params), dev<float>(t->stats), dev<float>(L.extra),
                      t->settings.extent, t->iteration, t->batch ? t->batch->views : 1u, dev<float>(L.features));
    *features = L.features;
  });
}

extern "C" dio_buf* dio_trainer_extra_stats(dio_trainer* t) { return t ? t->learned.extra : nullptr; }
extern "C" dio_buf* dio_trainer_footprint(dio_trainer* t) { return t ? t->learned.footprint : nullptr; }

extern "C" int dio_trainer_batch_frames(dio_trainer* t, uint32_t* views, uint32_t capacity, uint32_t* count,
                                        dio_buf** splats, uint32_t* stride, dio_buf** pixels) {
  return guard_read([&] {
    need(t, "trainer");
    if (!t->batch) invalid("not a batch trainer");
    const BatchState* b = t->batch;
    const uint32_t used = static_cast<uint32_t>(b->step_views.size());
    if (count) *count = used;
    if (views)
      for (uint32_t k = 0; k < std::min(capacity, used); k++) views[k] = b->step_views[k];
    if (splats) *splats = b->splats;
    if (stride) *stride = b->n;
    if (pixels) *pixels = b->pixel_color;
  });
}

// ================================================================= tuning

namespace {

void check_tuning(const dio_trainer* t, const dio_tuning& k) {
  check_launch_tuning(k);
  if (k.persistent && t->batch) invalid("a batch trainer launches a workgroup per tile");
  if (t->batch && k.tile_size != 16) invalid("a batch trainer rasterizes 16-pixel tiles");
  if (t->learned.mode == 1 && k.tile_size != 16) invalid("a trainer logging samples rasterizes 16-pixel tiles");
  if (k.segment && t->batch) invalid("a batch trainer replays whole lists");
  if (k.per_tile_depth && t->batch) invalid("a batch trainer sorts its splats by depth");
}

// The autotuner's candidates: every valid (depth order, tile size, persistent,
// segment, digit width, splat batch); segments of 1024 pairs are left to explicit
// settings.
std::vector<dio_tuning> tuning_candidates(const dio_trainer* t) {
  std::vector<dio_tuning> out;
  for (uint32_t per_tile : {0u, 1u, 2u})
    for (uint32_t tile : {16u, 8u, 32u})
      for (uint32_t persistent : {0u, 1u})
        for (uint32_t segment : {0u, 256u, 512u})
          for (uint32_t bits : {8u, 4u})
            for (uint32_t batch : {64u, 128u, 256u}) {
              if ((t->batch && tile != 16) || (t->learned.mode == 1 && tile != 16) || (tile == 32 && batch == 256))
                continue;
              if (persistent && (tile != 16 || t->batch)) continue;
              if (segment && (tile != 16 || persistent || t->batch)) continue;
              if (per_tile && t->batch) continue;
              if (per_tile == 2 && bits != 8) continue;  // binning by counts sorts no pairs by radix
              out.push_back(dio_tuning{tile, batch, bits, persistent, segment, per_tile});
            }
  return out;
}

void apply_tuning(dio_trainer* t, const dio_tuning& k) {
  check_tuning(t, k);
  const bool changed = k.splat_batch != t->tuning.splat_batch || k.digit_bits != t->tuning.digit_bits ||
                       k.tile_size != t->tuning.tile_size || k.persistent != t->tuning.persistent ||
                       k.segment != t->tuning.segment || k.per_tile_depth != t->tuning.per_tile_depth;
  t->tuning = k;
  if (changed) t->generation++;  // a captured batch graph bakes the launches in
}

}  // namespace

extern "C" int dio_trainer_set_tuning(dio_trainer* t, const dio_tuning* tuning) {
  return guard([&] {
    need(t, "trainer");
    need(tuning, "tuning");
    apply_tuning(t, *tuning);
  });
}

extern "C" int dio_trainer_get_tuning(const dio_trainer* t, dio_tuning* out) {
  return guard_read([&] {
    need(t, "trainer");
    need(out, "out");
    *out = t->tuning;
  });
}

extern "C" int dio_trainer_autotune(dio_trainer* t, uint32_t steps, dio_tuning_result* results, uint32_t capacity,
                                    uint32_t* count) {
  return guard([&] {
    need(t, "trainer");
    if (steps == 0) invalid("the autotuner needs at least one timed step per candidate");
    hipStream_t s = t->ctx->stream;
    hipEvent_t e0, e1;
    DIO_HIP(hipEventCreate(&e0));
    DIO_HIP(hipEventCreate(&e1));
    std::vector<dio_tuning_result> found;
    // Mean ms per step of `candidate` over `n` steps after one untimed step.
    auto time = [&](const dio_tuning& candidate, uint32_t n) {
      apply_tuning(t, candidate);
      trainer_optimize(t);  // untimed: captures a batch graph, warms the launches
      trainer_maintain(t);
      const bool replay = t->batch && t->batch->exec && !t->due_densify;
      DIO_HIP(hipEventRecord(e0, s));
      for (uint32_t k = 0; k < n; k++) {
        if (replay) {
          // A batch trainer replays its step graph on the same views, so every
          // candidate is timed on the same work.
          DIO_HIP(hipGraphLaunch(t->batch->exec, s));
        } else {
          trainer_optimize(t);
          trainer_maintain(t);
        }
      }
      DIO_HIP(hipEventRecord(e1, s));
      DIO_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      DIO_HIP(hipEventElapsedTime(&ms, e0, e1));
      return static_cast<double>(ms) / n;
    };
    try {
      for (const dio_tuning& candidate : tuning_candidates(t)) found.push_back({candidate, time(candidate, steps)});
      // A single-view trainer's steps each take the next training view, whose pair
      // count (and so its cost) varies by a few percent: the four fastest are timed
      // again over three times the steps, and those times replace their first ones.
      if (!t->batch && found.size() > 1) {
        std::vector<size_t> order(found.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = k;
        std::sort(order.begin(), order.end(),
                  [&](size_t a, size_t b) { return found[a].ms_per_step < found[b].ms_per_step; });
        for (size_t k = 0; k < std::min<size_t>(4, order.size()); k++)
          found[order[k]].ms_per_step = time(found[order[k]].tuning, 3 * steps);
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const auto best = std::min_element(found.begin(), found.end(),
                                       [](const dio_tuning_result& a, const dio_tuning_result& b) {
                                         return a.ms_per_step < b.ms_per_step;
                                       });
    apply_tuning(t, best->tuning);
    if (count) *count = static_cast<uint32_t>(found.size());
    if (results)
      for (uint32_t k = 0; k < capacity && k < found.size(); k++) results[k] = found[k];
  });
}
