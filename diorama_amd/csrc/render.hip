// The forward and backward renderer kernels.
//
// Tile binning sorts once per Gaussian and once per pair instead of one 48-bit
// key sort over every pair: Gaussians are radix-sorted by depth (32-bit keys),
// their pairs are written out in that order, and a stable sort by tile index
// alone (ceil(log2 tiles) bits) then leaves each tile's list front to back with
// equal depths in splat order. That is exactly the order of the reference's
// stable sort of (tile, depth) keys written in splat order
// (src/shaders/tiles.wgsl:36-59, src/radixSort.ts:103-104), so the lists and
// ranges are bit-identical while the pair traffic drops from 6 passes to 1-2.
#include <map>
#include <mutex>

#include "gaussian_math.hpp"
#include "block_copy.hpp"
#include "raster_tile.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
using raster::kGrads;

using ShBlock = RecordCopy<kThreads, kShRest, kThreads>;
using GradBlock = RecordCopy<kThreads, kGrads, kThreads>;

// ------------------------------------------------------------ preprocess --

// preprocess.wgsl:20-87 plus count_tiles (tiles.wgsl:23-34) fused: one thread per
// Gaussian writes its splat, its tile count and its depth-sort key. A workgroup
// stages its 256 Gaussians' SH records in LDS when any of them is visible.
__global__ __launch_bounds__(kThreads) void preprocess_kernel(const float4* __restrict__ gaussians,
                                                               const float* __restrict__ sh, uint32_t sh_degree,
                                                               uint32_t n, CameraArgs cam, float4* __restrict__ splats,
                                                               uint32_t* __restrict__ tile_counts,
                                                               uint32_t* __restrict__ depth_keys,
                                                               uint32_t* __restrict__ depth_values,
                                                               float2* __restrict__ vis, uint2* __restrict__ rects) {
  __shared__ __align__(16) float s_sh[kThreads * kShRest];
  const uint32_t first = blockIdx.x * kThreads, i = first + threadIdx.x;
  const bool active = i < n;
  float g[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    float4 v = active ? gaussians[i * 4 + q] : make_float4(0.f, 0.f, 0.f, 1.f);
    g[4 * q] = v.x; g[4 * q + 1] = v.y; g[4 * q + 2] = v.z; g[4 * q + 3] = v.w;
  }
  Projection p;
  const bool visible = active && project(g, cam, p);
  const int count = sh ? sh_count(sh_degree) : 0;
  if (count > 0 && __syncthreads_or(visible)) {
    const float* src = sh + static_cast<size_t>(first) * kShRest;
    if (n - first >= kThreads) {
      ShBlock c;
      c.load(src);
      c.store(s_sh);
    } else {
      block_copy<kThreads>(s_sh, src, (n - first) * kShRest);
    }
  }
  __syncthreads();
  if (!active) return;
  SplatRecord rec;
  splat_record(g, p, visible, cam, s_sh + threadIdx.x * kShRest, count, rec);
  splats[i * 3] = rec.s0;
  splats[i * 3 + 1] = rec.s1;
  splats[i * 3 + 2] = rec.s2;
  tile_counts[i] = rec.touched;
  depth_keys[i] = rec.key;
  depth_values[i] = i;
  if (vis) vis[i] = make_float2(rec.s0.z, rec.s0.w);
  if (rects) rects[i] = rec.rect;
}

// count_tiles (tiles.wgsl:23-34) for splats that did not come from the preprocess,
// in bins of (1 << shift) x (1 << shift) 16-px tiles.
__global__ __launch_bounds__(kThreads) void count_tiles_kernel(const float4* __restrict__ splats, uint32_t n,
                                                                uint32_t tiles_x, uint32_t tiles_y, uint32_t shift,
                                                                uint32_t* __restrict__ tile_counts,
                                                                uint32_t* __restrict__ depth_keys,
                                                                uint32_t* __restrict__ depth_values,
                                                                uint2* __restrict__ rects) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float4 s0 = splats[i * 3];
  uint32_t touched = 0, key = 0xffffffffu;
  uint2 rect = make_uint2(0u, 0u);
  if (s0.w > 0.f) {
    uint32_t r[4];
    bin_rect(s0.x, s0.y, s0.w, tiles_x, tiles_y, shift, r);
    touched = (r[2] - r[0]) * (r[3] - r[1]);
    key = __float_as_uint(s0.z);
    rect = pack_rect(r);
  }
  tile_counts[i] = touched;
  depth_keys[i] = key;
  depth_values[i] = i;
  if (rects) rects[i] = rect;
}

// duplicate (tiles.wgsl:36-59), load-balanced over pairs rather than splats: each
// workgroup writes 2048-slot stretches of the pair buffers (one, unless the pairs
// outgrow the grid sized for the expected count). For a stretch it finds the
// depth-ordered splat owning its first slot with a 64-ary search of the offsets
// (one wavefront, a few rounds), then deals its slots out to threads from LDS
// windows of 256 splats. The nearest splats cover the most tiles and sit together
// in depth order, so splat-parallel work would pile onto a few waves.
constexpr uint32_t kDupPairs = 2048;

// kBatched: splat s belongs to view s / per_view of `views`, whose tile grid and
// tile_base replace tiles_x, tiles_y and 0. Pairs go to bins of (1 << shift) x
// (1 << shift) 16-px tiles (tiles_x, tiles_y count 16-px tiles).
template <bool kBatched>
__global__ __launch_bounds__(kThreads) void duplicate_kernel(const float4* __restrict__ splats,
                                                              const uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ offsets,
                                                              const uint32_t* __restrict__ tile_counts, uint32_t n,
                                                              uint32_t tiles_x, uint32_t tiles_y, uint32_t shift,
                                                              const ViewDesc* __restrict__ views, uint32_t per_view,
                                                              uint32_t capacity,
                                                              const uint32_t* __restrict__ pair_total,
                                                              uint32_t* __restrict__ pair_tiles,
                                                              uint32_t* __restrict__ pair_splats) {
  __shared__ uint32_t s_start[kThreads], s_count[kThreads], s_x0[kThreads], s_y0[kThreads], s_w[kThreads],
      s_splat[kThreads], s_row[kThreads], s_base[kThreads];
  __shared__ uint32_t s_first;
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const uint32_t total = min(*pair_total, capacity);  // slots past the capacity are dropped
  for (uint32_t p0 = blockIdx.x * kDupPairs; p0 < total; p0 += gridDim.x * kDupPairs) {
    const uint32_t p_end = min(p0 + kDupPairs, total);
    if (tid < kWave) {
      // The last j with offsets[j] <= p0 (offsets never decrease; offsets[0] = 0).
      uint32_t lo = 0, hi = n;
      while (hi - lo > 1) {
        const uint32_t stride = ceil_div(hi - lo, kWave);
        const uint32_t probe = lo + lane * stride;
        const uint64_t ok = __ballot(probe < hi && offsets[probe] <= p0);
        lo += static_cast<uint32_t>(kWave - 1 - __builtin_clzll(ok)) * stride;
        hi = min(hi, lo + stride);
      }
      if (lane == 0) s_first = lo;
    }
    __syncthreads();
    uint32_t j_base = s_first, p_lo = p0;
    while (p_lo < p_end && j_base < n) {
      const uint32_t j = j_base + tid;
      uint32_t start = 0xffffffffu, count = 0, splat = 0, r[4] = {0, 0, 0, 0};
      uint32_t row_tiles = tiles_x, base = 0;
      if (j < n) {
        splat = order ? order[j] : j;
        start = offsets[j];
        count = tile_counts[splat];
        if (count) {
          uint32_t grid_y = tiles_y;
          if (kBatched) {
            const ViewDesc& view = views[splat / per_view];
            row_tiles = view.tiles_x;
            grid_y = view.tiles_y;
            base = view.tile_base;
          }
          const float4 s0 = splats[static_cast<size_t>(splat) * 3];
          bin_rect(s0.x, s0.y, s0.w, row_tiles, grid_y, shift, r);
          row_tiles = (row_tiles + (1u << shift) - 1u) >> shift;
        }
      }
      s_start[tid] = start;
      s_count[tid] = count;
      s_x0[tid] = r[0];
      s_y0[tid] = r[1];
      s_w[tid] = r[2] - r[0];
      s_splat[tid] = splat;
      s_row[tid] = row_tiles;
      s_base[tid] = base;
      __syncthreads();
      const int valid = static_cast<int>(min(static_cast<uint32_t>(kThreads), n - j_base));
      const uint32_t window_end = s_start[valid - 1] + s_count[valid - 1];
      const uint32_t p_hi = min(p_end, window_end);
      for (uint32_t p = p_lo + tid; p < p_hi; p += kThreads) {
        int lo = 0, hi = valid - 1;  // the last entry starting at or before p
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_start[mid] <= p) lo = mid; else hi = mid - 1;
        }
        const uint32_t k = p - s_start[lo], width = s_w[lo];
        const uint32_t row = k / width;
        pair_tiles[p] = s_base[lo] + (s_y0[lo] + row) * s_row[lo] + s_x0[lo] + (k - row * width);
        pair_splats[p] = s_splat[lo];
      }
      p_lo = max(p_lo, p_hi);
      j_base += kThreads;
      __syncthreads();
    }
  }
}

// find_ranges (tiles.wgsl:61-84): each tile's [start, end) run of the sorted pairs,
// (0, 0) for a tile without pairs, a thread per pair (grid-stride). The pair at a change of
// tile writes the end of the previous tile, the start of its own and the empty
// tiles between them (the first and last pairs also the tiles before and after
// them), so no clearing pass is needed.
__global__ __launch_bounds__(kThreads) void find_ranges_kernel(const uint32_t* __restrict__ pair_tiles,
                                                                const uint32_t* __restrict__ pair_total,
                                                                uint32_t capacity, uint32_t tiles,
                                                                uint2* __restrict__ ranges, bool cleared) {
  const uint32_t count = min(*pair_total, capacity);
  const uint32_t first = blockIdx.x * kThreads + threadIdx.x, stride = gridDim.x * kThreads;
  const uint2 empty = make_uint2(0u, 0u);
  // With `cleared` every range is already empty: only the runs' ends are written.
  if (count == 0) {
    if (!cleared)
      for (uint32_t t = first; t < tiles; t += stride) ranges[t] = empty;
    return;
  }
  for (uint32_t i = first; i < count; i += stride) {
    const uint32_t tile = min(pair_tiles[i], tiles - 1);
    if (i == 0) {
      if (!cleared)
        for (uint32_t t = 0; t < tile; t++) ranges[t] = empty;
      ranges[tile].x = 0;
    } else {
      const uint32_t previous = min(pair_tiles[i - 1], tiles - 1);
      if (tile != previous) {
        ranges[previous].y = i;
        if (!cleared)
          for (uint32_t t = previous + 1; t < tile; t++) ranges[t] = empty;
        ranges[tile].x = i;
      }
    }
    if (i == count - 1) {
      ranges[tile].y = count;
      if (!cleared)
        for (uint32_t t = tile + 1; t < tiles; t++) ranges[t] = empty;
    }
  }
}

// The rasterizer's launch order: tiles by falling pair count (in steps of
// kOrderStep pairs; ties in any order), so that the heaviest tiles start first and
// the light ones fill in behind them instead of the centre tiles of a view, which
// row-major order dispatches halfway through, finishing last. One workgroup
// counting-sorts the tiles, a thread per bucket.
constexpr uint32_t kOrderThreads = kOrderBuckets;

__global__ __launch_bounds__(kOrderThreads) void tile_order_kernel(const uint2* __restrict__ ranges, uint32_t tiles,
                                                                   uint32_t* __restrict__ order,
                                                                   uint32_t* __restrict__ clear) {
  __shared__ uint32_t s_next[kOrderThreads];
  __shared__ uint32_t s_wave[kOrderThreads / 64];
  const uint32_t tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  if (clear && tid <= kPartClasses) clear[tid] = 0;
  auto bucket = [&](uint32_t tile) {  // 0 for the heaviest
    const uint2 r = ranges[tile];
    return kOrderThreads - 1 - min((r.y - r.x) / kOrderStep, kOrderThreads - 1);
  };
  s_next[tid] = 0;
  __syncthreads();
  for (uint32_t t = tid; t < tiles; t += kOrderThreads) atomicAdd(&s_next[bucket(t)], 1u);
  __syncthreads();
  // Exclusive scan of the bucket sizes: within each wave, then over the wave totals.
  const uint32_t mine = s_next[tid];
  uint32_t x = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= static_cast<uint32_t>(o)) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; w++) before += s_wave[w];
  s_next[tid] = before + x - mine;
  __syncthreads();
  for (uint32_t t = tid; t < tiles; t += kOrderThreads) order[atomicAdd(&s_next[bucket(t)], 1u)] = t;
}

// -------------------------------------------------------------- rasterize --
// Three workgroup shapes, one per tile size; results do not depend on the shape
// (only pixel_last's list positions do, with 32-px bins):
// * 16: a 256-thread workgroup per 16-px tile, wave w on quadrant w;
// * 8: a 64-thread workgroup (one wave) per 8x8 quadrant of the 16-px lists. An
//   8-px tile's list culled by the exact 16-px rectangle is its 16-px tile's list,
//   so the quadrant reads that list instead of a copy of it;
// * 32: a 1024-thread workgroup per 32-px bin, wave w on quadrant w % 4 of the
//   bin's 16-px tile w / 4, culling the bin's list to that tile's.
// `order`, if given, maps workgroups to tiles or bins (heaviest first).

// rasterize.wgsl:31-100: front-to-back blend.
template <int kPer>
__global__ __launch_bounds__(256) void rasterize_kernel(const float4* __restrict__ splats,
                                                         const uint32_t* __restrict__ pair_splats,
                                                         const uint2* __restrict__ ranges,
                                                         const uint32_t* __restrict__ order, uint32_t width,
                                                         uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
                                                         float bg_b, uint32_t* __restrict__ target,
                                                         float4* __restrict__ pixel_color,
                                                         uint32_t* __restrict__ pixel_last, Segments seg) {
  const uint32_t tile = order ? order[blockIdx.x] : blockIdx.x;
  raster::rasterize_tile<kPer>(splats, pair_splats, ranges[tile], tile % tiles_x, tile / tiles_x, width, height, bg_r,
                               bg_g, bg_b, target, pixel_color, pixel_last, seg);
}

// Persistent workgroups (16-px tiles): a grid of as many workgroups as the GPU holds
// at once, each taking the next entry of `order` from work[0] until none is left.
// A workgroup counts itself out in work[1] after its last take, so the last one out
// knows every take is done and leaves both counters at zero for the next launch.
__device__ inline bool take_tile(uint32_t* work, uint32_t tiles, uint32_t& k) {
  __shared__ uint32_t s_k;
  __syncthreads();  // the previous tile is done and s_k has been read
  if (threadIdx.x == 0) s_k = atomicAdd(work, 1u);
  __syncthreads();
  k = s_k;
  if (k < tiles) return true;
  if (threadIdx.x == 0 && atomicAdd(work + 1, 1u) == gridDim.x - 1) {
    __hip_atomic_store(work, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(work + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return false;
}

template <int kPer>
__global__ __launch_bounds__(256) void rasterize_persistent_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ order, uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
    float bg_b, uint32_t* __restrict__ target, float4* __restrict__ pixel_color, uint32_t* __restrict__ pixel_last,
    uint32_t tiles, uint32_t* __restrict__ work) {
  uint32_t k;
  while (take_tile(work, tiles, k)) {
    const uint32_t tile = order ? order[k] : k;
    raster::rasterize_tile<kPer>(splats, pair_splats, ranges[tile], tile % tiles_x, tile / tiles_x, width, height,
                                 bg_r, bg_g, bg_b, target, pixel_color, pixel_last);
  }
}

template <int kPer>
__global__ __launch_bounds__(64) void rasterize_quad_kernel(const float4* __restrict__ splats,
                                                            const uint32_t* __restrict__ pair_splats,
                                                            const uint2* __restrict__ ranges,
                                                            const uint32_t* __restrict__ order, uint32_t width,
                                                            uint32_t height, uint32_t tiles_x, float bg_r,
                                                            float bg_g, float bg_b, uint32_t* __restrict__ target,
                                                            float4* __restrict__ pixel_color,
                                                            uint32_t* __restrict__ pixel_last) {
  __shared__ raster::SplatRecords s_records[kPer];
  const uint32_t tile = order ? order[blockIdx.x >> 2] : blockIdx.x >> 2;
  raster::rasterize_wave<kPer, false>(splats, pair_splats, ranges[tile],
                                      raster::WaveTile{tile % tiles_x, tile / tiles_x, blockIdx.x & 3u}, width, height,
                                      bg_r, bg_g, bg_b, target, pixel_color, pixel_last, s_records);
}

template <int kPer>
__global__ __launch_bounds__(1024) void rasterize_bin_kernel(const float4* __restrict__ splats,
                                                             const uint32_t* __restrict__ pair_splats,
                                                             const uint2* __restrict__ ranges,
                                                             const uint32_t* __restrict__ order, uint32_t width,
                                                             uint32_t height, uint32_t bins_x, float bg_r,
                                                             float bg_g, float bg_b, uint32_t* __restrict__ target,
                                                             float4* __restrict__ pixel_color,
                                                             uint32_t* __restrict__ pixel_last) {
  __shared__ raster::SplatRecords s_records[16][kPer];
  const uint32_t bin = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t wave = threadIdx.x / kWave, sub = wave >> 2;
  const uint32_t sx = (bin % bins_x) * 2 + (sub & 1), sy = (bin / bins_x) * 2 + (sub >> 1);
  if (sx * kTile >= width || sy * kTile >= height) return;  // past the frame's edge (no barriers follow)
  raster::rasterize_wave<kPer, true>(splats, pair_splats, ranges[bin], raster::WaveTile{sx, sy, wave & 3u}, width,
                                     height, bg_r, bg_g, bg_b, target, pixel_color, pixel_last, s_records[wave]);
}

// rasterizeBackward.wgsl:49-163.
template <int kPer>
__global__ __launch_bounds__(256) void rasterize_backward_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ order, uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
    float bg_b, const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads) {
  __shared__ raster::BackwardShared<kPer> shared;
  const uint32_t tile = order ? order[blockIdx.x] : blockIdx.x;
  raster::rasterize_backward_tile<kPer>(splats, pair_splats, ranges[tile], tile % tiles_x, tile / tiles_x, width,
                                        height, bg_r, bg_g, bg_b, pixel_color, pixel_last, pixel_grads, splat_grads,
                                        shared);
}

template <int kPer>
__global__ __launch_bounds__(256) void rasterize_backward_persistent_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ order, uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
    float bg_b, const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads, uint32_t tiles,
    uint32_t* __restrict__ work) {
  __shared__ raster::BackwardShared<kPer> shared;
  uint32_t k;
  while (take_tile(work, tiles, k)) {
    const uint32_t tile = order ? order[k] : k;
    raster::rasterize_backward_tile<kPer>(splats, pair_splats, ranges[tile], tile % tiles_x, tile / tiles_x, width,
                                          height, bg_r, bg_g, bg_b, pixel_color, pixel_last, pixel_grads, splat_grads,
                                          shared);
  }
}

// The segmented backward: each workgroup replays the items blockIdx.x,
// blockIdx.x + gridDim.x, ... of `items` (see Segments), with the footprint terms
// too when kFoot.
template <int kPer, bool kFoot>
__global__ __launch_bounds__(256) void rasterize_backward_segment_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    Segments seg, uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g, float bg_b,
    const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads, const float* __restrict__ pixel_loss,
    float* __restrict__ footprint) {
  __shared__ raster::BackwardShared<kPer, 4, false, kFoot> shared;
  const uint2* items = reinterpret_cast<const uint2*>(seg.items);
  uint32_t count[1 + kPartClasses], total = 0;
#pragma unroll
  for (uint32_t c = 0; c <= kPartClasses; c++) {
    count[c] = min(seg.counts[c], c ? seg.part_room : seg.whole_room);
    total += count[c];
  }
  for (uint32_t i = blockIdx.x; i < total; i += gridDim.x) {
    // Item i in class order: whole segments, then the short parts, longest first.
    uint32_t k = i, c = 0;
    while (k >= count[c]) k -= count[c++];
    const uint2 item = items[c ? seg.whole_room + (c - 1) * seg.part_room + k : k];
    const uint2 run = ranges[item.x];
    const uint32_t lo = max(run.x, item.y << seg.shift), hi = min(run.y, (item.y << seg.shift) + (1u << seg.shift));
    raster::rasterize_backward_group<kPer, 4, false, kFoot>(
        splats, pair_splats, run,
        raster::WaveTile{item.x % tiles_x, item.x / tiles_x, static_cast<uint32_t>(threadIdx.x / kWave)}, width,
        height, bg_r, bg_g, bg_b, pixel_color, pixel_last, pixel_grads, splat_grads, shared, pixel_loss, footprint,
        raster::Part{lo, hi, reinterpret_cast<const float4*>(seg.ckpt), seg.shift});
  }
}

// The same with the footprint terms (see raster::Replay) added into `footprint`.
template <int kPer>
__global__ __launch_bounds__(256) void rasterize_backward_foot_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ order, uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
    float bg_b, const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads, const float* __restrict__ pixel_loss,
    float* __restrict__ footprint) {
  __shared__ raster::BackwardShared<kPer, 4, false, true> shared;
  const uint32_t tile = order ? order[blockIdx.x] : blockIdx.x;
  raster::rasterize_backward_group<kPer, 4, false, true>(
      splats, pair_splats, ranges[tile],
      raster::WaveTile{tile % tiles_x, tile / tiles_x, static_cast<uint32_t>(threadIdx.x / kWave)}, width, height,
      bg_r, bg_g, bg_b, pixel_color, pixel_last, pixel_grads, splat_grads, shared, pixel_loss, footprint);
}

template <int kPer>
__global__ __launch_bounds__(64) void rasterize_backward_quad_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ order, uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
    float bg_b, const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads) {
  __shared__ raster::BackwardShared<kPer, 1, false> shared;
  const uint32_t tile = order ? order[blockIdx.x >> 2] : blockIdx.x >> 2;
  raster::rasterize_backward_group<kPer, 1, false>(
      splats, pair_splats, ranges[tile], raster::WaveTile{tile % tiles_x, tile / tiles_x, blockIdx.x & 3u}, width,
      height, bg_r, bg_g, bg_b, pixel_color, pixel_last, pixel_grads, splat_grads, shared);
}

// Waves whose tile lies past the frame's edge have no pixels and no hits, but join
// the workgroup's barriers.
template <int kPer>
__global__ __launch_bounds__(1024) void rasterize_backward_bin_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ order, uint32_t width, uint32_t height, uint32_t bins_x, float bg_r, float bg_g,
    float bg_b, const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads) {
  __shared__ raster::BackwardShared<kPer, 16, true> shared;
  const uint32_t bin = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t wave = threadIdx.x / kWave, sub = wave >> 2;
  const uint32_t sx = (bin % bins_x) * 2 + (sub & 1), sy = (bin / bins_x) * 2 + (sub >> 1);
  raster::rasterize_backward_group<kPer, 16, true>(splats, pair_splats, ranges[bin],
                                                   raster::WaveTile{sx, sy, wave & 3u}, width, height, bg_r, bg_g,
                                                   bg_b, pixel_color, pixel_last, pixel_grads, splat_grads, shared);
}

// ----------------------------------------------------- preprocess backward --

// preprocessBackward.wgsl:27-181: recompute the projection, then carry the
// splat's 9 gradients back to position, opacity, scale, rotation, colour and SH.
__global__ __launch_bounds__(kThreads) void preprocess_backward_kernel(
    const float4* __restrict__ gaussians, const float* __restrict__ sh, uint32_t sh_degree, uint32_t n,
    CameraArgs cam, const float4* __restrict__ splats, const float* __restrict__ splat_grads,
    float4* __restrict__ gaussian_grads, float* __restrict__ sh_grads) {
  // SH coefficients in, then their gradients out, staged through the same LDS block.
  __shared__ __align__(16) float s_sh[kThreads * kShRest];
  __shared__ __align__(16) float s_grad[kThreads * kGrads];
  const uint32_t first = blockIdx.x * kThreads, i = first + threadIdx.x;
  const uint32_t block_n = min(static_cast<uint32_t>(kThreads), n - first);
  const bool active = i < n;
  const bool visible = active && splats[i * 3].w != 0.f;
  const int count = sh_count(sh_degree);
  const bool any = __syncthreads_or(visible);
  if (any) {
    const float* src_sh = sh + static_cast<size_t>(first) * kShRest;
    const float* src_grad = splat_grads + static_cast<size_t>(first) * kGrads;
    if (block_n == kThreads) {
      ShBlock c_sh;
      GradBlock c_grad;
      if (count > 0) c_sh.load(src_sh);
      c_grad.load(src_grad);
      if (count > 0) c_sh.store(s_sh);
      c_grad.store(s_grad);
    } else {
      if (count > 0) block_copy<kThreads>(s_sh, src_sh, block_n * kShRest);
      block_copy<kThreads>(s_grad, src_grad, block_n * kGrads);
    }
  }
  __syncthreads();
  // A thread only touches its own SH row, so gradients overwrite coefficients in place.
  float* row = s_sh + threadIdx.x * kShRest;
  float4 out0 = make_float4(0.f, 0.f, 0.f, 0.f), out1 = out0, out2 = out0, out3 = out0;
  if (visible) {
    float g[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float4 v = gaussians[i * 4 + q];
      g[4 * q] = v.x; g[4 * q + 1] = v.y; g[4 * q + 2] = v.z; g[4 * q + 3] = v.w;
    }
    float gr[kGrads];
#pragma unroll
    for (int v = 0; v < kGrads; v++) gr[v] = s_grad[threadIdx.x * kGrads + v];
    Projection p;
    project(g, cam, p);
    GaussianGrad o;
    gaussian_backward(g, p, cam, gr, row, count, o);
    // The SH gradients replace the coefficients in this thread's LDS row.
    for (int k = 0; k < static_cast<int>(kShRest) / 3; k++) {
      const bool used = k < count;
      row[3 * k] = used ? o.basis[k] * o.raw_grad[0] : 0.f;
      row[3 * k + 1] = used ? o.basis[k] * o.raw_grad[1] : 0.f;
      row[3 * k + 2] = used ? o.basis[k] * o.raw_grad[2] : 0.f;
    }
    out0 = o.out0;
    out1 = o.out1;
    out2 = o.out2;
    out3 = o.out3;
  } else if (active) {
    for (int k = 0; k < static_cast<int>(kShRest); k++) row[k] = 0.f;
  }
  if (active) {
    gaussian_grads[i * 4] = out0;
    gaussian_grads[i * 4 + 1] = out1;
    gaussian_grads[i * 4 + 2] = out2;
    gaussian_grads[i * 4 + 3] = out3;
  }
  __syncthreads();
  float* dst = sh_grads + static_cast<size_t>(first) * kShRest;
  if (block_n == kThreads) {
    ShBlock c;
    c.load(s_sh);
    c.store(dst);
  } else {
    block_copy<kThreads>(dst, s_sh, block_n * kShRest);
  }
}

}  // namespace

void preprocess(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, uint32_t n,
                const CameraArgs& cam, float* splats, uint32_t* tile_counts, uint32_t* depth_keys,
                uint32_t* depth_values, float* vis, uint32_t* rects) {
  if (n == 0) return;
  preprocess_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(gaussians), sh, sh_degree, n, cam, reinterpret_cast<float4*>(splats),
      tile_counts, depth_keys, depth_values, reinterpret_cast<float2*>(vis), reinterpret_cast<uint2*>(rects));
  DIO_LAUNCHED("preprocess");
}

void count_tiles(hipStream_t s, const float* splats, uint32_t n, uint32_t width, uint32_t height,
                 uint32_t* tile_counts, uint32_t* depth_keys, uint32_t* depth_values, uint32_t bin_shift,
                 uint32_t* rects) {
  if (n == 0) return;
  count_tiles_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(reinterpret_cast<const float4*>(splats), n,
                                                                 ceil_div(width, kTile), ceil_div(height, kTile),
                                                                 bin_shift, tile_counts, depth_keys, depth_values,
                                                                 reinterpret_cast<uint2*>(rects));
  DIO_LAUNCHED("count_tiles");
}

// The grid for an expected pair count: the hint (0: the capacity) with an eighth more
// and one more workgroup, at most the capacity's.
static uint32_t hinted_groups(uint32_t capacity, uint32_t pair_hint, uint32_t per_group) {
  const uint64_t want = pair_hint ? static_cast<uint64_t>(pair_hint) + pair_hint / 8 + per_group : capacity;
  return std::max(1u, ceil_div(static_cast<uint32_t>(std::min<uint64_t>(want, capacity)), per_group));
}

void duplicate(hipStream_t s, const float* splats, const uint32_t* order, const uint32_t* offsets,
               const uint32_t* tile_counts, uint32_t n, uint32_t width, uint32_t height, uint32_t capacity,
               const uint32_t* pair_total, uint32_t* pair_tiles, uint32_t* pair_splats, uint32_t pair_hint,
               uint32_t bin_shift) {
  if (n == 0 || capacity == 0) return;
  duplicate_kernel<false><<<hinted_groups(capacity, pair_hint, kDupPairs), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(splats), order, offsets, tile_counts, n, ceil_div(width, kTile),
      ceil_div(height, kTile), bin_shift, nullptr, 0, capacity, pair_total, pair_tiles, pair_splats);
  DIO_LAUNCHED("duplicate");
}

void duplicate_batch(hipStream_t s, const float* splats, const uint32_t* order, const uint32_t* offsets,
                     const uint32_t* tile_counts, const BatchArgs& b, uint32_t capacity, const uint32_t* pair_total,
                     uint32_t* pair_tiles, uint32_t* pair_splats) {
  const uint32_t n = b.view_count * b.per_view;
  if (n == 0 || capacity == 0) return;
  duplicate_kernel<true><<<ceil_div(capacity, kDupPairs), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(splats), order, offsets, tile_counts, n, 0, 0, 0, b.views, b.per_view, capacity,
      pair_total, pair_tiles, pair_splats);
  DIO_LAUNCHED("duplicate_batch");
}

void find_ranges(hipStream_t s, const uint32_t* pair_tiles, const uint32_t* pair_total, uint32_t capacity,
                 uint32_t tiles, uint32_t* ranges, uint32_t* order, uint32_t pair_hint, uint32_t* clear,
                 bool sparse) {
  if (tiles == 0) return;
  if (sparse) fill_u32(s, ranges, 0u, 2 * static_cast<size_t>(tiles));
  find_ranges_kernel<<<hinted_groups(capacity, pair_hint, kThreads), kThreads, 0, s>>>(
      pair_tiles, pair_total, capacity, tiles, reinterpret_cast<uint2*>(ranges), sparse);
  DIO_LAUNCHED("find_ranges");
  if (order) {
    tile_order_kernel<<<1, kOrderThreads, 0, s>>>(reinterpret_cast<const uint2*>(ranges), tiles, order, clear);
    DIO_LAUNCHED("tile_order");
  } else if (clear) {
    fill_u32(s, clear, 0u, 1 + kPartClasses);
  }
}

// Workgroups of `kernel` (256 threads, `lds` dynamic LDS) the current GPU holds at
// once, looked up once per (device, kernel).
template <class K>
uint32_t resident_groups(K kernel, size_t lds) {
  static std::mutex lock;
  static std::map<std::pair<int, const void*>, uint32_t> known;
  int device = 0;
  check_hip(hipGetDevice(&device), "hipGetDevice");
  const auto key = std::make_pair(device, reinterpret_cast<const void*>(kernel));
  std::lock_guard<std::mutex> guard(lock);
  auto it = known.find(key);
  if (it != known.end()) return it->second;
  int cus = 0, per_cu = 0;
  check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
  check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds), "hipOccupancy");
  const uint32_t groups = static_cast<uint32_t>(std::max(1, cus * per_cu));
  known[key] = groups;
  return groups;
}

void rasterize(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
               const uint32_t* order, uint32_t width, uint32_t height, const float* background, uint32_t* target,
               float* pixel_color, uint32_t* pixel_last, uint32_t splat_batch, uint32_t tile_size, uint32_t* work,
               const Segments* seg) {
  const uint32_t tx = ceil_div(width, kTile), ty = ceil_div(height, kTile);
  if (tx * ty == 0) return;
  if (seg && (tile_size != 16 || work)) throw Error(-1, "the segmented backward runs on 16-pixel tiles, one workgroup each");
  if (work) {
    if (tile_size != 16) throw Error(-1, "persistent workgroups rasterize 16-pixel tiles");
    auto persist = [&](auto kernel) {
      const uint32_t resident = resident_groups(kernel, 0);
      kernel<<<std::min(resident, tx * ty), 256, 0, s>>>(
          reinterpret_cast<const float4*>(splats), pair_splats, reinterpret_cast<const uint2*>(ranges), order, width,
          height, tx, background[0], background[1], background[2], target, reinterpret_cast<float4*>(pixel_color),
          pixel_last, tx * ty, work);
    };
    switch (splat_batch) {
      case 64: persist(rasterize_persistent_kernel<1>); break;
      case 128: persist(rasterize_persistent_kernel<2>); break;
      case 256: persist(rasterize_persistent_kernel<4>); break;
      default: throw Error(-1, "splat batch must be 64, 128 or 256");
    }
    DIO_LAUNCHED("rasterize_persistent");
    return;
  }
  const uint32_t bx = ceil_div(tx, 2u), by = ceil_div(ty, 2u);
  auto launch = [&](auto kernel, uint32_t groups, uint32_t threads, uint32_t row) {
    kernel<<<groups, threads, 0, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                      reinterpret_cast<const uint2*>(ranges), order, width, height, row, background[0],
                                      background[1], background[2], target, reinterpret_cast<float4*>(pixel_color),
                                      pixel_last);
  };
  auto tiles16 = [&](auto kernel) {
    kernel<<<tx * ty, 256, 0, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                   reinterpret_cast<const uint2*>(ranges), order, width, height, tx, background[0],
                                   background[1], background[2], target, reinterpret_cast<float4*>(pixel_color),
                                   pixel_last, seg ? *seg : Segments{});
  };
  const uint32_t key = tile_size * 1000 + splat_batch;
  switch (key) {
    case 16064: tiles16(rasterize_kernel<1>); break;
    case 16128: tiles16(rasterize_kernel<2>); break;
    case 16256: tiles16(rasterize_kernel<4>); break;
    case 8064: launch(rasterize_quad_kernel<1>, 4 * tx * ty, 64, tx); break;
    case 8128: launch(rasterize_quad_kernel<2>, 4 * tx * ty, 64, tx); break;
    case 8256: launch(rasterize_quad_kernel<4>, 4 * tx * ty, 64, tx); break;
    case 32064: launch(rasterize_bin_kernel<1>, bx * by, 1024, bx); break;
    case 32128: launch(rasterize_bin_kernel<2>, bx * by, 1024, bx); break;
    default: throw Error(-1, "tile size 8, 16 or 32 with splat batch 64, 128 or 256 (32: 64 or 128)");
  }
  DIO_LAUNCHED("rasterize");
}

// The 16-px backward runs fastest with at most 5 workgroups (20 waves) per CU:
// measured at 64 splats a round, 7 took 371 µs, 6 369 µs and 5 365.5 µs (its 72
// VGPRs would let 7 in). Dynamic LDS pads a workgroup to just over a sixth of the
// CU's 160 KB.
constexpr uint32_t kBackwardGroupsPerCu = 5;

// The segmented backward measured the same at 4 to 8 workgroups per CU (323-327 µs
// with 512-pair segments) and keeps the same pad.
template <int kPer, bool kFoot = false>
constexpr size_t backward_lds_pad() {
  constexpr size_t group = 160 * 1024 / (kBackwardGroupsPerCu + 1) + 1;
  constexpr size_t used = sizeof(raster::BackwardShared<kPer, 4, false, kFoot>);
  return used >= group ? 0 : group - used;
}

void rasterize_backward(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
                        const uint32_t* order, uint32_t width, uint32_t height, const float* background,
                        const float* pixel_color, const uint32_t* pixel_last, const float* pixel_grads,
                        float* splat_grads, uint32_t splat_batch, uint32_t tile_size, const float* pixel_loss,
                        float* footprint, uint32_t* work, const Segments* seg) {
  const uint32_t tx = ceil_div(width, kTile), ty = ceil_div(height, kTile);
  if (tx * ty == 0) return;
  if (seg) {
    if (tile_size != 16 || work) throw Error(-1, "the segmented backward runs on 16-pixel tiles, one workgroup each");
    if (seg->shift < kSegmentShiftMin || seg->shift > kSegmentShiftMax)
      throw Error(-1, "segment length must be 256, 512 or 1024 pairs");
    auto run = [&](auto kernel, size_t pad) {
      kernel<<<std::max(seg->grid, 1u), 256, pad, s>>>(
          reinterpret_cast<const float4*>(splats), pair_splats, reinterpret_cast<const uint2*>(ranges), *seg, width,
          height, tx, background[0], background[1], background[2], reinterpret_cast<const float4*>(pixel_color),
          pixel_last, reinterpret_cast<const float4*>(pixel_grads), splat_grads, pixel_loss, footprint);
    };
    const uint32_t key = splat_batch * 2 + (footprint ? 1 : 0);
    switch (key) {
      case 128: run(rasterize_backward_segment_kernel<1, false>, backward_lds_pad<1>()); break;
      case 256: run(rasterize_backward_segment_kernel<2, false>, backward_lds_pad<2>()); break;
      case 512: run(rasterize_backward_segment_kernel<4, false>, backward_lds_pad<4>()); break;
      case 129: run(rasterize_backward_segment_kernel<1, true>, backward_lds_pad<1, true>()); break;
      case 257: run(rasterize_backward_segment_kernel<2, true>, backward_lds_pad<2, true>()); break;
      case 513: run(rasterize_backward_segment_kernel<4, true>, backward_lds_pad<4, true>()); break;
      default: throw Error(-1, "splat batch must be 64, 128 or 256");
    }
    DIO_LAUNCHED("rasterize_backward_segment");
    return;
  }
  if (work && !footprint) {
    if (tile_size != 16) throw Error(-1, "persistent workgroups rasterize 16-pixel tiles");
    auto persist = [&](auto kernel, size_t pad) {
      const uint32_t resident = resident_groups(kernel, pad);
      kernel<<<std::min(resident, tx * ty), 256, pad, s>>>(
          reinterpret_cast<const float4*>(splats), pair_splats, reinterpret_cast<const uint2*>(ranges), order, width,
          height, tx, background[0], background[1], background[2], reinterpret_cast<const float4*>(pixel_color),
          pixel_last, reinterpret_cast<const float4*>(pixel_grads), splat_grads, tx * ty, work);
    };
    switch (splat_batch) {
      case 64: persist(rasterize_backward_persistent_kernel<1>, backward_lds_pad<1>()); break;
      case 128: persist(rasterize_backward_persistent_kernel<2>, backward_lds_pad<2>()); break;
      case 256: persist(rasterize_backward_persistent_kernel<4>, backward_lds_pad<4>()); break;
      default: throw Error(-1, "splat batch must be 64, 128 or 256");
    }
    DIO_LAUNCHED("rasterize_backward_persistent");
    return;
  }
  if (footprint) {
    if (tile_size != 16) throw Error(-1, "the footprint terms are gathered with 16-pixel tiles");
    auto foot = [&](auto kernel, size_t pad) {
      kernel<<<tx * ty, 256, pad, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                       reinterpret_cast<const uint2*>(ranges), order, width, height, tx,
                                       background[0], background[1], background[2],
                                       reinterpret_cast<const float4*>(pixel_color), pixel_last,
                                       reinterpret_cast<const float4*>(pixel_grads), splat_grads, pixel_loss,
                                       footprint);
    };
    switch (splat_batch) {
      case 64: foot(rasterize_backward_foot_kernel<1>, backward_lds_pad<1, true>()); break;
      case 128: foot(rasterize_backward_foot_kernel<2>, backward_lds_pad<2, true>()); break;
      case 256: foot(rasterize_backward_foot_kernel<4>, backward_lds_pad<4, true>()); break;
      default: throw Error(-1, "splat batch must be 64, 128 or 256");
    }
    DIO_LAUNCHED("rasterize_backward_foot");
    return;
  }
  const uint32_t bx = ceil_div(tx, 2u), by = ceil_div(ty, 2u);
  auto launch = [&](auto kernel, uint32_t groups, uint32_t threads, uint32_t row, size_t pad) {
    kernel<<<groups, threads, pad, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                        reinterpret_cast<const uint2*>(ranges), order, width, height, row,
                                        background[0], background[1], background[2],
                                        reinterpret_cast<const float4*>(pixel_color), pixel_last,
                                        reinterpret_cast<const float4*>(pixel_grads), splat_grads);
  };
  const uint32_t key = tile_size * 1000 + splat_batch;
  switch (key) {
    case 16064: launch(rasterize_backward_kernel<1>, tx * ty, 256, tx, backward_lds_pad<1>()); break;
    case 16128: launch(rasterize_backward_kernel<2>, tx * ty, 256, tx, backward_lds_pad<2>()); break;
    case 16256: launch(rasterize_backward_kernel<4>, tx * ty, 256, tx, backward_lds_pad<4>()); break;
    case 8064: launch(rasterize_backward_quad_kernel<1>, 4 * tx * ty, 64, tx, 0); break;
    case 8128: launch(rasterize_backward_quad_kernel<2>, 4 * tx * ty, 64, tx, 0); break;
    case 8256: launch(rasterize_backward_quad_kernel<4>, 4 * tx * ty, 64, tx, 0); break;
    case 32064: launch(rasterize_backward_bin_kernel<1>, bx * by, 1024, bx, 0); break;
    case 32128: launch(rasterize_backward_bin_kernel<2>, bx * by, 1024, bx, 0); break;
    default: throw Error(-1, "tile size 8, 16 or 32 with splat batch 64, 128 or 256 (32: 64 or 128)");
  }
  DIO_LAUNCHED("rasterize_backward");
}

void preprocess_backward(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, uint32_t n,
                         const CameraArgs& cam, const float* splats, const float* splat_grads,
                         float* gaussian_grads, float* sh_grads) {
  if (n == 0) return;
  preprocess_backward_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(gaussians), sh, sh_degree, n, cam, reinterpret_cast<const float4*>(splats),
      splat_grads, reinterpret_cast<float4*>(gaussian_grads), sh_grads);
  DIO_LAUNCHED("preprocess_backward");
}

}  // namespace dio
