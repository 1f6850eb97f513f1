// Each tile's pair list put in front-to-back order after the pairs were binned by
// tile, as a second way to reach the pair order of src/tileRasterizer.ts:127-192:
// the reference sorts the splats by depth (radixSort.wgsl over the depth keys),
// duplicates them in that order (tiles.wgsl:36-59) and sorts the pairs stably by
// tile, so a tile's list runs by depth, ties in splat order. Here the pairs are
// duplicated in splat order and sorted stably by tile, so each list arrives in
// splat order, and one workgroup per list sorts it stably by depth key. The
// result is the same list, without a global sort of every splat's depth.
//
// A list of at most 2,048 pairs is sorted in LDS: 4 waves hold 8 rows of 64 keys
// each (wave-major, so element order is (wave, row, lane)), and each pass over an
// 8-bit digit ranks the keys stably per wave with match masks (8 ballots per key),
// offsets the ranks by the earlier waves' and digits' counts and scatters through
// LDS. Passes over a digit that every key of the list shares are skipped (they
// would not move anything). A longer list runs the same passes streamed through
// global memory in 2,048-pair chunks, ping-ponging between the list and a scratch
// list of the same layout: per pass one sweep counts the digits, and a second
// sweep ranks each chunk in order and writes it behind the pairs already placed.
// Keys are never stored: each pass reads them again as depth_keys[splat].
#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kRows = 8;
constexpr uint32_t kChunk = kThreads * kRows;  // 2048
constexpr int kRadix = 256;
constexpr uint32_t kBuckets = 2048;
constexpr uint32_t kBucketsPerThread = kBuckets / kThreads;  // 8
constexpr uint32_t kBucketLimit = 16;

struct OrderLds {
  uint32_t key[kChunk];
  uint32_t value[kChunk];
  uint32_t wave_count[kWaves][kRadix];  // per wave and digit: count, then where the wave's run starts
  uint32_t chunk_start[kRadix];         // where each digit's run starts in the chunk
  uint32_t run_start[kRadix];           // long lists: where each digit's run continues in the output
  uint32_t bucket[kBuckets];             // short lists: pairs per bucket, then where each bucket starts
  uint32_t wave_totals[kWaves];
  uint32_t reduce[2][kWaves];
  uint32_t crowded;  // short lists: set when a bucket holds more than kBucketLimit pairs
};

__device__ inline uint32_t element(uint32_t rows, int r) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  return (static_cast<uint32_t>(wave) * rows + r) * kWave + lane;
}

// Exclusive prefix over the 256 threads of `x` (thread t holds digit t's value).
__device__ inline uint32_t block_exclusive(OrderLds& L, uint32_t x) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  uint32_t inclusive = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(inclusive, o, kWave);
    if (lane >= o) inclusive += y;
  }
  if (lane == kWave - 1) L.wave_totals[wave] = inclusive;
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) before += w < wave ? L.wave_totals[w] : 0u;
  return before + inclusive - x;
}

// Stable positions of one chunk's keys ordered by the digit at `shift`. `n` keys
// are valid, laid out in `rows` rows per wave. Returns each key's position in the
// chunk in pos[r]; leaves in L.chunk_start where each digit's run starts in the
// chunk and returns the chunk's count of digit threadIdx.x.
__device__ inline uint32_t rank_chunk(OrderLds& L, const uint32_t (&key)[kRows], uint32_t n, uint32_t rows,
                                      uint32_t shift, uint32_t (&pos)[kRows]) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint64_t lower_lanes = (1ull << lane) - 1;
#pragma unroll
  for (int w = 0; w < kWaves; w++) L.wave_count[w][tid] = 0;
  __syncthreads();
  uint32_t rank[kRows], digit[kRows];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (static_cast<uint32_t>(r) >= rows) break;
    const bool valid = element(rows, r) < n;
    const uint32_t d = (key[r] >> shift) & (kRadix - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t set = __ballot(bit);
      peers &= bit ? set : ~set;
    }
    const uint64_t before = peers & lower_lanes;
    const uint32_t so_far = valid ? L.wave_count[wave][d] : 0u;
    rank[r] = so_far + __popcll(before);
    digit[r] = d;
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0) L.wave_count[wave][d] = so_far + __popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  uint32_t count[kWaves], total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    count[w] = L.wave_count[w][tid];
    total += count[w];
  }
  const uint32_t start = block_exclusive(L, total);
  L.chunk_start[tid] = start;
  uint32_t run = start;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    L.wave_count[w][tid] = run;  // only thread tid reads or writes column tid here
    run += count[w];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (static_cast<uint32_t>(r) >= rows) break;
    pos[r] = L.wave_count[wave][digit[r]] + rank[r];
  }
  return total;
}

enum class Reduce { kOr, kMin, kMax };

// x combined over the workgroup.
template <Reduce kOp>
__device__ inline uint32_t block_reduce(OrderLds& L, uint32_t x) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  auto op = [](uint32_t a, uint32_t b) { return kOp == Reduce::kOr ? a | b : kOp == Reduce::kMin ? min(a, b) : max(a, b); };
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) x = op(x, __shfl_xor(x, o, kWave));
  if (lane == 0) L.reduce[0][wave] = x;
  __syncthreads();
  uint32_t all = L.reduce[0][0];
#pragma unroll
  for (int w = 1; w < kWaves; w++) all = op(all, L.reduce[0][w]);
  __syncthreads();  // L.reduce is written again by the next call
  return all;
}

// The workgroup's minimum of `low` and maximum of `high`, in one exchange.
__device__ inline void block_min_max(OrderLds& L, uint32_t& low, uint32_t& high) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    low = min(low, __shfl_xor(low, o, kWave));
    high = max(high, __shfl_xor(high, o, kWave));
  }
  if (lane == 0) {
    L.reduce[0][wave] = low;
    L.reduce[1][wave] = high;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    low = min(low, L.reduce[0][w]);
    high = max(high, L.reduce[1][w]);
  }
  __syncthreads();  // L.reduce is written again by the next call
}

// Loads chunk [first, first + n) of `list` (n <= kChunk) in the (wave, row, lane)
// layout, with each pair's depth key.
__device__ inline void load_chunk(const uint32_t* __restrict__ list, const uint32_t* __restrict__ depth_keys,
                                  uint32_t first, uint32_t n, uint32_t rows, uint32_t (&key)[kRows],
                                  uint32_t (&value)[kRows]) {
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const uint32_t e = element(rows, r);
    value[r] = static_cast<uint32_t>(r) < rows && e < n ? list[first + e] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const uint32_t e = element(rows, r);
    key[r] = static_cast<uint32_t>(r) < rows && e < n ? depth_keys[value[r]] : 0u;
  }
}

// Radix passes over the 8-bit digits where `diff` has bits, in LDS, on a list of at
// most kChunk pairs held in registers; writes the list back in place.
__device__ void radix_in_lds(OrderLds& L, uint32_t* __restrict__ list, uint32_t lo, uint32_t n, uint32_t rows,
                             uint32_t diff, uint32_t (&key)[kRows], uint32_t (&value)[kRows]) {
  uint32_t pos[kRows];
  // One LDS copy serves every pass: a pass reads its keys back before the next
  // pass's first barrier, and scatters only after it.
  for (uint32_t shift = 0; shift < 32; shift += 8) {
    if (((diff >> shift) & (kRadix - 1)) == 0) continue;
    rank_chunk(L, key, n, rows, shift, pos);
#pragma unroll
    for (int r = 0; r < kRows; r++) {
      if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
        L.key[pos[r]] = key[r];
        L.value[pos[r]] = value[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRows; r++) {
      const uint32_t e = element(rows, r);
      if (static_cast<uint32_t>(r) < rows && e < n) {
        key[r] = L.key[e];
        value[r] = L.value[e];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const uint32_t e = element(rows, r);
    if (static_cast<uint32_t>(r) < rows && e < n) list[lo + e] = value[r];
  }
}

// (key, splat) before (key', splat'): depth first, then splat order. The splats of
// one list are distinct, so this is a strict order.
__device__ inline bool before(uint32_t k, uint32_t v, uint32_t k2, uint32_t v2) {
  return k < k2 || (k == k2 && v < v2);
}

// A list of at most kChunk pairs, sorted in LDS and written back in place. The
// keys go to kBuckets buckets by a map that never decreases with the key (a float
// scale of key - min over the list's key range); each pair's place inside its
// bucket, whose pairs the LDS atomics put in no particular order, is the number of
// the bucket's pairs before it in (key, splat) order. A list arrives in splat
// order, so (key, splat) order is the stable depth order. A bucket of more than
// kBucketLimit pairs (depths packed closer than the buckets resolve, or many equal
// depths) sends the list through radix passes instead.
__device__ void order_in_lds(OrderLds& L, const uint32_t* __restrict__ depth_keys, uint32_t* __restrict__ list,
                             uint32_t lo, uint32_t n) {
  const int tid = threadIdx.x;
  const uint32_t rows = ceil_div(n, kThreads);
  uint32_t key[kRows], value[kRows];
  load_chunk(list, depth_keys, lo, n, rows, key, value);
  uint32_t low = ~0u, high = 0u;
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      low = min(low, key[r]);
      high = max(high, key[r]);
    }
  }
#pragma unroll
  for (uint32_t b = 0; b < kBucketsPerThread; b++) L.bucket[tid * kBucketsPerThread + b] = 0;
  if (tid == 0) L.crowded = 0;
  block_min_max(L, low, high);  // its barriers also order the clearing above
  if (low == high) return;  // every key the same: the list is already in order
  const float scale = static_cast<float>(kBuckets) / (static_cast<float>(high - low) + 1.f);
  uint32_t bucket[kRows], slot[kRows];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      bucket[r] = min(kBuckets - 1, static_cast<uint32_t>(static_cast<float>(key[r] - low) * scale));
      slot[r] = atomicAdd(&L.bucket[bucket[r]], 1u);
    }
  }
  __syncthreads();
  uint32_t count[kBucketsPerThread], mine = 0, largest = 0;
#pragma unroll
  for (uint32_t b = 0; b < kBucketsPerThread; b++) {
    count[b] = L.bucket[tid * kBucketsPerThread + b];
    mine += count[b];
    largest = max(largest, count[b]);
  }
  if (largest > kBucketLimit) L.crowded = 1;
  uint32_t run = block_exclusive(L, mine);  // its barrier also publishes L.crowded
  if (L.crowded) {
    uint32_t diff = 0;  // the key bits that differ somewhere in the list
#pragma unroll
    for (int r = 0; r < kRows; r++)
      if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) diff |= key[r] ^ low;
    radix_in_lds(L, list, lo, n, rows, block_reduce<Reduce::kOr>(L, diff), key, value);
    return;
  }
#pragma unroll
  for (uint32_t b = 0; b < kBucketsPerThread; b++) {
    L.bucket[tid * kBucketsPerThread + b] = run;  // each thread rewrites only its own buckets
    run += count[b];
  }
  __syncthreads();
  uint32_t start[kRows], end[kRows];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      start[r] = L.bucket[bucket[r]];
      end[r] = bucket[r] + 1 < kBuckets ? L.bucket[bucket[r] + 1] : n;
      L.key[start[r] + slot[r]] = key[r];
      L.value[start[r] + slot[r]] = value[r];
    }
  }
  __syncthreads();
  // Each pair's place in its bucket: how many of the bucket's pairs come before it.
  // The bucket starts are no longer read, so their words take the ordered list.
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      uint32_t place = start[r];
      for (uint32_t j = start[r]; j < end[r]; j++) place += before(L.key[j], L.value[j], key[r], value[r]) ? 1u : 0u;
      L.bucket[place] = value[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const uint32_t e = element(rows, r);
    if (static_cast<uint32_t>(r) < rows && e < n) list[lo + e] = L.bucket[e];
  }
}

// A list longer than kChunk: the same passes streamed through `list` and `scratch`
// (the same pair positions of another pair-sized buffer).
__device__ void order_streamed(OrderLds& L, const uint32_t* __restrict__ depth_keys, uint32_t* list,
                               uint32_t* scratch, uint32_t lo, uint32_t n) {
  const int tid = threadIdx.x;
  uint32_t key[kRows], value[kRows], pos[kRows];
  const uint32_t key0 = depth_keys[list[lo]];
  uint32_t diff = 0;
  for (uint32_t c = 0; c < n; c += kChunk) {
    const uint32_t m = min(kChunk, n - c);
    load_chunk(list, depth_keys, lo + c, m, kRows, key, value);
#pragma unroll
    for (int r = 0; r < kRows; r++)
      if (element(kRows, r) < m) diff |= key[r] ^ key0;
  }
  diff = block_reduce<Reduce::kOr>(L, diff);
  uint32_t* src = list;
  uint32_t* dst = scratch;
  for (uint32_t shift = 0; shift < 32; shift += 8) {
    if (((diff >> shift) & (kRadix - 1)) == 0) continue;
    // Sweep 1: where each digit's run starts in the output.
    L.run_start[tid] = 0;
    __syncthreads();
    for (uint32_t c = 0; c < n; c += kChunk) {
      const uint32_t m = min(kChunk, n - c);
      load_chunk(src, depth_keys, lo + c, m, kRows, key, value);
#pragma unroll
      for (int r = 0; r < kRows; r++)
        if (element(kRows, r) < m) atomicAdd(&L.run_start[(key[r] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    const uint32_t start = block_exclusive(L, L.run_start[tid]);
    __syncthreads();
    L.run_start[tid] = start;
    // Sweep 2: chunks in order, each digit's pairs behind those placed before.
    for (uint32_t c = 0; c < n; c += kChunk) {
      const uint32_t m = min(kChunk, n - c);
      load_chunk(src, depth_keys, lo + c, m, kRows, key, value);
      const uint32_t total = rank_chunk(L, key, m, kRows, shift, pos);
#pragma unroll
      for (int r = 0; r < kRows; r++) {
        if (element(kRows, r) < m) {
          const uint32_t d = (key[r] >> shift) & (kRadix - 1);
          dst[lo + L.run_start[d] + pos[r] - L.chunk_start[d]] = value[r];
        }
      }
      __syncthreads();
      L.run_start[tid] += total;
    }
    // The workgroup's own stores are seen by its later loads after a barrier (all
    // its waves share one CU's L1, which stores write through).
    __syncthreads();
    uint32_t* t = src;
    src = dst;
    dst = t;
  }
  if (src != list) {
    for (uint32_t k = tid; k < n; k += kThreads) list[lo + k] = src[lo + k];
  }
}

__global__ __launch_bounds__(kThreads) void depth_order_kernel(const uint32_t* __restrict__ depth_keys,
                                                                const uint2* __restrict__ ranges,
                                                                const uint32_t* __restrict__ tile_order,
                                                                uint32_t* list, uint32_t* scratch) {
  __shared__ OrderLds L;
  const uint32_t tile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  const uint2 range = ranges[tile];
  const uint32_t n = range.y - range.x;
  if (n <= 1) return;
  if (n <= kChunk)
    order_in_lds(L, depth_keys, list, range.x, n);
  else
    order_streamed(L, depth_keys, list, scratch, range.x, n);
}

// A frame whose pairs overflow the capacity keeps, as the reference does
// (tiles.wgsl:47-58), the pair slots below the capacity in depth order: splats by
// (depth key, splat), each splat's tiles row by row. With pairs written in splat
// order that is a cut in (key, splat) order: every splat before the cut splat keeps
// all its pairs, the cut splat its first `keep`, the rest none. One workgroup finds
// the cut (the key 8 bits at a time, weighting each splat by its tile count, then
// the splat among those of that key), rewrites each splat's tile count to the pairs
// it keeps and rescans the slot offsets, so the duplicate writes exactly the kept
// pairs. Frames that fit return at once.
constexpr int kCutThreads = 1024;
constexpr int kCutWaves = kCutThreads / kWave;
constexpr int kCutItems = 4;  // consecutive splats per thread per sweep step
constexpr uint32_t kCutStep = kCutThreads * kCutItems;

struct CutLds {
  uint32_t hist[kRadix];
  uint32_t wave_sum[kCutWaves];
  uint32_t fixed, target, cut, keep;
};

// Exclusive prefix of x over the 1024 threads; *total gets the sum.
__device__ inline uint32_t cut_exclusive(CutLds& L, uint32_t x, uint32_t* total) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  uint32_t inclusive = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(inclusive, o, kWave);
    if (lane >= o) inclusive += y;
  }
  if (lane == kWave - 1) L.wave_sum[wave] = inclusive;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kCutWaves; w++) {
    const uint32_t v = L.wave_sum[w];
    before += w < wave ? v : 0u;
    all += v;
  }
  __syncthreads();  // wave_sum is written again by the next call
  *total = all;
  return before + inclusive - x;
}

__global__ __launch_bounds__(kCutThreads) void pair_cut_kernel(const uint32_t* __restrict__ depth_keys,
                                                                uint32_t* __restrict__ counts,
                                                                uint32_t* __restrict__ offsets, uint32_t n,
                                                                const uint32_t* __restrict__ pair_total,
                                                                uint32_t capacity) {
  if (*pair_total <= capacity) return;
  __shared__ CutLds L;
  const int tid = threadIdx.x;
  // 1. The cut splat's key, most significant digit first: the digit whose splats hold
  //    the target-th of the remaining slots.
  uint32_t fixed = 0, fixed_mask = 0, target = capacity;
  for (int shift = 24; shift >= 0 && capacity; shift -= 8) {
    if (tid < kRadix) L.hist[tid] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kCutStep) {
      uint32_t c[kCutItems], k[kCutItems];
#pragma unroll
      for (int i = 0; i < kCutItems; i++) {
        const uint32_t j = base + tid * kCutItems + i;
        c[i] = j < n ? counts[j] : 0u;
        k[i] = j < n ? depth_keys[j] : 0u;
      }
#pragma unroll
      for (int i = 0; i < kCutItems; i++)
        if (c[i] && (k[i] & fixed_mask) == fixed) atomicAdd(&L.hist[(k[i] >> shift) & (kRadix - 1)], c[i]);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t before = 0, d = 0;
      for (; d + 1 < kRadix && before + L.hist[d] < target; d++) before += L.hist[d];
      L.fixed = fixed | (d << shift);
      L.target = target - before;
    }
    __syncthreads();
    fixed = L.fixed;
    target = L.target;
    fixed_mask |= static_cast<uint32_t>(kRadix - 1) << shift;
  }
  // 2. The cut splat among the splats of that key, in splat order, and how many of
  //    its pairs it keeps.
  if (tid == 0) {
    L.cut = n;
    L.keep = 0;
  }
  __syncthreads();
  uint32_t running = 0;
  for (uint32_t base = 0; base < n && capacity && running < target; base += kCutStep) {
    uint32_t w[kCutItems], mine = 0;
#pragma unroll
    for (int i = 0; i < kCutItems; i++) {
      const uint32_t j = base + tid * kCutItems + i;
      w[i] = j < n && depth_keys[j] == fixed ? counts[j] : 0u;
      mine += w[i];
    }
    uint32_t step_total;
    uint32_t before = running + cut_exclusive(L, mine, &step_total);
#pragma unroll
    for (int i = 0; i < kCutItems; i++) {
      if (w[i] && before < target && target <= before + w[i]) {
        L.cut = base + tid * kCutItems + i;
        L.keep = target - before;
      }
      before += w[i];
    }
    running += step_total;
  }
  __syncthreads();
  const uint32_t cut = L.cut, keep = L.keep;
  // 3. Kept pairs per splat, and the slot offsets over them.
  running = 0;
  for (uint32_t base = 0; base < n; base += kCutStep) {
    uint32_t kept[kCutItems], mine = 0;
#pragma unroll
    for (int i = 0; i < kCutItems; i++) {
      const uint32_t j = base + tid * kCutItems + i;
      const uint32_t c = j < n ? counts[j] : 0u, k = j < n ? depth_keys[j] : 0u;
      kept[i] = c == 0 || capacity == 0 ? 0u : k < fixed || (k == fixed && j < cut) ? c : j == cut ? keep : 0u;
      mine += kept[i];
    }
    uint32_t step_total;
    uint32_t at = running + cut_exclusive(L, mine, &step_total);
#pragma unroll
    for (int i = 0; i < kCutItems; i++) {
      const uint32_t j = base + tid * kCutItems + i;
      if (j < n) {
        counts[j] = kept[i];
        offsets[j] = at;
      }
      at += kept[i];
    }
    running += step_total;
  }
}

}  // namespace

void pair_cut(hipStream_t s, const uint32_t* depth_keys, uint32_t* tile_counts, uint32_t* offsets, uint32_t n,
              const uint32_t* pair_total, uint32_t capacity) {
  if (n == 0) return;
  pair_cut_kernel<<<1, kCutThreads, 0, s>>>(depth_keys, tile_counts, offsets, n, pair_total, capacity);
  DIO_LAUNCHED("pair_cut");
}

void depth_order(hipStream_t s, const uint32_t* depth_keys, const uint32_t* ranges, const uint32_t* tile_order,
                 uint32_t tiles, uint32_t* list, uint32_t* scratch) {
  if (tiles == 0) return;
  depth_order_kernel<<<tiles, kThreads, 0, s>>>(depth_keys, reinterpret_cast<const uint2*>(ranges), tile_order, list,
                                                scratch);
  DIO_LAUNCHED("depth_order");
}

}  // namespace dio
