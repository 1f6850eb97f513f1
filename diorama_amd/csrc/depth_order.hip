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
#include <type_traits>

#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kRows = 8;
constexpr uint32_t kChunk = kThreads * kRows;  // 2048
constexpr int kRadix = 256;
constexpr uint32_t kBuckets = 2048;  // for a list of kChunk pairs; in proportion for other sizes
constexpr uint32_t kBucketLimit = 32;
// The long-list kernel sorts lists of up to kLongChunk pairs in LDS.
constexpr int kLongRows = 16;
constexpr uint32_t kLongChunk = kThreads * kLongRows;  // 4096
template <int R>
constexpr uint32_t buckets_for = kBuckets * R / kRows;

template <int R>
struct OrderLds {
  uint2 pair[kThreads * R];  // (depth key, splat): one 8-byte LDS access a pair
  uint32_t wave_count[kWaves][kRadix];  // per wave and digit: count, then where the wave's run starts
  uint32_t chunk_start[kRadix];         // where each digit's run starts in the chunk
  uint32_t run_start[kRadix];           // long lists: where each digit's run continues in the output
  uint32_t bucket[buckets_for<R>];       // in-LDS lists: pairs per bucket, then where each bucket starts
  uint32_t wave_totals[kWaves];
  uint32_t reduce[2][kWaves];
  uint32_t crowded;  // short lists: set when a bucket holds more than kBucketLimit pairs
};

__device__ inline uint32_t element(uint32_t rows, int r) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  return (static_cast<uint32_t>(wave) * rows + r) * kWave + lane;
}

// Exclusive prefix over the 256 threads of `x` (thread t holds digit t's value).
template <int R>
__device__ inline uint32_t block_exclusive(OrderLds<R>& L, uint32_t x) {
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
template <int R>
__device__ inline uint32_t rank_chunk(OrderLds<R>& L, const uint32_t (&key)[R], uint32_t n, uint32_t rows,
                                      uint32_t shift, uint32_t (&pos)[R]) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint64_t lower_lanes = (1ull << lane) - 1;
#pragma unroll
  for (int w = 0; w < kWaves; w++) L.wave_count[w][tid] = 0;
  __syncthreads();
  uint32_t rank[R], digit[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
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
  for (int r = 0; r < R; r++) {
    if (static_cast<uint32_t>(r) >= rows) break;
    pos[r] = L.wave_count[wave][digit[r]] + rank[r];
  }
  return total;
}

enum class Reduce { kOr, kMin, kMax };

// x combined over the workgroup.
template <Reduce kOp, int R>
__device__ inline uint32_t block_reduce(OrderLds<R>& L, uint32_t x) {
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
template <int R>
__device__ inline void block_min_max(OrderLds<R>& L, uint32_t& low, uint32_t& high) {
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

// Loads chunk [first, first + n) of a list (n <= kThreads x R) in the (wave, row, lane)
// layout with each pair's depth key: from `keyed` ((key, splat) pairs) when the
// pairs carry their keys, else from `list` (splats) and depth_keys[splat].
template <int R>
__device__ inline void load_chunk(const uint32_t* __restrict__ list, const uint2* __restrict__ keyed,
                                  const uint32_t* __restrict__ depth_keys, uint32_t first, uint32_t n, uint32_t rows,
                                  uint32_t (&key)[R], uint32_t (&value)[R]) {
  if (keyed) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t e = element(rows, r);
      const uint2 p = static_cast<uint32_t>(r) < rows && e < n ? keyed[first + e] : make_uint2(0u, 0u);
      key[r] = p.x;
      value[r] = p.y;
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = element(rows, r);
    value[r] = static_cast<uint32_t>(r) < rows && e < n ? list[first + e] : 0u;
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = element(rows, r);
    key[r] = static_cast<uint32_t>(r) < rows && e < n ? depth_keys[value[r]] : 0u;
  }
}

// Radix passes in LDS on a list of at most kThreads x R pairs held in registers, over the
// 8-bit digits where `value_diff` has bits (splat indices, first) and then where
// `diff` has bits (depth keys): stable passes, so the list ends in (key, splat)
// order whatever order it arrived in. Writes the list back in place.
template <int R>
__device__ void radix_in_lds(OrderLds<R>& L, uint32_t* __restrict__ list, uint32_t lo, uint32_t n, uint32_t rows,
                             uint32_t value_diff, uint32_t diff, uint32_t (&key)[R], uint32_t (&value)[R]) {
  uint32_t pos[R];
  // One LDS copy serves every pass: a pass reads its keys back before the next
  // pass's first barrier, and scatters only after it.
  for (uint32_t p = 0; p < 8; p++) {
    const bool by_value = p < 4;
    const uint32_t shift = (p & 3u) * 8u;
    if ((((by_value ? value_diff : diff) >> shift) & (kRadix - 1)) == 0) continue;
    uint32_t digits[R];
#pragma unroll
    for (int r = 0; r < R; r++) digits[r] = by_value ? value[r] : key[r];
    rank_chunk(L, digits, n, rows, shift, pos);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
        L.pair[pos[r]] = make_uint2(key[r], value[r]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint32_t e = element(rows, r);
      if (static_cast<uint32_t>(r) < rows && e < n) {
        const uint2 p = L.pair[e];
        key[r] = p.x;
        value[r] = p.y;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = element(rows, r);
    if (static_cast<uint32_t>(r) < rows && e < n) list[lo + e] = value[r];
  }
}

// The splat-index bits that differ somewhere in a list held in registers (0 for a
// list that arrives in splat order, which needs no passes over them).
template <int R>
__device__ inline uint32_t value_bits(OrderLds<R>& L, bool ordered, uint32_t first, uint32_t n, uint32_t rows,
                                      const uint32_t (&value)[R]) {
  if (ordered) return 0u;
  uint32_t diff = 0;
#pragma unroll
  for (int r = 0; r < R; r++)
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) diff |= value[r] ^ first;
  return block_reduce<Reduce::kOr>(L, diff);
}

// (key, splat) before (key', splat'): depth first, then splat order. The splats of
// one list are distinct, so this is a strict order.
__device__ inline bool before(uint32_t k, uint32_t v, uint32_t k2, uint32_t v2) {
  return k < k2 || (k == k2 && v < v2);
}

// A list of at most kThreads x R pairs, sorted in LDS and written back in place. The
// keys go to buckets_for<R> buckets by a map that never decreases with the key (a float
// scale of key - min over the list's key range); each pair's place inside its
// bucket, whose pairs the LDS atomics put in no particular order, is the number of
// the bucket's pairs before it in (key, splat) order, the stable depth order of a
// list in splat order, whatever order the list arrives in. Returns false, writing
// nothing, for a list this cannot order: a bucket of more than kBucketLimit pairs
// (depths packed closer than the buckets resolve, or many equal depths) or a list
// of one depth that does not arrive in splat order; order_by_digits takes those.
// The list comes from `keyed` when given (never `ordered` then) and always leaves
// in `list`.
template <int R>
__device__ bool order_in_lds(OrderLds<R>& L, const uint32_t* __restrict__ depth_keys, uint32_t* __restrict__ list,
                             const uint2* __restrict__ keyed, uint32_t lo, uint32_t n, bool ordered) {
  constexpr uint32_t kPer = buckets_for<R> / kThreads;  // buckets a thread counts
  const int tid = threadIdx.x;
  const uint32_t rows = ceil_div(n, kThreads);
  uint32_t key[R], value[R];
  load_chunk(list, keyed, depth_keys, lo, n, rows, key, value);
  uint32_t low = ~0u, high = 0u;
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      low = min(low, key[r]);
      high = max(high, key[r]);
    }
  }
#pragma unroll
  for (uint32_t b = 0; b < kPer; b++) L.bucket[tid * kPer + b] = 0;
  if (tid == 0) L.crowded = 0;
  block_min_max(L, low, high);  // its barriers also order the clearing above
  if (low == high) return ordered;  // every key the same: splat order is the order
  const float scale = static_cast<float>(buckets_for<R>) / (static_cast<float>(high - low) + 1.f);
  uint32_t bucket[R], slot[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      bucket[r] = min(buckets_for<R> - 1, static_cast<uint32_t>(static_cast<float>(key[r] - low) * scale));
      slot[r] = atomicAdd(&L.bucket[bucket[r]], 1u);
    }
  }
  __syncthreads();
  uint32_t count[kPer], mine = 0, largest = 0;
#pragma unroll
  for (uint32_t b = 0; b < kPer; b++) {
    count[b] = L.bucket[tid * kPer + b];
    mine += count[b];
    largest = max(largest, count[b]);
  }
  if (largest > kBucketLimit) L.crowded = 1;
  uint32_t run = block_exclusive(L, mine);  // its barrier also publishes L.crowded
  if (L.crowded) return false;
#pragma unroll
  for (uint32_t b = 0; b < kPer; b++) {
    L.bucket[tid * kPer + b] = run;  // each thread rewrites only its own buckets
    run += count[b];
  }
  __syncthreads();
  uint32_t start[R], end[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      start[r] = L.bucket[bucket[r]];
      end[r] = bucket[r] + 1 < buckets_for<R> ? L.bucket[bucket[r] + 1] : n;
      L.pair[start[r] + slot[r]] = make_uint2(key[r], value[r]);
    }
  }
  __syncthreads();
  // Each pair's place in its bucket: how many of the bucket's pairs come before it.
  // The bucket starts are no longer read, so their words take the ordered list.
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) {
      uint32_t place = start[r];
      for (uint32_t j = start[r]; j < end[r]; j++) {
        const uint2 p = L.pair[j];
        place += before(p.x, p.y, key[r], value[r]) ? 1u : 0u;
      }
      L.bucket[place] = value[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = element(rows, r);
    if (static_cast<uint32_t>(r) < rows && e < n) list[lo + e] = L.bucket[e];
  }
  return true;
}

// A list of at most kThreads x R pairs that order_in_lds left: radix passes in LDS over
// the splat-index digits that differ in it (none if `ordered`), then the depth-key
// digits that differ.
template <int R>
__device__ void order_by_digits(OrderLds<R>& L, const uint32_t* __restrict__ depth_keys, uint32_t* __restrict__ list,
                                const uint2* __restrict__ keyed, uint32_t lo, uint32_t n, bool ordered) {
  const uint32_t rows = ceil_div(n, kThreads);
  uint32_t key[R], value[R];
  load_chunk(list, keyed, depth_keys, lo, n, rows, key, value);
  // One pair of the list: the reference its bits differ from.
  const uint32_t first = keyed ? keyed[lo].y : list[lo];
  const uint32_t key0 = keyed ? keyed[lo].x : depth_keys[first];
  uint32_t diff = 0;
#pragma unroll
  for (int r = 0; r < R; r++)
    if (static_cast<uint32_t>(r) < rows && element(rows, r) < n) diff |= key[r] ^ key0;
  diff = block_reduce<Reduce::kOr>(L, diff);
  radix_in_lds(L, list, lo, n, rows, value_bits(L, ordered, first, n, rows, value), diff, key, value);
}

// A list longer than kThreads x R: the same passes streamed through the list and a
// scratch list (the same pair positions of another pair-sized buffer), over the
// splat indices first unless the list is `ordered`. kKeyed: the list is `keyed`
// ((key, splat) pairs, ping-ponged with an 8-byte scratch) and its splats leave in
// `list`; else it is `list` itself (splats, ping-ponged with a 4-byte scratch).
template <int R, bool kKeyed>
__device__ void order_streamed(OrderLds<R>& L, const uint32_t* __restrict__ depth_keys, uint32_t* list, uint2* keyed,
                               void* scratch, uint32_t lo, uint32_t n, bool ordered) {
  using Elem = typename std::conditional<kKeyed, uint2, uint32_t>::type;
  constexpr uint32_t kN = kThreads * R;  // pairs a chunk
  const int tid = threadIdx.x;
  uint32_t key[R], value[R], pos[R];
  Elem* src = kKeyed ? reinterpret_cast<Elem*>(keyed) : reinterpret_cast<Elem*>(list);
  Elem* dst = static_cast<Elem*>(scratch);
  auto load = [&](const Elem* from, uint32_t first, uint32_t m) {
    load_chunk(kKeyed ? nullptr : reinterpret_cast<const uint32_t*>(from),
               kKeyed ? reinterpret_cast<const uint2*>(from) : nullptr, depth_keys, first, m, R, key, value);
  };
  uint32_t value0, key0;
  if (kKeyed) {
    const uint2 p = keyed[lo];
    key0 = p.x;
    value0 = p.y;
  } else {
    value0 = list[lo];
    key0 = depth_keys[value0];
  }
  uint32_t diff = 0, value_diff = 0;
  for (uint32_t c = 0; c < n; c += kN) {
    const uint32_t m = min(kN, n - c);
    load(src, lo + c, m);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (element(R, r) < m) {
        diff |= key[r] ^ key0;
        value_diff |= value[r] ^ value0;
      }
    }
  }
  diff = block_reduce<Reduce::kOr>(L, diff);
  value_diff = ordered ? 0u : block_reduce<Reduce::kOr>(L, value_diff);
  for (uint32_t p = 0; p < 8; p++) {
    const bool by_value = p < 4;
    const uint32_t shift = (p & 3u) * 8u;
    if ((((by_value ? value_diff : diff) >> shift) & (kRadix - 1)) == 0) continue;
    // Sweep 1: where each digit's run starts in the output.
    L.run_start[tid] = 0;
    __syncthreads();
    for (uint32_t c = 0; c < n; c += kN) {
      const uint32_t m = min(kN, n - c);
      load(src, lo + c, m);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (element(R, r) < m)
          atomicAdd(&L.run_start[((by_value ? value[r] : key[r]) >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    const uint32_t start = block_exclusive(L, L.run_start[tid]);
    __syncthreads();
    L.run_start[tid] = start;
    // Sweep 2: chunks in order, each digit's pairs behind those placed before.
    for (uint32_t c = 0; c < n; c += kN) {
      const uint32_t m = min(kN, n - c);
      load(src, lo + c, m);
      uint32_t digits[R];
#pragma unroll
      for (int r = 0; r < R; r++) digits[r] = by_value ? value[r] : key[r];
      const uint32_t total = rank_chunk(L, digits, m, R, shift, pos);
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (element(R, r) < m) {
          const uint32_t d = (digits[r] >> shift) & (kRadix - 1);
          const uint32_t at = lo + L.run_start[d] + pos[r] - L.chunk_start[d];
          if constexpr (kKeyed)
            dst[at] = make_uint2(key[r], value[r]);
          else
            dst[at] = value[r];
        }
      }
      __syncthreads();
      L.run_start[tid] += total;
    }
    // The workgroup's own stores are seen by its later loads after a barrier (all
    // its waves share one CU's L1, which stores write through).
    __syncthreads();
    Elem* t = src;
    src = dst;
    dst = t;
  }
  if constexpr (kKeyed) {
    for (uint32_t k = tid; k < n; k += kThreads) list[lo + k] = src[lo + k].y;
  } else {
    if (src != list)
      for (uint32_t k = tid; k < n; k += kThreads) list[lo + k] = src[lo + k];
  }
}

// `ordered`: every list arrives in splat order (pairs written in splat order and
// sorted stably by tile); otherwise the lists are in any order (binned by count).
// keyed (nullable): the lists as (depth key, splat) pairs, at the same positions,
// sorted into `list`. One workgroup a tile sorts the lists of at most kChunk pairs
// in LDS and queues those its buckets cannot order (`queued` counts the tiles in
// `entries`); depth_order_long_kernel takes the queued lists and the longer ones,
// whose registers would otherwise limit this kernel's occupancy.
__global__ __launch_bounds__(kThreads) void depth_order_kernel(const uint32_t* __restrict__ depth_keys,
                                                                const uint2* __restrict__ ranges,
                                                                const uint32_t* __restrict__ tile_order,
                                                                uint32_t* list, const uint2* keyed, bool ordered,
                                                                uint32_t* __restrict__ queued,
                                                                uint32_t* __restrict__ entries) {
  __shared__ OrderLds<kRows> L;
  const uint32_t tile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  const uint2 range = ranges[tile];
  const uint32_t n = range.y - range.x;
  if (n == 0 || n > kChunk) return;
  if (n == 1) {
    if (keyed && threadIdx.x == 0) list[range.x] = keyed[range.x].y;
    return;
  }
  if (!order_in_lds(L, depth_keys, list, keyed, range.x, n, ordered) && threadIdx.x == 0) {
    const uint32_t k = atomicAdd(queued, 1u);
    if (k < gridDim.x) entries[k] = tile;  // one entry a tile at most: always true
  }
}

// The queued lists (radix passes in LDS), then the lists longer than kChunk: in
// LDS up to kLongChunk pairs, streamed through `scratch` (8 bytes a pair when keyed,
// else 4) beyond. With `tile_order` (tiles by falling pair count in steps of
// kOrderStep, which divides kChunk) a workgroup walks the order from its own index
// and stops at its first list shorter than kChunk: every later list is shorter too,
// so a grid of a few workgroups covers the long lists. Without it, one workgroup a
// tile. Workgroup 0 clears `clear`, the count of the queue the next call fills.
__global__ __launch_bounds__(kThreads) void depth_order_long_kernel(const uint32_t* __restrict__ depth_keys,
                                                                     const uint2* __restrict__ ranges,
                                                                     const uint32_t* __restrict__ tile_order,
                                                                     uint32_t tiles, uint32_t* list, uint2* keyed,
                                                                     void* scratch, bool ordered,
                                                                     const uint32_t* __restrict__ queued,
                                                                     const uint32_t* __restrict__ entries,
                                                                     uint32_t* __restrict__ clear) {
  static_assert(kChunk % kOrderStep == 0, "the early stop needs whole order steps below kChunk");
  __shared__ OrderLds<kLongRows> L;
  if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0;
  const uint32_t count = min(*queued, tiles);
  for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
    const uint32_t tile = entries[k];
    if (tile >= tiles) continue;
    const uint2 range = ranges[tile];
    order_by_digits(L, depth_keys, list, keyed, range.x, range.y - range.x, ordered);
    __syncthreads();  // the next list reuses L
  }
  for (uint32_t i = blockIdx.x; i < tiles; i += gridDim.x) {
    const uint2 range = ranges[tile_order ? tile_order[i] : i];
    const uint32_t n = range.y - range.x;
    if (n < kChunk && tile_order) break;
    if (n <= kChunk) continue;
    if (n <= kLongChunk) {
      if (!order_in_lds(L, depth_keys, list, keyed, range.x, n, ordered)) {
        __syncthreads();  // L is read again below
        order_by_digits(L, depth_keys, list, keyed, range.x, n, ordered);
      }
    } else if (keyed) {
      order_streamed<kLongRows, true>(L, depth_keys, list, keyed, scratch, range.x, n, ordered);
    } else {
      order_streamed<kLongRows, false>(L, depth_keys, list, nullptr, scratch, range.x, n, ordered);
    }
    __syncthreads();  // the next list reuses L
  }
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

// The cut of a frame whose pairs overflow `capacity`: the depth key of the splat
// holding the capacity-th pair slot in (key, splat) order (`fixed`), that splat
// (`cut`) and how many of its pairs it keeps (`keep`). Every thread of the
// workgroup gets the same three.
__device__ void find_cut(CutLds& L, const uint32_t* __restrict__ depth_keys, const uint32_t* __restrict__ counts,
                         uint32_t n, uint32_t capacity, uint32_t& fixed_out, uint32_t& cut_out, uint32_t& keep_out) {
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
  fixed_out = fixed;
  cut_out = L.cut;
  keep_out = L.keep;
}

// The pairs splat j keeps under the cut: all before the cut in (key, splat) order,
// `keep` at the cut, none after (none at all at capacity 0).
__device__ inline uint32_t kept_pairs(uint32_t c, uint32_t k, uint32_t j, uint32_t capacity, uint32_t fixed,
                                      uint32_t cut, uint32_t keep) {
  return c == 0 || capacity == 0 ? 0u : k < fixed || (k == fixed && j < cut) ? c : j == cut ? keep : 0u;
}

__global__ __launch_bounds__(kCutThreads) void pair_cut_kernel(const uint32_t* __restrict__ depth_keys,
                                                                uint32_t* __restrict__ counts,
                                                                uint32_t* __restrict__ offsets, uint32_t n,
                                                                const uint32_t* __restrict__ pair_total,
                                                                uint32_t capacity) {
  if (*pair_total <= capacity) return;
  __shared__ CutLds L;
  const int tid = threadIdx.x;
  uint32_t fixed, cut, keep;
  find_cut(L, depth_keys, counts, n, capacity, fixed, cut, keep);
  // 3. Kept pairs per splat, and the slot offsets over them.
  uint32_t running = 0;
  for (uint32_t base = 0; base < n; base += kCutStep) {
    uint32_t kept[kCutItems], mine = 0;
#pragma unroll
    for (int i = 0; i < kCutItems; i++) {
      const uint32_t j = base + tid * kCutItems + i;
      kept[i] = j < n ? kept_pairs(counts[j], depth_keys[j], j, capacity, fixed, cut, keep) : 0u;
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

// ------------------------------------------------------ binning by counts --
// The pairs binned by tile without a sort (per_tile_depth 2): the splats are cut
// into chunks, one per workgroup, and
// 1. bin_count: each chunk counts its pairs per tile in LDS. A splat covers a
//    rectangle of tiles, so it adds +1 / -1 at the rectangle's four corners of a
//    per-tile difference grid (four LDS atomics however many tiles it covers), and
//    a 2D prefix sum turns the grid into counts: one row of the (chunk x tile)
//    count matrix per chunk, and the chunk's pair total.
// 2. bin_cut: one workgroup adds the chunk totals into the frame's pair count and,
//    when the pairs overflow the capacity, cuts them exactly as pair_cut does and
//    counts the matrix again from the kept pairs.
// 3. bin_scan: per tile, the exclusive prefix of its counts over the chunks, and its
//    total.
// 4. bin_scatter: each workgroup places its tiles' bases (a prefix over the tile
//    totals in LDS) plus its own chunk's prefixes in LDS cursors and writes each
//    pair to a slot taken from its tile's cursor. Workgroup 0 also writes the ranges
//    and the rasterizer's tile order.
// Within a tile the pairs land in no particular order; depth_order then sorts each
// list by (depth key, splat), the order of the reference's lists.
constexpr uint32_t kBinRound = 2048;  // splats a chunk takes at least (below kBinMaxGroups chunks)
constexpr uint32_t kBinMaxGroups = 512;
constexpr int kScanColumns = 64, kScanBlocks = 16;       // bin_scan: 1024 threads
constexpr uint32_t kScanRows = kBinMaxGroups / kScanBlocks;
// bin_count: 16 waves a chunk, kCountItems splats a thread per round.
constexpr int kCountThreads = 1024;
constexpr int kCountWaves = kCountThreads / kWave;
constexpr uint32_t kCountItems = 2;
// bin_scatter: 16 waves a chunk (two chunks a CU, 8 waves a SIMD), each wave taking
// windows of 64 splats, kScatterItems windows' loads in flight.
constexpr int kScatterThreads = 1024;
constexpr int kScatterWaves = kScatterThreads / kWave;
constexpr uint32_t kScatterItems = 2;

// The chunk workgroup b takes. Workgroup b runs on XCD b mod 8, and each tile's run
// holds the chunks' pairs in chunk order, a few pairs per chunk: with the chunks of
// one XCD adjacent (groups a multiple of 8), a run is written through one XCD's L2
// in whole lines but for 8 seams, where consecutive chunks on different XCDs would
// leave a partly written line in two L2s at almost every pair.
__device__ inline uint32_t chunk_of(uint32_t b, uint32_t groups) {
  return groups % 8 == 0 ? (b % 8) * (groups / 8) + b / 8 : b;
}

struct BinShape {
  uint32_t tiles_x, tiles_y, shift;  // 16-px tiles and the bin shift
  uint32_t bx, by, tiles;            // bins
};

__device__ inline BinShape bin_shape(uint32_t width, uint32_t height, uint32_t shift) {
  BinShape b;
  b.tiles_x = ceil_div(width, kTile);
  b.tiles_y = ceil_div(height, kTile);
  b.shift = shift;
  b.bx = (b.tiles_x + (1u << shift) - 1u) >> shift;
  b.by = (b.tiles_y + (1u << shift) - 1u) >> shift;
  b.tiles = b.bx * b.by;
  return b;
}

// A splat's bin rectangle (pack_rect) and its pair count, at most the rectangle's
// area; a rectangle past the grid (a stale one) counts no pairs.
__device__ inline uint32_t splat_bins(uint2 packed, uint32_t count, const BinShape& b, uint32_t r[4]) {
  unpack_rect(packed, r);
  if (r[2] > b.bx || r[3] > b.by || r[2] <= r[0] || r[3] <= r[1]) return 0u;
  return min(count, (r[2] - r[0]) * (r[3] - r[1]));
}

// Loads of inputs read once (non-temporal: they do not take L2 lines from the pairs
// being written).
__device__ inline uint32_t stream_load(const uint32_t* p) { return __builtin_nontemporal_load(p); }
__device__ inline uint2 stream_load(const uint2* p) {
  const uint64_t v = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(p));
  return make_uint2(static_cast<uint32_t>(v), static_cast<uint32_t>(v >> 32));
}

// +1 over the bins [x0, x1) x [y0, y1) of a bx x by difference grid: four corner
// updates (corners past the grid's last row or column fall away).
__device__ inline void add_rect(uint32_t* grid, const BinShape& b, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  if (x1 <= x0 || y1 <= y0) return;
  atomicAdd(&grid[y0 * b.bx + x0], 1u);
  if (x1 < b.bx) atomicAdd(&grid[y0 * b.bx + x1], ~0u);
  if (y1 < b.by) {
    atomicAdd(&grid[y1 * b.bx + x0], ~0u);
    if (x1 < b.bx) atomicAdd(&grid[y1 * b.bx + x1], 1u);
  }
}

// The first c bins of rectangle r, row by row (the order tiles.wgsl:47-58 writes
// them): its whole rows, then the start of the next.
__device__ inline void add_pairs(uint32_t* grid, const BinShape& b, const uint32_t r[4], uint32_t c) {
  const uint32_t w = r[2] - r[0];
  const uint32_t rows = c / w, rest = c - rows * w;
  add_rect(grid, b, r[0], r[1], r[2], r[1] + rows);
  add_rect(grid, b, r[0], r[1] + rows, r[0] + rest, r[1] + rows + 1);
}

// The difference grid summed into counts: rows a wave each (64 bins a step), then
// columns a thread each. Ends with a barrier.
template <int kN>
__device__ inline void grid_prefix(uint32_t* grid, const BinShape& b) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  for (uint32_t y = wave; y < b.by; y += kN / kWave) {
    uint32_t carry = 0;
    for (uint32_t x0 = 0; x0 < b.bx; x0 += kWave) {
      const uint32_t x = x0 + lane;
      uint32_t v = x < b.bx ? grid[y * b.bx + x] : 0u;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
      }
      v += carry;
      if (x < b.bx) grid[y * b.bx + x] = v;
      carry = __shfl(v, kWave - 1, kWave);
    }
  }
  __syncthreads();
  for (uint32_t x = tid; x < b.bx; x += kN) {
    uint32_t run = 0;
    for (uint32_t y = 0; y < b.by; y++) {
      run += grid[y * b.bx + x];
      grid[y * b.bx + x] = run;
    }
  }
  __syncthreads();
}

// Sum of x over the workgroup (kN threads), in `slots` (kN / 64 words); every thread
// gets it. Starts and ends with its own barriers.
template <int kN>
__device__ inline uint32_t group_sum(uint32_t* slots, uint32_t x) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  __syncthreads();
  if (lane == 0) slots[wave] = x;
  __syncthreads();
  uint32_t all = 0;
#pragma unroll
  for (int w = 0; w < kN / kWave; w++) all += slots[w];
  return all;
}

__global__ __launch_bounds__(kCountThreads) void bin_count_kernel(const uint2* __restrict__ rects,
                                                                 const uint32_t* __restrict__ counts, uint32_t n,
                                                                 uint32_t chunk, uint32_t width, uint32_t height,
                                                                 uint32_t shift, uint32_t* __restrict__ matrix,
                                                                 uint32_t* __restrict__ partial) {
  __shared__ uint32_t grid[kBinMaxTiles];
  __shared__ uint32_t slots[kCountWaves];
  const int tid = threadIdx.x;
  const BinShape b = bin_shape(width, height, shift);
  for (uint32_t t = tid; t < b.tiles; t += kCountThreads) grid[t] = 0;
  __syncthreads();
  const uint32_t g = chunk_of(blockIdx.x, gridDim.x);
  const uint32_t first = min(n, g * chunk), last = min(first + chunk, n);
  uint32_t sum = 0;
  for (uint32_t base = first; base < last; base += kCountThreads * kCountItems) {
    uint32_t c[kCountItems];
    uint2 rect[kCountItems];
#pragma unroll
    for (uint32_t k = 0; k < kCountItems; k++) {
      const uint32_t j = base + k * kCountThreads + tid;
      c[k] = j < last ? stream_load(counts + j) : 0u;
      rect[k] = j < last ? stream_load(rects + j) : make_uint2(0u, 0u);
    }
#pragma unroll
    for (uint32_t k = 0; k < kCountItems; k++) {
      uint32_t r[4];
      const uint32_t kept = splat_bins(rect[k], c[k], b, r);
      if (kept) {
        add_pairs(grid, b, r, kept);
        sum += kept;
      }
    }
  }
  __syncthreads();
  grid_prefix<kCountThreads>(grid, b);
  uint32_t* row = matrix + static_cast<size_t>(g) * b.tiles;
  for (uint32_t t = tid; t < b.tiles; t += kCountThreads) row[t] = grid[t];
  sum = group_sum<kCountThreads>(slots, sum);
  if (tid == 0) partial[g] = sum;
}

// The frame's pair count (pair_total[0], uncut) from the chunk totals; on overflow
// the cut, the count matrix counted again from the kept pairs and the tile counts
// rewritten to them. One workgroup of kCutThreads.
__global__ __launch_bounds__(kCutThreads) void bin_cut_kernel(const uint2* __restrict__ rects,
                                                               const uint32_t* __restrict__ depth_keys,
                                                               uint32_t* __restrict__ counts, uint32_t n,
                                                               uint32_t chunk, uint32_t groups, uint32_t width,
                                                               uint32_t height, uint32_t shift,
                                                               uint32_t* __restrict__ matrix,
                                                               const uint32_t* __restrict__ partial,
                                                               uint32_t* __restrict__ pair_total, uint32_t capacity) {
  __shared__ CutLds L;
  __shared__ uint32_t grid[kBinMaxTiles];
  __shared__ uint32_t slots[kCutWaves];
  const int tid = threadIdx.x;
  uint32_t mine = 0;
  for (uint32_t g = tid; g < groups; g += kCutThreads) mine += partial[g];
  const uint32_t total = group_sum<kCutThreads>(slots, mine);
  if (tid == 0) pair_total[0] = total;
  if (total <= capacity) return;
  uint32_t fixed, cut, keep;
  find_cut(L, depth_keys, counts, n, capacity, fixed, cut, keep);
  const BinShape b = bin_shape(width, height, shift);
  // Rows again, from the kept pairs, before the counts are rewritten.
  for (uint32_t g = 0; g < groups; g++) {
    for (uint32_t t = tid; t < b.tiles; t += kCutThreads) grid[t] = 0;
    __syncthreads();
    const uint32_t first = min(n, g * chunk), last = min(first + chunk, n);
    for (uint32_t j = first + tid; j < last; j += kCutThreads) {
      const uint32_t kept = kept_pairs(counts[j], depth_keys[j], j, capacity, fixed, cut, keep);
      if (kept) {
        uint32_t r[4];
        const uint32_t c = splat_bins(rects[j], kept, b, r);
        if (c) add_pairs(grid, b, r, c);
      }
    }
    __syncthreads();
    grid_prefix<kCutThreads>(grid, b);
    uint32_t* row = matrix + static_cast<size_t>(g) * b.tiles;
    for (uint32_t t = tid; t < b.tiles; t += kCutThreads) row[t] = grid[t];
    __syncthreads();
  }
  for (uint32_t j = tid; j < n; j += kCutThreads) counts[j] = kept_pairs(counts[j], depth_keys[j], j, capacity, fixed, cut, keep);
}

// Per tile (a column of the count matrix): the exclusive prefix of its counts over
// the chunks, in place, and its total. A workgroup takes 64 tiles; each of its 16
// row blocks sums its rows, the blocks' sums are scanned in LDS, and each block
// writes its prefixes back.
__global__ __launch_bounds__(kScanColumns * kScanBlocks) void bin_scan_kernel(uint32_t* __restrict__ matrix,
                                                                               uint32_t groups, uint32_t tiles,
                                                                               uint32_t* __restrict__ totals) {
  __shared__ uint32_t block_sum[kScanBlocks][kScanColumns];
  const int tid = threadIdx.x, c = tid % kScanColumns, blk = tid / kScanColumns;
  const uint32_t t = blockIdx.x * kScanColumns + c;
  const uint32_t rows = ceil_div(groups, kScanBlocks), r0 = blk * rows;
  uint32_t v[kScanRows];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanRows; k++) {
    v[k] = 0;
    if (k < rows && r0 + k < groups && t < tiles) v[k] = matrix[static_cast<size_t>(r0 + k) * tiles + t];
    sum += v[k];
  }
  block_sum[blk][c] = sum;
  __syncthreads();
  uint32_t run = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kScanBlocks; k++) {
    const uint32_t x = block_sum[k][c];
    run += k < blk ? x : 0u;
    total += x;
  }
#pragma unroll
  for (uint32_t k = 0; k < kScanRows; k++) {
    if (k < rows && r0 + k < groups && t < tiles) {
      matrix[static_cast<size_t>(r0 + k) * tiles + t] = run;
      run += v[k];
    }
  }
  if (blk == 0 && t < tiles) totals[t] = total;
}

struct ScatterLds {
  uint32_t cursor[kBinMaxTiles];
  uint32_t start[kScatterWaves][kWave], x0[kScatterWaves][kWave], y0[kScatterWaves][kWave], w[kScatterWaves][kWave],
      key[kScatterWaves][kWave];
  uint32_t next[kOrderBuckets];  // workgroup 0: the tile order's buckets
  uint32_t slots[kScatterWaves];
};

// Exclusive prefix of x over the workgroup; brackets its use of `slots` with barriers.
__device__ inline uint32_t bin_exclusive(uint32_t* slots, uint32_t x) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint32_t inclusive = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(inclusive, o, kWave);
    if (lane >= o) inclusive += y;
  }
  __syncthreads();
  if (lane == kWave - 1) slots[wave] = inclusive;
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (int w = 0; w < kScatterWaves; w++) before += w < wave ? slots[w] : 0u;
  return before + inclusive - x;
}

constexpr uint32_t kTilesPerThread = kBinMaxTiles / kScatterThreads;  // 8

// Workgroup 0's share: the ranges, the rasterizer's launch order (tiles by falling
// pair count in steps of kOrderStep, ties in any order) and the cleared item counts.
// tot[k] is tile tid + k kScatterThreads's total; L.cursor holds the tiles' bases.
__device__ void write_ranges_and_order(ScatterLds& L, const uint32_t (&tot)[kTilesPerThread], uint32_t tiles,
                                       uint2* __restrict__ ranges, uint32_t* __restrict__ order,
                                       uint32_t* __restrict__ clear) {
  const int tid = threadIdx.x;
  if (clear && tid <= static_cast<int>(kPartClasses)) clear[tid] = 0;
#pragma unroll
  for (uint32_t k = 0; k < kTilesPerThread; k++) {
    const uint32_t t = tid + k * kScatterThreads;
    if (t < tiles) ranges[t] = make_uint2(L.cursor[t], L.cursor[t] + tot[k]);
  }
  if (!order) return;
  auto bucket = [](uint32_t total) { return kOrderBuckets - 1 - min(total / kOrderStep, kOrderBuckets - 1); };
  for (uint32_t k = tid; k < kOrderBuckets; k += kScatterThreads) L.next[k] = 0;
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kTilesPerThread; k++)
    if (tid + k * kScatterThreads < tiles) atomicAdd(&L.next[bucket(tot[k])], 1u);
  __syncthreads();
  constexpr uint32_t kPer = kOrderBuckets / kScatterThreads;
  uint32_t c[kPer], mine = 0;
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    c[k] = L.next[tid * kPer + k];
    mine += c[k];
  }
  uint32_t run = bin_exclusive(L.slots, mine);
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    L.next[tid * kPer + k] = run;
    run += c[k];
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kTilesPerThread; k++) {
    const uint32_t t = tid + k * kScatterThreads;
    if (t < tiles) order[atomicAdd(&L.next[bucket(tot[k])], 1u)] = t;
  }
}

__global__ __launch_bounds__(kScatterThreads) void bin_scatter_kernel(const uint2* __restrict__ rects,
                                                                   const uint32_t* __restrict__ counts,
                                                                   const uint32_t* __restrict__ depth_keys, uint32_t n,
                                                                   uint32_t chunk, uint32_t width, uint32_t height,
                                                                   uint32_t shift, const uint32_t* __restrict__ matrix,
                                                                   const uint32_t* __restrict__ totals,
                                                                   uint2* __restrict__ keyed,
                                                                   uint2* __restrict__ ranges,
                                                                   uint32_t* __restrict__ order,
                                                                   uint32_t* __restrict__ clear) {
  __shared__ ScatterLds L;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const BinShape b = bin_shape(width, height, shift);
  const uint32_t g = chunk_of(blockIdx.x, gridDim.x);
  const uint32_t* row = matrix + static_cast<size_t>(g) * b.tiles;
  // 1. The tile totals and this chunk's prefixes, tile tid + k kScatterThreads in slot k,
  //    every load in flight at once.
  uint32_t tot[kTilesPerThread], mine_row[kTilesPerThread];
#pragma unroll
  for (uint32_t k = 0; k < kTilesPerThread; k++) {
    const uint32_t t = tid + k * kScatterThreads;
    tot[k] = t < b.tiles ? stream_load(totals + t) : 0u;
    mine_row[k] = t < b.tiles ? stream_load(row + t) : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < kTilesPerThread; k++)
    if (tid + k * kScatterThreads < b.tiles) L.cursor[tid + k * kScatterThreads] = tot[k];
  __syncthreads();
  // Where each tile's run starts: an exclusive prefix of the totals, a thread taking
  // consecutive tiles in LDS.
  const uint32_t per = ceil_div(b.tiles, kScatterThreads), t0 = min(b.tiles, tid * per), t1 = min(b.tiles, t0 + per);
  uint32_t mine = 0;
  for (uint32_t t = t0; t < t1; t++) mine += L.cursor[t];
  uint32_t run = bin_exclusive(L.slots, mine);  // its barriers order the reads above before the writes below
  for (uint32_t t = t0; t < t1; t++) {
    const uint32_t total = L.cursor[t];
    L.cursor[t] = run;
    run += total;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    write_ranges_and_order(L, tot, b.tiles, ranges, order, clear);
    __syncthreads();
  }
  // 2. This chunk's slots in each tile: behind the earlier chunks' pairs.
#pragma unroll
  for (uint32_t k = 0; k < kTilesPerThread; k++)
    if (tid + k * kScatterThreads < b.tiles) L.cursor[tid + k * kScatterThreads] += mine_row[k];
  __syncthreads();
  // 3. Each wave takes windows of 64 consecutive splats (kScatterItems of them loaded at
  //    once) and deals their pairs out to its lanes: pair q of a window belongs to
  //    the last splat whose first pair is at or before q.
  const uint32_t first = min(n, g * chunk), last = min(first + chunk, n);
  for (uint32_t base = first + wave * kWave; base < last; base += kScatterWaves * kWave * kScatterItems) {
    uint32_t c[kScatterItems], key[kScatterItems];
    uint2 rect[kScatterItems];
#pragma unroll
    for (uint32_t k = 0; k < kScatterItems; k++) {
      const uint32_t j = base + k * kScatterWaves * kWave + lane;
      c[k] = j < last ? stream_load(counts + j) : 0u;
      rect[k] = j < last ? stream_load(rects + j) : make_uint2(0u, 0u);
      key[k] = j < last ? stream_load(depth_keys + j) : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kScatterItems; k++) {
      const uint32_t window = base + k * kScatterWaves * kWave;
      if (window >= last) break;
      uint32_t r[4];
      const uint32_t kept = splat_bins(rect[k], c[k], b, r);
      uint32_t inclusive = kept;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(inclusive, o, kWave);
        if (lane >= o) inclusive += y;
      }
      const uint32_t pairs = __shfl(inclusive, kWave - 1, kWave);
      if (pairs == 0) continue;
      L.start[wave][lane] = inclusive - kept;
      L.x0[wave][lane] = r[0];
      L.y0[wave][lane] = r[1];
      L.w[wave][lane] = r[2] - r[0];
      L.key[wave][lane] = key[k];
      __builtin_amdgcn_wave_barrier();
      for (uint32_t q = lane; q < pairs; q += kWave) {
        uint32_t lo = 0;  // the last lane whose first pair is at or before q
#pragma unroll
        for (uint32_t step = kWave / 2; step > 0; step >>= 1)
          if (L.start[wave][lo + step] <= q) lo += step;
        const uint32_t k2 = q - L.start[wave][lo], width_ = L.w[wave][lo];
        const uint32_t rowi = k2 / width_;
        const uint32_t tile = (L.y0[wave][lo] + rowi) * b.bx + L.x0[wave][lo] + (k2 - rowi * width_);
        keyed[atomicAdd(&L.cursor[tile], 1u)] = make_uint2(L.key[wave][lo], window + lo);
      }
      __builtin_amdgcn_wave_barrier();
    }
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
                 uint32_t tiles, uint32_t* list, void* scratch, bool ordered, uint32_t* queue, uint32_t parity,
                 uint32_t* keyed) {
  if (tiles == 0) return;
  // Two queues, used by turns: this call fills queue `parity` (cleared by the call
  // before) and clears the other one for the next call.
  uint32_t* queued = queue + (parity & 1u);
  uint32_t* entries = queue + 2 + static_cast<size_t>(parity & 1u) * tiles;
  depth_order_kernel<<<tiles, kThreads, 0, s>>>(depth_keys, reinterpret_cast<const uint2*>(ranges), tile_order, list,
                                                reinterpret_cast<const uint2*>(keyed), ordered, queued, entries);
  DIO_LAUNCHED("depth_order");
  // With the order, the long lists are the first tiles of it: 256 workgroups take
  // them (each stops at its first short list) and the queued ones.
  const uint32_t long_groups = tile_order ? std::min<uint32_t>(tiles, 256) : tiles;
  depth_order_long_kernel<<<long_groups, kThreads, 0, s>>>(
      depth_keys, reinterpret_cast<const uint2*>(ranges), tile_order, tiles, list, reinterpret_cast<uint2*>(keyed),
      scratch, ordered, queued, entries, queue + ((parity + 1u) & 1u));
  DIO_LAUNCHED("depth_order_long");
}

BinLayout bin_layout(uint32_t n) {
  BinLayout l;
  l.groups = std::max<uint32_t>(1, std::min(kBinMaxGroups, ceil_div(n, kBinRound)));
  if (l.groups > 8) l.groups = (l.groups + 7) / 8 * 8;  // whole XCD shares (kBinMaxGroups is one too)
  l.chunk = n == 0 ? 0 : ceil_div(n, l.groups);
  return l;
}

void bin_count(hipStream_t s, const uint32_t* rects, const uint32_t* depth_keys, uint32_t* tile_counts, uint32_t n,
               uint32_t width, uint32_t height, uint32_t shift, uint32_t capacity, uint32_t* pair_total,
               const BinWorkspace& ws) {
  const uint32_t tiles = ((ceil_div(width, kTile) + (1u << shift) - 1u) >> shift) *
                         ((ceil_div(height, kTile) + (1u << shift) - 1u) >> shift);
  if (tiles == 0 || tiles > kBinMaxTiles) throw Error(-2, "binning by counts takes 1 to 8192 tiles");
  const BinLayout l = bin_layout(n);
  if (l.groups > ws.groups || tiles > ws.tiles) throw Error(-2, "bin workspace too small");
  bin_count_kernel<<<l.groups, kCountThreads, 0, s>>>(reinterpret_cast<const uint2*>(rects), tile_counts, n, l.chunk,
                                                    width, height, shift, ws.matrix, ws.partial);
  DIO_LAUNCHED("bin_count");
  bin_cut_kernel<<<1, kCutThreads, 0, s>>>(reinterpret_cast<const uint2*>(rects), depth_keys, tile_counts, n, l.chunk,
                                           l.groups, width, height, shift, ws.matrix, ws.partial, pair_total,
                                           capacity);
  DIO_LAUNCHED("bin_cut");
  bin_scan_kernel<<<ceil_div(tiles, kScanColumns), kScanColumns * kScanBlocks, 0, s>>>(ws.matrix, l.groups, tiles,
                                                                                       ws.totals);
  DIO_LAUNCHED("bin_scan");
}

void bin_scatter(hipStream_t s, const uint32_t* rects, const uint32_t* tile_counts, const uint32_t* depth_keys,
                 uint32_t n, uint32_t width, uint32_t height, uint32_t shift, const BinWorkspace& ws, uint32_t* keyed,
                 uint32_t* ranges, uint32_t* order, uint32_t* clear) {
  const BinLayout l = bin_layout(n);
  bin_scatter_kernel<<<l.groups, kScatterThreads, 0, s>>>(reinterpret_cast<const uint2*>(rects), tile_counts, depth_keys,
                                                      n, l.chunk, width, height, shift, ws.matrix, ws.totals,
                                                      reinterpret_cast<uint2*>(keyed), reinterpret_cast<uint2*>(ranges),
                                                      order, clear);
  DIO_LAUNCHED("bin_scatter");
}

}  // namespace dio
