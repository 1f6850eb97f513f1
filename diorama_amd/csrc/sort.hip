// Stable LSD radix sort with 8-bit (or 4-bit) digits over u32 keys, or u64 keys stored as
// (low word, high word) pairs, each key carrying a u32 value: the drop-in for
// src/radixSort.ts:27-142 / src/shaders/radixSort.wgsl:44-120 (same result: a
// stable sort by key bits [begin, end)), built as a onesweep sort for wave64:
// * one histogram kernel counts every pass's digits at once (one read of the keys);
// * then one kernel per pass: a 4096-key tile per workgroup (4 waves, each owning
//   1,024 consecutive keys) ranks its keys stably with wavefront "match" masks
//   (8 ballots per key) and per-wave digit counts, then a per-digit prefix across
//   the 4 waves (two barriers per tile), learns where its run of each digit
//   starts by per-digit look-back over earlier tiles, and scatters through LDS so
//   each digit's run is written contiguously.
// The look-back is two-level. Tiles form groups of 16; each tile adds its counts
// into its group's sum (with a count of tiles added in the high bits), reads its
// own group's earlier tiles, then whole earlier groups (their inclusive prefix
// once a group's last tile has published it, else their completed sum). When
// every tile of a pass runs at once (a million keys is 269 tiles) that is a few
// round trips, where a window over single tiles needs one per 16 tiles.
// Tiles are handed out by an atomic ticket, so a tile only waits on tiles that
// started before it. Per pass the keys are read and written once.
#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = kSortTile / kThreads;  // 16
constexpr int kWaves = kThreads / kWave;
constexpr uint32_t kMaxPasses = 16;  // 64 key bits in 4-bit digits
// The global digit counts are kept in one copy per XCD (workgroup b runs on XCD
// b mod 8): every workgroup adds its counts with device-scope atomics, and one
// shared copy made each counter take an atomic from every workgroup in turn.
constexpr uint32_t kHistCopies = 8, kHistStride = kMaxPasses * 256;
constexpr uint32_t kHistDone = 32;  // the histogram's finish counter among the tickets
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
// Group sums: tiles added so far above bit 24, their counts (at most 16 x 4096) below.
constexpr uint32_t kGroupShift = 24;
constexpr uint32_t kGroupSumMask = (1u << kGroupShift) - 1;

// One pass's look-back words: per tile, per group sum, per group prefix.
struct PassState {
  uint32_t* state;
  uint32_t* group_sum;
  uint32_t* group_prefix;
};

// Keys of one width with digits of kBits (8 or 4) bits; a digit never straddles
// the two words of a 64-bit key.
template <bool kWide, int kBits>
struct KeyOps;

template <int kBits>
struct KeyOps<false, kBits> {
  using T = uint32_t;
  __device__ static uint32_t digit(T k, uint32_t shift) { return (k >> shift) & ((1u << kBits) - 1u); }
  __device__ static T invalid() { return 0xffffffffu; }
};

template <int kBits>
struct KeyOps<true, kBits> {
  using T = uint2;  // (low word, high word)
  __device__ static uint32_t digit(T k, uint32_t shift) {
    return (shift >= 32 ? k.y >> (shift - 32) : k.x >> shift) & ((1u << kBits) - 1u);
  }
  __device__ static T invalid() { return make_uint2(0xffffffffu, 0xffffffffu); }
};

__device__ inline uint32_t key_count(const uint32_t* count, uint32_t capacity) {
  return count ? min(*count, capacity) : capacity;
}

// Device-coherent word access (agent-scope atomics bypass the per-XCD L2).
__device__ inline void publish(uint32_t* state, uint32_t word) {
  __hip_atomic_store(state, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t peek(uint32_t* state) {
  return __hip_atomic_load(state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Counts the digits of every pass in one read of the keys (LDS histograms added
// into the global ones), a 4096-key tile at a time with all 16 loads per thread
// in flight together, clears the first pass's look-back words and every pass
// ticket, and ends with the digit offsets of every pass.
template <bool kWide, int kBits>
__global__ __launch_bounds__(kThreads) void histogram_kernel(const typename KeyOps<kWide, kBits>::T* __restrict__ keys,
                                                              const uint32_t* count, uint32_t capacity,
                                                              uint32_t begin_bit, uint32_t passes,
                                                              uint32_t* __restrict__ hist, uint32_t* __restrict__ offsets,
                                                              uint32_t* __restrict__ state0,
                                                              uint32_t* __restrict__ group_sum0,
                                                              uint32_t* __restrict__ group_prefix0,
                                                              uint32_t* __restrict__ tickets) {
  using Ops = KeyOps<kWide, kBits>;
  using Key = typename Ops::T;
  constexpr int kRadix = 1 << kBits;
  __shared__ uint32_t local[kMaxPasses][kRadix];
  const int tid = threadIdx.x;
  const uint32_t n = key_count(count, capacity);
  const uint32_t tiles = ceil_div(n, kSortTile), groups = ceil_div(tiles, kSortGroup);
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t k = blockIdx.x * kThreads + tid; k < tiles * 256u; k += stride) state0[k] = 0;
  for (uint32_t k = blockIdx.x * kThreads + tid; k < groups * 256u; k += stride) {
    group_sum0[k] = 0;
    group_prefix0[k] = 0;
  }
  if (blockIdx.x == 0 && tid < static_cast<int>(kMaxPasses)) tickets[tid] = 0;
  if (tid < kRadix)
    for (uint32_t p = 0; p < passes; p++) local[p][tid] = 0;
  __syncthreads();
  // Plain LDS atomics: lanes hitting the same digit serialise in the LDS (a hot
  // exponent byte costs up to 64 cycles an instruction), which is still far
  // cheaper than agreeing on one count per digit with 8 ballots per key first.
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    Key key[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      const uint32_t i = tile * kSortTile + k * kThreads + tid;
      key[k] = i < n ? keys[i] : Ops::invalid();
    }
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      if (tile * kSortTile + k * kThreads + tid < n)
        for (uint32_t p = 0; p < passes; p++) atomicAdd(&local[p][Ops::digit(key[k], begin_bit + kBits * p)], 1u);
    }
  }
  __syncthreads();
  uint32_t* mine = hist + (blockIdx.x % kHistCopies) * kHistStride;
  if (tid < kRadix)
    for (uint32_t p = 0; p < passes; p++)
      if (local[p][tid]) atomicAdd(&mine[p * 256 + tid], local[p][tid]);
  // The last workgroup to finish sums the copies into each pass's digit offsets
  // (exclusive scan over the digits) and zeroes the copies and its counter. The
  // copies only ever see device-scope atomics, and every workgroup's have completed
  // before its finish is counted.
  __shared__ bool s_last;
  __shared__ uint32_t wave_totals[kWaves];
  wait_memory();
  __syncthreads();
  if (tid == 0) s_last = atomicAdd(&tickets[kHistDone], 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  for (uint32_t p = 0; p < passes; p++) {
    uint32_t c = 0;
#pragma unroll
    for (uint32_t x = 0; x < kHistCopies; x++)
      c += peek(&hist[x * kHistStride + p * 256 + tid]);
    uint32_t inclusive = c;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(inclusive, o, kWave);
      if (lane >= o) inclusive += y;
    }
    if (lane == kWave - 1) wave_totals[wave] = inclusive;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) before += w < wave ? wave_totals[w] : 0u;
    offsets[p * 256 + tid] = before + inclusive - c;
#pragma unroll
    for (uint32_t x = 0; x < kHistCopies; x++) publish(&hist[x * kHistStride + p * 256 + tid], 0u);
    __syncthreads();  // wave_totals is rewritten by the next pass
  }
  if (tid == 0) publish(&tickets[kHistDone], 0u);
}

template <bool kWide, int kBits>
__global__ __launch_bounds__(kThreads) void onesweep_kernel(const typename KeyOps<kWide, kBits>::T* __restrict__ keys_in,
                                                             const uint32_t* __restrict__ values_in,
                                                             typename KeyOps<kWide, kBits>::T* __restrict__ keys_out,
                                                             uint32_t* __restrict__ values_out, const uint32_t* count,
                                                             uint32_t capacity, uint32_t shift,
                                                             const uint32_t* __restrict__ digit_start, PassState now,
                                                             PassState next, uint32_t* ticket, uint32_t eager,
                                                             const uint32_t* __restrict__ gather_src,
                                                             uint32_t* __restrict__ gather_out) {
  using Ops = KeyOps<kWide, kBits>;
  using Key = typename Ops::T;
  constexpr int kRadix = 1 << kBits;  // thread d < kRadix owns digit d
  // The per-wave digit counts are dead once each key's rank is offset by them (a
  // barrier follows), before the first key is staged: the two share their LDS, which
  // lets three workgroups of 64-bit keys fit on a CU instead of two.
  union Shared {
    struct {
      uint32_t count[kWaves][kRadix];
      uint32_t offset[kWaves][kRadix];
    } waves;
    struct {
      Key keys[kSortTile];
      uint32_t values[kSortTile];
    } staged;
  };
  __shared__ Shared shared;
  auto& wave_count = shared.waves.count;
  auto& wave_offset = shared.waves.offset;
  auto& staged_keys = shared.staged.keys;
  auto& staged_values = shared.staged.values;
  __shared__ uint32_t running[kRadix];
  __shared__ uint32_t local_start[kRadix];
  __shared__ uint32_t run_start[kRadix];
  __shared__ uint32_t wave_totals[kWaves];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  if (tid == 0) {
    // The grid covers the capacity. Workgroups from `eager` on first read the key
    // count and take no ticket past the last tile, so the idle ones do not queue
    // on the ticket (one same-address atomic each). Tiles 0 .. tiles-1 still all
    // get a ticket, whatever `eager` is.
    const bool wanted =
        blockIdx.x < eager || blockIdx.x < ceil_div(key_count(count, capacity), static_cast<uint32_t>(kSortTile));
    s_tile = wanted ? atomicAdd(ticket, 1u) : ~0u;
  }
  if (tid < kRadix) {
    running[tid] = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) wave_count[w][tid] = 0;
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t n = key_count(count, capacity);
  const uint32_t tiles = ceil_div(n, kSortTile);
  if (tile >= tiles) return;
  const uint32_t groups = ceil_div(tiles, kSortGroup);
  if (next.state) {  // the next pass's look-back words
    next.state[tile * 256 + tid] = 0;
    if (tile < groups) {
      next.group_sum[tile * 256 + tid] = 0;
      next.group_prefix[tile * 256 + tid] = 0;
    }
  }
  const uint32_t base = tile * kSortTile;
  const uint64_t lower_lanes = (1ull << lane) - 1;

  // Each wave owns 1,024 consecutive keys of the tile (16 rows of 64), so index
  // order is (wave, row, lane) and a wave can rank its keys on its own.
  Key key[kItems];
  uint32_t value[kItems];
  uint32_t rank[kItems];
  const uint32_t wave_base = base + wave * (kItems * kWave);
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const uint32_t i = wave_base + k * kWave + lane;
    const bool valid = i < n;
    key[k] = valid ? keys_in[i] : Ops::invalid();
    value[k] = valid ? values_in[i] : 0u;
  }

  // Stable ranks within the wave: lanes with the same digit agree through match
  // masks, and the wave's own per-digit counts carry across its rows (one wave
  // reads and writes its row of LDS in program order, so no barrier is needed).
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const bool valid = wave_base + k * kWave + lane < n;
    const uint32_t d = Ops::digit(key[k], shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kBits; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t set = __ballot(bit);
      peers &= bit ? set : ~set;
    }
    const uint64_t before = peers & lower_lanes;
    const uint32_t so_far = valid ? wave_count[wave][d] : 0u;
    rank[k] = so_far + __popcll(before);
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0) wave_count[wave][d] = so_far + __popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // Per digit: the earlier waves' counts offset each wave's ranks.
  if (tid < kRadix) {
    uint32_t r = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      const uint32_t c = wave_count[w][tid];
      wave_offset[w][tid] = r;
      r += c;
    }
    running[tid] = r;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; k++) rank[k] += wave_offset[wave][Ops::digit(key[k], shift)];

  // This tile's count of digit `tid`: publish it and add it to its group's sum,
  // then add up what came before. Tiles come in groups of 16: the tile reads its
  // own group's earlier tiles (nearest first, back to one that knows its
  // inclusive prefix) and then whole earlier groups (a group's inclusive prefix
  // where its last tile has published one, else its completed sum). When every
  // tile runs at once this takes a few round trips instead of one per 16 tiles.
  const uint32_t mine = tid < kRadix ? running[tid] : 0u;
  uint32_t exclusive = 0;
  if (tid < kRadix) {
    const uint32_t group = tile / kSortGroup, first = group * kSortGroup;
    uint32_t* const state = now.state;
    publish(state + tile * 256 + tid, (tile == 0 ? kFlagPrefix : kFlagAggregate) | mine);
    atomicAdd(now.group_sum + group * 256 + tid, (1u << kGroupShift) | mine);
    bool found = tile == 0;
    if (!found) {
      const uint32_t m = tile - first;  // earlier tiles of the own group
      uint32_t word[kSortGroup - 1];
#pragma unroll
      for (uint32_t k = 0; k < kSortGroup - 1; k++) word[k] = k < m ? peek(state + (tile - 1 - k) * 256 + tid) : 0u;
#pragma unroll
      for (uint32_t k = 0; k < kSortGroup - 1; k++)
        if (k < m)
          while ((word[k] >> 30) == 0) word[k] = peek(state + (tile - 1 - k) * 256 + tid);
#pragma unroll
      for (uint32_t k = 0; k < kSortGroup - 1; k++) {
        if (k < m && !found) {
          exclusive += word[k] & kValueMask;
          found = (word[k] >> 30) == 2;
        }
      }
    }
    constexpr int kWindow = 8;  // earlier groups read per round trip
    for (int32_t g0 = static_cast<int32_t>(group) - 1; !found && g0 >= 0; g0 -= kWindow) {
      uint32_t inclusive[kWindow], sum[kWindow];
#pragma unroll
      for (int k = 0; k < kWindow; k++) {
        const int32_t g = g0 - k;
        inclusive[k] = g >= 0 ? peek(now.group_prefix + g * 256 + tid) : 0u;
        sum[k] = g >= 0 ? peek(now.group_sum + g * 256 + tid) : 0u;
      }
#pragma unroll
      for (int k = 0; k < kWindow; k++) {
        const int32_t g = g0 - k;
        if (found || g < 0) continue;
        // an earlier group is whole (16 tiles): wait until all of them have added in
        while ((inclusive[k] >> 30) != 2 && (sum[k] >> kGroupShift) < kSortGroup) {
          inclusive[k] = peek(now.group_prefix + g * 256 + tid);
          sum[k] = peek(now.group_sum + g * 256 + tid);
        }
        if ((inclusive[k] >> 30) == 2) {
          exclusive += inclusive[k] & kValueMask;
          found = true;
        } else {
          exclusive += sum[k] & kGroupSumMask;
        }
      }
    }
    if (tile != 0) publish(state + tile * 256 + tid, kFlagPrefix | (exclusive + mine));
    if (tile == min(first + kSortGroup, tiles) - 1)
      publish(now.group_prefix + group * 256 + tid, kFlagPrefix | (exclusive + mine));
    run_start[tid] = digit_start[tid] + exclusive;
  }

  // Where each digit's run starts inside the tile.
  {
    uint32_t x = mine;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      uint32_t y = __shfl_up(x, o, kWave);
      if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wave_totals[wave] = x;
    __syncthreads();
    uint32_t before_wave = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) before_wave += w < wave ? wave_totals[w] : 0u;
    if (tid < kRadix) local_start[tid] = before_wave + x - mine;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    if (wave_base + k * kWave + lane < n) {
      const uint32_t pos = local_start[Ops::digit(key[k], shift)] + rank[k];
      staged_keys[pos] = key[k];
      staged_values[pos] = value[k];
    }
  }
  __syncthreads();
  const uint32_t tile_n = min(static_cast<uint32_t>(kSortTile), n - base);
  if (gather_out && tile_n == kSortTile) {
    // A full tile of a pass that also writes gather_src[value] beside each value:
    // every gather load of a thread is issued before its first store, and nothing
    // is conditional, so no load waits inside a branch of its own store.
    uint32_t dst[kItems], gathered[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      const uint32_t p = k * kThreads + tid;
      const Key kk = staged_keys[p];
      const uint32_t d = Ops::digit(kk, shift);
      const uint32_t v = staged_values[p];
      dst[k] = run_start[d] + p - local_start[d];
      gathered[k] = gather_src[v];
      keys_out[dst[k]] = kk;
      values_out[dst[k]] = v;
    }
#pragma unroll
    for (int k = 0; k < kItems; k++) gather_out[dst[k]] = gathered[k];
    return;
  }
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const uint32_t p = k * kThreads + tid;
    if (p < tile_n) {
      const Key kk = staged_keys[p];
      const uint32_t d = Ops::digit(kk, shift);
      const uint32_t dst = run_start[d] + p - local_start[d];
      const uint32_t v = staged_values[p];
      keys_out[dst] = kk;
      values_out[dst] = v;
      if (gather_out) gather_out[dst] = gather_src[v];
    }
  }
}

}  // namespace

size_t sort_workspace_bytes(uint32_t capacity) {
  const uint32_t tiles = capacity == 0 ? 1 : ceil_div(capacity, kSortTile);
  const uint32_t groups = ceil_div(tiles, kSortGroup);
  // digit counts per XCD, digit offsets for every pass, two tile and two pairs of group
  // look-back arrays, pass tickets
  return kHistCopies * kHistStride * 4 + kHistStride * 4 + 2 * static_cast<size_t>(tiles) * 256 * 4 +
         4 * static_cast<size_t>(groups) * 256 * 4 + 256;
}

SortWorkspace sort_workspace(void* base, uint32_t capacity) {
  SortWorkspace ws;
  ws.tiles = capacity == 0 ? 1 : ceil_div(capacity, kSortTile);
  ws.groups = ceil_div(ws.tiles, kSortGroup);
  char* p = static_cast<char*>(base);
  auto take = [&p](size_t words) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    p += words * 4;
    return q;
  };
  ws.hist = take(kHistCopies * kHistStride);
  ws.offsets = take(kHistStride);
  for (int k = 0; k < 2; k++) ws.state[k] = take(static_cast<size_t>(ws.tiles) * 256);
  for (int k = 0; k < 2; k++) ws.group_sum[k] = take(static_cast<size_t>(ws.groups) * 256);
  for (int k = 0; k < 2; k++) ws.group_prefix[k] = take(static_cast<size_t>(ws.groups) * 256);
  ws.tickets = take(64);
  return ws;
}

namespace {

template <int kBits>
bool radix_sort_bits(hipStream_t s, bool wide, void* keys, uint32_t* values, void* alt_keys, uint32_t* alt_values,
                     const uint32_t* count, uint32_t capacity, uint32_t begin_bit, uint32_t end_bit,
                     const SortWorkspace& ws, const uint32_t* gather_src, uint32_t* gather_out,
                     uint32_t count_hint) {
  const uint32_t tiles = ceil_div(capacity, kSortTile);
  const uint32_t passes = (end_bit - begin_bit + kBits - 1) / kBits;
  if (passes > kMaxPasses) throw Error(-2, "radix sort: too many passes");
  // With a device-side count, the tiles expected from the host's hint (any value
  // gives the same result; the grids still cover the capacity).
  const uint32_t eager =
      count && count_hint ? std::min(tiles, ceil_div(std::min(count_hint, capacity), kSortTile)) : tiles;
  const uint32_t hist_blocks = std::max<uint32_t>(1, std::min<uint32_t>(eager, 1024));
  if (wide)
    histogram_kernel<true, kBits><<<hist_blocks, kThreads, 0, s>>>(static_cast<const uint2*>(keys), count, capacity,
                                                                    begin_bit, passes, ws.hist, ws.offsets, ws.state[0],
                                                                    ws.group_sum[0], ws.group_prefix[0], ws.tickets);
  else
    histogram_kernel<false, kBits><<<hist_blocks, kThreads, 0, s>>>(static_cast<const uint32_t*>(keys), count,
                                                                     capacity, begin_bit, passes, ws.hist, ws.offsets,
                                                                     ws.state[0],
                                                                     ws.group_sum[0], ws.group_prefix[0], ws.tickets);
  DIO_LAUNCHED("radix histogram");
  void* kin = keys;
  uint32_t* vin = values;
  void* kout = alt_keys;
  uint32_t* vout = alt_values;
  for (uint32_t p = 0; p < passes; p++) {
    const uint32_t shift = begin_bit + kBits * p;
    const PassState now{ws.state[p % 2], ws.group_sum[p % 2], ws.group_prefix[p % 2]};
    const PassState next = p + 1 < passes
                               ? PassState{ws.state[(p + 1) % 2], ws.group_sum[(p + 1) % 2], ws.group_prefix[(p + 1) % 2]}
                               : PassState{nullptr, nullptr, nullptr};
    const bool last = p + 1 == passes;
    const uint32_t* const src = last ? gather_src : nullptr;
    uint32_t* const out = last ? gather_out : nullptr;
    if (wide)
      onesweep_kernel<true, kBits><<<tiles, kThreads, 0, s>>>(static_cast<const uint2*>(kin), vin,
                                                               static_cast<uint2*>(kout), vout, count, capacity, shift,
                                                               ws.offsets + p * 256, now, next, ws.tickets + p, eager, src,
                                                               out);
    else
      onesweep_kernel<false, kBits><<<tiles, kThreads, 0, s>>>(static_cast<const uint32_t*>(kin), vin,
                                                                static_cast<uint32_t*>(kout), vout, count, capacity,
                                                                shift, ws.offsets + p * 256, now, next, ws.tickets + p, eager,
                                                                src, out);
    DIO_LAUNCHED("radix onesweep");
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  return passes % 2 == 1;
}

}  // namespace

bool radix_sort(hipStream_t s, bool wide, void* keys, uint32_t* values, void* alt_keys, uint32_t* alt_values,
                const uint32_t* count, uint32_t capacity, uint32_t begin_bit, uint32_t end_bit,
                const SortWorkspace& ws, uint32_t digit_bits, const uint32_t* gather_src, uint32_t* gather_out,
                uint32_t count_hint) {
  if (digit_bits != 8 && digit_bits != 4) throw Error(-1, "radix digit width must be 4 or 8 bits");
  if (begin_bit % digit_bits != 0) throw Error(-2, "radix sort: begin bit must be a multiple of the digit width");
  if (end_bit > (wide ? 64u : 32u)) throw Error(-2, "radix sort: end bit past the key");
  if (capacity >= (1u << 30)) throw Error(-2, "radix sort: at most 2^30 keys");
  if (!gather_src != !gather_out) throw Error(-1, "radix sort: a gather needs both its source and its output");
  if (gather_out && end_bit <= begin_bit) throw Error(-1, "radix sort: a gather needs at least one pass");
  if (capacity == 0 || end_bit <= begin_bit) return false;
  if (ceil_div(capacity, kSortTile) > ws.tiles) throw Error(-2, "sort workspace too small");
  return digit_bits == 8
             ? radix_sort_bits<8>(s, wide, keys, values, alt_keys, alt_values, count, capacity, begin_bit, end_bit, ws,
                                    gather_src, gather_out, count_hint)
             : radix_sort_bits<4>(s, wide, keys, values, alt_keys, alt_values, count, capacity, begin_bit, end_bit, ws,
                                    gather_src, gather_out, count_hint);
}

}  // namespace dio
