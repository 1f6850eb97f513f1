// Stable LSD radix sort with 8-bit digits over u32 keys, or u64 keys stored as
// (low word, high word) pairs, each key carrying a u32 value. The drop-in for
// src/radixSort.ts:27-142 / src/shaders/radixSort.wgsl:44-120: the same passes
// (digit-major block counts, an exclusive scan of them, a stable scatter), built
// for wave64:
// * a 4096-key tile per workgroup (256 threads x 16 keys, coalesced stripes);
// * stable ranks from wavefront "match" masks (8 ballots per key) plus a
//   per-digit prefix across the 4 waves, instead of 256-bit LDS masks;
// * the tile is reordered in LDS first, so each digit's run is written out as one
//   contiguous stretch of global memory;
// * the count scan is the one-pass decoupled look-back scan.
#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = kSortTile / kThreads;  // 16
constexpr int kWaves = kThreads / kWave;

template <bool kWide>
struct KeyOps;

template <>
struct KeyOps<false> {
  using T = uint32_t;
  __device__ static uint32_t digit(T k, uint32_t shift) { return (k >> shift) & 0xffu; }
  __device__ static T invalid() { return 0xffffffffu; }
};

template <>
struct KeyOps<true> {
  using T = uint2;  // (low word, high word)
  __device__ static uint32_t digit(T k, uint32_t shift) {
    return shift >= 32 ? (k.y >> (shift - 32)) & 0xffu : (k.x >> shift) & 0xffu;
  }
  __device__ static T invalid() { return make_uint2(0xffffffffu, 0xffffffffu); }
};

__device__ inline uint32_t key_count(const uint32_t* count, uint32_t capacity) {
  return count ? min(*count, capacity) : capacity;
}

template <bool kWide>
__global__ __launch_bounds__(kThreads) void histogram_kernel(const typename KeyOps<kWide>::T* __restrict__ keys,
                                                              const uint32_t* count, uint32_t capacity,
                                                              uint32_t shift, uint32_t* __restrict__ counts,
                                                              uint32_t* counts_len) {
  __shared__ uint32_t hist[kWaves][256];
  const int tid = threadIdx.x, wave = tid / kWave;
  const uint32_t n = key_count(count, capacity);
  const uint32_t tiles = ceil_div(n, kSortTile);
  if (blockIdx.x == 0 && tid == 0) *counts_len = 256u * tiles;
  if (blockIdx.x >= tiles) return;
#pragma unroll
  for (int w = 0; w < kWaves; w++) hist[w][tid] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kThreads + tid;
    if (i < n) atomicAdd(&hist[wave][KeyOps<kWide>::digit(keys[i], shift)], 1u);
  }
  __syncthreads();
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) total += hist[w][tid];
  counts[tid * tiles + blockIdx.x] = total;
}

template <bool kWide>
__global__ __launch_bounds__(kThreads) void scatter_kernel(const typename KeyOps<kWide>::T* __restrict__ keys_in,
                                                            const uint32_t* __restrict__ values_in,
                                                            typename KeyOps<kWide>::T* __restrict__ keys_out,
                                                            uint32_t* __restrict__ values_out, const uint32_t* count,
                                                            uint32_t capacity, uint32_t shift,
                                                            const uint32_t* __restrict__ counts) {
  using Ops = KeyOps<kWide>;
  using Key = typename Ops::T;
  __shared__ uint32_t wave_count[kWaves][256];
  __shared__ uint32_t wave_offset[kWaves][256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t local_start[256];
  __shared__ uint32_t global_start[256];
  __shared__ uint32_t wave_totals[kWaves];
  __shared__ Key staged_keys[kSortTile];
  __shared__ uint32_t staged_values[kSortTile];

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t n = key_count(count, capacity);
  const uint32_t tiles = ceil_div(n, kSortTile);
  if (blockIdx.x >= tiles) return;
  const uint32_t base = blockIdx.x * kSortTile;
  global_start[tid] = counts[tid * tiles + blockIdx.x];
  running[tid] = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) wave_count[w][tid] = 0;
  const uint64_t lower_lanes = (1ull << lane) - 1;

  Key key[kItems];
  uint32_t value[kItems];
  uint32_t rank[kItems];
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kThreads + tid;
    bool valid = i < n;
    key[k] = valid ? keys_in[i] : Ops::invalid();
    value[k] = valid ? values_in[i] : 0u;
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const bool valid = base + k * kThreads + tid < n;
    const uint32_t d = Ops::digit(key[k], shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t set = __ballot(bit);
      peers &= bit ? set : ~set;
    }
    const uint64_t before = peers & lower_lanes;
    if (valid && before == 0) wave_count[wave][d] = __popcll(peers);
    __syncthreads();
    {
      uint32_t r = running[tid];
#pragma unroll
      for (int w = 0; w < kWaves; w++) {
        uint32_t c = wave_count[w][tid];
        wave_offset[w][tid] = r;
        r += c;
        wave_count[w][tid] = 0;
      }
      running[tid] = r;
    }
    __syncthreads();
    rank[k] = wave_offset[wave][d] + __popcll(before);
  }

  // Where each digit's run starts inside the tile: an exclusive scan of the counts.
  {
    const uint32_t c = running[tid];
    uint32_t x = c;
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
    local_start[tid] = before_wave + x - c;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    if (base + k * kThreads + tid < n) {
      const uint32_t pos = local_start[Ops::digit(key[k], shift)] + rank[k];
      staged_keys[pos] = key[k];
      staged_values[pos] = value[k];
    }
  }
  __syncthreads();
  const uint32_t tile_n = min(static_cast<uint32_t>(kSortTile), n - base);
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const uint32_t p = k * kThreads + tid;
    if (p < tile_n) {
      const Key kk = staged_keys[p];
      const uint32_t d = Ops::digit(kk, shift);
      const uint32_t dst = global_start[d] + p - local_start[d];
      keys_out[dst] = kk;
      values_out[dst] = staged_values[p];
    }
  }
}

}  // namespace

size_t sort_workspace_bytes(uint32_t capacity) {
  uint32_t tiles = capacity == 0 ? 1 : ceil_div(capacity, kSortTile);
  size_t counts = static_cast<size_t>(256) * tiles * 4;
  return counts + 256 + scan_workspace_bytes(256u * tiles);
}

SortWorkspace sort_workspace(void* base, uint32_t capacity) {
  SortWorkspace ws;
  ws.tiles = capacity == 0 ? 1 : ceil_div(capacity, kSortTile);
  char* p = static_cast<char*>(base);
  ws.counts = reinterpret_cast<uint32_t*>(p);
  p += static_cast<size_t>(256) * ws.tiles * 4;
  ws.counts_len = reinterpret_cast<uint32_t*>(p);
  p += 256;
  ws.scan = scan_workspace(p, 256u * ws.tiles);
  return ws;
}

bool radix_sort(hipStream_t s, bool wide, void* keys, uint32_t* values, void* alt_keys, uint32_t* alt_values,
                const uint32_t* count, uint32_t capacity, uint32_t begin_bit, uint32_t end_bit,
                const SortWorkspace& ws) {
  if (begin_bit % 8 != 0) throw Error(-2, "radix sort: begin bit must be a multiple of 8");
  if (end_bit > (wide ? 64u : 32u)) throw Error(-2, "radix sort: end bit past the key");
  if (capacity == 0) return false;
  const uint32_t tiles = ceil_div(capacity, kSortTile);
  if (tiles > ws.tiles) throw Error(-2, "sort workspace too small");
  void* kin = keys;
  uint32_t* vin = values;
  void* kout = alt_keys;
  uint32_t* vout = alt_values;
  bool flipped = false;
  for (uint32_t shift = begin_bit; shift < end_bit; shift += 8) {
    if (wide) {
      histogram_kernel<true><<<tiles, kThreads, 0, s>>>(static_cast<const uint2*>(kin), count, capacity, shift,
                                                         ws.counts, ws.counts_len);
    } else {
      histogram_kernel<false><<<tiles, kThreads, 0, s>>>(static_cast<const uint32_t*>(kin), count, capacity, shift,
                                                          ws.counts, ws.counts_len);
    }
    DIO_LAUNCHED("radix histogram");
    scan_exclusive(s, ws.counts, ws.counts, ws.counts_len, 256u * tiles, ws.counts_len + 1, ws.scan);
    if (wide) {
      scatter_kernel<true><<<tiles, kThreads, 0, s>>>(static_cast<const uint2*>(kin), vin, static_cast<uint2*>(kout),
                                                       vout, count, capacity, shift, ws.counts);
    } else {
      scatter_kernel<false><<<tiles, kThreads, 0, s>>>(static_cast<const uint32_t*>(kin), vin,
                                                        static_cast<uint32_t*>(kout), vout, count, capacity, shift,
                                                        ws.counts);
    }
    DIO_LAUNCHED("radix scatter");
    void* tk = kin;
    kin = kout;
    kout = tk;
    uint32_t* tv = vin;
    vin = vout;
    vout = tv;
    flipped = !flipped;
  }
  return flipped;
}

}  // namespace dio
