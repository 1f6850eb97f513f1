// Exclusive prefix sum of u32 values in one pass over HBM: each workgroup scans a
// 4096-value tile in registers (16 values per thread, wave64 shuffles across the
// threads) and learns the total of every tile before it by decoupled look-back:
// a tile publishes its own sum ("aggregate") at once and its inclusive prefix
// once known, and one wavefront reads up to 64 predecessors per step.
// Tiles are handed out by an atomic ticket, so a tile only ever waits for tiles
// that started before it. The last workgroup to finish clears the tile states,
// the ticket and the finish counter for the next call. Replaces the Blelloch block scan + recursive levels of
// src/shaders/prefixSum.wgsl:30-91 / src/prefixSum.ts:59-98 (same results: an
// exclusive scan of min(count, capacity) values, total in element 0 of `total`).
#include <mutex>

#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
// Tiles of 256 threads x 16 values (4096), or for large scans 256 x 64 (16384): a
// large scan's cost is otherwise the tile tickets, as every tile takes one from the
// same word and those atomics complete one after another.
constexpr int kItemsSmall = kScanTile / kThreads;  // 16
constexpr int kItemsLarge = 64, kThreadsLarge = 256;
constexpr uint32_t kLargeScan = 1u << 20;  // values from which the large tiles pay
constexpr uint64_t kFlagAggregate = 1ull << 32;
constexpr uint64_t kFlagPrefix = 2ull << 32;
constexpr uint32_t kSelfCleanTiles = 1024;

__device__ inline uint32_t wave_inclusive_scan(uint32_t x, int lane) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  return x;
}

// The same inclusive scan across the wave with DPP moves (no LDS crossbar): the
// sums within each row of 16 lanes by row shifts, then lane 15 of each row added
// into the next row and lane 31 into rows 2 and 3 by row broadcasts. Lanes without
// a source in a move add the 0 of `old`.
__device__ inline uint32_t wave_scan_dpp(uint32_t x) {
  int v = static_cast<int>(x);
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
  return static_cast<uint32_t>(v);
}

__device__ inline uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

__device__ inline void publish(uint64_t* state, uint32_t tile, uint64_t word) {
  __hip_atomic_store(state + tile, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t peek(const uint64_t* state, int32_t tile) {
  return __hip_atomic_load(const_cast<uint64_t*>(state) + tile, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The exclusive prefix of `tile`: sums of the aggregates back to the first
// published inclusive prefix. Wave 0 only. A step reads kLook x 64 predecessors
// with all its loads in flight (item k of lane l is predecessor k 64 + l back):
// when every tile of a large grid starts at once, few of them have published a
// prefix yet, and a step of 64 would take one memory round trip per 64 tiles.
constexpr int kLook = 8;
__device__ uint32_t look_back(const uint64_t* state, uint32_t tile, int lane) {
  uint32_t exclusive = 0;
  int32_t base = static_cast<int32_t>(tile) - 1;
  for (;;) {
    uint64_t word[kLook];
#pragma unroll
    for (int k = 0; k < kLook; k++) {
      const int32_t j = base - k * kWave - lane;
      word[k] = j >= 0 ? peek(state, j) : kFlagPrefix;
    }
#pragma unroll
    for (int k = 0; k < kLook; k++) {
      const int32_t j = base - k * kWave - lane;
      for (;;) {
        const bool waiting = (word[k] >> 32) == 0;
        if (!__any(waiting)) break;
        if (waiting) word[k] = peek(state, j);
      }
    }
    // The nearest predecessor with a published prefix: the first item with any,
    // and the first lane of that item.
    int item = kLook, first = kWave - 1;
#pragma unroll
    for (int k = kLook - 1; k >= 0; k--) {
      const uint64_t prefixes = __ballot((word[k] >> 32) == 2);
      if (prefixes) {
        item = k;
        first = __ffsll(static_cast<long long>(prefixes)) - 1;
      }
    }
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kLook; k++)
      v += k < item || (k == item && lane <= first) ? static_cast<uint32_t>(word[k]) : 0u;
    exclusive += wave_sum(v);
    if (item < kLook) return exclusive;
    base -= kLook * kWave;
  }
}

// Counts this workgroup finished with the tile states (lane 0 of a wave); true for
// the last one of the grid. Without a counter (a large grid, cleared by the host
// after the kernel instead) it is never the last.
__device__ inline bool finish(uint32_t* done) {
  if (!done) return false;
  wait_memory();  // this wave's publishes have completed
  return atomicAdd(done, 1u) == gridDim.x - 1;
}

// One tile of the scan (tile < tiles): kItems / 4 blocks of 1024 values, thread tid
// holding values [tid 4, tid 4 + 4) of every block (one 16-byte load per block,
// coalesced across the workgroup). Wave 0 counts the workgroup finished with the tile
// states once it has published its inclusive prefix, before the outputs are
// written; `last` tells whether it was the last of the grid.
template <int kThreads, int kItems>
__device__ void scan_tile(const uint32_t* in, uint32_t* out, uint32_t n,
                          uint32_t tiles, uint32_t tile, uint32_t* total, uint64_t* state, uint32_t* done,
                          bool& last) {
  constexpr int kBlocks = kItems / 4, kWaves = kThreads / kWave;
  constexpr uint32_t kBlock = kThreads * 4;
  __shared__ bool s_last;
  __shared__ uint32_t s_wave[kBlocks][kWaves];
  __shared__ uint32_t s_prefix;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile_base = tile * (kThreads * kItems);
  const bool full = tile_base + kThreads * kItems <= n;
  uint4 v[kBlocks];
  const uint4* src = reinterpret_cast<const uint4*>(in + tile_base);
  if (full) {
#pragma unroll
    for (int q = 0; q < kBlocks; q++) v[q] = src[q * kThreads + tid];
  } else {
    // The last tile: whole 16-byte loads where they fit, one value at a time only for
    // the one that straddles the end.
#pragma unroll
    for (int q = 0; q < kBlocks; q++) {
      const uint32_t i = tile_base + q * kBlock + tid * 4;
      if (i + 4 <= n) {
        v[q] = src[q * kThreads + tid];
      } else {
        v[q] = make_uint4(0u, 0u, 0u, 0u);
        if (i < n) v[q].x = in[i];
        if (i + 1 < n) v[q].y = in[i + 1];
        if (i + 2 < n) v[q].z = in[i + 2];
      }
    }
  }
  // Per block: the values before the thread's four within its wave (its wave's
  // scan of the threads' sums), and each wave's total in LDS.
  uint32_t before[kBlocks];
#pragma unroll
  for (int q = 0; q < kBlocks; q++) {
    const uint32_t sum = v[q].x + v[q].y + v[q].z + v[q].w;
    const uint32_t inclusive = wave_scan_dpp(sum);
    if (lane == kWave - 1) s_wave[q][wave] = inclusive;
    before[q] = inclusive - sum;
  }
  __syncthreads();
  // Then the values before them in tile order (block, then wave, then lane), and the
  // tile's aggregate.
  uint32_t aggregate = 0;
#pragma unroll
  for (int q = 0; q < kBlocks; q++) {
    uint32_t earlier_waves = 0, block_total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      const uint32_t t = s_wave[q][w];
      earlier_waves += w < wave ? t : 0u;
      block_total += t;
    }
    before[q] += aggregate + earlier_waves;
    aggregate += block_total;
  }
  // The outputs less the tile's prefix, so that only they stay live across the
  // look-back.
#pragma unroll
  for (int q = 0; q < kBlocks; q++) {
    const uint32_t a = before[q], b = a + v[q].x, c = b + v[q].y;
    v[q] = make_uint4(a, b, c, c + v[q].z);
  }

  if (wave == 0) {
    uint32_t exclusive = 0;
    if (tile == 0) {
      if (lane == 0) publish(state, 0, kFlagPrefix | aggregate);
    } else {
      if (lane == 0) publish(state, tile, kFlagAggregate | aggregate);
      exclusive = look_back(state, tile, lane);
      if (lane == 0) publish(state, tile, kFlagPrefix | (exclusive + aggregate));
    }
    if (lane == 0) {
      s_prefix = exclusive;
      if (tile == tiles - 1) *total = exclusive + aggregate;
      s_last = finish(done);
    }
  }
  __syncthreads();
  last = s_last;
  const uint32_t prefix = s_prefix;
  if (full) {
    uint4* dst = reinterpret_cast<uint4*>(out + tile_base);
#pragma unroll
    for (int q = 0; q < kBlocks; q++)
      dst[q * kThreads + tid] = make_uint4(v[q].x + prefix, v[q].y + prefix, v[q].z + prefix, v[q].w + prefix);
  } else {
#pragma unroll
    for (int q = 0; q < kBlocks; q++) {
      const uint32_t i = tile_base + q * kBlock + tid * 4;
      if (i < n) out[i] = v[q].x + prefix;
      if (i + 1 < n) out[i + 1] = v[q].y + prefix;
      if (i + 2 < n) out[i + 2] = v[q].z + prefix;
      if (i + 3 < n) out[i + 3] = v[q].w + prefix;
    }
  }
}

template <int kThreads, int kItems>
__global__ __launch_bounds__(kThreads) void scan_kernel(const uint32_t* in, uint32_t* out, const uint32_t* count, uint32_t capacity,
                                                         uint32_t* total, uint64_t* state, uint32_t* ticket,
                                                         uint32_t* done) {
  __shared__ uint32_t s_tile;
  __shared__ bool s_idle_last;
  const int tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t n = count ? min(*count, capacity) : capacity;
  const uint32_t tiles = n == 0 ? 1 : ceil_div(n, kThreads * kItems);
  // The states only see device-scope atomics, and a workgroup counts as finished
  // after its look-back reads and once its publishes have completed, so the last
  // one can clear them.
  bool last;
  if (tile < tiles) {
    scan_tile<kThreads, kItems>(in, out, n, tiles, tile, total, state, done, last);
  } else {
    if (tid == 0) s_idle_last = finish(done);
    __syncthreads();
    last = s_idle_last;
  }
  if (!last) return;
  for (uint32_t k = tid; k < tiles; k += kThreads) publish(state, k, 0);
  if (tid == 0) {
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void fill_kernel(uint32_t* p, uint32_t value, size_t count) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < count; i += stride) p[i] = value;
}

// The allocation-free scan's workspace: one look-back state per 4096-value tile of
// up to kStaticScanMaxValues values, then the ticket and the finish counter. Every
// GPU the library runs on holds its own copy for the life of the process. Host code
// reaches it only through the address hipGetSymbolAddress gives for the current
// device (the array's name on the host is a stand-in, not a device address).
constexpr uint32_t kStaticScanTiles = kStaticScanMaxValues / kScanTile;
static __device__ __align__(256) uint64_t scratch[kStaticScanTiles + 1];

}  // namespace

ScanWorkspace static_scan_workspace(int device, uint32_t capacity) {
  if (capacity > kStaticScanMaxValues) {
    throw Error(-1, "the allocation-free scan takes at most " + std::to_string(kStaticScanMaxValues) + " values");
  }
  constexpr int kDevices = 64;
  static std::mutex lock;
  static void* address[kDevices] = {};
  if (device < 0 || device >= kDevices) throw Error(-1, "no such device");
  void* base;
  {
    std::lock_guard<std::mutex> guard(lock);
    if (!address[device]) {
      int current = 0;
      check_hip(hipGetDevice(&current), "hipGetDevice");
      if (current != device) check_hip(hipSetDevice(device), "hipSetDevice");
      void* p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (current != device) (void)hipSetDevice(current);
      check_hip(e, "hipGetSymbolAddress(scratch)");
      address[device] = p;
    }
    base = address[device];
  }
  return scan_workspace(base, capacity);
}

void fill_u32(hipStream_t s, uint32_t* p, uint32_t value, size_t count) {
  if (count == 0) return;
  uint32_t blocks = ceil_div(count, 256);
  if (blocks > 4096) blocks = 4096;
  fill_kernel<<<blocks, 256, 0, s>>>(p, value, count);
  DIO_LAUNCHED("fill");
}

size_t scan_workspace_bytes(uint32_t capacity) {
  uint32_t tiles = capacity == 0 ? 1 : ceil_div(capacity, kScanTile);
  return static_cast<size_t>(tiles) * 8 + 256;
}

ScanWorkspace scan_workspace(void* base, uint32_t capacity) {
  ScanWorkspace ws;
  ws.tiles = capacity == 0 ? 1 : ceil_div(capacity, kScanTile);
  ws.tile_state = static_cast<uint64_t*>(base);
  ws.ticket = reinterpret_cast<uint32_t*>(static_cast<char*>(base) + static_cast<size_t>(ws.tiles) * 8);
  ws.done = ws.ticket + 1;
  return ws;
}

void scan_exclusive(hipStream_t s, const uint32_t* in, uint32_t* out, const uint32_t* count, uint32_t capacity,
                    uint32_t* total, const ScanWorkspace& ws) {
  if ((capacity == 0 ? 1 : ceil_div(capacity, kScanTile)) > ws.tiles) throw Error(-2, "scan workspace too small");
  const bool large = capacity >= kLargeScan;
  const uint32_t tiles =
      capacity == 0 ? 1 : ceil_div(capacity, large ? kThreadsLarge * kItemsLarge : kThreads * kItemsSmall);
  // Up to kSelfCleanTiles workgroups the last one clears the workspace; past that,
  // and for large tiles (whose scans are bound by same-address atomics: the tickets),
  // one more such atomic per workgroup costs more than a clearing launch.
  const bool self_clean = !large && tiles <= kSelfCleanTiles;
  uint32_t* done = self_clean ? ws.done : nullptr;
  if (large)
    scan_kernel<kThreadsLarge, kItemsLarge><<<tiles, kThreadsLarge, 0, s>>>(in, out, count, capacity, total,
                                                                             ws.tile_state, ws.ticket, done);
  else
    scan_kernel<kThreads, kItemsSmall><<<tiles, kThreads, 0, s>>>(in, out, count, capacity, total, ws.tile_state,
                                                                   ws.ticket, done);
  DIO_LAUNCHED("scan");
  // The states, the ticket and the finish counter are one run of words.
  if (!self_clean) fill_u32(s, reinterpret_cast<uint32_t*>(ws.tile_state), 0u, static_cast<size_t>(ws.tiles) * 2 + 2);
}

}  // namespace dio
