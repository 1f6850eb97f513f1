// Initial Gaussian sizes: for each structure-from-motion point, the mean squared
// distance to its 3 nearest other points (src/shaders/nearestNeighbors.wgsl:12-44,
// src/loadDataset.ts:126-149), then the starting raw parameters
// (src/loadDataset.ts:87-110).
//
// The reference scans all n points for every point (n^2 distances). Here the
// points are put in Morton order (30-bit codes over each point's rank along each
// axis, all four orderings by the radix sort of sort.hip) and cut into chunks of
// 64 with bounding boxes, and chunks of 64 chunks with theirs. A wave owns 64 consecutive sorted queries and visits the boxes outward
// from its own chunk, skipping every box whose nearest possible distance cannot
// beat any of its queries' third-best. Skipping is exact: the box bound is built
// from the same correctly rounded operations as a distance (differences, then
// fma(dz, dz, fma(dy, dy, dx * dx))), and each of them is monotone, so no point
// inside a skipped box can have a smaller computed distance. The three smallest
// distances are a set of values, so the visiting order does not change them: the
// result is bit-identical to the brute-force scan with the same distance formula.
//
// Non-finite points never improve a finite point's neighbours (their distances
// are NaN or infinite, which the reference's `d2 < best` rejects); they are sorted
// last, stand in as +inf when a chunk is read, and their own result is the
// reference's (3.4e38 * 3) / 3 = inf.
#include <cmath>

#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr uint32_t kChunk = 64;      // points per chunk: one wave of queries, one box
constexpr uint32_t kSuper = 64;      // chunks per super-chunk box
constexpr float kFar = 3.4e38f;      // the reference's starting "best" (nearestNeighbors.wgsl:20)
constexpr uint32_t kNonFinite = 1u << 30;

__device__ inline bool finite3(float x, float y, float z) { return isfinite(x) && isfinite(y) && isfinite(z); }

// A float's bits mapped so unsigned order is numeric order.
__device__ inline uint32_t order_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return b & 0x80000000u ? ~b : b | 0x80000000u;
}

__device__ inline float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// The distance formula, fixed: no contraction choices left to the compiler.
__device__ inline float distance2(float dx, float dy, float dz) {
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// The smallest distance the formula can give from q to any point of [lo, hi].
__device__ inline float box_bound(float4 q, float4 lo, float4 hi) {
  const float ex = fmaxf(fmaxf(lo.x - q.x, q.x - hi.x), 0.f);
  const float ey = fmaxf(fmaxf(lo.y - q.y, q.y - hi.y), 0.f);
  const float ez = fmaxf(fmaxf(lo.z - q.z, q.z - hi.z), 0.f);
  return distance2(ex, ey, ez);
}

// finite[0] += how many points have three finite coordinates.
__global__ __launch_bounds__(256) void count_finite_kernel(const float* __restrict__ pts, uint32_t n,
                                                           uint32_t* __restrict__ finite) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool ok = false;
  if (i < n)
    ok = finite3(pts[3 * static_cast<size_t>(i)], pts[3 * static_cast<size_t>(i) + 1],
                 pts[3 * static_cast<size_t>(i) + 2]);
  const uint32_t count = static_cast<uint32_t>(__popcll(__ballot(ok)));
  if ((threadIdx.x & 63) == 0 && count) atomicAdd(finite, count);
}

// Sort keys along one axis: the coordinate's order key, and the largest key for
// non-finite points so they sort last.
__global__ __launch_bounds__(256) void axis_keys_kernel(const float* __restrict__ pts, uint32_t n, uint32_t axis,
                                                        uint32_t* __restrict__ keys, uint32_t* __restrict__ values) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* p = pts + 3 * static_cast<size_t>(i);
  keys[i] = finite3(p[0], p[1], p[2]) ? order_key(p[axis]) : 0xffffffffu;
  values[i] = i;
}

__device__ inline uint32_t spread10(uint32_t v) {  // 10 bits to every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Morton codes over per-axis ranks rather than coordinates, so outliers and
// tight clusters do not squeeze the points into a few cells: the point at sorted
// position r along `axis` gets cell r * 1024 / finite on that axis.
__global__ __launch_bounds__(256) void rank_cells_kernel(const uint32_t* __restrict__ order, uint32_t n,
                                                         const uint32_t* __restrict__ finite, uint32_t axis,
                                                         uint32_t* __restrict__ code) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const uint32_t f = *finite, i = order[r];
  if (r >= f) {
    code[i] = kNonFinite;
    return;
  }
  const uint32_t cell = static_cast<uint32_t>(static_cast<uint64_t>(r) * 1024 / f);
  code[i] = (axis == 0 ? 0u : code[i]) | (spread10(cell) << axis);
}

__global__ __launch_bounds__(256) void iota_kernel(uint32_t n, uint32_t* __restrict__ values) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) values[i] = i;
}

// One wave per chunk: the chunk's points in sorted order (non-finite ones as +inf)
// and its box over the finite ones (empty: lo +inf, hi -inf).
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ pts, const uint32_t* __restrict__ order,
                                                     uint32_t n, const uint32_t* __restrict__ finite_count,
                                                     float4* __restrict__ sorted, float4* __restrict__ chunk_lo,
                                                     float4* __restrict__ chunk_hi, uint32_t chunks) {
  const uint32_t chunk = blockIdx.x * 4 + threadIdx.x / 64;
  if (chunk >= chunks) return;
  const uint32_t i = chunk * kChunk + (threadIdx.x & 63);
  const uint32_t finite = *finite_count;
  const float inf = __uint_as_float(0x7f800000u);
  float4 p = make_float4(inf, inf, inf, 0.f);
  const bool ok = i < finite;
  if (ok) {
    const size_t src = 3 * static_cast<size_t>(order[i]);
    p = make_float4(pts[src], pts[src + 1], pts[src + 2], 0.f);
  }
  if (i < n) sorted[i] = p;
  const float4 lo = make_float4(wave_min(ok ? p.x : inf), wave_min(ok ? p.y : inf), wave_min(ok ? p.z : inf), 0.f);
  const float4 hi = make_float4(wave_max(ok ? p.x : -inf), wave_max(ok ? p.y : -inf), wave_max(ok ? p.z : -inf), 0.f);
  if ((threadIdx.x & 63) == 0) {
    chunk_lo[chunk] = lo;
    chunk_hi[chunk] = hi;
  }
}

__global__ __launch_bounds__(256) void super_kernel(const float4* __restrict__ chunk_lo,
                                                    const float4* __restrict__ chunk_hi, uint32_t chunks,
                                                    float4* __restrict__ super_lo, float4* __restrict__ super_hi,
                                                    uint32_t supers) {
  const uint32_t s = blockIdx.x * 4 + threadIdx.x / 64;
  if (s >= supers) return;
  const uint32_t c = s * kSuper + (threadIdx.x & 63);
  const float inf = __uint_as_float(0x7f800000u);
  float4 lo = make_float4(inf, inf, inf, 0.f), hi = make_float4(-inf, -inf, -inf, 0.f);
  if (c < chunks) {
    lo = chunk_lo[c];
    hi = chunk_hi[c];
  }
  lo = make_float4(wave_min(lo.x), wave_min(lo.y), wave_min(lo.z), 0.f);
  hi = make_float4(wave_max(hi.x), wave_max(hi.y), wave_max(hi.z), 0.f);
  if ((threadIdx.x & 63) == 0) {
    super_lo[s] = lo;
    super_hi[s] = hi;
  }
}

struct Best {
  float x = kFar, y = kFar, z = kFar;  // the three smallest so far, smallest first
  // The reference's insertion (nearestNeighbors.wgsl:27-40) as min/max; d2 is
  // never NaN here (finite query, finite or +inf candidate).
  __device__ void insert(float d2) {
    z = fminf(z, fmaxf(y, d2));
    y = fminf(y, fmaxf(x, d2));
    x = fminf(x, d2);
  }
};

// Every lane's distances to the 64 points of chunk `c` (wave-private LDS slot).
template <bool kSelf>
__device__ inline void scan_chunk(const float4* __restrict__ sorted, uint32_t n, uint32_t c, float4 q, Best& best,
                                  float4* slot, uint32_t lane) {
  const float inf = __uint_as_float(0x7f800000u);
  const uint32_t i = c * kChunk + lane;
  slot[lane] = i < n ? sorted[i] : make_float4(inf, inf, inf, 0.f);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 8
  for (uint32_t j = 0; j < kChunk; j++) {
    const float4 p = slot[j];
    float d2 = distance2(p.x - q.x, p.y - q.y, p.z - q.z);
    if (kSelf && j == lane) d2 = inf;  // the reference skips the point itself by index
    best.insert(d2);
  }
  // The slot is rewritten by the next chunk only after every lane has read it.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline bool any_lane(bool v) { return __ballot(v) != 0; }

__global__ __launch_bounds__(256) void nearest_kernel(const float4* __restrict__ sorted,
                                                      const uint32_t* __restrict__ order, uint32_t n,
                                                      const uint32_t* __restrict__ finite_count,
                                                      const float4* __restrict__ chunk_lo,
                                                      const float4* __restrict__ chunk_hi, uint32_t chunks,
                                                      const float4* __restrict__ super_lo,
                                                      const float4* __restrict__ super_hi, uint32_t supers,
                                                      float* __restrict__ mean_distance2) {
  __shared__ float4 s_slot[4][kChunk];
  const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const uint32_t own = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  if (own >= chunks) return;
  const uint32_t i = own * kChunk + lane;
  const uint32_t finite = *finite_count;
  const float inf = __uint_as_float(0x7f800000u);
  const bool active = i < finite;
  if (own * kChunk >= finite) {  // only non-finite points here
    if (i < n) mean_distance2[order[i]] = inf;
    return;
  }
  const float4 q = active ? sorted[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  Best best;
  if (!active) best.z = -inf;  // never asks for a box
  float4* slot = s_slot[wave];
  scan_chunk<true>(sorted, n, own, q, best, slot, lane);
  const uint32_t own_super = own / kSuper;
  // Super-chunks outward from the wave's own: own, +1, -1, +2, -2, ...
  for (uint32_t step = 0; step < 2 * supers; step++) {
    const uint32_t k = (step + 1) / 2;
    uint32_t s;
    if (step % 2 == 0) {
      if (own_super + k >= supers) continue;
      s = own_super + k;
    } else {
      if (k > own_super) continue;
      s = own_super - k;
    }
    if (!any_lane(box_bound(q, super_lo[s], super_hi[s]) < best.z)) continue;
    const uint32_t first = s * kSuper, last = min(first + kSuper, chunks);
    for (uint32_t c = first; c < last; c++) {
      if (c == own) continue;
      if (!any_lane(box_bound(q, chunk_lo[c], chunk_hi[c]) < best.z)) continue;
      scan_chunk<false>(sorted, n, c, q, best, slot, lane);
    }
  }
  if (i < n) mean_distance2[order[i]] = active ? (best.x + best.y + best.z) / 3.f : inf;
}

constexpr double kShC0 = 0.28209479177387814;

// src/loadDataset.ts:93-110, one thread per output float: positions, log scale
// 0.5 log(max(d2, 1e-7)) on all axes, opacity logit log(0.1 / 0.9), quaternion
// (1, 0, 0, 0), dc (rgb / 255 - 0.5) / SH_C0; the arithmetic in double, rounded once.
__global__ __launch_bounds__(256) void initial_gaussians_kernel(const float* __restrict__ pts,
                                                                const uint8_t* __restrict__ colors,
                                                                const float* __restrict__ d2, uint32_t n,
                                                                float* __restrict__ params) {
  const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= static_cast<size_t>(n) * kFloatsPerGaussian) return;
  const size_t i = e / kFloatsPerGaussian;
  const uint32_t k = e % kFloatsPerGaussian;
  float v = 0.f;
  if (k < 3) {
    v = pts[3 * i + k];
  } else if (k == 3) {
    v = static_cast<float>(log(0.1 / 0.9));
  } else if (k < 7) {
    v = static_cast<float>(0.5 * log(fmax(static_cast<double>(d2[i]), 1e-7)));
  } else if (k == 8) {
    v = 1.f;
  } else if (k >= 12 && k < 15) {
    v = static_cast<float>((colors[3 * i + (k - 12)] / 255.0 - 0.5) / kShC0);
  }
  params[e] = v;
}


// The workspace carved into its arrays (each 256-byte aligned); with a null base
// only `bytes` is meaningful.
struct NeighborsLayout {
  uint32_t chunks, supers;
  uint32_t *finite, *code, *keys, *values, *alt_keys, *alt_values;
  void* sort;
  float4 *sorted, *chunk_lo, *chunk_hi, *super_lo, *super_hi;
  size_t bytes;
  NeighborsLayout(void* base, uint32_t n) {
    chunks = ceil_div(n, kChunk);
    supers = ceil_div(chunks, kSuper);
    char* p = static_cast<char*>(base);
    size_t at = 0;
    auto take = [&](size_t b) {
      void* q = p ? p + at : nullptr;
      at += (b + 255) & ~static_cast<size_t>(255);
      return q;
    };
    finite = static_cast<uint32_t*>(take(4));
    code = static_cast<uint32_t*>(take(4 * static_cast<size_t>(n)));
    keys = static_cast<uint32_t*>(take(4 * static_cast<size_t>(n)));
    values = static_cast<uint32_t*>(take(4 * static_cast<size_t>(n)));
    alt_keys = static_cast<uint32_t*>(take(4 * static_cast<size_t>(n)));
    alt_values = static_cast<uint32_t*>(take(4 * static_cast<size_t>(n)));
    sort = take(sort_workspace_bytes(n));
    sorted = static_cast<float4*>(take(16 * static_cast<size_t>(n)));
    chunk_lo = static_cast<float4*>(take(16 * static_cast<size_t>(chunks)));
    chunk_hi = static_cast<float4*>(take(16 * static_cast<size_t>(chunks)));
    super_lo = static_cast<float4*>(take(16 * static_cast<size_t>(supers)));
    super_hi = static_cast<float4*>(take(16 * static_cast<size_t>(supers)));
    bytes = at;
  }
};

}  // namespace

size_t neighbors_workspace_bytes(uint32_t n) { return NeighborsLayout(nullptr, n).bytes; }

void mean_neighbor_distance2(hipStream_t s, const float* pts, uint32_t n, float* out, void* workspace) {
  if (n == 0) return;
  const NeighborsLayout w(workspace, n);
  const uint32_t chunks = w.chunks, supers = w.supers;
  const SortWorkspace ws = sort_workspace(w.sort, n);
  const uint32_t blocks = ceil_div(n, 256);
  fill_u32(s, w.finite, 0u, 1);
  count_finite_kernel<<<blocks, 256, 0, s>>>(pts, n, w.finite);
  DIO_LAUNCHED("count_finite_kernel");
  for (uint32_t axis = 0; axis < 3; axis++) {
    axis_keys_kernel<<<blocks, 256, 0, s>>>(pts, n, axis, w.keys, w.values);
    DIO_LAUNCHED("axis_keys_kernel");
    const bool alt = radix_sort(s, false, w.keys, w.values, w.alt_keys, w.alt_values, nullptr, n, 0, 32, ws);
    rank_cells_kernel<<<blocks, 256, 0, s>>>(alt ? w.alt_values : w.values, n, w.finite, axis, w.code);
    DIO_LAUNCHED("rank_cells_kernel");
  }
  iota_kernel<<<blocks, 256, 0, s>>>(n, w.values);
  DIO_LAUNCHED("iota_kernel");
  const bool alt = radix_sort(s, false, w.code, w.values, w.alt_keys, w.alt_values, nullptr, n, 0, 31, ws);
  const uint32_t* order = alt ? w.alt_values : w.values;
  gather_kernel<<<ceil_div(chunks, 4), 256, 0, s>>>(pts, order, n, w.finite, w.sorted, w.chunk_lo, w.chunk_hi,
                                                    chunks);
  DIO_LAUNCHED("gather_kernel");
  super_kernel<<<ceil_div(supers, 4), 256, 0, s>>>(w.chunk_lo, w.chunk_hi, chunks, w.super_lo, w.super_hi, supers);
  DIO_LAUNCHED("super_kernel");
  nearest_kernel<<<ceil_div(chunks, 4), 256, 0, s>>>(w.sorted, order, n, w.finite, w.chunk_lo, w.chunk_hi, chunks,
                                                     w.super_lo, w.super_hi, supers, out);
  DIO_LAUNCHED("nearest_kernel");
}

void initial_gaussians(hipStream_t s, const float* pts, const uint8_t* colors, const float* d2, uint32_t n,
                       float* params, float* sh) {
  if (n == 0) return;
  const size_t total = static_cast<size_t>(n) * kFloatsPerGaussian;
  initial_gaussians_kernel<<<static_cast<uint32_t>((total + 255) / 256), 256, 0, s>>>(pts, colors, d2, n, params);
  DIO_LAUNCHED("initial_gaussians_kernel");
  fill_u32(s, reinterpret_cast<uint32_t*>(sh), 0u, static_cast<size_t>(n) * kShRest);
}

}  // namespace dio
