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
#include "gaussian_math.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;

// Copies `count` floats between global memory and LDS with the whole workgroup,
// 16 bytes per lane per step (both sides 16-byte aligned): the 45-float SH records
// are 180 bytes apart, so per-thread access would touch a new cache line per lane.
__device__ inline void block_copy(float* __restrict__ dst, const float* __restrict__ src, uint32_t count) {
  const uint32_t quads = count / 4;
  for (uint32_t q = threadIdx.x; q < quads; q += kThreads)
    reinterpret_cast<float4*>(dst)[q] = reinterpret_cast<const float4*>(src)[q];
  for (uint32_t k = quads * 4 + threadIdx.x; k < count; k += kThreads) dst[k] = src[k];
}

// ------------------------------------------------------------ preprocess --

// preprocess.wgsl:20-87 plus count_tiles (tiles.wgsl:23-34) fused: one thread per
// Gaussian writes its splat, its tile count and its depth-sort key. A workgroup
// stages its 256 Gaussians' SH records in LDS when any of them is visible.
__global__ __launch_bounds__(kThreads) void preprocess_kernel(const float4* __restrict__ gaussians,
                                                               const float* __restrict__ sh, uint32_t sh_degree,
                                                               uint32_t n, CameraArgs cam, float4* __restrict__ splats,
                                                               uint32_t* __restrict__ tile_counts,
                                                               uint32_t* __restrict__ depth_keys,
                                                               uint32_t* __restrict__ depth_values) {
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
  if (count > 0 && __syncthreads_or(visible))
    block_copy(s_sh, sh + static_cast<size_t>(first) * kShRest, min(static_cast<uint32_t>(kThreads), n - first) * kShRest);
  __syncthreads();
  if (!active) return;
  float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0, s2 = s0;
  uint32_t touched = 0, key = 0xffffffffu;
  if (visible) {
    float color[3] = {g[12], g[13], g[14]};
    if (count > 0) {
      float cp[3];
      camera_position(cam, cp);
      float ox = g[0] - cp[0], oy = g[1] - cp[1], oz = g[2] - cp[2];
      float l = sqrtf(ox * ox + oy * oy + oz * oz);
      float b[15];
      sh_basis(ox / l, oy / l, oz / l, b);
      const float* c = s_sh + threadIdx.x * kShRest;
      for (int k = 0; k < count; k++) {
        color[0] += b[k] * c[3 * k];
        color[1] += b[k] * c[3 * k + 1];
        color[2] += b[k] * c[3 * k + 2];
      }
    }
    s0 = make_float4(p.mean[0], p.mean[1], p.depth, p.radius);
    s1 = make_float4(p.c / p.det, -p.b / p.det, p.a / p.det, 0.f);
    s2 = make_float4(fmaxf(color[0], 0.f), fmaxf(color[1], 0.f), fmaxf(color[2], 0.f), g[3]);
    const uint32_t tiles_x = ceil_div(static_cast<uint32_t>(cam.viewport[0]), kTile);
    const uint32_t tiles_y = ceil_div(static_cast<uint32_t>(cam.viewport[1]), kTile);
    uint32_t r[4];
    tile_rect(p.mean[0], p.mean[1], p.radius, tiles_x, tiles_y, r);
    touched = (r[2] - r[0]) * (r[3] - r[1]);
    key = __float_as_uint(p.depth);
  }
  splats[i * 3] = s0;
  splats[i * 3 + 1] = s1;
  splats[i * 3 + 2] = s2;
  tile_counts[i] = touched;
  depth_keys[i] = key;
  depth_values[i] = i;
}

// count_tiles (tiles.wgsl:23-34) for splats that did not come from the preprocess.
__global__ __launch_bounds__(kThreads) void count_tiles_kernel(const float4* __restrict__ splats, uint32_t n,
                                                                uint32_t tiles_x, uint32_t tiles_y,
                                                                uint32_t* __restrict__ tile_counts,
                                                                uint32_t* __restrict__ depth_keys,
                                                                uint32_t* __restrict__ depth_values) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float4 s0 = splats[i * 3];
  uint32_t touched = 0, key = 0xffffffffu;
  if (s0.w > 0.f) {
    uint32_t r[4];
    tile_rect(s0.x, s0.y, s0.w, tiles_x, tiles_y, r);
    touched = (r[2] - r[0]) * (r[3] - r[1]);
    key = __float_as_uint(s0.z);
  }
  tile_counts[i] = touched;
  depth_keys[i] = key;
  depth_values[i] = i;
}

// duplicate (tiles.wgsl:36-59), load-balanced over pairs rather than splats: each
// workgroup writes a fixed 2048-slot stretch of the pair buffers. It finds the
// depth-ordered splat owning its first slot with a 64-ary search of the offsets
// (one wavefront, a few rounds), then deals its slots out to threads from LDS
// windows of 256 splats. The nearest splats cover the most tiles and sit together
// in depth order, so splat-parallel work would pile onto a few waves.
constexpr uint32_t kDupPairs = 2048;

__global__ __launch_bounds__(kThreads) void duplicate_kernel(const float4* __restrict__ splats,
                                                              const uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ offsets,
                                                              const uint32_t* __restrict__ tile_counts, uint32_t n,
                                                              uint32_t tiles_x, uint32_t tiles_y, uint32_t capacity,
                                                              const uint32_t* __restrict__ pair_total,
                                                              uint32_t* __restrict__ pair_tiles,
                                                              uint32_t* __restrict__ pair_splats) {
  __shared__ uint32_t s_start[kThreads], s_count[kThreads], s_x0[kThreads], s_y0[kThreads], s_w[kThreads],
      s_splat[kThreads];
  __shared__ uint32_t s_first;
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const uint32_t total = min(*pair_total, capacity);  // slots past the capacity are dropped
  const uint32_t p0 = blockIdx.x * kDupPairs;
  if (p0 >= total) return;
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
    if (j < n) {
      splat = order[j];
      start = offsets[j];
      count = tile_counts[splat];
      if (count) {
        const float4 s0 = splats[splat * 3];
        tile_rect(s0.x, s0.y, s0.w, tiles_x, tiles_y, r);
      }
    }
    s_start[tid] = start;
    s_count[tid] = count;
    s_x0[tid] = r[0];
    s_y0[tid] = r[1];
    s_w[tid] = r[2] - r[0];
    s_splat[tid] = splat;
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
      pair_tiles[p] = (s_y0[lo] + row) * tiles_x + s_x0[lo] + (k - row * width);
      pair_splats[p] = s_splat[lo];
    }
    p_lo = max(p_lo, p_hi);
    j_base += kThreads;
    __syncthreads();
  }
}

// find_ranges (tiles.wgsl:64-84): tile boundaries in the sorted pairs. Ranges
// are cleared first, so tiles without pairs read [0, 0).
__global__ __launch_bounds__(kThreads) void find_ranges_kernel(const uint32_t* __restrict__ pair_tiles,
                                                                const uint32_t* __restrict__ pair_total,
                                                                uint32_t capacity, uint2* __restrict__ ranges) {
  const uint32_t count = min(*pair_total, capacity);
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= count) return;
  const uint32_t tile = pair_tiles[i];
  if (i == 0) {
    ranges[tile].x = 0;
  } else {
    const uint32_t previous = pair_tiles[i - 1];
    if (tile != previous) {
      ranges[previous].y = i;
      ranges[tile].x = i;
    }
  }
  if (i == count - 1) ranges[tile].y = count;
}

// -------------------------------------------------------------- rasterize --
//
// Both rasterizer kernels run one workgroup per 16x16 tile as four independent
// wavefronts, one per 8x8 quadrant (one pixel per lane). A wave stages 64 of the
// tile's splats at a time in registers (one per lane) and broadcasts each to the
// whole wave with v_readlane into scalar registers, so the blend loop reads no
// LDS and the forward pass needs no workgroup barrier at all. Each staged splat
// also gets the bounding box of its alpha >= 1/255 ellipse; a wave only visits
// the splats whose box meets its quadrant (a ballot mask), which changes no
// result: outside the box every pixel would skip the splat anyway.

constexpr int kChunk = kWave;
constexpr int kGrads = 9;

__device__ inline float lane_value(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}


// One splat of a chunk, staged in a lane's registers.
struct Staged {
  float mx, my, ca, cb, cc, op, r, g, b;
  float x0, x1, y0, y1;  // box holding every pixel centre where opacity * G >= 1/255
  uint32_t index;
};

__device__ inline void stage(const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, uint32_t k,
                             Staged& s) {
  s.index = pair_splats[k];
  const float4 a = splats[s.index * 3], b = splats[s.index * 3 + 1], c = splats[s.index * 3 + 2];
  s.mx = a.x; s.my = a.y;
  s.ca = b.x; s.cb = b.y; s.cc = b.z;
  s.r = c.x; s.g = c.y; s.b = c.z; s.op = c.w;
  // opacity * exp(-q) >= 1/255 needs q <= t = ln(255 opacity), with
  // q = (a dx² + c dy²)/2 + b dx dy. That ellipse's half-widths are sqrt(2t Σxx)
  // and sqrt(2t Σyy), Σ = conic⁻¹. Widened by 1% and 0.01 px against rounding.
  const float det = s.ca * s.cc - s.cb * s.cb;
  const float t = __logf(255.f * s.op) * 1.01f + 0.01f;
  float ex = sqrtf(2.f * t * s.cc / det) + 0.01f, ey = sqrtf(2.f * t * s.ca / det) + 0.01f;
  ex = !(det > 0.f) ? INFINITY : (!(t > 0.f) ? -INFINITY : ex);  // degenerate: never cull; too faint: always
  ey = !(det > 0.f) ? INFINITY : (!(t > 0.f) ? -INFINITY : ey);
  s.x0 = s.mx - ex;
  s.x1 = s.mx + ex;
  s.y0 = s.my - ey;
  s.y1 = s.my + ey;
}

__device__ inline bool meets(const Staged& s, float bx0, float bx1, float by0, float by1) {
  return s.x1 >= bx0 && s.x0 <= bx1 && s.y1 >= by0 && s.y0 <= by1;
}

// rasterize.wgsl:31-100: front-to-back blend; a wave stops once all its pixels
// are nearly opaque.
__global__ __launch_bounds__(256) void rasterize_kernel(const float4* __restrict__ splats,
                                                         const uint32_t* __restrict__ pair_splats,
                                                         const uint2* __restrict__ ranges, uint32_t width,
                                                         uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
                                                         float bg_b, uint32_t* __restrict__ target,
                                                         float4* __restrict__ pixel_color,
                                                         uint32_t* __restrict__ pixel_last) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint32_t tile = blockIdx.x;
  const uint32_t qx = (tile % tiles_x) * kTile + (wave & 1) * 8, qy = (tile / tiles_x) * kTile + (wave >> 1) * 8;
  const uint32_t px = qx + (lane & 7), py = qy + (lane >> 3);
  const bool inside = px < width && py < height;
  const float cx = static_cast<float>(px) + 0.5f, cy = static_cast<float>(py) + 0.5f;
  const float bx0 = static_cast<float>(qx) + 0.5f, bx1 = bx0 + 7.f;
  const float by0 = static_cast<float>(qy) + 0.5f, by1 = by0 + 7.f;
  const uint2 run = ranges[tile];
  float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f;
  bool done = !inside;
  uint32_t last = 0;
  for (uint32_t base = run.x; base < run.y; base += kChunk) {
    if (__all(done)) break;
    Staged s;
    bool hit = false;
    if (base + lane < run.y) {
      stage(splats, pair_splats, base + lane, s);
      hit = meets(s, bx0, bx1, by0, by1);
    }
    uint64_t mask = __ballot(hit);
    while (mask) {
      const int j = __builtin_ctzll(mask);
      mask &= mask - 1;
      const float mx = lane_value(s.mx, j), my = lane_value(s.my, j);
      const float ca = lane_value(s.ca, j), cbb = lane_value(s.cb, j), cc = lane_value(s.cc, j);
      const float op = lane_value(s.op, j);
      const float sr = lane_value(s.r, j), sg = lane_value(s.g, j), sb = lane_value(s.b, j);
      if (!done) {
        const float dx = cx - mx, dy = cy - my;
        const float power = -0.5f * (ca * dx * dx + cc * dy * dy) - cbb * dx * dy;
        if (power <= 0.f) {
          const float alpha = fminf(0.99f, op * __expf(power));
          if (alpha >= 1.f / 255.f) {
            const float remaining = T * (1.f - alpha);
            if (remaining < 0.0001f) {
              done = true;
            } else {
              const float w = alpha * T;
              cr += sr * w;
              cg += sg * w;
              cb += sb * w;
              T = remaining;
              last = base + j - run.x + 1;
            }
          }
        }
      }
    }
  }
  if (inside) {
    const float r = cr + T * bg_r, g = cg + T * bg_g, b = cb + T * bg_b;
    const uint32_t index = py * width + px;
    pixel_color[index] = make_float4(r, g, b, T);
    pixel_last[index] = last;
    if (target) {
      auto u8 = [](float v) { return static_cast<uint32_t>(fminf(fmaxf(v, 0.f), 1.f) * 255.f + 0.5f); };
      target[index] = u8(r) | (u8(g) << 8) | (u8(b) << 16) | (255u << 24);
    }
  }
}

// ------------------------------------------------------ rasterize backward --

template <int kCtrl>
__device__ inline float dpp_add(float v) {
  return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xf, 0xf,
                                                                   false));
}

// Sum over each 16-lane row with DPP; every lane of the row gets the row total.
__device__ inline float row_total(float v) {
  v = dpp_add<0xb1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4e>(v);   // quad_perm [2,3,0,1]: quad sums
  v = dpp_add<0x141>(v);  // row_half_mirror: 8-lane sums
  v = dpp_add<0x140>(v);  // row_mirror: row sums
  return v;
}

// Four splats' per-lane values (a, b, c, d) to four wave totals at once: a
// butterfly across lane halves keeps a|c, b|d by half, one across 16-lane rows
// keeps a|b|c|d by row, so one row reduction finishes all four: row r of the
// wave ends with splat r's total.
__device__ inline float four_totals(float a, float b, float c, float d) {
  const int lane = __lane_id();
  const bool lower = lane < 32;
  const float lo = (lower ? a : c) + __shfl_xor(lower ? c : a, 32, kWave);
  const float hi = (lower ? b : d) + __shfl_xor(lower ? d : b, 32, kWave);
  const bool even_row = (lane & 16) == 0;
  return row_total((even_row ? lo : hi) + __shfl_xor(even_row ? hi : lo, 16, kWave));
}

__device__ inline uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kWave)));
  return v;
}

// The per-pixel replay state of the backward pass (rasterizeBackward.wgsl:161-177).
struct Replay {
  float T, final_T, bg_grad, gr, gg, gb;
  float behind_r, behind_g, behind_b;
  float next_alpha, next_r, next_g, next_b;
};

// One splat's 9 gradient shares at one pixel (rasterizeBackward.wgsl:199-230),
// written without branches: a pixel the splat does not blend into (power > 0,
// alpha < 1/255, or past the pixel's last splat) runs the same instructions with
// alpha = 0, which leaves T, the colour behind and the next splat as they were and
// zeroes every share. The blended lanes compute exactly the branchy form.
__device__ inline bool replay_splat(Replay& st, bool blended, float cx, float cy, float mx, float my, float ca,
                                    float cbb, float cc, float op, float sr, float sg, float sb, float share[kGrads]) {
  const float dx = cx - mx, dy = cy - my;
  const float power = -0.5f * (ca * dx * dx + cc * dy * dy) - cbb * dx * dy;
  const float G = __expf(power);
  const float clamped = fminf(0.99f, op * G);
  const bool valid = blended && power <= 0.f && clamped >= 1.f / 255.f;
  const float alpha = valid ? clamped : 0.f;
  const float keep = __builtin_amdgcn_rcpf(1.f - alpha);  // one hardware reciprocal for both divisions; 1 when idle
  st.T *= keep;
  const float na = valid ? st.next_alpha : 0.f;
  st.behind_r = na * st.next_r + (1.f - na) * st.behind_r;
  st.behind_g = na * st.next_g + (1.f - na) * st.behind_g;
  st.behind_b = na * st.next_b + (1.f - na) * st.behind_b;
  st.next_alpha = valid ? alpha : st.next_alpha;
  st.next_r = valid ? sr : st.next_r;
  st.next_g = valid ? sg : st.next_g;
  st.next_b = valid ? sb : st.next_b;
  const float alpha_grad = st.T * ((sr - st.behind_r) * st.gr + (sg - st.behind_g) * st.gg + (sb - st.behind_b) * st.gb) -
                           st.final_T * keep * st.bg_grad;
  const float power_grad = valid ? alpha_grad * op * G : 0.f;
  share[0] = power_grad * (ca * dx + cbb * dy);
  share[1] = power_grad * (cbb * dx + cc * dy);
  share[2] = power_grad * (-0.5f * dx * dx);
  share[3] = power_grad * (-dx * dy);
  share[4] = power_grad * (-0.5f * dy * dy);
  share[5] = valid ? alpha_grad * G : 0.f;
  const float at = alpha * st.T;
  share[6] = at * st.gr;
  share[7] = at * st.gg;
  share[8] = at * st.gb;
  return valid;
}

// rasterizeBackward.wgsl:49-163: each pixel replays its list back to front from
// its last blended splat. A wave takes the splats that touch its quadrant four at
// a time, reduces their 9 gradients across its 64 pixels together, and its row
// leaders park the totals in LDS; per 64-splat chunk the four quadrant waves add
// theirs and one float atomic per (splat, tile, gradient) goes out.
__global__ __launch_bounds__(256) void rasterize_backward_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g, float bg_b,
    const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads) {
  __shared__ float s_sums[4][kGrads][kChunk];
  __shared__ uint32_t s_index[kChunk];
  __shared__ uint32_t s_last[4];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile = blockIdx.x;
  const uint32_t qx = (tile % tiles_x) * kTile + (wave & 1) * 8, qy = (tile / tiles_x) * kTile + (wave >> 1) * 8;
  const uint32_t px = qx + (lane & 7), py = qy + (lane >> 3);
  const bool inside = px < width && py < height;
  const float cx = static_cast<float>(px) + 0.5f, cy = static_cast<float>(py) + 0.5f;
  const float bx0 = static_cast<float>(qx) + 0.5f, bx1 = bx0 + 7.f;
  const float by0 = static_cast<float>(qy) + 0.5f, by1 = by0 + 7.f;
  const uint2 run = ranges[tile];
  Replay st{};
  uint32_t last = 0;
  if (inside) {
    const uint32_t index = py * width + px;
    st.final_T = pixel_color[index].w;
    last = pixel_last[index];
    const float4 pg = pixel_grads[index];
    st.gr = pg.x; st.gg = pg.y; st.gb = pg.z;
  }
  st.T = st.final_T;
  st.bg_grad = bg_r * st.gr + bg_g * st.gg + bg_b * st.gb;
  const uint32_t my_last = wave_max(last);
  if (lane == 0) s_last[wave] = my_last;
  __syncthreads();
  const uint32_t tile_last = max(max(s_last[0], s_last[1]), max(s_last[2], s_last[3]));
  const int row = lane >> 4;
  const bool row_leader = (lane & 15) == 0;

  for (int c = static_cast<int>(ceil_div(tile_last, kChunk)) - 1; c >= 0; c--) {
    const uint32_t base = run.x + static_cast<uint32_t>(c) * kChunk;
    Staged s;
    bool hit = false;
    s.index = 0;
    if (base + lane < run.y) {
      stage(splats, pair_splats, base + lane, s);
      hit = base + lane - run.x < my_last && meets(s, bx0, bx1, by0, by1);
    }
    if (wave == 0) s_index[lane] = s.index;
#pragma unroll
    for (int v = 0; v < kGrads; v++) s_sums[wave][v][lane] = 0.f;
    uint64_t mask = __ballot(hit);
    while (mask) {
      // Up to four splats, back to front; missing ones contribute zeros.
      float share[4][kGrads];
      int js[4];
      bool any = false;
#pragma unroll
      for (int g = 0; g < 4; g++) {
        js[g] = -1;
        bool blended = false;
        float mx = 0.f, my = 0.f, ca = 0.f, cbb = 0.f, cc = 0.f, op = 0.f, sr = 0.f, sg = 0.f, sb = 0.f;
        if (mask) {
          const int j = kWave - 1 - __builtin_clzll(mask);
          mask &= ~(1ull << j);
          js[g] = j;
          mx = lane_value(s.mx, j); my = lane_value(s.my, j);
          ca = lane_value(s.ca, j); cbb = lane_value(s.cb, j); cc = lane_value(s.cc, j);
          op = lane_value(s.op, j);
          sr = lane_value(s.r, j); sg = lane_value(s.g, j); sb = lane_value(s.b, j);
          blended = base + j - run.x < last;
        }
        any |= replay_splat(st, blended, cx, cy, mx, my, ca, cbb, cc, op, sr, sg, sb, share[g]);
      }
      if (!__any(any)) continue;
      const int mine = row == 0 ? js[0] : (row == 1 ? js[1] : (row == 2 ? js[2] : js[3]));
#pragma unroll
      for (int v = 0; v < kGrads; v++) {
        const float total = four_totals(share[0][v], share[1][v], share[2][v], share[3][v]);
        if (row_leader && mine >= 0) s_sums[wave][v][mine] = total;
      }
    }
    __syncthreads();
    for (int e = tid; e < kGrads * kChunk; e += 256) {
      const int v = e / kChunk, slot = e % kChunk;
      const float total = s_sums[0][v][slot] + s_sums[1][v][slot] + s_sums[2][v][slot] + s_sums[3][v][slot];
      if (total != 0.f) unsafeAtomicAdd(splat_grads + static_cast<size_t>(s_index[slot]) * kGrads + v, total);
    }
    __syncthreads();
  }
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
    if (count > 0) block_copy(s_sh, sh + static_cast<size_t>(first) * kShRest, block_n * kShRest);
    block_copy(s_grad, splat_grads + static_cast<size_t>(first) * kGrads, block_n * kGrads);
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
  const float d = p.depth, d2 = d * d, d3 = d2 * d;
  const float a = p.a, b = p.b, c = p.c, det = p.det, det2 = det * det;
  // 1. conic -> 2D covariance
  const float ga = (-c * c * gr[2] + b * c * gr[3] - b * b * gr[4]) / det2;
  const float gb = (2.f * b * c * gr[2] - (det + 2.f * b * b) * gr[3] + 2.f * a * b * gr[4]) / det2;
  const float gc = (-b * b * gr[2] + a * b * gr[3] - a * a * gr[4]) / det2;
  const float G2[2][2] = {{ga, 0.5f * gb}, {0.5f * gb, gc}};
  // 2. dL/dcov3d = JWᵀ G JW, dL/dJ = 2 G JW cov3d Wᵀ
  float cov3d_grad[3][3];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++)
      cov3d_grad[r][k] = p.jw[0][r] * (G2[0][0] * p.jw[0][k] + G2[0][1] * p.jw[1][k]) +
                         p.jw[1][r] * (G2[1][0] * p.jw[0][k] + G2[1][1] * p.jw[1][k]);
  float gjw[2][3], gjwc[2][3], j_grad[2][3];
  for (int u = 0; u < 2; u++)
    for (int k = 0; k < 3; k++) gjw[u][k] = G2[u][0] * p.jw[0][k] + G2[u][1] * p.jw[1][k];
  for (int u = 0; u < 2; u++)
    for (int k = 0; k < 3; k++)
      gjwc[u][k] = gjw[u][0] * p.cov3d[0][k] + gjw[u][1] * p.cov3d[1][k] + gjw[u][2] * p.cov3d[2][k];
  const float* v = cam.view;
  for (int u = 0; u < 2; u++)
    for (int k = 0; k < 3; k++)  // (X Wᵀ)[u][k] = Σ_e X[u][e] W[k][e], W[k][e] = view[e * 4 + k]
      j_grad[u][k] = 2.f * (gjwc[u][0] * v[k] + gjwc[u][1] * v[4 + k] + gjwc[u][2] * v[8 + k]);
  // 3. J -> view-space centre, where its evaluation point was not clamped
  const float free_x = fabsf(p.ratio[0]) > p.limit[0] ? 0.f : 1.f;
  const float free_y = fabsf(p.ratio[1]) > p.limit[1] ? 0.f : 1.f;
  const float fx = p.focal[0], fy = p.focal[1];
  float t_grad[3];
  t_grad[0] = j_grad[0][2] * free_x * fx / d2;
  t_grad[1] = j_grad[1][2] * free_y * -fy / d2;
  const float depth_grad = j_grad[0][0] * (-fx / d2) + j_grad[1][1] * (fy / d2) +
                           j_grad[0][2] * (-(1.f + free_x) * fx * p.tx / d3) +
                           j_grad[1][2] * ((1.f + free_y) * fy * p.ty / d3);
  t_grad[2] = -depth_grad;
  // 4. the 2D centre through the projection
  const float ndc_x = gr[0] * 0.5f * cam.viewport[0], ndc_y = gr[1] * -0.5f * cam.viewport[1];
  const float cw = p.clip[3];
  const float clip_grad[4] = {ndc_x / cw, ndc_y / cw, 0.f, -(ndc_x * p.clip[0] + ndc_y * p.clip[1]) / (cw * cw)};
  const float* P = cam.proj;
  for (int k = 0; k < 3; k++)
    t_grad[k] += P[k * 4] * clip_grad[0] + P[k * 4 + 1] * clip_grad[1] + P[k * 4 + 2] * clip_grad[2] +
                 P[k * 4 + 3] * clip_grad[3];
  // 5. colour through the SH, masked where the forward clamp cut it
  float cp[3];
  camera_position(cam, cp);
  const float ox = g[0] - cp[0], oy = g[1] - cp[1], oz = g[2] - cp[2];
  const float dist = sqrtf(ox * ox + oy * oy + oz * oz);
  const float dx = ox / dist, dy = oy / dist, dz = oz / dist;
  const int count = sh_count(sh_degree);
  float basis[15];
  sh_basis(dx, dy, dz, basis);
  float* coeff = row;
  float raw[3] = {g[12], g[13], g[14]};
  for (int k = 0; k < count; k++) {
    raw[0] += basis[k] * coeff[3 * k];
    raw[1] += basis[k] * coeff[3 * k + 1];
    raw[2] += basis[k] * coeff[3 * k + 2];
  }
  const float raw_grad[3] = {raw[0] < 0.f ? 0.f : gr[6], raw[1] < 0.f ? 0.f : gr[7], raw[2] < 0.f ? 0.f : gr[8]};
  float bgx[15], bgy[15], bgz[15];
  sh_basis_gradient(dx, dy, dz, bgx, bgy, bgz);
  float dgx = 0.f, dgy = 0.f, dgz = 0.f;
  for (int k = 0; k < static_cast<int>(kShRest) / 3; k++) {
    if (k < count) {
      const float dot = coeff[3 * k] * raw_grad[0] + coeff[3 * k + 1] * raw_grad[1] + coeff[3 * k + 2] * raw_grad[2];
      dgx += bgx[k] * dot;
      dgy += bgy[k] * dot;
      dgz += bgz[k] * dot;
      coeff[3 * k] = basis[k] * raw_grad[0];
      coeff[3 * k + 1] = basis[k] * raw_grad[1];
      coeff[3 * k + 2] = basis[k] * raw_grad[2];
    } else {
      coeff[3 * k] = coeff[3 * k + 1] = coeff[3 * k + 2] = 0.f;
    }
  }
  const float dd = dx * dgx + dy * dgy + dz * dgz;
  // 6. position: t = W p + translation
  float pos_grad[3];
  const float dir[3] = {dx, dy, dz}, dirg[3] = {dgx, dgy, dgz};
  for (int e = 0; e < 3; e++)
    pos_grad[e] = v[e * 4] * t_grad[0] + v[e * 4 + 1] * t_grad[1] + v[e * 4 + 2] * t_grad[2] +
                  (dirg[e] - dir[e] * dd) / dist;
  // 7. M = R S -> scale and rotation
  float m_grad[3][3];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++)
      m_grad[r][k] = 2.f * (cov3d_grad[r][0] * p.m[0][k] + cov3d_grad[r][1] * p.m[1][k] + cov3d_grad[r][2] * p.m[2][k]);
  float scale_grad[3], rg[3][3];
  for (int k = 0; k < 3; k++) {
    scale_grad[k] = m_grad[0][k] * p.r[0][k] + m_grad[1][k] * p.r[1][k] + m_grad[2][k] * p.r[2][k];
    for (int r = 0; r < 3; r++) rg[r][k] = m_grad[r][k] * g[4 + k];
  }
  const float qw = p.q[0], qx = p.q[1], qy = p.q[2], qz = p.q[3];
  float qg[4];
  qg[0] = 2.f * (-qz * rg[0][1] + qy * rg[0][2] + qz * rg[1][0] - qx * rg[1][2] - qy * rg[2][0] + qx * rg[2][1]);
  qg[1] = 2.f * (qy * rg[0][1] + qz * rg[0][2] + qy * rg[1][0] - 2.f * qx * rg[1][1] - qw * rg[1][2] +
                 qz * rg[2][0] + qw * rg[2][1] - 2.f * qx * rg[2][2]);
  qg[2] = 2.f * (-2.f * qy * rg[0][0] + qx * rg[0][1] + qw * rg[0][2] + qx * rg[1][0] + qz * rg[1][2] -
                 qw * rg[2][0] + qz * rg[2][1] - 2.f * qy * rg[2][2]);
  qg[3] = 2.f * (-2.f * qz * rg[0][0] - qw * rg[0][1] + qx * rg[0][2] + qw * rg[1][0] - 2.f * qz * rg[1][1] +
                 qy * rg[1][2] + qx * rg[2][0] + qy * rg[2][1]);
  const float qdot = qw * qg[0] + qx * qg[1] + qy * qg[2] + qz * qg[3];
  const float inv_l = 1.f / p.q_length;
  out0 = make_float4(pos_grad[0], pos_grad[1], pos_grad[2], gr[5]);
  out1 = make_float4(scale_grad[0], scale_grad[1], scale_grad[2], 0.f);
  out2 = make_float4((qg[0] - qw * qdot) * inv_l, (qg[1] - qx * qdot) * inv_l, (qg[2] - qy * qdot) * inv_l,
                     (qg[3] - qz * qdot) * inv_l);
  out3 = make_float4(raw_grad[0], raw_grad[1], raw_grad[2], 0.f);
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
  block_copy(sh_grads + static_cast<size_t>(first) * kShRest, s_sh, block_n * kShRest);
}

}  // namespace

void preprocess(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, uint32_t n,
                const CameraArgs& cam, float* splats, uint32_t* tile_counts, uint32_t* depth_keys,
                uint32_t* depth_values) {
  if (n == 0) return;
  preprocess_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(gaussians), sh, sh_degree, n, cam, reinterpret_cast<float4*>(splats),
      tile_counts, depth_keys, depth_values);
  DIO_LAUNCHED("preprocess");
}

void count_tiles(hipStream_t s, const float* splats, uint32_t n, uint32_t width, uint32_t height,
                 uint32_t* tile_counts, uint32_t* depth_keys, uint32_t* depth_values) {
  if (n == 0) return;
  count_tiles_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(reinterpret_cast<const float4*>(splats), n,
                                                                 ceil_div(width, kTile), ceil_div(height, kTile),
                                                                 tile_counts, depth_keys, depth_values);
  DIO_LAUNCHED("count_tiles");
}

void duplicate(hipStream_t s, const float* splats, const uint32_t* order, const uint32_t* offsets,
               const uint32_t* tile_counts, uint32_t n, uint32_t width, uint32_t height, uint32_t capacity,
               const uint32_t* pair_total, uint32_t* pair_tiles, uint32_t* pair_splats) {
  if (n == 0 || capacity == 0) return;
  duplicate_kernel<<<ceil_div(capacity, kDupPairs), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(splats), order, offsets, tile_counts, n, ceil_div(width, kTile),
      ceil_div(height, kTile), capacity, pair_total, pair_tiles, pair_splats);
  DIO_LAUNCHED("duplicate");
}

void find_ranges(hipStream_t s, const uint32_t* pair_tiles, const uint32_t* pair_total, uint32_t capacity,
                 uint32_t tiles, uint32_t* ranges) {
  fill_u32(s, ranges, 0u, static_cast<size_t>(tiles) * 2);
  if (capacity == 0) return;
  find_ranges_kernel<<<ceil_div(capacity, kThreads), kThreads, 0, s>>>(pair_tiles, pair_total, capacity,
                                                                        reinterpret_cast<uint2*>(ranges));
  DIO_LAUNCHED("find_ranges");
}

void rasterize(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
               uint32_t width, uint32_t height, const float* background, uint32_t* target, float* pixel_color,
               uint32_t* pixel_last) {
  const uint32_t tx = ceil_div(width, kTile), ty = ceil_div(height, kTile);
  if (tx * ty == 0) return;
  rasterize_kernel<<<tx * ty, kThreads, 0, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                                reinterpret_cast<const uint2*>(ranges), width, height, tx,
                                                background[0], background[1], background[2], target,
                                                reinterpret_cast<float4*>(pixel_color), pixel_last);
  DIO_LAUNCHED("rasterize");
}

void rasterize_backward(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
                        uint32_t width, uint32_t height, const float* background, const float* pixel_color,
                        const uint32_t* pixel_last, const float* pixel_grads, float* splat_grads) {
  const uint32_t tx = ceil_div(width, kTile), ty = ceil_div(height, kTile);
  if (tx * ty == 0) return;
  rasterize_backward_kernel<<<tx * ty, kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(splats), pair_splats, reinterpret_cast<const uint2*>(ranges), width, height, tx,
      background[0], background[1], background[2], reinterpret_cast<const float4*>(pixel_color), pixel_last,
      reinterpret_cast<const float4*>(pixel_grads), splat_grads);
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
