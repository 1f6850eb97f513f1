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

// ------------------------------------------------------------ preprocess --

// preprocess.wgsl:20-87 plus count_tiles (tiles.wgsl:23-34) fused: one thread per
// Gaussian writes its splat, its tile count and its depth-sort key.
__global__ __launch_bounds__(kThreads) void preprocess_kernel(const float4* __restrict__ gaussians,
                                                               const float* __restrict__ sh, uint32_t sh_degree,
                                                               uint32_t n, CameraArgs cam, float4* __restrict__ splats,
                                                               uint32_t* __restrict__ tile_counts,
                                                               uint32_t* __restrict__ depth_keys,
                                                               uint32_t* __restrict__ depth_values) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float g[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    float4 v = gaussians[i * 4 + q];
    g[4 * q] = v.x; g[4 * q + 1] = v.y; g[4 * q + 2] = v.z; g[4 * q + 3] = v.w;
  }
  Projection p;
  const bool visible = project(g, cam, p);
  float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0, s2 = s0;
  uint32_t touched = 0, key = 0xffffffffu;
  if (visible) {
    float color[3] = {g[12], g[13], g[14]};
    const int count = sh ? sh_count(sh_degree) : 0;
    if (count > 0) {
      float cp[3];
      camera_position(cam, cp);
      float ox = g[0] - cp[0], oy = g[1] - cp[1], oz = g[2] - cp[2];
      float l = sqrtf(ox * ox + oy * oy + oz * oz);
      float b[15];
      sh_basis(ox / l, oy / l, oz / l, b);
      const float* c = sh + static_cast<size_t>(i) * kShRest;
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

// duplicate (tiles.wgsl:36-59), one wavefront per 64 depth-ordered splats: the
// wave's pairs are dealt out one per lane, so a splat covering many tiles does
// not serialise on one thread, and consecutive lanes write consecutive slots.
__global__ __launch_bounds__(kThreads) void duplicate_kernel(const float4* __restrict__ splats,
                                                              const uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ offsets,
                                                              const uint32_t* __restrict__ tile_counts, uint32_t n,
                                                              uint32_t tiles_x, uint32_t tiles_y, uint32_t capacity,
                                                              uint32_t* __restrict__ pair_tiles,
                                                              uint32_t* __restrict__ pair_splats) {
  __shared__ uint32_t s_end[kThreads];  // inclusive prefix of counts within the wave
  __shared__ uint32_t s_x0[kThreads], s_y0[kThreads], s_w[kThreads], s_splat[kThreads];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t j = blockIdx.x * kThreads + tid;
  uint32_t count = 0, splat = 0, r[4] = {0, 0, 0, 0};
  if (j < n) {
    splat = order[j];
    count = tile_counts[splat];
    if (count) {
      const float4 s0 = splats[splat * 3];
      tile_rect(s0.x, s0.y, s0.w, tiles_x, tiles_y, r);
    }
  }
  uint32_t inclusive = count;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(inclusive, o, kWave);
    if (lane >= o) inclusive += y;
  }
  const uint32_t wave_total = __shfl(inclusive, kWave - 1, kWave);
  s_end[tid] = inclusive;
  s_x0[tid] = r[0];
  s_y0[tid] = r[1];
  s_w[tid] = r[2] - r[0];
  s_splat[tid] = splat;
  __syncthreads();
  if (wave_total == 0) return;
  // Slots of this wave's pairs start at the offset of its first splat.
  const uint32_t wave_base = offsets[blockIdx.x * kThreads + wave * kWave];
  const int w0 = wave * kWave;
  for (uint32_t q = lane; q < wave_total; q += kWave) {
    int lo = 0, hi = kWave - 1;  // first lane whose inclusive end exceeds q
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (s_end[w0 + mid] > q) hi = mid; else lo = mid + 1;
    }
    const uint32_t own_count = (lo == 0 ? s_end[w0] : s_end[w0 + lo] - s_end[w0 + lo - 1]);
    const uint32_t k = q - (s_end[w0 + lo] - own_count);
    const uint32_t width = s_w[w0 + lo];
    const uint32_t tx = s_x0[w0 + lo] + k % width, ty = s_y0[w0 + lo] + k / width;
    const uint32_t slot = wave_base + q;
    if (slot < capacity) {
      pair_tiles[slot] = ty * tiles_x + tx;
      pair_splats[slot] = s_splat[w0 + lo];
    }
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

constexpr int kBatch = 256;

// rasterize.wgsl:31-100: one workgroup per 16x16 tile, splats staged 256 at a time
// in LDS, front-to-back blend per pixel, and the whole tile stops once every
// pixel is nearly opaque (one barrier-count per batch).
__global__ __launch_bounds__(kThreads) void rasterize_kernel(const float4* __restrict__ splats,
                                                              const uint32_t* __restrict__ pair_splats,
                                                              const uint2* __restrict__ ranges, uint32_t width,
                                                              uint32_t height, uint32_t tiles_x, float bg_r, float bg_g,
                                                              float bg_b, uint32_t* __restrict__ target,
                                                              float4* __restrict__ pixel_color,
                                                              uint32_t* __restrict__ pixel_last) {
  __shared__ float4 s_geom[kBatch];   // mean x, mean y, conic a, conic b
  __shared__ float4 s_shade[kBatch];  // conic c, opacity, r, g
  __shared__ float s_blue[kBatch];
  const int tid = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const uint32_t px = (tile % tiles_x) * kTile + (tid % kTile);
  const uint32_t py = (tile / tiles_x) * kTile + (tid / kTile);
  const bool inside = px < width && py < height;
  const float cx = static_cast<float>(px) + 0.5f, cy = static_cast<float>(py) + 0.5f;
  const uint2 run = ranges[tile];
  float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f;
  bool done = !inside;
  uint32_t last = 0;
  for (uint32_t start = run.x; start < run.y; start += kBatch) {
    if (__syncthreads_count(done) == kThreads) break;
    const uint32_t k = start + tid;
    if (k < run.y) {
      const uint32_t idx = pair_splats[k];
      const float4 a = splats[idx * 3], b = splats[idx * 3 + 1], c = splats[idx * 3 + 2];
      s_geom[tid] = make_float4(a.x, a.y, b.x, b.y);
      s_shade[tid] = make_float4(b.z, c.w, c.x, c.y);
      s_blue[tid] = c.z;
    }
    __syncthreads();
    const int batch = static_cast<int>(min(static_cast<uint32_t>(kBatch), run.y - start));
    if (!done) {
      for (int j = 0; j < batch; j++) {
        const float4 ge = s_geom[j];
        const float4 sh = s_shade[j];
        const float dx = cx - ge.x, dy = cy - ge.y;
        const float power = -0.5f * (ge.z * dx * dx + sh.x * dy * dy) - ge.w * dx * dy;
        if (power > 0.f) continue;
        const float alpha = fminf(0.99f, sh.y * __expf(power));
        if (alpha < 1.f / 255.f) continue;
        const float remaining = T * (1.f - alpha);
        if (remaining < 0.0001f) {
          done = true;
          break;
        }
        const float w = alpha * T;
        cr += sh.z * w;
        cg += sh.w * w;
        cb += s_blue[j] * w;
        T = remaining;
        last = start + j - run.x + 1;
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

constexpr int kBackBatch = 64;
constexpr int kGrads = 9;

__device__ inline float wave_sum(float x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// rasterizeBackward.wgsl:49-163: each pixel replays its list back to front from
// its last blended splat; a splat's 9 gradients are summed across each wavefront
// with shuffles (skipped when no lane of the wave touched it), the 4 wave sums
// are added in LDS once per 64-splat batch, and one hardware float atomic per
// (splat, tile, gradient) goes to global memory.
__global__ __launch_bounds__(kThreads) void rasterize_backward_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    uint32_t width, uint32_t height, uint32_t tiles_x, float bg_r, float bg_g, float bg_b,
    const float4* __restrict__ pixel_color, const uint32_t* __restrict__ pixel_last,
    const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads) {
  __shared__ float4 s_geom[kBackBatch];   // mean x, mean y, conic a, conic b
  __shared__ float4 s_shade[kBackBatch];  // conic c, opacity, r, g
  __shared__ float s_blue[kBackBatch];
  __shared__ uint32_t s_index[kBackBatch];
  __shared__ float s_sums[kThreads / kWave][kBackBatch][kGrads];
  __shared__ uint32_t s_max_last;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile = blockIdx.x;
  const uint32_t px = (tile % tiles_x) * kTile + (tid % kTile);
  const uint32_t py = (tile / tiles_x) * kTile + (tid / kTile);
  const bool inside = px < width && py < height;
  const float cx = static_cast<float>(px) + 0.5f, cy = static_cast<float>(py) + 0.5f;
  const uint2 run = ranges[tile];
  float final_T = 0.f, gr = 0.f, gg = 0.f, gb = 0.f;
  uint32_t last = 0;
  if (inside) {
    const uint32_t index = py * width + px;
    final_T = pixel_color[index].w;
    last = pixel_last[index];
    const float4 pg = pixel_grads[index];
    gr = pg.x; gg = pg.y; gb = pg.z;
  }
  if (tid == 0) s_max_last = 0;
  __syncthreads();
  if (last) atomicMax(&s_max_last, last);
  __syncthreads();
  const uint32_t end = run.x + s_max_last;
  const float bg_grad = bg_r * gr + bg_g * gg + bg_b * gb;
  float T = final_T;
  float behind_r = 0.f, behind_g = 0.f, behind_b = 0.f;
  float next_alpha = 0.f, next_r = 0.f, next_g = 0.f, next_b = 0.f;

  for (uint32_t batch_end = end; batch_end > run.x;) {
    const uint32_t batch_start = batch_end - run.x > kBackBatch ? batch_end - kBackBatch : run.x;
    const int count = static_cast<int>(batch_end - batch_start);
    __syncthreads();
    if (tid < count) {
      const uint32_t idx = pair_splats[batch_start + tid];
      const float4 a = splats[idx * 3], b = splats[idx * 3 + 1], c = splats[idx * 3 + 2];
      s_geom[tid] = make_float4(a.x, a.y, b.x, b.y);
      s_shade[tid] = make_float4(b.z, c.w, c.x, c.y);
      s_blue[tid] = c.z;
      s_index[tid] = idx;
    }
    __syncthreads();
    for (int slot = count - 1; slot >= 0; slot--) {
      float share[kGrads];
#pragma unroll
      for (int v = 0; v < kGrads; v++) share[v] = 0.f;
      bool contributes = false;
      if (batch_start + slot - run.x < last) {
        const float4 ge = s_geom[slot];
        const float4 sh = s_shade[slot];
        const float dx = cx - ge.x, dy = cy - ge.y;
        const float power = -0.5f * (ge.z * dx * dx + sh.x * dy * dy) - ge.w * dx * dy;
        if (power <= 0.f) {
          const float G = __expf(power);
          const float alpha = fminf(0.99f, sh.y * G);
          if (alpha >= 1.f / 255.f) {
            contributes = true;
            T /= 1.f - alpha;
            const float cr = sh.z, cg = sh.w, cb = s_blue[slot];
            behind_r = next_alpha * next_r + (1.f - next_alpha) * behind_r;
            behind_g = next_alpha * next_g + (1.f - next_alpha) * behind_g;
            behind_b = next_alpha * next_b + (1.f - next_alpha) * behind_b;
            next_alpha = alpha;
            next_r = cr; next_g = cg; next_b = cb;
            const float alpha_grad = T * ((cr - behind_r) * gr + (cg - behind_g) * gg + (cb - behind_b) * gb) -
                                     final_T / (1.f - alpha) * bg_grad;
            const float power_grad = alpha_grad * sh.y * G;
            share[0] = power_grad * (ge.z * dx + ge.w * dy);
            share[1] = power_grad * (ge.w * dx + sh.x * dy);
            share[2] = power_grad * (-0.5f * dx * dx);
            share[3] = power_grad * (-dx * dy);
            share[4] = power_grad * (-0.5f * dy * dy);
            share[5] = alpha_grad * G;
            const float at = alpha * T;
            share[6] = at * gr;
            share[7] = at * gg;
            share[8] = at * gb;
          }
        }
      }
      if (__any(contributes)) {
#pragma unroll
        for (int v = 0; v < kGrads; v++) share[v] = wave_sum(share[v]);
      }
      if (lane < kGrads) {
        float mine = 0.f;
#pragma unroll
        for (int v = 0; v < kGrads; v++) mine = lane == v ? share[v] : mine;
        s_sums[wave][slot][lane] = mine;
      }
    }
    __syncthreads();
    for (int e = tid; e < count * kGrads; e += kThreads) {
      const int slot = e / kGrads, v = e % kGrads;
      float total = 0.f;
#pragma unroll
      for (int w = 0; w < kThreads / kWave; w++) total += s_sums[w][slot][v];
      if (total != 0.f) unsafeAtomicAdd(splat_grads + static_cast<size_t>(s_index[slot]) * kGrads + v, total);
    }
    batch_end = batch_start;
  }
}

// ----------------------------------------------------- preprocess backward --

// preprocessBackward.wgsl:27-181: recompute the projection, then carry the
// splat's 9 gradients back to position, opacity, scale, rotation, colour and SH.
__global__ __launch_bounds__(kThreads) void preprocess_backward_kernel(
    const float4* __restrict__ gaussians, const float* __restrict__ sh, uint32_t sh_degree, uint32_t n,
    CameraArgs cam, const float4* __restrict__ splats, const float* __restrict__ splat_grads,
    float4* __restrict__ gaussian_grads, float* __restrict__ sh_grads) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float* sg = sh_grads + static_cast<size_t>(i) * kShRest;
  const float radius = splats[i * 3].w;
  if (radius == 0.f) {
#pragma unroll
    for (int q = 0; q < 4; q++) gaussian_grads[i * 4 + q] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < static_cast<int>(kShRest); k++) sg[k] = 0.f;
    return;
  }
  float g[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    float4 v = gaussians[i * 4 + q];
    g[4 * q] = v.x; g[4 * q + 1] = v.y; g[4 * q + 2] = v.z; g[4 * q + 3] = v.w;
  }
  float gr[kGrads];
#pragma unroll
  for (int v = 0; v < kGrads; v++) gr[v] = splat_grads[static_cast<size_t>(i) * kGrads + v];
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
  const float* coeff = sh + static_cast<size_t>(i) * kShRest;
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
      sg[3 * k] = basis[k] * raw_grad[0];
      sg[3 * k + 1] = basis[k] * raw_grad[1];
      sg[3 * k + 2] = basis[k] * raw_grad[2];
      const float dot = coeff[3 * k] * raw_grad[0] + coeff[3 * k + 1] * raw_grad[1] + coeff[3 * k + 2] * raw_grad[2];
      dgx += bgx[k] * dot;
      dgy += bgy[k] * dot;
      dgz += bgz[k] * dot;
    } else {
      sg[3 * k] = 0.f;
      sg[3 * k + 1] = 0.f;
      sg[3 * k + 2] = 0.f;
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
  gaussian_grads[i * 4] = make_float4(pos_grad[0], pos_grad[1], pos_grad[2], gr[5]);
  gaussian_grads[i * 4 + 1] = make_float4(scale_grad[0], scale_grad[1], scale_grad[2], 0.f);
  gaussian_grads[i * 4 + 2] = make_float4((qg[0] - qw * qdot) * inv_l, (qg[1] - qx * qdot) * inv_l,
                                          (qg[2] - qy * qdot) * inv_l, (qg[3] - qz * qdot) * inv_l);
  gaussian_grads[i * 4 + 3] = make_float4(raw_grad[0], raw_grad[1], raw_grad[2], 0.f);
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
               uint32_t* pair_tiles, uint32_t* pair_splats) {
  if (n == 0) return;
  duplicate_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(reinterpret_cast<const float4*>(splats), order, offsets,
                                                               tile_counts, n, ceil_div(width, kTile),
                                                               ceil_div(height, kTile), capacity, pair_tiles,
                                                               pair_splats);
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
