// Multi-view batch kernels: several views of one Gaussian set rendered, scored
// and back-propagated in the same launches, with their gradients summed for one
// Adam step. Each view is the plain method's frame (the per-tile bodies are the
// single-frame ones in raster_tile.hpp and loss_tile.hpp); what changes is the
// addressing. Tiles of all views form one tile space (a view's tiles start at its
// tile_base), so the binning sort, the ranges and every tile kernel run once per
// batch; per-Gaussian kernels loop over the views, reading each Gaussian and its
// SH record once per batch instead of once per view.
//
// The launches depend only on the batch's capacity (views, Gaussians, largest
// tile space), never on which views it holds: the view table sits in device
// memory and tiles past the batch's tile count return at once, so one captured
// graph serves every step.
#include "block_copy.hpp"
#include "gaussian_math.hpp"
#include "loss_tile.hpp"
#include "raster_tile.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
using raster::kGrads;

// A block's 256 SH records between global memory and LDS: every load in flight at
// once for a full block, one at a time for the partial last block.
__device__ inline void stage_sh(float* __restrict__ dst, const float* __restrict__ src, uint32_t block_n) {
  if (block_n == kThreads) {
    RecordCopy<kThreads, kShRest, kThreads> c;
    c.load(src);
    c.store(dst);
  } else {
    block_copy<kThreads>(dst, src, block_n * kShRest);
  }
}

// The view owning a tile of the batch's tile space (views are in tile order).
__device__ inline const ViewDesc& view_of_tile(const BatchArgs& b, uint32_t tile) {
  uint32_t v = 0;
  for (uint32_t k = 1; k < b.view_count; k++) v += b.views[k].tile_base <= tile ? 1u : 0u;
  return b.views[v];
}

// preprocess.wgsl:20-87 + tiles.wgsl:23-34 for every view: a workgroup reads its
// 256 Gaussians and their SH records once, then writes each view's splats, tile
// counts and depth keys at v N + i.
__global__ __launch_bounds__(kThreads) void preprocess_batch_kernel(const float4* __restrict__ gaussians,
                                                                     const float* __restrict__ sh, uint32_t sh_degree,
                                                                     BatchArgs b, float4* __restrict__ splats,
                                                                     uint32_t* __restrict__ tile_counts,
                                                                     uint32_t* __restrict__ depth_keys,
                                                                     uint32_t* __restrict__ depth_values) {
  __shared__ __align__(16) float s_sh[kThreads * kShRest];
  const uint32_t n = b.per_view;
  const uint32_t first = blockIdx.x * kThreads, i = first + threadIdx.x;
  const bool active = i < n;
  float g[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    float4 v = active ? gaussians[i * 4 + q] : make_float4(0.f, 0.f, 0.f, 1.f);
    g[4 * q] = v.x; g[4 * q + 1] = v.y; g[4 * q + 2] = v.z; g[4 * q + 3] = v.w;
  }
  const int count = sh ? sh_count(sh_degree) : 0;
  if (count > 0) stage_sh(s_sh, sh + static_cast<size_t>(first) * kShRest, min(static_cast<uint32_t>(kThreads), n - first));
  __syncthreads();
  if (!active) return;
  const float* row = s_sh + threadIdx.x * kShRest;
  for (uint32_t v = 0; v < b.view_count; v++) {
    const CameraArgs& cam = b.views[v].cam;
    Projection p;
    const bool visible = project(g, cam, p);
    SplatRecord rec;
    splat_record(g, p, visible, cam, row, count, rec);
    const size_t o = static_cast<size_t>(v) * n + i;
    splats[o * 3] = rec.s0;
    splats[o * 3 + 1] = rec.s1;
    splats[o * 3 + 2] = rec.s2;
    tile_counts[o] = rec.touched;
    depth_keys[o] = rec.key;
    depth_values[o] = static_cast<uint32_t>(o);
  }
}

template <int kPer>
__global__ __launch_bounds__(256) void rasterize_batch_kernel(const float4* __restrict__ splats,
                                                               const uint32_t* __restrict__ pair_splats,
                                                               const uint2* __restrict__ ranges, BatchArgs b,
                                                               float bg_r, float bg_g, float bg_b,
                                                               float4* __restrict__ pixel_color,
                                                               uint32_t* __restrict__ pixel_last) {
  const uint32_t tile = blockIdx.x;
  if (tile >= b.header->tiles) return;
  const ViewDesc& view = view_of_tile(b, tile);
  const uint32_t local = tile - view.tile_base;
  raster::rasterize_tile<kPer>(splats, pair_splats, ranges[tile], local % view.tiles_x, local / view.tiles_x,
                               view.width, view.height, bg_r, bg_g, bg_b, nullptr, pixel_color + view.pixel_base,
                               pixel_last + view.pixel_base);
}

template <int kPer>
__global__ __launch_bounds__(256) void rasterize_backward_batch_kernel(
    const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, const uint2* __restrict__ ranges,
    BatchArgs b, float bg_r, float bg_g, float bg_b, const float4* __restrict__ pixel_color,
    const uint32_t* __restrict__ pixel_last, const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads) {
  __shared__ raster::BackwardShared<kPer> shared;
  const uint32_t tile = blockIdx.x;
  if (tile >= b.header->tiles) return;
  const ViewDesc& view = view_of_tile(b, tile);
  const uint32_t local = tile - view.tile_base;
  raster::rasterize_backward_tile<kPer>(splats, pair_splats, ranges[tile], local % view.tiles_x, local / view.tiles_x,
                                  view.width, view.height, bg_r, bg_g, bg_b, pixel_color + view.pixel_base,
                                  pixel_last + view.pixel_base, pixel_grads + view.pixel_base, splat_grads, shared);
}

__global__ __launch_bounds__(256) void ssim_partials_batch_kernel(BatchArgs b, const float4* __restrict__ rendered,
                                                                   loss::Window win, float4* __restrict__ partials) {
  __shared__ loss::PartialsShared shared;
  const uint32_t tile = blockIdx.x;
  if (tile >= b.header->tiles) return;
  const ViewDesc& view = view_of_tile(b, tile);
  const uint32_t local = tile - view.tile_base;
  loss::ssim_partials_tile(local % view.tiles_x, local / view.tiles_x, view.width, view.height,
                           rendered + view.pixel_base, view.target, win,
                           partials + static_cast<size_t>(view.pixel_base) * 3, shared);
}

__global__ __launch_bounds__(256) void loss_gradient_batch_kernel(BatchArgs b, const float4* __restrict__ rendered,
                                                                   loss::Window win,
                                                                   const float4* __restrict__ partials,
                                                                   float4* __restrict__ pixel_grads,
                                                                   float* __restrict__ pixel_loss) {
  __shared__ loss::GradientShared shared;
  const uint32_t tile = blockIdx.x;
  if (tile >= b.header->tiles) return;
  const ViewDesc& view = view_of_tile(b, tile);
  const uint32_t local = tile - view.tile_base;
  loss::loss_gradient_tile(local % view.tiles_x, local / view.tiles_x, view.width, view.height,
                           rendered + view.pixel_base, view.target, win,
                           partials + static_cast<size_t>(view.pixel_base) * 3, view.loss_scale,
                           pixel_grads + view.pixel_base, pixel_loss + view.pixel_base, shared);
}

// preprocessBackward.wgsl:27-181 summed over the views, with accumulate_stats
// (adam.wgsl:125-134) for every view that sees the Gaussian (and, when `extra` is
// set, the learned densification's statistics): a thread keeps its
// Gaussian's 16 + 45 gradient sums in registers across the views.
__global__ __launch_bounds__(kThreads) void preprocess_backward_batch_kernel(
    const float4* __restrict__ gaussians, const float* __restrict__ sh, uint32_t sh_degree, BatchArgs b,
    const float4* __restrict__ splats, const float* __restrict__ splat_grads, float4* __restrict__ gaussian_grads,
    float* __restrict__ sh_grads, float4* __restrict__ stats, float4* __restrict__ extra) {
  __shared__ __align__(16) float s_sh[kThreads * kShRest];
  const uint32_t n = b.per_view;
  const uint32_t first = blockIdx.x * kThreads, i = first + threadIdx.x;
  const uint32_t block_n = min(static_cast<uint32_t>(kThreads), n - first);
  const bool active = i < n;
  const int count = sh_count(sh_degree);
  if (count > 0) stage_sh(s_sh, sh + static_cast<size_t>(first) * kShRest, block_n);
  __syncthreads();
  float* row = s_sh + threadIdx.x * kShRest;
  float4 out0 = make_float4(0.f, 0.f, 0.f, 0.f), out1 = out0, out2 = out0, out3 = out0;
  float shg[kShRest];
#pragma unroll
  for (int k = 0; k < static_cast<int>(kShRest); k++) shg[k] = 0.f;
  if (active) {
    float g[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float4 v = gaussians[i * 4 + q];
      g[4 * q] = v.x; g[4 * q + 1] = v.y; g[4 * q + 2] = v.z; g[4 * q + 3] = v.w;
    }
    float4 st = stats[i];
    float4 ex = extra ? extra[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t v = 0; v < b.view_count; v++) {
      const size_t o = static_cast<size_t>(v) * n + i;
      const float4 s0 = splats[o * 3];
      const float radius = s0.w;
      if (radius == 0.f) continue;  // not visible in this view
      const ViewDesc& view = b.views[v];
      float gr[kGrads];
#pragma unroll
      for (int k = 0; k < kGrads; k++) gr[k] = splat_grads[o * kGrads + k];
      Projection p;
      project(g, view.cam, p);
      GaussianGrad gg;
      gaussian_backward(g, p, view.cam, gr, row, count, gg);
      out0.x += gg.out0.x; out0.y += gg.out0.y; out0.z += gg.out0.z; out0.w += gg.out0.w;
      out1.x += gg.out1.x; out1.y += gg.out1.y; out1.z += gg.out1.z;
      out2.x += gg.out2.x; out2.y += gg.out2.y; out2.z += gg.out2.z; out2.w += gg.out2.w;
      out3.x += gg.out3.x; out3.y += gg.out3.y; out3.z += gg.out3.z;
#pragma unroll
      for (int k = 0; k < static_cast<int>(kShRest) / 3; k++) {
        if (k < count) {
          shg[3 * k] += gg.basis[k] * gg.raw_grad[0];
          shg[3 * k + 1] += gg.basis[k] * gg.raw_grad[1];
          shg[3 * k + 2] += gg.basis[k] * gg.raw_grad[2];
        }
      }
      const float gx = gr[0] * view.cam.viewport[0] * 0.5f, gy = gr[1] * view.cam.viewport[1] * 0.5f;
      const float length = sqrtf(gx * gx + gy * gy);
      st = make_float4(st.x + length, st.y + 1.f, fmaxf(st.z, radius), 0.f);
      ex = make_float4(fmaxf(ex.x, length), ex.y + s0.z, ex.z + gx, ex.w + gy);
    }
    stats[i] = st;
    if (extra) extra[i] = ex;
    gaussian_grads[i * 4] = out0;
    gaussian_grads[i * 4 + 1] = out1;
    gaussian_grads[i * 4 + 2] = out2;
    gaussian_grads[i * 4 + 3] = out3;
  }
  __syncthreads();  // every thread is done reading its coefficients
#pragma unroll
  for (int k = 0; k < static_cast<int>(kShRest); k++) row[k] = shg[k];
  __syncthreads();
  stage_sh(sh_grads + static_cast<size_t>(first) * kShRest, s_sh, block_n);
}

}  // namespace

void preprocess_batch(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree, const BatchArgs& b,
                      float* splats, uint32_t* tile_counts, uint32_t* depth_keys, uint32_t* depth_values) {
  if (b.per_view == 0 || b.view_count == 0) return;
  preprocess_batch_kernel<<<ceil_div(b.per_view, kThreads), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(gaussians), sh, sh_degree, b, reinterpret_cast<float4*>(splats), tile_counts,
      depth_keys, depth_values);
  DIO_LAUNCHED("preprocess_batch");
}

void rasterize_batch(hipStream_t s, const float* splats, const uint32_t* pair_splats, const uint32_t* ranges,
                     const BatchArgs& b, const float* background, float* pixel_color, uint32_t* pixel_last,
                     uint32_t splat_batch) {
  if (b.max_tiles == 0) return;
  auto launch = [&](auto kernel) {
    kernel<<<b.max_tiles, 256, 0, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                       reinterpret_cast<const uint2*>(ranges), b, background[0], background[1],
                                       background[2], reinterpret_cast<float4*>(pixel_color), pixel_last);
  };
  switch (splat_batch) {
    case 64: launch(rasterize_batch_kernel<1>); break;
    case 128: launch(rasterize_batch_kernel<2>); break;
    case 256: launch(rasterize_batch_kernel<4>); break;
    default: throw Error(-1, "splat batch must be 64, 128 or 256");
  }
  DIO_LAUNCHED("rasterize_batch");
}

void training_loss_batch(hipStream_t s, const BatchArgs& b, const float* rendered, float* partials,
                         float* pixel_grads, float* pixel_loss) {
  if (b.max_tiles == 0) return;
  const loss::Window win = loss::window_weights();
  ssim_partials_batch_kernel<<<b.max_tiles, 256, 0, s>>>(b, reinterpret_cast<const float4*>(rendered), win,
                                                          reinterpret_cast<float4*>(partials));
  DIO_LAUNCHED("ssim_partials_batch");
  loss_gradient_batch_kernel<<<b.max_tiles, 256, 0, s>>>(b, reinterpret_cast<const float4*>(rendered), win,
                                                          reinterpret_cast<const float4*>(partials),
                                                          reinterpret_cast<float4*>(pixel_grads), pixel_loss);
  DIO_LAUNCHED("loss_gradient_batch");
}

void rasterize_backward_batch(hipStream_t s, const float* splats, const uint32_t* pair_splats,
                              const uint32_t* ranges, const BatchArgs& b, const float* background,
                              const float* pixel_color, const uint32_t* pixel_last, const float* pixel_grads,
                              float* splat_grads, uint32_t splat_batch) {
  if (b.max_tiles == 0) return;
  auto launch = [&](auto kernel) {
    kernel<<<b.max_tiles, 256, 0, s>>>(reinterpret_cast<const float4*>(splats), pair_splats,
                                       reinterpret_cast<const uint2*>(ranges), b, background[0], background[1],
                                       background[2], reinterpret_cast<const float4*>(pixel_color), pixel_last,
                                       reinterpret_cast<const float4*>(pixel_grads), splat_grads);
  };
  switch (splat_batch) {
    case 64: launch(rasterize_backward_batch_kernel<1>); break;
    case 128: launch(rasterize_backward_batch_kernel<2>); break;
    case 256: launch(rasterize_backward_batch_kernel<4>); break;
    default: throw Error(-1, "splat batch must be 64, 128 or 256");
  }
  DIO_LAUNCHED("rasterize_backward_batch");
}

void preprocess_backward_batch(hipStream_t s, const float* gaussians, const float* sh, uint32_t sh_degree,
                               const BatchArgs& b, const float* splats, const float* splat_grads,
                               float* gaussian_grads, float* sh_grads, float* stats, float* extra) {
  if (b.per_view == 0 || b.view_count == 0) return;
  preprocess_backward_batch_kernel<<<ceil_div(b.per_view, kThreads), kThreads, 0, s>>>(
      reinterpret_cast<const float4*>(gaussians), sh, sh_degree, b, reinterpret_cast<const float4*>(splats),
      splat_grads, reinterpret_cast<float4*>(gaussian_grads), sh_grads, reinterpret_cast<float4*>(stats),
      reinterpret_cast<float4*>(extra));
  DIO_LAUNCHED("preprocess_backward_batch");
}

}  // namespace dio
