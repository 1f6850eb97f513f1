// The training loss kernels for one image: one workgroup per 16x16 tile, the
// bodies in loss_tile.hpp.
#include "loss_tile.hpp"

namespace dio {
namespace {

using loss::Window;
using loss::kT;

__global__ __launch_bounds__(256) void ssim_partials_kernel(uint32_t width, uint32_t height,
                                                             const float4* __restrict__ rendered,
                                                             const uint32_t* __restrict__ target, Window win,
                                                             float4* __restrict__ partials,
                                                             uint32_t* __restrict__ clear, size_t clear_words) {
  __shared__ loss::PartialsShared shared;
  loss::ssim_partials_tile(blockIdx.x, blockIdx.y, width, height, rendered, target, win, partials, shared);
  if (clear) {
    const size_t threads = static_cast<size_t>(gridDim.x) * gridDim.y * 256;
    const size_t id = (static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x) * 256 + threadIdx.x;
    const size_t quads = clear_words / 4;
    for (size_t k = id; k < quads; k += threads) reinterpret_cast<uint4*>(clear)[k] = make_uint4(0u, 0u, 0u, 0u);
    for (size_t k = quads * 4 + id; k < clear_words; k += threads) clear[k] = 0u;
  }
}

__global__ __launch_bounds__(256) void loss_gradient_kernel(uint32_t width, uint32_t height,
                                                             const float4* __restrict__ rendered,
                                                             const uint32_t* __restrict__ target, Window win,
                                                             const float4* __restrict__ partials, float scale,
                                                             float4* __restrict__ pixel_grads,
                                                             float* __restrict__ pixel_loss) {
  __shared__ loss::GradientShared shared;
  loss::loss_gradient_tile(blockIdx.x, blockIdx.y, width, height, rendered, target, win, partials, scale, pixel_grads,
                           pixel_loss, shared);
}

}  // namespace

void training_loss(hipStream_t s, uint32_t width, uint32_t height, const float* rendered, const uint32_t* target,
                   float* partials, float* pixel_grads, float* pixel_loss, uint32_t* clear, size_t clear_words) {
  if (width == 0 || height == 0) return;
  const Window win = loss::window_weights();
  const dim3 grid(ceil_div(width, kT), ceil_div(height, kT));
  ssim_partials_kernel<<<grid, 256, 0, s>>>(width, height, reinterpret_cast<const float4*>(rendered), target, win,
                                            reinterpret_cast<float4*>(partials), clear, clear_words);
  DIO_LAUNCHED("ssim_partials");
  const float scale = static_cast<float>(1.0 / (3.0 * static_cast<double>(width) * height));
  loss_gradient_kernel<<<grid, 256, 0, s>>>(width, height, reinterpret_cast<const float4*>(rendered), target, win,
                                            reinterpret_cast<const float4*>(partials), scale,
                                            reinterpret_cast<float4*>(pixel_grads), pixel_loss);
  DIO_LAUNCHED("loss_gradient");
}

}  // namespace dio
