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
                                                             float4* __restrict__ partials) {
  __shared__ loss::PartialsShared shared;
  loss::ssim_partials_tile(blockIdx.x, blockIdx.y, width, height, rendered, target, win, partials, shared);
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
                   float* partials, float* pixel_grads, float* pixel_loss) {
  if (width == 0 || height == 0) return;
  const Window win = loss::window_weights();
  const dim3 grid(ceil_div(width, kT), ceil_div(height, kT));
  ssim_partials_kernel<<<grid, 256, 0, s>>>(width, height, reinterpret_cast<const float4*>(rendered), target, win,
                                            reinterpret_cast<float4*>(partials));
  DIO_LAUNCHED("ssim_partials");
  const float scale = static_cast<float>(1.0 / (3.0 * static_cast<double>(width) * height));
  loss_gradient_kernel<<<grid, 256, 0, s>>>(width, height, reinterpret_cast<const float4*>(rendered), target, win,
                                            reinterpret_cast<const float4*>(partials), scale,
                                            reinterpret_cast<float4*>(pixel_grads), pixel_loss);
  DIO_LAUNCHED("loss_gradient");
}

}  // namespace dio
