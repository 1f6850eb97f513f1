// The training loss 0.8 * L1 + 0.2 * (1 - SSIM) and dL/dpixel in two tiled
// kernels instead of the reference's seven full-image passes
// (src/shaders/loss.wgsl:52-121, src/trainingLoss.ts:74-107):
// 1. per 16x16 tile, load the render and the 8-bit target with a 5-pixel halo into
//    LDS, blur the five moment maps separably on chip, and write each pixel's SSIM
//    and its three partial derivatives (12 floats per pixel);
// 2. per 16x16 tile, blur the partials (again from an LDS halo) and combine them
//    with the L1 term into dL/dpixel and the pixel's loss share.
// Blurs use the reference's normalised 11-tap sigma=1.5 window with zeros past
// the image edges.
#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kT = 16;               // output tile side
constexpr int kR = 5;                // window radius
constexpr int kSpan = kT + 2 * kR;   // 26: tile plus halo
constexpr float kLambda = 0.2f;
constexpr float kC1 = 0.0001f;
constexpr float kC2 = 0.0009f;

struct Window {
  float w[kR + 1];
};

__device__ inline float unorm8(uint32_t v, int channel) {
  return static_cast<float>((v >> (8 * channel)) & 0xffu) / 255.f;
}

__global__ __launch_bounds__(256) void ssim_partials_kernel(uint32_t width, uint32_t height,
                                                             const float4* __restrict__ rendered,
                                                             const uint32_t* __restrict__ target, Window win,
                                                             float4* __restrict__ partials) {
  __shared__ float s_x[kSpan][kSpan];
  __shared__ float s_y[kSpan][kSpan];
  __shared__ float s_h[5][kSpan][kT];
  const int tid = threadIdx.x;
  const int lx = tid % kT, ly = tid / kT;
  const int x0 = blockIdx.x * kT - kR, y0 = blockIdx.y * kT - kR;
  const int px = blockIdx.x * kT + lx, py = blockIdx.y * kT + ly;
  float out[3][4];
  for (int ch = 0; ch < 3; ch++) {
    for (int e = tid; e < kSpan * kSpan; e += 256) {
      const int rx = e % kSpan, ry = e / kSpan;
      const int gx = x0 + rx, gy = y0 + ry;
      float x = 0.f, y = 0.f;
      if (gx >= 0 && gy >= 0 && gx < static_cast<int>(width) && gy < static_cast<int>(height)) {
        const size_t p = static_cast<size_t>(gy) * width + gx;
        const float4 r = rendered[p];
        x = ch == 0 ? r.x : (ch == 1 ? r.y : r.z);
        y = unorm8(target[p], ch);
      }
      s_x[ry][rx] = x;
      s_y[ry][rx] = y;
    }
    __syncthreads();
    // Horizontal blur of x, y, x², y², xy for every halo row and the tile's columns.
    for (int e = tid; e < kSpan * kT; e += 256) {
      const int col = e % kT, row = e / kT;
      float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = -kR; k <= kR; k++) {
        const float w = win.w[k < 0 ? -k : k];
        const float x = s_x[row][col + kR + k], y = s_y[row][col + kR + k];
        m[0] += w * x;
        m[1] += w * y;
        m[2] += w * (x * x);
        m[3] += w * (y * y);
        m[4] += w * (x * y);
      }
#pragma unroll
      for (int q = 0; q < 5; q++) s_h[q][row][col] = m[q];
    }
    __syncthreads();
    float mu[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = -kR; k <= kR; k++) {
      const float w = win.w[k < 0 ? -k : k];
#pragma unroll
      for (int q = 0; q < 5; q++) mu[q] += w * s_h[q][ly + kR + k][lx];
    }
    const float mx = mu[0], my = mu[1];
    const float vx = mu[2] - mx * mx, vy = mu[3] - my * my, cxy = mu[4] - mx * my;
    const float a1 = 2.f * mx * my + kC1, a2 = 2.f * cxy + kC2;
    const float b1 = mx * mx + my * my + kC1, b2 = vx + vy + kC2;
    const float s = a1 * a2 / (b1 * b2);
    out[ch][0] = s * (2.f * my / a1 - 2.f * my / a2 - 2.f * mx / b1 + 2.f * mx / b2);
    out[ch][1] = -s / b2;
    out[ch][2] = 2.f * s / a2;
    out[ch][3] = s;
    __syncthreads();
  }
  if (px < static_cast<int>(width) && py < static_cast<int>(height)) {
    const size_t p = static_cast<size_t>(py) * width + px;
    // Per pixel: (∂μ r,g,b, s r), (∂E[x²] r,g,b, s g), (∂E[xy] r,g,b, s b).
    partials[p * 3] = make_float4(out[0][0], out[1][0], out[2][0], out[0][3]);
    partials[p * 3 + 1] = make_float4(out[0][1], out[1][1], out[2][1], out[1][3]);
    partials[p * 3 + 2] = make_float4(out[0][2], out[1][2], out[2][2], out[2][3]);
  }
}

__global__ __launch_bounds__(256) void loss_gradient_kernel(uint32_t width, uint32_t height,
                                                             const float4* __restrict__ rendered,
                                                             const uint32_t* __restrict__ target, Window win,
                                                             const float4* __restrict__ partials, float scale,
                                                             float4* __restrict__ pixel_grads,
                                                             float* __restrict__ pixel_loss) {
  __shared__ float s_p[3][kSpan][kSpan];
  __shared__ float s_h[3][kSpan][kT];
  const int tid = threadIdx.x;
  const int lx = tid % kT, ly = tid / kT;
  const int x0 = blockIdx.x * kT - kR, y0 = blockIdx.y * kT - kR;
  const int px = blockIdx.x * kT + lx, py = blockIdx.y * kT + ly;
  const bool inside = px < static_cast<int>(width) && py < static_cast<int>(height);
  const size_t me = inside ? static_cast<size_t>(py) * width + px : 0;
  float4 x4 = make_float4(0.f, 0.f, 0.f, 0.f), s4 = x4;
  uint32_t t8 = 0;
  if (inside) {
    x4 = rendered[me];
    t8 = target[me];
    s4 = make_float4(partials[me * 3].w, partials[me * 3 + 1].w, partials[me * 3 + 2].w, 0.f);
  }
  float grad[3], share = 0.f;
  for (int ch = 0; ch < 3; ch++) {
    for (int e = tid; e < kSpan * kSpan; e += 256) {
      const int rx = e % kSpan, ry = e / kSpan;
      const int gx = x0 + rx, gy = y0 + ry;
      float v0 = 0.f, v1 = 0.f, v2 = 0.f;
      if (gx >= 0 && gy >= 0 && gx < static_cast<int>(width) && gy < static_cast<int>(height)) {
        const size_t p = (static_cast<size_t>(gy) * width + gx) * 3;
        const float4 a = partials[p], b = partials[p + 1], c = partials[p + 2];
        v0 = ch == 0 ? a.x : (ch == 1 ? a.y : a.z);
        v1 = ch == 0 ? b.x : (ch == 1 ? b.y : b.z);
        v2 = ch == 0 ? c.x : (ch == 1 ? c.y : c.z);
      }
      s_p[0][ry][rx] = v0;
      s_p[1][ry][rx] = v1;
      s_p[2][ry][rx] = v2;
    }
    __syncthreads();
    for (int e = tid; e < kSpan * kT; e += 256) {
      const int col = e % kT, row = e / kT;
      float m[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int k = -kR; k <= kR; k++) {
        const float w = win.w[k < 0 ? -k : k];
#pragma unroll
        for (int q = 0; q < 3; q++) m[q] += w * s_p[q][row][col + kR + k];
      }
#pragma unroll
      for (int q = 0; q < 3; q++) s_h[q][row][col] = m[q];
    }
    __syncthreads();
    float b[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = -kR; k <= kR; k++) {
      const float w = win.w[k < 0 ? -k : k];
#pragma unroll
      for (int q = 0; q < 3; q++) b[q] += w * s_h[q][ly + kR + k][lx];
    }
    const float x = ch == 0 ? x4.x : (ch == 1 ? x4.y : x4.z);
    const float y = unorm8(t8, ch);
    const float s = ch == 0 ? s4.x : (ch == 1 ? s4.y : s4.z);
    const float diff = x - y;
    const float sign = diff > 0.f ? 1.f : (diff < 0.f ? -1.f : 0.f);
    const float ssim_grad = b[0] + 2.f * x * b[1] + y * b[2];
    grad[ch] = ((1.f - kLambda) * sign - kLambda * ssim_grad) * scale;
    share += (1.f - kLambda) * fabsf(diff) - kLambda * s;
    __syncthreads();
  }
  if (inside) {
    pixel_grads[me] = make_float4(grad[0], grad[1], grad[2], 0.f);
    pixel_loss[me] = share * scale;
  }
}

}  // namespace

void training_loss(hipStream_t s, uint32_t width, uint32_t height, const float* rendered, const uint32_t* target,
                   float* partials, float* pixel_grads, float* pixel_loss) {
  if (width == 0 || height == 0) return;
  Window win;
  double raw[kR + 1], total = 0;
  for (int k = 0; k <= kR; k++) raw[k] = std::exp(-static_cast<double>(k * k) / (2 * 1.5 * 1.5));
  total = raw[0];
  for (int k = 1; k <= kR; k++) total += 2 * raw[k];
  for (int k = 0; k <= kR; k++) win.w[k] = static_cast<float>(raw[k] / total);
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
