// Held-out image quality on the device: the PSNR and SSIM of src/metrics.ts:5-72
// on renders clamped to [0, 1] and rounded to 8 bits as the reference measures
// them (src/trainer.ts:371-386; the error_heatmap tile sums of
// src/shaders/inspect.wgsl:117-144 for the PSNR alone).
//
// One workgroup per 16x16 tile writes the tile's sum of squared 8-bit errors as
// an integer (exact: at most 256 x 3 x 255^2 < 2^32) and, when asked, its sum of
// per-pixel SSIM. The SSIM blurs the five moment maps with the 11-tap sigma 1.5
// window, zero past the image edges, from a halo tile in LDS, in double: the
// reference computes in double (storing its blurred maps as float32).
#include <cmath>

#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kT = 16, kR = 5, kSpan = kT + 2 * kR;

struct WindowD {
  double w[kR + 1];
};

// Math.round(min(1, max(0, x)) * 255): x * 255 is exact in double.
__device__ inline uint32_t to_8bit(float x) {
  const double v = fmin(1.0, fmax(0.0, static_cast<double>(x))) * 255.0;
  return static_cast<uint32_t>(floor(v + 0.5));
}

__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  const double total = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(256) void evaluate_kernel(const float4* __restrict__ rendered,
                                                       const uint32_t* __restrict__ target, uint32_t width,
                                                       uint32_t height, uint32_t tiles_x, int with_ssim, WindowD win,
                                                       uint32_t* __restrict__ sse, double* __restrict__ ssim) {
  __shared__ uint8_t s_x[kSpan][kSpan], s_y[kSpan][kSpan];
  __shared__ double s_h[5][kSpan][kT];
  __shared__ double s_red[4];
  const int tid = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  const int lx = tid % kT, ly = tid / kT;
  const int x0 = tx * kT - kR, y0 = ty * kT - kR;
  const int px = tx * kT + lx, py = ty * kT + ly;
  const bool inside = px < static_cast<int>(width) && py < static_cast<int>(height);
  const double c1 = 0.01 * 0.01, c2 = 0.03 * 0.03;
  uint32_t squared = 0;
  double structure = 0.0;
  for (int ch = 0; ch < 3; ch++) {
    for (int e = tid; e < kSpan * kSpan; e += 256) {
      const int rx = e % kSpan, ry = e / kSpan;
      const int gx = x0 + rx, gy = y0 + ry;
      uint32_t x = 0, y = 0;
      if (gx >= 0 && gy >= 0 && gx < static_cast<int>(width) && gy < static_cast<int>(height)) {
        const size_t p = static_cast<size_t>(gy) * width + gx;
        const float4 r = rendered[p];
        x = to_8bit(ch == 0 ? r.x : (ch == 1 ? r.y : r.z));
        y = (target[p] >> (8 * ch)) & 0xffu;
      }
      s_x[ry][rx] = static_cast<uint8_t>(x);
      s_y[ry][rx] = static_cast<uint8_t>(y);
    }
    __syncthreads();
    if (inside) {
      const int d = static_cast<int>(s_x[ly + kR][lx + kR]) - static_cast<int>(s_y[ly + kR][lx + kR]);
      squared += static_cast<uint32_t>(d * d);
    }
    if (with_ssim) {
      for (int e = tid; e < kSpan * kT; e += 256) {
        const int col = e % kT, row = e / kT;
        double m[5] = {0, 0, 0, 0, 0};
        for (int k = -kR; k <= kR; k++) {
          const double w = win.w[k < 0 ? -k : k];
          const double x = s_x[row][col + kR + k] / 255.0, y = s_y[row][col + kR + k] / 255.0;
          m[0] += w * x;
          m[1] += w * y;
          m[2] += w * (x * x);
          m[3] += w * (y * y);
          m[4] += w * (x * y);
        }
        for (int q = 0; q < 5; q++) s_h[q][row][col] = m[q];
      }
      __syncthreads();
      double mu[5] = {0, 0, 0, 0, 0};
      for (int k = -kR; k <= kR; k++) {
        const double w = win.w[k < 0 ? -k : k];
        for (int q = 0; q < 5; q++) mu[q] += w * s_h[q][ly + kR + k][lx];
      }
      const double mx = mu[0], my = mu[1];
      const double vx = mu[2] - mx * mx, vy = mu[3] - my * my, cxy = mu[4] - mx * my;
      if (inside) structure += ((2 * mx * my + c1) * (2 * cxy + c2)) / ((mx * mx + my * my + c1) * (vx + vy + c2));
    }
    __syncthreads();
  }
  const double total_squared = block_sum(static_cast<double>(squared), s_red);
  const double total_structure = with_ssim ? block_sum(structure, s_red) : 0.0;
  if (tid == 0) {
    sse[tile] = static_cast<uint32_t>(total_squared);
    if (with_ssim) ssim[tile] = total_structure;
  }
}

}  // namespace

uint32_t evaluate_tiles(uint32_t width, uint32_t height) { return ceil_div(width, kT) * ceil_div(height, kT); }

void evaluate_image(hipStream_t s, const float4* rendered, const uint32_t* target, uint32_t width, uint32_t height,
                    bool with_ssim, uint32_t* sse, double* ssim) {
  WindowD win;
  double raw[kR + 1], total = 0;
  // src/metrics.ts:49-53: exp(-(k - 5)^2 / (2 * 1.5^2)) over k = 0..10, normalised
  for (int k = 0; k <= kR; k++) raw[k] = std::exp(-static_cast<double>(k * k) / (2 * 1.5 * 1.5));
  for (int k = -kR; k <= kR; k++) total += raw[k < 0 ? -k : k];
  for (int k = 0; k <= kR; k++) win.w[k] = raw[k] / total;
  const uint32_t tiles_x = ceil_div(width, kT);
  evaluate_kernel<<<evaluate_tiles(width, height), 256, 0, s>>>(rendered, target, width, height, tiles_x,
                                                                with_ssim ? 1 : 0, win, sse, ssim);
  DIO_LAUNCHED("evaluate_kernel");
}

}  // namespace dio
