// The per-tile bodies of the training loss 0.8 * L1 + 0.2 * (1 - SSIM), shared by
// the single-image kernels (loss.hip) and the multi-view batch kernels (batch.hip)
// (src/shaders/loss.wgsl:52-121, src/trainingLoss.ts:74-107):
// 1. per 16x16 tile, load the render and the 8-bit target with a 5-pixel halo into
//    LDS, blur the five moment maps separably on chip, and write each pixel's SSIM
//    and its three partial derivatives (12 floats per pixel);
// 2. per 16x16 tile, blur the partials (again from an LDS halo) and combine them
//    with the L1 term into dL/dpixel and the pixel's loss share.
// Blurs use the reference's normalised 11-tap sigma=1.5 window with zeros past
// the image edges. Both bodies end on a barrier, so a workgroup can run another
// tile after them.
#pragma once
#include <cmath>
#include "dio_internal.hpp"

namespace dio {
namespace loss {

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

// A tile's halo as LDS rows of kSpan x kSpan values, padded to whole rounds of the
// 256 threads: every lane of every load round stores somewhere, so the loads carry
// no condition that would hold them back until their stores.
constexpr int kHalo = kSpan * kSpan;                 // 676
constexpr int kHaloRounds = (kHalo + 255) / 256;     // 3
constexpr int kHaloPad = kHaloRounds * 256;          // 768

// Where halo element e of a tile whose halo starts at (x0, y0) reads from (clamped
// into the image) and whether it lies inside it (outside, the value is zero).
struct HaloPixel {
  size_t index;
  bool inside;
};
__device__ inline HaloPixel halo_pixel(int e, int x0, int y0, uint32_t width, uint32_t height) {
  const int ec = min(e, kHalo - 1);
  const int gx = x0 + ec % kSpan, gy = y0 + ec / kSpan;
  const bool inside = gx >= 0 && gy >= 0 && gx < static_cast<int>(width) && gy < static_cast<int>(height);
  const int cx = min(max(gx, 0), static_cast<int>(width) - 1), cy = min(max(gy, 0), static_cast<int>(height) - 1);
  return {static_cast<size_t>(cy) * width + cx, inside};
}

// The horizontal blur's outputs (kSpan rows x kT columns) over two rounds of the 256
// threads: wave w's lane l takes row (w mod 8) + 8 (l / 16), column l mod 16, so the
// four rows a wave reads lie 8 x kSpan = 208 floats apart, 16 banks further round
// each time: a wave's 64 reads of one tap hit 64 different banks of the kSpan-pitch
// halo rows (with consecutive rows, 26 floats apart, two lanes shared a bank). Rows
// past kSpan (26..31) are idle lanes.
struct BlurSlot {
  int row, col;
  bool valid;
};
__device__ inline BlurSlot blur_slot(int e) {
  const int w = e / 64, l = e % 64;
  const int row = (w & 7) + 8 * (l / 16);
  return {row, l % 16, row < kSpan};
}
constexpr int kBlurSlots = 512;
// The horizontal blur's outputs stay rows of kT: the vertical blur reads four
// consecutive rows a wave (64 banks, no conflict), 55 reads a pixel against the
// horizontal blur's 5 writes an item, whose four rows 8 apart share banks.
__device__ inline int blur_index(int row, int col) { return row * kT + col; }

struct PartialsShared {
  float x[3][kHaloPad];  // per channel, row-major kSpan x kSpan
  float y[3][kHaloPad];
  float h[5][kSpan * kT];
};

// Pass 1 over the 16x16 tile (tile_x, tile_y) of a width x height image.
__device__ inline void ssim_partials_tile(uint32_t tile_x, uint32_t tile_y, uint32_t width, uint32_t height,
                                          const float4* __restrict__ rendered, const uint32_t* __restrict__ target,
                                          const Window& win, float4* __restrict__ partials, PartialsShared& sh) {
  auto& s_x = sh.x;
  auto& s_y = sh.y;
  auto& s_h = sh.h;
  const int tid = threadIdx.x;
  const int lx = tid % kT, ly = tid / kT;
  const int x0 = static_cast<int>(tile_x) * kT - kR, y0 = static_cast<int>(tile_y) * kT - kR;
  const int px = static_cast<int>(tile_x) * kT + lx, py = static_cast<int>(tile_y) * kT + ly;
  // The halo of all three channels, every load in flight at once.
  {
    float4 r[kHaloRounds];
    uint32_t t[kHaloRounds];
    bool in[kHaloRounds];
#pragma unroll
    for (int k = 0; k < kHaloRounds; k++) {
      const HaloPixel hp = halo_pixel(tid + k * 256, x0, y0, width, height);
      r[k] = rendered[hp.index];
      t[k] = target[hp.index];
      in[k] = hp.inside;
    }
#pragma unroll
    for (int k = 0; k < kHaloRounds; k++) {
      const int e = tid + k * 256;
      s_x[0][e] = in[k] ? r[k].x : 0.f;
      s_x[1][e] = in[k] ? r[k].y : 0.f;
      s_x[2][e] = in[k] ? r[k].z : 0.f;
#pragma unroll
      for (int ch = 0; ch < 3; ch++) s_y[ch][e] = in[k] ? unorm8(t[k], ch) : 0.f;
    }
  }
  __syncthreads();
  float out[3][4];
  // Kept rolled: unrolled, the three channels' blurs need 169 VGPRs instead of 45.
#pragma unroll 1
  for (int ch = 0; ch < 3; ch++) {
    // Horizontal blur of x, y, x², y², xy for every halo row and the tile's columns.
    for (int e = tid; e < kBlurSlots; e += 256) {
      const BlurSlot b = blur_slot(e);
      if (!b.valid) continue;
      const int col = b.col, row = b.row;
      float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = -kR; k <= kR; k++) {
        const float w = win.w[k < 0 ? -k : k];
        const float x = s_x[ch][row * kSpan + col + kR + k], y = s_y[ch][row * kSpan + col + kR + k];
        m[0] += w * x;
        m[1] += w * y;
        m[2] += w * (x * x);
        m[3] += w * (y * y);
        m[4] += w * (x * y);
      }
#pragma unroll
      for (int q = 0; q < 5; q++) s_h[q][blur_index(row, col)] = m[q];
    }
    __syncthreads();
    float mu[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = -kR; k <= kR; k++) {
      const float w = win.w[k < 0 ? -k : k];
#pragma unroll
      for (int q = 0; q < 5; q++) mu[q] += w * s_h[q][blur_index(ly + kR + k, lx)];
    }
    const float mx = mu[0], my = mu[1];
    const float vx = mu[2] - mx * mx, vy = mu[3] - my * my, cxy = mu[4] - mx * my;
    const float a1 = 2.f * mx * my + kC1, a2 = 2.f * cxy + kC2;
    const float b1 = mx * mx + my * my + kC1, b2 = vx + vy + kC2;
    // s / a2 taken as a1 / (b1 b2): the reference's s (1 / a2) (loss.wgsl:103-106) is
    // 0 * inf = NaN where a2 = 2 σxy + C2 rounds to exactly 0; a1, b1, b2 stay >= C1.
    const float s_over_a2 = a1 / (b1 * b2);
    const float s = s_over_a2 * a2;
    out[ch][0] = s * (2.f * my / a1 - 2.f * mx / b1 + 2.f * mx / b2) - 2.f * my * s_over_a2;
    out[ch][1] = -s / b2;
    out[ch][2] = 2.f * s_over_a2;
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

struct GradientShared {
  float p[9][kHaloPad];  // (∂μ, ∂E[x²], ∂E[xy]) x channel, row-major kSpan x kSpan
  float h[3][kSpan * kT];
};

// Pass 2 over one tile: dL/dpixel (scaled by `scale` = 1 / (3 pixels)) and each
// pixel's loss share.
__device__ inline void loss_gradient_tile(uint32_t tile_x, uint32_t tile_y, uint32_t width, uint32_t height,
                                          const float4* __restrict__ rendered, const uint32_t* __restrict__ target,
                                          const Window& win, const float4* __restrict__ partials, float scale,
                                          float4* __restrict__ pixel_grads, float* __restrict__ pixel_loss,
                                          GradientShared& sh) {
  auto& s_p = sh.p;
  auto& s_h = sh.h;
  const int tid = threadIdx.x;
  const int lx = tid % kT, ly = tid / kT;
  const int x0 = static_cast<int>(tile_x) * kT - kR, y0 = static_cast<int>(tile_y) * kT - kR;
  const int px = static_cast<int>(tile_x) * kT + lx, py = static_cast<int>(tile_y) * kT + ly;
  const bool inside = px < static_cast<int>(width) && py < static_cast<int>(height);
  const size_t me = inside ? static_cast<size_t>(py) * width + px : 0;
  float4 x4 = make_float4(0.f, 0.f, 0.f, 0.f), s4 = x4;
  uint32_t t8 = 0;
  if (inside) {
    x4 = rendered[me];
    t8 = target[me];
    s4 = make_float4(partials[me * 3].w, partials[me * 3 + 1].w, partials[me * 3 + 2].w, 0.f);
  }
  // The partials' halo for all three channels, every load in flight at once.
  {
    float4 v[kHaloRounds][3];
    bool in[kHaloRounds];
#pragma unroll
    for (int k = 0; k < kHaloRounds; k++) {
      const HaloPixel hp = halo_pixel(tid + k * 256, x0, y0, width, height);
#pragma unroll
      for (int q = 0; q < 3; q++) v[k][q] = partials[hp.index * 3 + q];
      in[k] = hp.inside;
    }
#pragma unroll
    for (int k = 0; k < kHaloRounds; k++) {
      const int e = tid + k * 256;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        s_p[q * 3][e] = in[k] ? v[k][q].x : 0.f;
        s_p[q * 3 + 1][e] = in[k] ? v[k][q].y : 0.f;
        s_p[q * 3 + 2][e] = in[k] ? v[k][q].z : 0.f;
      }
    }
  }
  __syncthreads();
  float grad[3], share = 0.f;
  for (int ch = 0; ch < 3; ch++) {
    for (int e = tid; e < kBlurSlots; e += 256) {
      const BlurSlot bs = blur_slot(e);
      if (!bs.valid) continue;
      const int col = bs.col, row = bs.row;
      float m[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int k = -kR; k <= kR; k++) {
        const float w = win.w[k < 0 ? -k : k];
#pragma unroll
        for (int q = 0; q < 3; q++) m[q] += w * s_p[q * 3 + ch][row * kSpan + col + kR + k];
      }
#pragma unroll
      for (int q = 0; q < 3; q++) s_h[q][blur_index(row, col)] = m[q];
    }
    __syncthreads();
    float b[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = -kR; k <= kR; k++) {
      const float w = win.w[k < 0 ? -k : k];
#pragma unroll
      for (int q = 0; q < 3; q++) b[q] += w * s_h[q][blur_index(ly + kR + k, lx)];
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


// The window of src/trainingLoss.ts:7-11, normalised in double on the host.
inline Window window_weights() {
  Window win;
  double raw[kR + 1], total = 0;
  for (int k = 0; k <= kR; k++) raw[k] = std::exp(-static_cast<double>(k * k) / (2 * 1.5 * 1.5));
  total = raw[0];
  for (int k = 1; k <= kR; k++) total += 2 * raw[k];
  for (int k = 0; k <= kR; k++) win.w[k] = static_cast<float>(raw[k] / total);
  return win;
}

}  // namespace loss
}  // namespace dio
