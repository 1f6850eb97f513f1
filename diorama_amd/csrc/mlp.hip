// The learned densification proposal: a 12 -> 64 -> 64 -> 4 MLP over per-Gaussian
// statistics whose outputs are keep / clone / split / prune logits. It has no
// counterpart in the reference; the plain threshold rule (densify.wgsl:72-85)
// supplies its training labels and stays its fallback.
//
// Weights are stored in BF16 and every product accumulates in FP32. Training keeps
// an FP32 master copy for Adam and re-rounds the BF16 copy after each step. The
// layer is small (5,252 parameters, 10.5 KB in BF16), so every workgroup holds all
// of it in LDS as FP32.
//
// Layout of the 5,252 parameters: W1 [12][64], b1 [64], W2 [64][64], b2 [64],
// W3 [64][4], b3 [4], row-major with the input index first.
#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr int kIn = kMlpInputs, kHidden = 64, kOut = 4;
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHidden, kW2 = kB1 + kHidden, kB2 = kW2 + kHidden * kHidden,
              kW3 = kB2 + kHidden, kB3 = kW3 + kHidden * kOut, kParams = kB3 + kOut;
static_assert(kParams == static_cast<int>(kMlpParams), "the MLP layout");
constexpr int kThreads = 256;

__device__ inline float bf16_to_float(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Round to nearest even (finite inputs).
__device__ inline uint16_t float_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// The 12 inputs of one Gaussian, each brought to order one:
//  0 log(1 + mean view-space gradient / 2e-4)   1 the same of the largest one
//  2 opacity                                     3 log(largest scale / extent)
//  4 log(smallest scale / extent)                5 log(largest / smallest scale)
//  6 luminance of the SH DC colour               7 log(1 + views that saw it)
//  8 log(1 + age in views seen)                 9 log(1 + largest screen radius)
// 10 log(mean depth)                            11 gradient-direction variance,
//    1 - |sum of gradients| / sum of their lengths.
__global__ __launch_bounds__(kThreads) void features_kernel(uint32_t n, const float* __restrict__ params,
                                                             const float4* __restrict__ stats,
                                                             const float4* __restrict__ extra, float extent,
                                                             uint32_t iteration, float views_per_step,
                                                             float* __restrict__ features) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float* p = params + static_cast<size_t>(i) * kFloatsPerGaussian;
  const float4 s = stats[i], e = extra[i];
  const float views = s.y;
  const float mean_grad = views > 0.f ? s.x / views : 0.f;
  const float sx = __expf(p[4]), sy = __expf(p[5]), sz = __expf(p[6]);
  const float big = fmaxf(sx, fmaxf(sy, sz)), small = fminf(sx, fminf(sy, sz));
  const float c0 = 0.5f + 0.28209479177387814f * p[12], c1 = 0.5f + 0.28209479177387814f * p[13],
              c2 = 0.5f + 0.28209479177387814f * p[14];
  const float age = fmaxf(static_cast<float>(iteration) - p[15], 0.f) * views_per_step;
  const float depth = views > 0.f ? e.y / views : 1.f;
  const float pull = sqrtf(e.z * e.z + e.w * e.w);
  float f[kIn];
  f[0] = logf(1.f + mean_grad / 2e-4f);
  f[1] = logf(1.f + e.x / 2e-4f);
  f[2] = 1.f / (1.f + __expf(-p[3]));
  f[3] = logf(big / extent);
  f[4] = logf(small / extent);
  f[5] = logf(big / small);
  f[6] = 0.299f * c0 + 0.587f * c1 + 0.114f * c2;
  f[7] = logf(1.f + views);
  f[8] = logf(1.f + age);
  f[9] = logf(1.f + s.z);
  f[10] = logf(fmaxf(depth, 1e-3f));
  f[11] = s.x > 0.f ? fmaxf(1.f - pull / s.x, 0.f) : 0.f;
#pragma unroll
  for (int k = 0; k < kIn; k++) features[static_cast<size_t>(i) * kIn + k] = f[k];
}

// The network on the matrix cores: v_mfma_f32_16x16x16_bf16, BF16 operands and FP32
// accumulation. A wave takes 16 samples at a time. Each layer of the forward pass
// runs transposed, out^T = W^T in^T, so the samples are the N dimension: the
// accumulator a lane holds (rows 4 quad + r, r = 0..3, of an output block, for
// sample `col`, where col = lane % 16 and quad = lane / 16) is exactly the B operand
// the next layer needs for that block of its K dimension (k = 4 quad + r, the same
// sample), so the layers chain in registers with no exchange between lanes. The
// same holds for the backward pass's deltas, d2^T = W3 dz^T and d1^T = W2 d2^T.
// Weight tiles are A operands held in registers; element r of lane (col, quad) is
// row col, k = 4 quad + r of its tile. The standardised inputs, both hidden layers
// and the deltas are rounded to BF16 (nearest even) as the operands they become;
// the bias gradients sum the deltas before that rounding.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline bf16x4 pack_bf16(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = static_cast<short>(float_to_bf16(a));
  v[1] = static_cast<short>(float_to_bf16(b));
  v[2] = static_cast<short>(float_to_bf16(c));
  v[3] = static_cast<short>(float_to_bf16(d));
  return v;
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

constexpr int kBlocks = kHidden / 16;  // 16-wide blocks of a hidden layer

// The forward pass's weights as A operands and its biases in the accumulator layout.
struct ForwardTiles {
  bf16x4 w1[kBlocks], w2[kBlocks][kBlocks], w3[kBlocks];
  float b1[kBlocks][4], b2[kBlocks][4], b3[4];
  float mean[4], scale[4];  // the standardisation of inputs 4 quad .. 4 quad + 3

  __device__ void load(const uint16_t* __restrict__ weights, const float* __restrict__ norm, int col, int quad) {
#pragma unroll
    for (int m = 0; m < kBlocks; m++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int k = 4 * quad + r;
        w1[m][r] = k < kIn ? static_cast<short>(weights[kW1 + k * kHidden + m * 16 + col]) : short(0);
        b1[m][r] = bf16_to_float(weights[kB1 + m * 16 + 4 * quad + r]);
        b2[m][r] = bf16_to_float(weights[kB2 + m * 16 + 4 * quad + r]);
#pragma unroll
        for (int ks = 0; ks < kBlocks; ks++)
          w2[m][ks][r] = static_cast<short>(weights[kW2 + (ks * 16 + k) * kHidden + m * 16 + col]);
        w3[m][r] = col < kOut ? static_cast<short>(weights[kW3 + (m * 16 + k) * kOut + col]) : short(0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = 4 * quad + r;
      b3[r] = bf16_to_float(weights[kB3 + r]);
      mean[r] = k < kIn ? norm[k] : 0.f;
      scale[r] = k < kIn ? norm[kIn + k] : 0.f;
    }
  }

  // One group's forward pass from a lane's 4 raw inputs (zeros past input 11): the
  // BF16 operands of each layer and the logits (rows 0..3 = classes, in lanes of
  // quad 0).
  __device__ void run(const float x[4], bf16x4& in, bf16x4 (&h1)[kBlocks], bf16x4 (&h2)[kBlocks], f32x4& z) const {
    in = pack_bf16((x[0] - mean[0]) * scale[0], (x[1] - mean[1]) * scale[1], (x[2] - mean[2]) * scale[2],
                   (x[3] - mean[3]) * scale[3]);
#pragma unroll
    for (int m = 0; m < kBlocks; m++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(w1[m], in, acc);
      h1[m] = pack_bf16(fmaxf(acc[0] + b1[m][0], 0.f), fmaxf(acc[1] + b1[m][1], 0.f),
                        fmaxf(acc[2] + b1[m][2], 0.f), fmaxf(acc[3] + b1[m][3], 0.f));
    }
#pragma unroll
    for (int m = 0; m < kBlocks; m++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kBlocks; ks++) acc = mfma(w2[m][ks], h1[ks], acc);
      h2[m] = pack_bf16(fmaxf(acc[0] + b2[m][0], 0.f), fmaxf(acc[1] + b2[m][1], 0.f),
                        fmaxf(acc[2] + b2[m][2], 0.f), fmaxf(acc[3] + b2[m][3], 0.f));
    }
    z = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kBlocks; ks++) z = mfma(w3[ks], h2[ks], z);
#pragma unroll
    for (int r = 0; r < 4; r++) z[r] += b3[r];
  }
};

// A lane's 4 inputs of `sample` (inputs 4 quad .. 4 quad + 3; zeros past input 11 or
// past the samples).
__device__ inline void load_inputs(const float* __restrict__ features, uint32_t sample, bool valid, int quad,
                                   float x[4]) {
  x[0] = x[1] = x[2] = x[3] = 0.f;
  if (valid && quad < kIn / 4) {
    const float4 v = *reinterpret_cast<const float4*>(features + static_cast<size_t>(sample) * kIn + 4 * quad);
    x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
  }
}

// Inference: a wave takes groups of 16 Gaussians: 4 MFMAs for layer 1 (12 inputs
// padded to 16), 16 for layer 2 and 4 for layer 3 (4 classes padded to 16).
__global__ __launch_bounds__(kThreads) void mlp_forward_kernel(uint32_t n, const float* __restrict__ features,
                                                                const uint16_t* __restrict__ weights,
                                                                const float* __restrict__ norm,
                                                                float* __restrict__ logits) {
  const int lane = threadIdx.x & (kWave - 1);
  const int col = lane % 16, quad = lane / 16;
  ForwardTiles net;
  net.load(weights, norm, col, quad);
  const uint32_t groups = ceil_div(n, 16u);
  const uint32_t wave = (blockIdx.x * kThreads + threadIdx.x) / kWave, waves = gridDim.x * (kThreads / kWave);
  for (uint32_t g = wave; g < groups; g += waves) {
    const uint32_t sample = g * 16 + col;
    float x[4];
    load_inputs(features, sample, sample < n, quad, x);
    bf16x4 in, h1[kBlocks], h2[kBlocks];
    f32x4 z;
    net.run(x, in, h1, h2, z);
    // Rows 0..3 of the logits tile (the classes) sit in lanes 0..15, one sample each.
    if (quad == 0 && sample < n)
      *reinterpret_cast<float4*>(logits + static_cast<size_t>(sample) * kOut) = make_float4(z[0], z[1], z[2], z[3]);
  }
}

// The sample a minibatch slot reads: every sample when `seed` is 0 (slot s reads
// sample s), else a hashed draw from the `count` samples.
__device__ inline uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

__device__ inline uint32_t sample_of(uint32_t slot, uint32_t count, uint32_t seed) {
  if (seed == 0) return slot;
  return pcg_hash(pcg_hash(seed) ^ slot) % count;
}

// Gradients of the mean class-weighted softmax cross-entropy over `batch` samples,
// added into `grads` (which the caller cleared), and the loss sum into loss[0]. A
// workgroup takes 64 samples, 16 per wave:
// 1. each wave runs its samples forward, forms dz = (p - onehot) w_class / batch,
//    and runs the deltas back, d2^T = (W3 dz^T) masked by ReLU and d1^T = (W2 d2^T)
//    masked; the BF16 operands (inputs, hidden layers, deltas) go to LDS sample by
//    sample, and the bias gradients' sums over the wave's samples to LDS atomics;
// 2. the weight gradients dW1 = x^T d1, dW2 = h1^T d2, dW3 = h2^T dz are 24 output
//    tiles of 16 x 16 over K = the 64 samples, six a wave, each read from LDS in the
//    layout of its operand; every tile goes out with one float atomic per element.
constexpr int kTrainWaveSamples = 16, kTrainSamples = kTrainWaveSamples * (kThreads / kWave);  // 64
constexpr int kPad = 16;  // inputs and classes padded to one K block

struct TrainShared {
  uint16_t x[kTrainSamples][kPad];      // BF16, by sample
  uint16_t h1[kTrainSamples][kHidden];
  uint16_t h2[kTrainSamples][kHidden];
  uint16_t d1[kTrainSamples][kHidden];  // dL/d(pre-activation of layer 1), BF16
  uint16_t d2[kTrainSamples][kHidden];  // dL/d(pre-activation of layer 2), BF16
  uint16_t dz[kTrainSamples][kPad];     // dL/dlogits, BF16
  float bias[kHidden + kHidden + kOut];  // the bias gradients' sums over the workgroup
};

// Sums of v over the 16 lanes of a quad (lanes with the same `quad`), in every lane.
__device__ inline float quad_sum(float v) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__global__ __launch_bounds__(kThreads) void mlp_train_kernel(uint32_t batch, uint32_t count, uint32_t seed,
                                                              const float* __restrict__ features,
                                                              const uint32_t* __restrict__ labels,
                                                              const uint16_t* __restrict__ weights,
                                                              const float* __restrict__ norm, float4 class_weight,
                                                              float* __restrict__ grads, float* __restrict__ loss) {
  __shared__ TrainShared sh;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int col = lane % 16, quad = lane / 16;
  for (int k = tid; k < kHidden + kHidden + kOut; k += kThreads) sh.bias[k] = 0.f;
  __syncthreads();  // the bias sums start at zero before any wave adds into them
  // Phase 1: this wave's 16 samples forward and back.
  {
    ForwardTiles net;
    net.load(weights, norm, col, quad);
    const uint32_t slot = blockIdx.x * kTrainSamples + wave * kTrainWaveSamples + col;
    const bool active = slot < batch;
    const uint32_t sample = active ? sample_of(slot, count, seed) : 0u;
    float x[4];
    load_inputs(features, sample, active, quad, x);
    bf16x4 in, h1[kBlocks], h2[kBlocks];
    f32x4 z;
    net.run(x, in, h1, h2, z);
    const int row = wave * kTrainWaveSamples + col;  // this lane's sample in LDS
    // Softmax cross-entropy in the lanes holding the logits (quad 0).
    float dz[4] = {0.f, 0.f, 0.f, 0.f}, sample_loss = 0.f;
    if (quad == 0 && active) {
      const uint32_t label = labels[sample];
      const float top = fmaxf(fmaxf(z[0], z[1]), fmaxf(z[2], z[3]));
      float total = 0.f;
#pragma unroll
      for (int c = 0; c < kOut; c++) total += __expf(z[c] - top);
      const float log_total = logf(total);
      const float cw = label == 0u ? class_weight.x : (label == 1u ? class_weight.y : (label == 2u ? class_weight.z
                                                                                                    : class_weight.w));
      const float inv_batch = 1.f / static_cast<float>(batch);
#pragma unroll
      for (int c = 0; c < kOut; c++)
        dz[c] = (__expf(z[c] - top - log_total) - (static_cast<uint32_t>(c) == label ? 1.f : 0.f)) * cw * inv_batch;
      const float zl = label == 0u ? z[0] : (label == 1u ? z[1] : (label == 2u ? z[2] : z[3]));
      sample_loss = cw * (top + log_total - zl);
    }
    const float wave_loss = quad_sum(sample_loss);
    if (lane == 0 && wave_loss != 0.f) atomicAdd(loss, wave_loss);
    // dz^T as a B operand: rows (classes) 4 quad + r, only quad 0 holds classes 0..3.
    const bf16x4 dzb = pack_bf16(dz[0], dz[1], dz[2], dz[3]);
    // d2^T = W3 dz^T (A tile m: rows = hidden m 16 + col, k = class 4 quad + r), masked.
    bf16x4 d2b[kBlocks];
#pragma unroll
    for (int m = 0; m < kBlocks; m++) {
      bf16x4 a;
#pragma unroll
      for (int r = 0; r < 4; r++)
        a[r] = quad == 0 ? static_cast<short>(weights[kW3 + (m * 16 + col) * kOut + r]) : short(0);
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(a, dzb, acc);
      float d[4];
#pragma unroll
      for (int r = 0; r < 4; r++)  // ReLU: h2 > 0 exactly where its pre-activation is
        d[r] = bf16_to_float(static_cast<uint16_t>(h2[m][r])) > 0.f ? acc[r] : 0.f;
      d2b[m] = pack_bf16(d[0], d[1], d[2], d[3]);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float s = quad_sum(d[r]);
        if (col == 0 && s != 0.f) atomicAdd(&sh.bias[kHidden + m * 16 + 4 * quad + r], s);
      }
      *reinterpret_cast<bf16x4*>(&sh.d2[row][m * 16 + 4 * quad]) = d2b[m];
      *reinterpret_cast<bf16x4*>(&sh.h2[row][m * 16 + 4 * quad]) = h2[m];
      *reinterpret_cast<bf16x4*>(&sh.h1[row][m * 16 + 4 * quad]) = h1[m];
    }
    // d1^T = W2 d2^T (A tile (mj, ks): rows = j = mj 16 + col, k = ks 16 + 4 quad + r), masked.
#pragma unroll
    for (int mj = 0; mj < kBlocks; mj++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kBlocks; ks++) {
        bf16x4 a;
#pragma unroll
        for (int r = 0; r < 4; r++)
          a[r] = static_cast<short>(weights[kW2 + (mj * 16 + col) * kHidden + ks * 16 + 4 * quad + r]);
        acc = mfma(a, d2b[ks], acc);
      }
      float d[4];
#pragma unroll
      for (int r = 0; r < 4; r++) d[r] = bf16_to_float(static_cast<uint16_t>(h1[mj][r])) > 0.f ? acc[r] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float s = quad_sum(d[r]);
        if (col == 0 && s != 0.f) atomicAdd(&sh.bias[mj * 16 + 4 * quad + r], s);
      }
      *reinterpret_cast<bf16x4*>(&sh.d1[row][mj * 16 + 4 * quad]) = pack_bf16(d[0], d[1], d[2], d[3]);
    }
    *reinterpret_cast<bf16x4*>(&sh.x[row][4 * quad]) = in;
    *reinterpret_cast<bf16x4*>(&sh.dz[row][4 * quad]) = dzb;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float s = quad_sum(dz[r]);
      if (lane == 0 && s != 0.f) atomicAdd(&sh.bias[2 * kHidden + r], s);
    }
  }
  __syncthreads();
  // Phase 2: weight-gradient tiles, dW[i][j] = sum over samples s of P[s][i] Q[s][j]:
  // A = P^T (row i = col, k = sample 4 quad + r of the K block), B = Q (k = sample
  // 4 quad + r, column j = col); the accumulator holds rows 4 quad + r, column col.
  for (int tile = wave; tile < 24; tile += kThreads / kWave) {
    const uint16_t* P;
    const uint16_t* Q;
    int p_stride, q_stride, ib, jb, rows, cols, base;
    if (tile < 4) {  // dW1: 16 (12) x 64
      P = &sh.x[0][0]; p_stride = kPad; Q = &sh.d1[0][0]; q_stride = kHidden;
      ib = 0; jb = tile; rows = kIn; cols = kHidden; base = kW1;
    } else if (tile < 20) {  // dW2: 64 x 64
      P = &sh.h1[0][0]; p_stride = kHidden; Q = &sh.d2[0][0]; q_stride = kHidden;
      ib = (tile - 4) / 4; jb = (tile - 4) % 4; rows = kHidden; cols = kHidden; base = kW2;
    } else {  // dW3: 64 x 16 (4)
      P = &sh.h2[0][0]; p_stride = kHidden; Q = &sh.dz[0][0]; q_stride = kPad;
      ib = tile - 20; jb = 0; rows = kHidden; cols = kOut; base = kW3;
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kTrainSamples / 16; ks++) {
      bf16x4 a, b;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int s = ks * 16 + 4 * quad + r;
        a[r] = static_cast<short>(P[s * p_stride + ib * 16 + col]);
        b[r] = static_cast<short>(Q[s * q_stride + jb * 16 + col]);
      }
      acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = ib * 16 + 4 * quad + r, j = jb * 16 + col;
      if (i < rows && j < cols && acc[r] != 0.f) atomicAdd(grads + base + i * cols + j, acc[r]);
    }
  }
  for (int k = tid; k < kHidden + kHidden + kOut; k += kThreads) {
    const float g = sh.bias[k];
    const int at = k < kHidden ? kB1 + k : (k < 2 * kHidden ? kB2 + (k - kHidden) : kB3 + (k - 2 * kHidden));
    if (g != 0.f) atomicAdd(grads + at, g);
  }
}

// Adam on the FP32 master weights (then their BF16 copy), clearing the gradients
// for the next step.
__global__ __launch_bounds__(kThreads) void mlp_adam_kernel(float* __restrict__ master, float2* __restrict__ moments,
                                                             float* __restrict__ grads, uint16_t* __restrict__ weights,
                                                             float lr, float inv_bias1, float inv_bias2) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= kParams) return;
  const float g = grads[e];
  float2 m = moments[e];
  m.x = 0.9f * m.x + 0.1f * g;
  m.y = 0.999f * m.y + 0.001f * g * g;
  moments[e] = m;
  const float w = master[e] - lr * (m.x * inv_bias1) / (sqrtf(m.y * inv_bias2) + 1e-8f);
  master[e] = w;
  weights[e] = float_to_bf16(w);
  grads[e] = 0.f;
}

__global__ __launch_bounds__(kThreads) void mlp_round_kernel(const float* __restrict__ master,
                                                              uint16_t* __restrict__ weights) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e < kParams) weights[e] = float_to_bf16(master[e]);
}

// Learned actions from the logits (keep, clone, split, prune): the most likely
// action where its probability reaches `confidence`, else the rule's code already
// in `codes`. The growth priority is P(clone) + P(split).
__global__ __launch_bounds__(kThreads) void learned_codes_kernel(uint32_t n, const float* __restrict__ logits,
                                                                  float confidence, uint32_t* __restrict__ codes,
                                                                  float* __restrict__ priority) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float z[kOut];
#pragma unroll
  for (int c = 0; c < kOut; c++) z[c] = logits[static_cast<size_t>(i) * kOut + c];
  float top = z[0];
  int best = 0;
#pragma unroll
  for (int c = 1; c < kOut; c++) {
    if (z[c] > top) {
      top = z[c];
      best = c;
    }
  }
  float total = 0.f, p[kOut];
#pragma unroll
  for (int c = 0; c < kOut; c++) total += p[c] = __expf(z[c] - top);
  const float inv = 1.f / total;
  constexpr uint32_t kCode[kOut] = {1u, 2u, 3u, 0u};  // class -> action code
  if (p[best] * inv >= confidence) codes[i] = kCode[best];
  priority[i] = (p[1] + p[2]) * inv;
}

// A pending sample: the 12 inputs, the rule's class, the Gaussian's first index
// after the densification, the loss over its footprint before it (the mean
// per-pixel loss share over the pixels it blended into, weighted by its blend
// weight) and whether it blended into any pixel at all (1 or 0). Loss shares may
// be negative (the SSIM term's constant is kept apart from them), so only the
// comparison of two footprint losses of one view size is meaningful.
constexpr int kPendingWords = 16;

__device__ inline bool footprint_loss(const float* footprint, uint32_t i, float& loss) {
  const float w = footprint[2 * static_cast<size_t>(i) + 1];
  loss = w > 0.f ? footprint[2 * static_cast<size_t>(i)] / w : 0.f;
  return w > 0.f;
}

__device__ inline uint32_t class_of_code(uint32_t code) { return code == 0u ? 3u : code - 1u; }

__global__ __launch_bounds__(kThreads) void log_pending_kernel(uint32_t n, const float* __restrict__ features,
                                                                const float* __restrict__ footprint,
                                                                const uint32_t* __restrict__ codes,
                                                                const uint32_t* __restrict__ outputs,
                                                                uint32_t step, uint32_t threshold,
                                                                uint32_t capacity, float* __restrict__ pending,
                                                                uint32_t* __restrict__ pending_count) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n || pcg_hash(pcg_hash(step) ^ i) > threshold) return;
  const uint32_t slot = atomicAdd(pending_count, 1u);
  if (slot >= capacity) return;
  float* out = pending + static_cast<size_t>(slot) * kPendingWords;
#pragma unroll
  for (int k = 0; k < kIn; k++) out[k] = features[static_cast<size_t>(i) * kIn + k];
  out[12] = __uint_as_float(class_of_code(codes[i]));
  out[13] = __uint_as_float(outputs[i]);
  float loss;
  out[15] = footprint_loss(footprint, i, loss) ? 1.f : 0.f;
  out[14] = loss;
}

// A pending sample enters the training set when its outcome held: the Gaussian
// was pruned (its footprint is gone, so its loss cannot be compared), or the loss
// over the footprint of its first descendant, in the views since, fell below the
// Gaussian's own before. A Gaussian that blended into no pixel before, or whose
// descendant blended into none since, gives no verdict.
__global__ __launch_bounds__(kThreads) void resolve_kernel(const uint32_t* __restrict__ pending_count,
                                                            uint32_t pending_capacity,
                                                            const float* __restrict__ pending,
                                                            const float* __restrict__ footprint, uint32_t n,
                                                            uint32_t store_capacity, float* __restrict__ store_features,
                                                            uint32_t* __restrict__ store_labels,
                                                            uint32_t* __restrict__ store_count) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= min(*pending_count, pending_capacity)) return;
  const float* in = pending + static_cast<size_t>(j) * kPendingWords;
  const uint32_t label = __float_as_uint(in[12]);
  bool held = label == 3u;
  if (!held) {
    const uint32_t o = __float_as_uint(in[13]);
    if (o >= n) return;
    float after;
    const bool seen = footprint_loss(footprint, o, after);
    held = in[15] != 0.f && seen && after < in[14];
  }
  if (!held) return;
  const uint32_t slot = atomicAdd(store_count, 1u);
  if (slot >= store_capacity) return;
#pragma unroll
  for (int k = 0; k < kIn; k++) store_features[static_cast<size_t>(slot) * kIn + k] = in[k];
  store_labels[slot] = label;
}

}  // namespace

void mlp_learned_codes(hipStream_t s, uint32_t n, const float* logits, float confidence, uint32_t* codes,
                       float* priority) {
  if (n == 0) return;
  learned_codes_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, logits, confidence, codes, priority);
  DIO_LAUNCHED("mlp_learned_codes");
}

void mlp_log_pending(hipStream_t s, uint32_t n, const float* features, const float* footprint, const uint32_t* codes,
                     const uint32_t* outputs, uint32_t step, uint32_t threshold, uint32_t capacity, float* pending,
                     uint32_t* pending_count) {
  if (n == 0) return;
  log_pending_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, features, footprint, codes, outputs, step,
                                                                threshold, capacity, pending, pending_count);
  DIO_LAUNCHED("mlp_log_pending");
}

void mlp_resolve(hipStream_t s, const uint32_t* pending_count, uint32_t pending_capacity, const float* pending,
                 const float* footprint, uint32_t n, uint32_t store_capacity, float* store_features,
                 uint32_t* store_labels, uint32_t* store_count) {
  if (pending_capacity == 0) return;
  resolve_kernel<<<ceil_div(pending_capacity, kThreads), kThreads, 0, s>>>(
      pending_count, pending_capacity, pending, footprint, n, store_capacity, store_features, store_labels,
      store_count);
  DIO_LAUNCHED("mlp_resolve");
}

void mlp_features(hipStream_t s, uint32_t n, const float* params, const float* stats, const float* extra,
                  float extent, uint32_t iteration, uint32_t views_per_step, float* features) {
  if (n == 0) return;
  features_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, params, reinterpret_cast<const float4*>(stats),
                                                              reinterpret_cast<const float4*>(extra), extent,
                                                              iteration, static_cast<float>(views_per_step),
                                                              features);
  DIO_LAUNCHED("mlp_features");
}

void mlp_forward(hipStream_t s, uint32_t n, const float* features, const uint16_t* weights, const float* norm,
                 float* logits) {
  if (n == 0) return;
  // A wave takes groups of 16 Gaussians; enough waves to fill the GPU a few times over.
  const uint32_t blocks = std::min<uint32_t>(ceil_div(ceil_div(n, 16u), kThreads / kWave), 2048);
  mlp_forward_kernel<<<blocks, kThreads, 0, s>>>(n, features, weights, norm, logits);
  DIO_LAUNCHED("mlp_forward");
}

void mlp_gradients(hipStream_t s, uint32_t batch, uint32_t count, uint32_t seed, const float* features,
                   const uint32_t* labels, const uint16_t* weights, const float* norm, const float class_weights[4],
                   float* grads, float* loss) {
  if (batch == 0) return;
  const float4 cw = make_float4(class_weights[0], class_weights[1], class_weights[2], class_weights[3]);
  mlp_train_kernel<<<ceil_div(batch, kTrainSamples), kThreads, 0, s>>>(
      batch, count, seed, features, labels, weights, norm, cw, grads, loss);
  DIO_LAUNCHED("mlp_train");
}

void mlp_adam(hipStream_t s, float* master, float* moments, float* grads, uint16_t* weights, float lr,
              float inv_bias1, float inv_bias2) {
  mlp_adam_kernel<<<ceil_div(kParams, kThreads), kThreads, 0, s>>>(master, reinterpret_cast<float2*>(moments), grads,
                                                                     weights, lr, inv_bias1, inv_bias2);
  DIO_LAUNCHED("mlp_adam");
}

void mlp_round(hipStream_t s, const float* master, uint16_t* weights) {
  mlp_round_kernel<<<ceil_div(kParams, kThreads), kThreads, 0, s>>>(master, weights);
  DIO_LAUNCHED("mlp_round");
}

}  // namespace dio
