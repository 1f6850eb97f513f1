// The optimizer stage: densification statistics, Adam with activation, opacity
// reset, and clone/split/prune by stream compaction (src/shaders/adam.wgsl:51-134,
// src/shaders/densify.wgsl:64-180).
//
// Adam runs one thread per (Gaussian, moment slot) rather than one thread per
// Gaussian: the 61 float2 moments of a Gaussian are contiguous, so a flat index
// over slots reads and writes them fully coalesced, which is what bounds this
// kernel (1,708 bytes per Gaussian per step).
#include "block_copy.hpp"
#include "gaussian_math.hpp"

namespace dio {
namespace {

constexpr int kThreads = 256;
constexpr float kBeta1 = 0.9f, kBeta2 = 0.999f, kEpsilon = 1e-15f;
constexpr float kResetLogit = -4.59511985013459f;  // logit(0.01)

// Hardware reciprocal and square root (about 1 ulp): Adam is a byte stream, and the
// IEEE-exact division and root sequences would make it instruction-bound.
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ inline float sigmoid(float x) { return fast_rcp(1.f + __expf(-x)); }

__device__ inline float activated(uint32_t slot, float raw) {
  if (slot == 3) return sigmoid(raw);
  if (slot >= 4 && slot < 7) return __expf(raw);
  if (slot >= 12 && slot < 15) return 0.5f + kShC0 * raw;
  return raw;
}

// Adam for 32 Gaussians per workgroup. Their parameters, SH coefficients and both
// gradient arrays are contiguous runs (64 and 180 bytes per Gaussian), so the
// workgroup moves them through LDS with 16-byte loads and stores; the 61 moment
// slots per Gaussian are read and written straight from global memory, one slot
// per lane (512 contiguous bytes per wave). Per slot the update is
// adam.wgsl:51-57 with the chain rule of adam.wgsl:85-105.
constexpr uint32_t kAdamGaussians = 32;

// One Adam slot (adam.wgsl:51-107): the chain rule from the activated gradient to
// the raw parameter, the moment update (m updated in place) and the step,
// lr (m / bc1) / (sqrt(v / bc2) + eps) with the two corrections folded on the host;
// returns the new raw value and sets `act` to its activated value (as activated()).
// Without branches: a wave holds lanes of different slots, so every branch would
// run masked. One exp serves the chain rule (exp(-x) of the opacity sigmoid,
// exp(x) of a scale) and one the activated value; each lane's arithmetic is that
// of the branchy form.
__device__ inline float adam_slot(uint32_t slot, float value, float grad, float2& m, AdamArgs a, float& act) {
  const bool opacity = slot == 3, scale = slot - 4u < 3u, color = slot - 12u < 3u;
  const float e = __expf(opacity ? -value : value);
  const float o = fast_rcp(1.f + e);
  grad *= opacity ? o * (1.f - o) : (scale ? e : (color ? kShC0 : 1.f));
  const float lr = slot < 3 ? a.position_lr
                   : opacity ? a.opacity_lr
                   : slot < 7 ? a.scale_lr
                   : slot < 12 ? a.rotation_lr
                   : slot < 16 ? a.color_lr
                               : a.sh_lr;
  m.x = kBeta1 * m.x + (1.f - kBeta1) * grad;
  m.y = kBeta2 * m.y + (1.f - kBeta2) * grad * grad;
  const float updated =
      value - lr * a.inv_bias_correction1 * m.x * fast_rcp(fast_sqrt(m.y * a.inv_bias_correction2) + kEpsilon);
  const float e2 = __expf(opacity ? -updated : updated);
  act = opacity ? fast_rcp(1.f + e2) : (scale ? e2 : (color ? 0.5f + kShC0 * updated : updated));
  return updated;
}

// Walks a workgroup's (Gaussian, slot) elements f = threadIdx.x + k x 256 without
// dividing: 256 = 4 x 61 + 12.
struct SlotWalk {
  uint32_t il, slot;
  __device__ explicit SlotWalk(uint32_t f) : il(f / kMomentSlots), slot(f % kMomentSlots) {}
  __device__ void next() {
    static_assert(kThreads == 4 * kMomentSlots + 12, "the step below assumes 256 threads");
    il += 4;
    slot += 12;
    if (slot >= kMomentSlots) {
      slot -= kMomentSlots;
      il += 1;
    }
  }
};

// kDeviceArgs: the arguments come from device memory (`args`) instead of `a`.
template <bool kDeviceArgs>
__global__ __launch_bounds__(kThreads) void adam_kernel(uint32_t n, float* __restrict__ params, float* __restrict__ sh,
                                                         float2* __restrict__ moments,
                                                         const float* __restrict__ gaussian_grads,
                                                         const float* __restrict__ sh_grads, AdamArgs a_value,
                                                         const AdamArgs* __restrict__ args,
                                                         float* __restrict__ gaussians) {
  const AdamArgs a = kDeviceArgs ? *args : a_value;
  __shared__ __align__(16) float s_p[kAdamGaussians * kFloatsPerGaussian];
  __shared__ __align__(16) float s_gp[kAdamGaussians * kFloatsPerGaussian];
  __shared__ __align__(16) float s_sh[kAdamGaussians * kShRest];
  __shared__ __align__(16) float s_gsh[kAdamGaussians * kShRest];
  __shared__ __align__(16) float s_act[kAdamGaussians * kFloatsPerGaussian];
  const uint32_t first = blockIdx.x * kAdamGaussians;
  const uint32_t count = min(kAdamGaussians, n - first);
  const float* src_p = params + first * kFloatsPerGaussian;
  const float* src_gp = gaussian_grads + first * kFloatsPerGaussian;
  const float* src_sh = sh + first * kShRest;
  const float* src_gsh = sh_grads + first * kShRest;
  if (count == kAdamGaussians) {
    RecordCopy<kAdamGaussians, kFloatsPerGaussian, kThreads> c_p, c_gp;
    RecordCopy<kAdamGaussians, kShRest, kThreads> c_sh, c_gsh;
    c_p.load(src_p);
    c_gp.load(src_gp);
    c_sh.load(src_sh);
    c_gsh.load(src_gsh);
    c_p.store(s_p);
    c_gp.store(s_gp);
    c_sh.store(s_sh);
    c_gsh.store(s_gsh);
  } else {
    block_copy<kThreads>(s_p, src_p, count * kFloatsPerGaussian);
    block_copy<kThreads>(s_gp, src_gp, count * kFloatsPerGaussian);
    block_copy<kThreads>(s_sh, src_sh, count * kShRest);
    block_copy<kThreads>(s_gsh, src_gsh, count * kShRest);
  }
  __syncthreads();
  // A thread's moments are all loaded before the first update; the padding slots
  // (7: kind, 15: birth stamp) are never trained and go back unchanged.
  constexpr uint32_t kRounds = (kAdamGaussians * kMomentSlots + kThreads - 1) / kThreads;
  const uint32_t total = count * kMomentSlots;
  float2* m_base = moments + first * kMomentSlots;
  float2 m[kRounds];
#pragma unroll
  for (uint32_t k = 0; k < kRounds; k++) m[k] = m_base[min(threadIdx.x + k * kThreads, total - 1)];
  SlotWalk w(threadIdx.x);
#pragma unroll
  for (uint32_t k = 0; k < kRounds; k++, w.next()) {
    const uint32_t f = threadIdx.x + k * kThreads;
    if (f >= total) break;
    const uint32_t i = w.il, slot = w.slot;
    if (slot == 7 || slot == 15) {
      s_act[i * kFloatsPerGaussian + slot] = 0.f;
    } else {
      float* value;
      float grad;
      if (slot < kFloatsPerGaussian) {
        value = s_p + i * kFloatsPerGaussian + slot;
        grad = s_gp[i * kFloatsPerGaussian + slot];
      } else {
        value = s_sh + i * kShRest + (slot - kFloatsPerGaussian);
        grad = s_gsh[i * kShRest + (slot - kFloatsPerGaussian)];
      }
      float act;
      const float updated = adam_slot(slot, *value, grad, m[k], a, act);
      *value = updated;
      if (slot < kFloatsPerGaussian) s_act[i * kFloatsPerGaussian + slot] = act;
    }
    m_base[f] = m[k];
  }
  __syncthreads();
  if (count == kAdamGaussians) {
    RecordCopy<kAdamGaussians, kFloatsPerGaussian, kThreads> c_p;
    RecordCopy<kAdamGaussians, kShRest, kThreads> c_sh;
    c_p.load(s_p);
    c_sh.load(s_sh);
    c_p.store(params + first * kFloatsPerGaussian);
    c_sh.store(sh + first * kShRest);
  } else {
    block_copy<kThreads>(params + first * kFloatsPerGaussian, s_p, count * kFloatsPerGaussian);
    block_copy<kThreads>(sh + first * kShRest, s_sh, count * kShRest);
  }
  if (gaussians) {
    // The pads (kind, birth) of the activated copy are left as they are.
    for (uint32_t e = threadIdx.x; e < count * kFloatsPerGaussian; e += kThreads) {
      const uint32_t slot = e % kFloatsPerGaussian;
      if (slot != 7 && slot != 15) gaussians[first * kFloatsPerGaussian + e] = s_act[e];
    }
  }
}

__global__ __launch_bounds__(kThreads) void activate_kernel(uint32_t n, const float* __restrict__ params,
                                                             float* __restrict__ gaussians) {
  const size_t f = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (f >= static_cast<size_t>(n) * kFloatsPerGaussian) return;
  const uint32_t slot = static_cast<uint32_t>(f % kFloatsPerGaussian);
  if (slot == 7 || slot == 15) return;
  gaussians[f] = activated(slot, params[f]);
}

// adam.wgsl:125-134. `extra` (optional, for the learned densification) gathers
// (largest gradient length, summed depth, summed gradient x, summed gradient y).
__global__ __launch_bounds__(kThreads) void stats_kernel(uint32_t n, const float4* __restrict__ splats,
                                                          const float* __restrict__ splat_grads, float vx, float vy,
                                                          float4* __restrict__ stats, float4* __restrict__ extra) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float4 s0 = splats[i * 3];
  const float radius = s0.w;
  if (radius == 0.f) return;
  const float gx = splat_grads[static_cast<size_t>(i) * 9] * vx * 0.5f;
  const float gy = splat_grads[static_cast<size_t>(i) * 9 + 1] * vy * 0.5f;
  const float length = sqrtf(gx * gx + gy * gy);
  const float4 s = stats[i];
  stats[i] = make_float4(s.x + length, s.y + 1.f, fmaxf(s.z, radius), 0.f);
  if (extra) {
    const float4 e = extra[i];
    extra[i] = make_float4(fmaxf(e.x, length), e.y + s0.z, e.z + gx, e.w + gy);
  }
}

// The plain trainer's step after the rasterizer's backward, in one pass: the
// preprocess backward (preprocessBackward.wgsl:27-187), the densification
// statistics (adam.wgsl:125-134) and Adam with activation (adam.wgsl:51-107),
// with the same arithmetic as preprocess_backward_kernel, stats_kernel and
// adam_kernel. A workgroup takes 64 Gaussians: their raw and activated
// parameters, SH and splat gradients come into LDS with 16-byte loads; one wave
// runs their backward (a Gaussian per lane) and keeps only what Adam needs, 16
// parameter gradients, 15 SH basis values and 3 colour gradients (an SH
// coefficient's gradient is basis x colour gradient); then all four waves run
// Adam over every (Gaussian, slot) with the moments streamed coalesced, and the
// updated parameters leave through LDS with 16-byte stores. The per-Gaussian
// gradients never reach HBM (about 0.5 KB a Gaussian saved against separate
// passes); every Gaussian, visible or not, gets its Adam step, as the reference.
// (Measured against variants: 32 Gaussians with the moments staged in LDS, and
// waves 1-3 loading their moments ahead while wave 0 runs the backward, were
// both slower. The backward phase wants 106 VGPRs, which allows 4 waves per
// SIMD; asking for 5, which the 31 KB of LDS also allows, costs a 36-byte spill
// in that phase and gains about 6% on the kernel.)
constexpr uint32_t kFusedGaussians = 64;
constexpr uint32_t kFusedRow = 35;  // 16 + 15 + 3, odd so lanes' rows start in different banks

template <bool kExtra, bool kPrep>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void backward_adam_kernel(
    uint32_t n, float* __restrict__ gaussians, float* __restrict__ sh, uint32_t sh_degree, CameraArgs cam,
    const float2* __restrict__ vis, const float* __restrict__ splat_grads, float* __restrict__ params,
    float2* __restrict__ moments, AdamArgs a, float4* __restrict__ stats, float4* __restrict__ extra,
    PrepArgs prep) {
  __shared__ __align__(16) float s_sh[kFusedGaussians * kShRest];
  __shared__ __align__(16) float s_raw[kFusedGaussians * kFloatsPerGaussian];
  __shared__ __align__(16) float s_act[kFusedGaussians * kFloatsPerGaussian];
  __shared__ __align__(16) float s_sg[kFusedGaussians * 9];
  __shared__ float s_rows[kFusedGaussians * kFusedRow];
  __shared__ float2 s_splat[kFusedGaussians];  // depth, radius
  __shared__ float4 s_stats[kFusedGaussians];
  __shared__ float4 s_extra[kExtra ? kFusedGaussians : 1];
  // Blocks run from the last Gaussians to the first: the next step's preprocess reads
  // the activated Gaussians and SH from the first ones on, so it starts on the lines
  // written last, which the memory-side cache still holds.
  const uint32_t first = (gridDim.x - 1 - blockIdx.x) * kFusedGaussians;
  const uint32_t count = min(kFusedGaussians, n - first);
  // Stage the block's inputs with all loads in flight together; wave 1 fetches the
  // splats' depth and radius (8 bytes each rather than their 48-byte records), wave 2
  // the statistics, wave 3 the extra statistics.
  const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const uint32_t mine = first + min(lane, count - 1);
  float4 side = make_float4(0.f, 0.f, 0.f, 0.f);
  if (wave == 1) {
    const float2 v = vis[mine];
    side = make_float4(0.f, 0.f, v.x, v.y);
  }
  if (wave == 2) side = stats[mine];
  if (kExtra && wave == 3) side = extra[mine];
  float* g_sh = sh + static_cast<size_t>(first) * kShRest;
  float* g_raw = params + static_cast<size_t>(first) * kFloatsPerGaussian;
  float* g_act = gaussians + static_cast<size_t>(first) * kFloatsPerGaussian;
  const float* g_sg = splat_grads + static_cast<size_t>(first) * 9;
  if (count == kFusedGaussians) {
    RecordCopy<kFusedGaussians, kShRest, kThreads> c_sh;
    RecordCopy<kFusedGaussians, kFloatsPerGaussian, kThreads> c_raw, c_act;
    RecordCopy<kFusedGaussians, 9, kThreads> c_sg;
    c_sh.load(g_sh);
    c_raw.load(g_raw);
    c_act.load(g_act);
    c_sg.load(g_sg);
    c_sh.store(s_sh);
    c_raw.store(s_raw);
    c_act.store(s_act);
    c_sg.store(s_sg);
  } else {
    block_copy<kThreads>(s_sh, g_sh, count * kShRest);
    block_copy<kThreads>(s_raw, g_raw, count * kFloatsPerGaussian);
    block_copy<kThreads>(s_act, g_act, count * kFloatsPerGaussian);
    block_copy<kThreads>(s_sg, g_sg, count * 9);
  }
  if (wave == 1) s_splat[lane] = make_float2(side.z, side.w);
  if (wave == 2) s_stats[lane] = side;
  if (kExtra && wave == 3) s_extra[lane] = side;
  __syncthreads();
  if (threadIdx.x < kFusedGaussians) {
    const uint32_t t = threadIdx.x, i = first + t;
    float* row = s_rows + t * kFusedRow;
    for (uint32_t k = 0; k < kFusedRow; k++) row[k] = 0.f;
    const float2 splat = t < count ? s_splat[t] : make_float2(0.f, 0.f);
    if (splat.y != 0.f) {
      const int used = sh_count(sh_degree);
      float g[16], gr[9];
#pragma unroll
      for (int k = 0; k < 16; k++) g[k] = s_act[t * kFloatsPerGaussian + k];
#pragma unroll
      for (int v = 0; v < 9; v++) gr[v] = s_sg[t * 9 + v];
      Projection p;
      project(g, cam, p);
      GaussianGrad o;
      gaussian_backward(g, p, cam, gr, s_sh + t * kShRest, used, o);
      const float out[16] = {o.out0.x, o.out0.y, o.out0.z, o.out0.w, o.out1.x, o.out1.y, o.out1.z, o.out1.w,
                             o.out2.x, o.out2.y, o.out2.z, o.out2.w, o.out3.x, o.out3.y, o.out3.z, o.out3.w};
#pragma unroll
      for (int k = 0; k < 16; k++) row[k] = out[k];
#pragma unroll
      for (int k = 0; k < 15; k++) row[16 + k] = k < used ? o.basis[k] : 0.f;
      row[31] = o.raw_grad[0];
      row[32] = o.raw_grad[1];
      row[33] = o.raw_grad[2];
      // the statistics, as stats_kernel
      const float gx = gr[0] * a.viewport[0] * 0.5f, gy = gr[1] * a.viewport[1] * 0.5f;
      const float length = sqrtf(gx * gx + gy * gy);
      const float4 st = s_stats[t];
      stats[i] = make_float4(st.x + length, st.y + 1.f, fmaxf(st.z, splat.y), 0.f);
      if (kExtra) {
        const float4 e = s_extra[t];
        extra[i] = make_float4(fmaxf(e.x, length), e.y + splat.x, e.z + gx, e.w + gy);
      }
    }
  }
  __syncthreads();
  // Adam over every (Gaussian, slot): a thread's moments are all loaded before the
  // first update. The padding slots (7: kind, 15: birth stamp) are never trained; their
  // moments go back unchanged so that every written line is whole.
  constexpr uint32_t kRounds = (kFusedGaussians * kMomentSlots + kThreads - 1) / kThreads;
  const uint32_t total = count * kMomentSlots;
  float2* m_base = moments + static_cast<size_t>(first) * kMomentSlots;
  float2 m[kRounds];
#pragma unroll
  for (uint32_t k = 0; k < kRounds; k++) m[k] = m_base[min(threadIdx.x + k * kThreads, total - 1)];
  SlotWalk w(threadIdx.x);
#pragma unroll
  for (uint32_t k = 0; k < kRounds; k++, w.next()) {
    const uint32_t f = threadIdx.x + k * kThreads;
    if (f >= total) break;
    const uint32_t il = w.il, slot = w.slot;
    if (slot != 7 && slot != 15) {
      const float* r = s_rows + il * kFusedRow;
      float* value;
      float grad;
      if (slot < kFloatsPerGaussian) {
        value = s_raw + il * kFloatsPerGaussian + slot;
        grad = r[slot];
      } else {
        const uint32_t j = slot - kFloatsPerGaussian, c = (j * 43u) >> 7;  // j / 3 for j < 45
        value = s_sh + il * kShRest + j;
        grad = r[16 + c] * r[31 + (j - 3 * c)];
      }
      float act;
      const float updated = adam_slot(slot, *value, grad, m[k], a, act);
      *value = updated;
      if (slot < kFloatsPerGaussian) s_act[il * kFloatsPerGaussian + slot] = act;
    }
    m_base[f] = m[k];
  }
  __syncthreads();
  if (kPrep && count == kFusedGaussians) {
    // Waves 1-3 write the block back while wave 0 runs the next step's preprocess
    // below; both only read LDS.
    if (threadIdx.x >= kWave) {
      RecordCopy<kFusedGaussians, kShRest, kThreads - kWave, kWave> c_sh;
      RecordCopy<kFusedGaussians, kFloatsPerGaussian, kThreads - kWave, kWave> c_raw, c_act;
      c_sh.load(s_sh);
      c_raw.load(s_raw);
      c_act.load(s_act);
      c_sh.store(g_sh);
      c_raw.store(g_raw);
      c_act.store(g_act);
    }
  } else if (count == kFusedGaussians) {
    RecordCopy<kFusedGaussians, kShRest, kThreads> c_sh;
    RecordCopy<kFusedGaussians, kFloatsPerGaussian, kThreads> c_raw, c_act;
    c_sh.load(s_sh);
    c_raw.load(s_raw);
    c_act.load(s_act);
    c_sh.store(g_sh);
    c_raw.store(g_raw);
    c_act.store(g_act);
  } else {
    block_copy<kThreads>(g_raw, s_raw, count * kFloatsPerGaussian);
    block_copy<kThreads>(g_sh, s_sh, count * kShRest);
    block_copy<kThreads>(g_act, s_act, count * kFloatsPerGaussian);
  }
  if (kPrep && threadIdx.x < count) {
    // The next step's preprocess, as preprocess_kernel computes it (project, then
    // splat_record), from the updated activated Gaussian and SH still in LDS: the
    // same values the step writes out above.
    const uint32_t t = threadIdx.x, i = first + t;
    float g[16];
#pragma unroll
    for (int k = 0; k < 16; k++) g[k] = s_act[t * kFloatsPerGaussian + k];
    Projection p;
    const bool visible = project(g, prep.cam, p);
    SplatRecord rec;
    splat_record(g, p, visible, prep.cam, s_sh + t * kShRest, sh_count(prep.sh_degree), rec);
    float4* out = reinterpret_cast<float4*>(prep.splats) + static_cast<size_t>(i) * 3;
    out[0] = rec.s0;
    out[1] = rec.s1;
    out[2] = rec.s2;
    prep.tile_counts[i] = rec.touched;
    prep.depth_keys[i] = rec.key;
    prep.depth_values[i] = i;
    reinterpret_cast<float2*>(prep.vis)[i] = make_float2(rec.s0.z, rec.s0.w);
    if (prep.rects) reinterpret_cast<uint2*>(prep.rects)[i] = rec.rect;
  }
}

__global__ __launch_bounds__(kThreads) void reset_opacity_kernel(uint32_t n, float* __restrict__ params,
                                                                  float2* __restrict__ moments,
                                                                  float* __restrict__ gaussians) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float* p = params + static_cast<size_t>(i) * kFloatsPerGaussian;
  const float logit = fminf(p[3], kResetLogit);
  p[3] = logit;
  moments[static_cast<size_t>(i) * kMomentSlots + 3] = make_float2(0.f, 0.f);
  if (gaussians) gaussians[static_cast<size_t>(i) * kFloatsPerGaussian + 3] = sigmoid(logit);
}

enum : uint32_t { kDrop = 0, kKeep = 1, kClone = 2, kSplit = 3 };
constexpr float kSplitShrink = 1.6f;

__device__ inline bool removed(const DensifyArgs& a, float opacity, float size) {
  return opacity < a.min_opacity || (a.max_size > 0.f && size > a.max_size);
}

// densify.wgsl:72-85: the reference clones, then splits, then prunes the result.
__device__ inline uint32_t rule_action(const float* params, const float4* stats, const DensifyArgs& a, uint32_t i) {
  const float* p = params + static_cast<size_t>(i) * kFloatsPerGaussian;
  const float4 s = stats[i];
  const float grad = s.y > 0.f ? s.x / s.y : 0.f;
  const float size = __expf(fmaxf(p[4], fmaxf(p[5], p[6])));
  const float opacity = sigmoid(p[3]);
  if (grad >= a.grad_threshold) {
    if (size <= a.clone_max_size) return removed(a, opacity, size) ? kDrop : kClone;
    return removed(a, opacity, size / kSplitShrink) ? kDrop : kSplit;
  }
  return removed(a, opacity, size) ? kDrop : kKeep;
}

// The rule's action, or the one chosen for Gaussian i beforehand (`forced`).
__device__ inline uint32_t action(const float* params, const float4* stats, const DensifyArgs& a,
                                  const uint32_t* forced, uint32_t i) {
  return forced ? forced[i] : rule_action(params, stats, a, i);
}

// The rule's actions, and each Gaussian's claim to grow (its mean view-space
// gradient) for a budget that cannot take every clone and split.
__global__ __launch_bounds__(kThreads) void rule_codes_kernel(uint32_t n, const float* __restrict__ params,
                                                               const float4* __restrict__ stats, DensifyArgs a,
                                                               uint32_t* __restrict__ codes,
                                                               float* __restrict__ priority) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  codes[i] = rule_action(params, stats, a, i);
  const float4 st = stats[i];
  priority[i] = st.y > 0.f ? st.x / st.y : 0.f;
}

// Growth candidates (clone or split) in order of priority, highest first, as sort
// keys (the other Gaussians sort last).
__global__ __launch_bounds__(kThreads) void growth_keys_kernel(uint32_t n, const uint32_t* __restrict__ codes,
                                                                const float* __restrict__ priority,
                                                                uint32_t* __restrict__ keys,
                                                                uint32_t* __restrict__ values) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = codes[i];
  const bool grows = c == kClone || c == kSplit;
  keys[i] = grows ? ~__float_as_uint(fmaxf(priority[i], 0.f)) : 0xffffffffu;
  values[i] = i;
}

// Growth past the first `allowed` candidates of that order falls back to keeping
// (or pruning, where the rule prunes the unchanged Gaussian).
__global__ __launch_bounds__(kThreads) void demote_kernel(uint32_t n, const uint32_t* __restrict__ sorted_keys,
                                                           const uint32_t* __restrict__ sorted_values,
                                                           uint32_t allowed, const float* __restrict__ params,
                                                           DensifyArgs a, uint32_t* __restrict__ codes) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n || k < allowed || sorted_keys[k] == 0xffffffffu) return;
  const uint32_t i = sorted_values[k];
  const float* p = params + static_cast<size_t>(i) * kFloatsPerGaussian;
  const float size = __expf(fmaxf(p[4], fmaxf(p[5], p[6])));
  codes[i] = removed(a, sigmoid(p[3]), size) ? kDrop : kKeep;
}

__global__ __launch_bounds__(kThreads) void densify_count_kernel(uint32_t n, const float* __restrict__ params,
                                                                  const float4* __restrict__ stats, DensifyArgs a,
                                                                  const uint32_t* __restrict__ forced,
                                                                  uint32_t* __restrict__ outputs,
                                                                  uint32_t* __restrict__ counters) {
  __shared__ uint32_t tally[4];
  if (threadIdx.x < 4) tally[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    const uint32_t act = action(params, stats, a, forced, i);
    outputs[i] = act == kDrop ? 0u : (act == kKeep ? 1u : 2u);
    atomicAdd(&tally[act], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 4 && tally[threadIdx.x]) atomicAdd(&counters[threadIdx.x], tally[threadIdx.x]);
}

__device__ inline uint32_t pcg(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

__device__ inline float unit(uint32_t h) { return (static_cast<float>(h >> 8) + 0.5f) / 16777216.f; }

// densify.wgsl's scatter as a stream compaction written in runs: the 256 inputs of
// a workgroup own one contiguous run of outputs (offsets from the scan of their
// output counts). Each input first lays out its outputs in LDS: which input each
// comes from, whether it keeps that input's moments (a kept Gaussian and a clone's
// first copy) or starts them at zero (a clone's second copy and both split
// children), and its 16 parameters (copied, or as the clone and split rules
// change them). Then the workgroup writes the run's parameters, SH and moments as
// consecutive words, each read from its output's source record, so stores are
// coalesced and loads nearly so (a kept Gaussian's record follows the one before).
constexpr uint32_t kOutputsPerGroup = 2 * kThreads;

__global__ __launch_bounds__(kThreads) void densify_scatter_kernel(
    uint32_t n, const float* __restrict__ params, const float4* __restrict__ stats,
    const uint32_t* __restrict__ outputs, const float* __restrict__ sh, const float2* __restrict__ moments,
    DensifyArgs a, const uint32_t* __restrict__ forced, float* __restrict__ new_params, float* __restrict__ new_sh,
    float2* __restrict__ new_moments) {
  __shared__ float s_params[kOutputsPerGroup * kFloatsPerGaussian];
  __shared__ uint32_t s_source[kOutputsPerGroup];  // input (within the group) | zero moments << 16
  __shared__ uint32_t s_first, s_end;
  const uint32_t first = blockIdx.x * kThreads, i = first + threadIdx.x;
  const uint32_t last = min(first + kThreads, n) - 1;
  if (threadIdx.x == 0) s_first = outputs[first];
  __syncthreads();
  const uint32_t o0 = s_first;
  if (i < n) {
    const uint32_t act = action(params, stats, a, forced, i);
    const uint32_t o = outputs[i] - o0;  // this input's first output within the run
    const float* p = params + static_cast<size_t>(i) * kFloatsPerGaussian;
    const uint32_t t = threadIdx.x;
    auto put = [&](uint32_t slot, bool zero_moments) {
      s_source[slot] = t | (zero_moments ? 1u << 16 : 0u);
      for (uint32_t k = 0; k < kFloatsPerGaussian; k++) s_params[slot * kFloatsPerGaussian + k] = p[k];
    };
    if (act == kKeep || act == kClone) put(o, false);
    if (act == kClone) {
      put(o + 1, true);
      s_params[(o + 1) * kFloatsPerGaussian + 7] = 1.f;
      s_params[(o + 1) * kFloatsPerGaussian + 15] = static_cast<float>(a.seed);
    }
    if (act == kSplit) {
      float r[3][3], q[4], l;
      rotation_matrix(p[8], p[9], p[10], p[11], r, q, &l);
      const float sx = __expf(p[4]), sy = __expf(p[5]), sz = __expf(p[6]);
      const float shrink = logf(kSplitShrink);
      for (uint32_t c = 0; c < 2; c++) {
        const uint32_t child = o + c;
        put(child, true);
        // normal3(pcg(seed ^ pcg(2i + c))) of densify.wgsl:270-279,305: the seed is hashed once more.
        const uint32_t h1 = pcg(pcg(a.seed ^ pcg(2u * i + c))), h2 = pcg(h1), h3 = pcg(h2), h4 = pcg(h3);
        // Box-Muller: three standard normal numbers from four hashes.
        const float tau = 6.283185307179586f;
        const float r1 = sqrtf(-2.f * logf(unit(h1))), r2 = sqrtf(-2.f * logf(unit(h3)));
        const float nx = r1 * cosf(tau * unit(h2)), ny = r1 * sinf(tau * unit(h2)), nz = r2 * cosf(tau * unit(h4));
        const float vx = sx * nx, vy = sy * ny, vz = sz * nz;
        float* out = s_params + child * kFloatsPerGaussian;
        for (int k = 0; k < 3; k++) {
          out[k] = p[k] + r[k][0] * vx + r[k][1] * vy + r[k][2] * vz;
          out[4 + k] = p[4 + k] - shrink;
        }
        out[7] = 2.f;
        out[15] = static_cast<float>(a.seed);
      }
    }
    if (i == last) s_end = o + (act == kDrop ? 0u : act == kKeep ? 1u : 2u);
  }
  __syncthreads();
  const uint32_t count = s_end;  // outputs of the group, at most 2 per input
  for (uint32_t f = threadIdx.x; f < count * kFloatsPerGaussian; f += kThreads)
    new_params[static_cast<size_t>(o0) * kFloatsPerGaussian + f] = s_params[f];
  for (uint32_t f = threadIdx.x; f < count * kShRest; f += kThreads) {
    const uint32_t slot = f / kShRest, k = f - slot * kShRest;
    const uint32_t src = first + (s_source[slot] & 0xffffu);
    new_sh[static_cast<size_t>(o0) * kShRest + f] = sh[static_cast<size_t>(src) * kShRest + k];
  }
  for (uint32_t f = threadIdx.x; f < count * kMomentSlots; f += kThreads) {
    const uint32_t slot = f / kMomentSlots, k = f - slot * kMomentSlots;
    const uint32_t source = s_source[slot];
    const uint32_t src = first + (source & 0xffffu);
    new_moments[static_cast<size_t>(o0) * kMomentSlots + f] =
        source >> 16 ? make_float2(0.f, 0.f) : moments[static_cast<size_t>(src) * kMomentSlots + k];
  }
}

}  // namespace

void accumulate_stats(hipStream_t s, uint32_t n, const float* splats, const float* splat_grads, const AdamArgs& a,
                      float* stats, float* extra) {
  if (n == 0) return;
  stats_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, reinterpret_cast<const float4*>(splats), splat_grads,
                                                          a.viewport[0], a.viewport[1],
                                                          reinterpret_cast<float4*>(stats),
                                                          reinterpret_cast<float4*>(extra));
  DIO_LAUNCHED("accumulate_stats");
}

void adam_step(hipStream_t s, uint32_t n, float* params, float* sh, float* moments, const float* gaussian_grads,
               const float* sh_grads, const AdamArgs& a, float* gaussians) {
  if (n == 0) return;
  if (static_cast<uint64_t>(n) * kMomentSlots >= (1ull << 32)) throw Error(-2, "adam: too many Gaussians");
  adam_kernel<false><<<ceil_div(n, kAdamGaussians), kThreads, 0, s>>>(
      n, params, sh, reinterpret_cast<float2*>(moments), gaussian_grads, sh_grads, a, nullptr, gaussians);
  DIO_LAUNCHED("adam");
}

void adam_step_device_args(hipStream_t s, uint32_t n, float* params, float* sh, float* moments,
                           const float* gaussian_grads, const float* sh_grads, const AdamArgs* a, float* gaussians) {
  if (n == 0) return;
  if (static_cast<uint64_t>(n) * kMomentSlots >= (1ull << 32)) throw Error(-2, "adam: too many Gaussians");
  adam_kernel<true><<<ceil_div(n, kAdamGaussians), kThreads, 0, s>>>(
      n, params, sh, reinterpret_cast<float2*>(moments), gaussian_grads, sh_grads, AdamArgs{}, a, gaussians);
  DIO_LAUNCHED("adam");
}

void activate(hipStream_t s, uint32_t n, const float* params, float* gaussians) {
  if (n == 0) return;
  activate_kernel<<<ceil_div(static_cast<size_t>(n) * kFloatsPerGaussian, kThreads), kThreads, 0, s>>>(n, params,
                                                                                                        gaussians);
  DIO_LAUNCHED("activate");
}

void reset_opacity(hipStream_t s, uint32_t n, float* params, float* moments, float* gaussians) {
  if (n == 0) return;
  reset_opacity_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, params, reinterpret_cast<float2*>(moments),
                                                                   gaussians);
  DIO_LAUNCHED("reset_opacity");
}

void densify_count(hipStream_t s, uint32_t n, const float* params, const float* stats, const DensifyArgs& a,
                   const uint32_t* forced, uint32_t* outputs, uint32_t* counters) {
  fill_u32(s, counters, 0u, 4);
  if (n == 0) return;
  densify_count_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, params, reinterpret_cast<const float4*>(stats), a,
                                                                  forced, outputs, counters);
  DIO_LAUNCHED("densify_count");
}

void densify_scatter(hipStream_t s, uint32_t n, const float* params, const float* stats, const uint32_t* outputs,
                     const float* sh, const float* moments, const DensifyArgs& a, const uint32_t* forced,
                     float* new_params, float* new_sh, float* new_moments) {
  if (n == 0) return;
  densify_scatter_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(
      n, params, reinterpret_cast<const float4*>(stats), outputs, sh, reinterpret_cast<const float2*>(moments), a,
      forced, new_params, new_sh, reinterpret_cast<float2*>(new_moments));
  DIO_LAUNCHED("densify_scatter");
}

void densify_rule_codes(hipStream_t s, uint32_t n, const float* params, const float* stats, const DensifyArgs& a,
                        uint32_t* codes, float* priority) {
  if (n == 0) return;
  rule_codes_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, params, reinterpret_cast<const float4*>(stats), a,
                                                               codes, priority);
  DIO_LAUNCHED("densify_rule_codes");
}

void densify_growth_keys(hipStream_t s, uint32_t n, const uint32_t* codes, const float* priority, uint32_t* keys,
                         uint32_t* values) {
  if (n == 0) return;
  growth_keys_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, codes, priority, keys, values);
  DIO_LAUNCHED("densify_growth_keys");
}

void densify_demote(hipStream_t s, uint32_t n, const uint32_t* sorted_keys, const uint32_t* sorted_values,
                    uint32_t allowed, const float* params, const DensifyArgs& a, uint32_t* codes) {
  if (n == 0) return;
  demote_kernel<<<ceil_div(n, kThreads), kThreads, 0, s>>>(n, sorted_keys, sorted_values, allowed, params, a, codes);
  DIO_LAUNCHED("densify_demote");
}

void backward_adam(hipStream_t s, uint32_t n, float* gaussians, float* sh, uint32_t sh_degree, const CameraArgs& cam,
                   const float* vis, const float* splat_grads, float* params, float* moments, const AdamArgs& a,
                   float* stats, float* extra, const PrepArgs* prep) {
  if (n == 0) return;
  auto launch = [&](auto kernel) {
    kernel<<<ceil_div(n, kFusedGaussians), kThreads, 0, s>>>(
        n, gaussians, sh, sh_degree, cam, reinterpret_cast<const float2*>(vis), splat_grads, params,
        reinterpret_cast<float2*>(moments), a, reinterpret_cast<float4*>(stats), reinterpret_cast<float4*>(extra),
        prep ? *prep : PrepArgs{});
  };
  if (extra)
    prep ? launch(backward_adam_kernel<true, true>) : launch(backward_adam_kernel<true, false>);
  else
    prep ? launch(backward_adam_kernel<false, true>) : launch(backward_adam_kernel<false, false>);
  DIO_LAUNCHED("backward_adam");
}

}  // namespace dio
