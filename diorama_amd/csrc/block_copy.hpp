// Workgroup copies between global memory and LDS. Per-Gaussian records (the
// 180-byte SH records, 64-byte Gaussians, 36-byte splat gradients) are moved by the
// whole workgroup 16 bytes per lane, so every load is coalesced; per-thread access
// would touch a new cache line per lane.
#pragma once
#include "dio_internal.hpp"

namespace dio {

// `count` floats with a workgroup of kBlock threads (both sides 16-byte aligned).
// The loop keeps one load per thread in flight: for a partial block of unknown size.
template <uint32_t kBlock>
__device__ inline void block_copy(float* __restrict__ dst, const float* __restrict__ src, uint32_t count) {
  const uint32_t quads = count / 4;
  for (uint32_t q = threadIdx.x; q < quads; q += kBlock)
    reinterpret_cast<float4*>(dst)[q] = reinterpret_cast<const float4*>(src)[q];
  for (uint32_t k = quads * 4 + threadIdx.x; k < count; k += kBlock) dst[k] = src[k];
}

// A run of kVec float4s whose size is known at compile time (a full block), copied
// in rounds of kBlock threads with the loads split from the stores: a kernel issues
// the loads of all its runs, then stores them, so a thread has every load in flight
// at once. In the last round the lanes past the run take element q mod kVec, which
// round 0 also copies (same data, same address), so no load or store is conditional
// and none can be sunk into a branch that waits for it.
// Threads kFirst .. kFirst + kBlock - 1 of the workgroup do the copy.
template <uint32_t kVec, uint32_t kBlock, uint32_t kFirst = 0>
struct RunCopy {
  static constexpr uint32_t kRounds = (kVec + kBlock - 1) / kBlock;
  float4 v[kRounds];
  __device__ static uint32_t index(uint32_t k) {
    const uint32_t q = (threadIdx.x - kFirst) + k * kBlock;
    return k + 1 == kRounds && kRounds * kBlock > kVec ? q % kVec : q;
  }
  __device__ void load(const float* __restrict__ src) {
#pragma unroll
    for (uint32_t k = 0; k < kRounds; k++) v[k] = reinterpret_cast<const float4*>(src)[index(k)];
  }
  __device__ void store(float* __restrict__ dst) const {
#pragma unroll
    for (uint32_t k = 0; k < kRounds; k++) reinterpret_cast<float4*>(dst)[index(k)] = v[k];
  }
};

// A full block of kRecords records of `kFloats` floats each.
template <uint32_t kRecords, uint32_t kFloats, uint32_t kBlock, uint32_t kFirst = 0>
struct RecordCopy : RunCopy<kRecords * kFloats / 4, kBlock, kFirst> {
  static_assert(kRecords * kFloats % 4 == 0, "a full block's records are whole float4s");
};

}  // namespace dio
