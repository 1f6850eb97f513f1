// The .ply body of trained Gaussians, packed and unpacked where the Gaussians
// live (HBM), so export and import move the file body over PCIe once and never
// transpose it on the host.
//
// Writer (src/plyWriter.ts:8-43): 62 float columns per vertex in the 3DGS order
// x y z nx ny nz f_dc_0..2 f_rest_0..44 opacity scale_0..2 rot_0..3, raw values
// (log scales, opacity logit, unnormalised quaternion), normals 0, and the 45
// higher-order coefficients channel-major (15 red, 15 green, 15 blue) where
// training keeps them coefficient-major (c1 rgb, c2 rgb, ...).
//
// Loader (src/plyLoader.ts:21-98): any float column order and SH degree 0-3,
// activated into the renderer's 16-float Gaussian (exp scales, sigmoid opacity,
// 0.5 + SH_C0 dc) with coefficient-major SH, zero past the file's degree. The
// activations run in double and round once to float, as the loader's
// Math.exp arithmetic stored into a Float32Array does.
#include "dio_internal.hpp"

namespace dio {
namespace {

constexpr double kShC0 = 0.28209479177387814;  // src/plyLoader.ts:9
constexpr uint32_t kPackRows = 128;             // Gaussians per pack workgroup

// Column c (0..61) of vertex g's row, from the staged raw params and SH.
__device__ inline float ply_column(const float* p, const float* sh, uint32_t c) {
  if (c < 3) return p[c];
  if (c < 6) return 0.f;                           // nx ny nz
  if (c < 9) return p[12 + (c - 6)];               // f_dc_*
  if (c < 9 + kShRest) {                           // f_rest_*: channel = j / 15, coefficient = j % 15
    const uint32_t j = c - 9;
    return sh[(j % 15) * 3 + j / 15];
  }
  if (c == 9 + kShRest) return p[3];               // opacity
  if (c < 13 + kShRest) return p[4 + (c - 55)];    // scale_*
  return p[8 + (c - 58)];                          // rot_*
}

__global__ __launch_bounds__(256) void ply_pack_kernel(const float* __restrict__ params, const float* __restrict__ sh,
                                                        uint32_t count, float* __restrict__ body) {
  __shared__ __align__(16) float s_p[kPackRows * kFloatsPerGaussian];
  __shared__ __align__(16) float s_sh[kPackRows * kShRest];
  const uint32_t first = blockIdx.x * kPackRows;
  const uint32_t rows = min(kPackRows, count - first);
  const size_t p0 = static_cast<size_t>(first) * kFloatsPerGaussian, s0 = static_cast<size_t>(first) * kShRest;
  // Both chunks start 16-byte aligned (128 x 64 B, 128 x 180 B) and are read as whole float4s where full.
  for (uint32_t q = threadIdx.x; q < rows * kFloatsPerGaussian / 4; q += 256)
    reinterpret_cast<float4*>(s_p)[q] = reinterpret_cast<const float4*>(params + p0)[q];
  const uint32_t sh_floats = rows * kShRest, sh_quads = sh_floats / 4;
  for (uint32_t q = threadIdx.x; q < sh_quads; q += 256)
    reinterpret_cast<float4*>(s_sh)[q] = reinterpret_cast<const float4*>(sh + s0)[q];
  for (uint32_t e = sh_quads * 4 + threadIdx.x; e < sh_floats; e += 256) s_sh[e] = sh[s0 + e];
  __syncthreads();
  float* out = body + static_cast<size_t>(first) * kPlyColumns;
  const uint32_t total = rows * kPlyColumns;
  for (uint32_t e = threadIdx.x; e < total; e += 256) {
    const uint32_t g = e / kPlyColumns, c = e % kPlyColumns;
    out[e] = ply_column(s_p + g * kFloatsPerGaussian, s_sh + g * kShRest, c);
  }
}

__global__ __launch_bounds__(256) void ply_unpack_kernel(const float* __restrict__ body, uint32_t count,
                                                          PlyColumns cols, uint32_t rows_per_block,
                                                          float* __restrict__ gaussians, float* __restrict__ sh) {
  extern __shared__ float s_rows[];  // rows_per_block x stride floats, at most 64 KiB
  const uint32_t first = blockIdx.x * rows_per_block;
  const uint32_t rows = min(rows_per_block, count - first);
  const size_t b0 = static_cast<size_t>(first) * cols.stride;
  for (uint32_t e = threadIdx.x; e < rows * cols.stride; e += 256) s_rows[e] = body[b0 + e];
  __syncthreads();
  float* g_out = gaussians + static_cast<size_t>(first) * kFloatsPerGaussian;
  for (uint32_t e = threadIdx.x; e < rows * kFloatsPerGaussian; e += 256) {
    const float* r = s_rows + (e / kFloatsPerGaussian) * cols.stride;
    const uint32_t k = e % kFloatsPerGaussian;
    float v = 0.f;  // pads 7 and 15
    if (k < 3) {
      v = r[cols.position[k]];
    } else if (k == 3) {
      v = static_cast<float>(1.0 / (1.0 + exp(-static_cast<double>(r[cols.opacity]))));
    } else if (k < 7) {
      v = static_cast<float>(exp(static_cast<double>(r[cols.scale[k - 4]])));
    } else if (k >= 8 && k < 12) {
      v = r[cols.rotation[k - 8]];
    } else if (k >= 12 && k < 15) {
      v = static_cast<float>(0.5 + kShC0 * static_cast<double>(r[cols.dc[k - 12]]));
    }
    g_out[e] = v;
  }
  float* s_out = sh + static_cast<size_t>(first) * kShRest;
  for (uint32_t e = threadIdx.x; e < rows * kShRest; e += 256) {
    const float* r = s_rows + (e / kShRest) * cols.stride;
    const uint32_t j = e % kShRest, c = j / 3, channel = j % 3;
    s_out[e] = c < cols.per_channel ? r[cols.rest[channel * cols.per_channel + c]] : 0.f;
  }
}

}  // namespace

void ply_pack(hipStream_t s, const float* params, const float* sh, uint32_t count, float* body) {
  if (count == 0) return;
  ply_pack_kernel<<<ceil_div(count, kPackRows), 256, 0, s>>>(params, sh, count, body);
  DIO_LAUNCHED("ply_pack_kernel");
}

uint32_t ply_unpack_rows(uint32_t stride) { return std::min<uint32_t>(256u, kPlyLdsFloats / stride); }

void ply_unpack(hipStream_t s, const float* body, uint32_t count, const PlyColumns& cols, float* gaussians, float* sh) {
  if (count == 0) return;
  const uint32_t rows = ply_unpack_rows(cols.stride);
  ply_unpack_kernel<<<ceil_div(count, rows), 256, rows * cols.stride * sizeof(float), s>>>(body, count, cols, rows,
                                                                                          gaussians, sh);
  DIO_LAUNCHED("ply_unpack_kernel");
}

}  // namespace dio
