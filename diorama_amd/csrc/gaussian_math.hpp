// Device math shared by the projection and its backward pass, so the two cannot
// drift apart: quaternion rotation, the SH basis of degrees 1-3 and its gradient
// (src/shaders/gaussianMath.wgsl:5-76), and the EWA projection of one Gaussian
// (src/shaders/preprocess.wgsl:28-68). All FP32.
#pragma once
#include "dio_internal.hpp"

namespace dio {

constexpr float kNearCull = 0.2f;
constexpr float kBlur = 0.3f;
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f, kShC2_1 = -1.0925484305920792f, kShC2_2 = 0.31539156525252005f,
                kShC2_3 = -1.0925484305920792f, kShC2_4 = 0.5462742152789498f;
constexpr float kShC3_0 = -0.5900435899266435f, kShC3_1 = 2.890611442640554f, kShC3_2 = -0.4570457994644658f,
                kShC3_3 = 0.3731763325901154f, kShC3_4 = -0.4570457994644658f, kShC3_5 = 1.445305721320277f,
                kShC3_6 = -0.5900435899266435f;

__device__ inline int sh_count(uint32_t degree) { return static_cast<int>((degree + 1) * (degree + 1) - 1); }

// r[row][col] of the normalised (w, x, y, z) quaternion.
__device__ inline void rotation_matrix(float qw, float qx, float qy, float qz, float r[3][3], float q[4],
                                       float* length) {
  float l = sqrtf(qw * qw + qx * qx + qy * qy + qz * qz);
  float w = qw / l, x = qx / l, y = qy / l, z = qz / l;
  r[0][0] = 1.f - 2.f * (y * y + z * z); r[0][1] = 2.f * (x * y - w * z); r[0][2] = 2.f * (x * z + w * y);
  r[1][0] = 2.f * (x * y + w * z); r[1][1] = 1.f - 2.f * (x * x + z * z); r[1][2] = 2.f * (y * z - w * x);
  r[2][0] = 2.f * (x * z - w * y); r[2][1] = 2.f * (y * z + w * x); r[2][2] = 1.f - 2.f * (x * x + y * y);
  q[0] = w; q[1] = x; q[2] = y; q[3] = z;
  *length = l;
}

__device__ inline void sh_basis(float x, float y, float z, float b[15]) {
  float xx = x * x, yy = y * y, zz = z * z;
  b[0] = -kShC1 * y; b[1] = kShC1 * z; b[2] = -kShC1 * x;
  b[3] = kShC2_0 * x * y; b[4] = kShC2_1 * y * z; b[5] = kShC2_2 * (2.f * zz - xx - yy);
  b[6] = kShC2_3 * x * z; b[7] = kShC2_4 * (xx - yy);
  b[8] = kShC3_0 * y * (3.f * xx - yy); b[9] = kShC3_1 * x * y * z;
  b[10] = kShC3_2 * y * (4.f * zz - xx - yy); b[11] = kShC3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy);
  b[12] = kShC3_4 * x * (4.f * zz - xx - yy); b[13] = kShC3_5 * z * (xx - yy);
  b[14] = kShC3_6 * x * (xx - 3.f * yy);
}

// d(basis k)/d(x, y, z) as three arrays.
__device__ inline void sh_basis_gradient(float x, float y, float z, float gx[15], float gy[15], float gz[15]) {
  float xx = x * x, yy = y * y, zz = z * z;
  gx[0] = 0; gy[0] = -kShC1; gz[0] = 0;
  gx[1] = 0; gy[1] = 0; gz[1] = kShC1;
  gx[2] = -kShC1; gy[2] = 0; gz[2] = 0;
  gx[3] = kShC2_0 * y; gy[3] = kShC2_0 * x; gz[3] = 0;
  gx[4] = 0; gy[4] = kShC2_1 * z; gz[4] = kShC2_1 * y;
  gx[5] = kShC2_2 * -2.f * x; gy[5] = kShC2_2 * -2.f * y; gz[5] = kShC2_2 * 4.f * z;
  gx[6] = kShC2_3 * z; gy[6] = 0; gz[6] = kShC2_3 * x;
  gx[7] = kShC2_4 * 2.f * x; gy[7] = kShC2_4 * -2.f * y; gz[7] = 0;
  gx[8] = kShC3_0 * 6.f * x * y; gy[8] = kShC3_0 * (3.f * xx - 3.f * yy); gz[8] = 0;
  gx[9] = kShC3_1 * y * z; gy[9] = kShC3_1 * x * z; gz[9] = kShC3_1 * x * y;
  gx[10] = kShC3_2 * -2.f * x * y; gy[10] = kShC3_2 * (4.f * zz - xx - 3.f * yy); gz[10] = kShC3_2 * 8.f * y * z;
  gx[11] = kShC3_3 * -6.f * x * z; gy[11] = kShC3_3 * -6.f * y * z; gz[11] = kShC3_3 * (6.f * zz - 3.f * xx - 3.f * yy);
  gx[12] = kShC3_4 * (4.f * zz - 3.f * xx - yy); gy[12] = kShC3_4 * -2.f * x * y; gz[12] = kShC3_4 * 8.f * x * z;
  gx[13] = kShC3_5 * 2.f * x * z; gy[13] = kShC3_5 * -2.f * y * z; gz[13] = kShC3_5 * (xx - yy);
  gx[14] = kShC3_6 * (3.f * xx - 3.f * yy); gy[14] = kShC3_6 * -6.f * x * y; gz[14] = 0;
}

// The projection of one Gaussian and every intermediate its backward pass needs.
struct Projection {
  float t[3], depth;
  float r[3][3], q[4], q_length;
  float m[3][3], cov3d[3][3];
  float focal[2], limit[2], ratio[2], tx, ty;
  float jw[2][3];
  float a, b, c, det;
  float clip[4];
  float mean[2];
  float radius;
};

// preprocess.wgsl:28-87. Returns false when culled (near plane, degenerate
// covariance or entirely off screen). Every field is filled in either way, so the
// backward pass can recompute a visible splat's intermediates unconditionally.
__device__ inline bool project(const float* g, const CameraArgs& cam, Projection& p) {
  const float* v = cam.view;
  for (int r = 0; r < 3; r++) p.t[r] = v[r] * g[0] + v[4 + r] * g[1] + v[8 + r] * g[2] + v[12 + r];
  p.depth = -p.t[2];
  rotation_matrix(g[8], g[9], g[10], g[11], p.r, p.q, &p.q_length);
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) p.m[r][k] = p.r[r][k] * g[4 + k];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) p.cov3d[r][c] = p.m[r][0] * p.m[c][0] + p.m[r][1] * p.m[c][1] + p.m[r][2] * p.m[c][2];
  p.focal[0] = cam.proj[0] * cam.viewport[0] * 0.5f;
  p.focal[1] = cam.proj[5] * cam.viewport[1] * 0.5f;
  const float d = p.depth;
  for (int k = 0; k < 2; k++) {
    p.limit[k] = 1.3f * cam.viewport[k] * 0.5f / p.focal[k];
    p.ratio[k] = p.t[k] / d;
  }
  p.tx = fminf(fmaxf(p.ratio[0], -p.limit[0]), p.limit[0]) * d;
  p.ty = fminf(fmaxf(p.ratio[1], -p.limit[1]), p.limit[1]) * d;
  const float j00 = p.focal[0] / d, j02 = p.focal[0] * p.tx / (d * d);
  const float j11 = -p.focal[1] / d, j12 = -p.focal[1] * p.ty / (d * d);
  // W[r][c] = view[c * 4 + r]; JW = J W.
  for (int c = 0; c < 3; c++) {
    p.jw[0][c] = j00 * v[c * 4 + 0] + j02 * v[c * 4 + 2];
    p.jw[1][c] = j11 * v[c * 4 + 1] + j12 * v[c * 4 + 2];
  }
  float tmp[2][3];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++)
      tmp[r][c] = p.jw[r][0] * p.cov3d[0][c] + p.jw[r][1] * p.cov3d[1][c] + p.jw[r][2] * p.cov3d[2][c];
  const float c00 = tmp[0][0] * p.jw[0][0] + tmp[0][1] * p.jw[0][1] + tmp[0][2] * p.jw[0][2];
  const float c01 = tmp[0][0] * p.jw[1][0] + tmp[0][1] * p.jw[1][1] + tmp[0][2] * p.jw[1][2];
  const float c11 = tmp[1][0] * p.jw[1][0] + tmp[1][1] * p.jw[1][1] + tmp[1][2] * p.jw[1][2];
  p.a = c00 + kBlur;
  p.b = c01;
  p.c = c11 + kBlur;
  p.det = p.a * p.c - p.b * p.b;
  const float mid = 0.5f * (p.a + p.c);
  const float lambda = mid + sqrtf(fmaxf(0.1f, mid * mid - p.det));
  p.radius = ceilf(3.f * sqrtf(lambda));
  const float* P = cam.proj;
  for (int r = 0; r < 4; r++) p.clip[r] = P[r] * p.t[0] + P[4 + r] * p.t[1] + P[8 + r] * p.t[2] + P[12 + r];
  p.mean[0] = (p.clip[0] / p.clip[3] * 0.5f + 0.5f) * cam.viewport[0];
  p.mean[1] = (p.clip[1] / p.clip[3] * -0.5f + 0.5f) * cam.viewport[1];
  if (p.depth <= kNearCull || p.det <= 0.f) return false;
  if (p.mean[0] + p.radius < 0.f || p.mean[1] + p.radius < 0.f || p.mean[0] - p.radius > cam.viewport[0] ||
      p.mean[1] - p.radius > cam.viewport[1])
    return false;
  return true;
}

__device__ inline void camera_position(const CameraArgs& cam, float cp[3]) {
  const float* v = cam.view;
  for (int c = 0; c < 3; c++) cp[c] = -(v[c * 4] * v[12] + v[c * 4 + 1] * v[13] + v[c * 4 + 2] * v[14]);
}

// One Gaussian's splat (preprocess.wgsl:20-87 + tiles.wgsl:23-34) from its
// projection `p`: the three float4 of the Splat2D record, the number of tiles (16
// px, or 32-px bins with cam.bin_shift 1) it touches in cam's viewport and its
// depth-sort key. An invisible
// Gaussian gets a zero record, no tiles and the largest key. `coeff` holds its
// `count` SH coefficient triples.
struct SplatRecord {
  float4 s0, s1, s2;
  uint32_t touched, key;
  uint2 rect;  // the tiles (bins) touched, packed: x0 | y0 << 16, x1 | y1 << 16
};


__device__ inline void splat_record(const float g[16], const Projection& p, bool visible, const CameraArgs& cam,
                                    const float* coeff, int count, SplatRecord& o) {
  o.s0 = o.s1 = o.s2 = make_float4(0.f, 0.f, 0.f, 0.f);
  o.touched = 0;
  o.key = 0xffffffffu;
  o.rect = make_uint2(0u, 0u);
  if (!visible) return;
  float color[3] = {g[12], g[13], g[14]};
  if (count > 0) {
    float cp[3];
    camera_position(cam, cp);
    float ox = g[0] - cp[0], oy = g[1] - cp[1], oz = g[2] - cp[2];
    float l = sqrtf(ox * ox + oy * oy + oz * oz);
    float b[15];
    sh_basis(ox / l, oy / l, oz / l, b);
#pragma unroll
    for (int k = 0; k < 15; k++) {  // unrolled with a guard: b stays in registers
      if (k < count) {
        color[0] += b[k] * coeff[3 * k];
        color[1] += b[k] * coeff[3 * k + 1];
        color[2] += b[k] * coeff[3 * k + 2];
      }
    }
  }
  o.s0 = make_float4(p.mean[0], p.mean[1], p.depth, p.radius);
  o.s1 = make_float4(p.c / p.det, -p.b / p.det, p.a / p.det, 0.f);
  o.s2 = make_float4(fmaxf(color[0], 0.f), fmaxf(color[1], 0.f), fmaxf(color[2], 0.f), g[3]);
  const uint32_t tiles_x = ceil_div(static_cast<uint32_t>(cam.viewport[0]), kTile);
  const uint32_t tiles_y = ceil_div(static_cast<uint32_t>(cam.viewport[1]), kTile);
  uint32_t r[4];
  bin_rect(p.mean[0], p.mean[1], p.radius, tiles_x, tiles_y, cam.bin_shift, r);
  o.touched = (r[2] - r[0]) * (r[3] - r[1]);
  o.key = __float_as_uint(p.depth);
  o.rect = pack_rect(r);
}

// One Gaussian's gradients for one view (preprocessBackward.wgsl:27-181): the
// splat's 9 gradients `gr` (mean2, conic3, opacity, rgb) carried back through the
// projection `p` (recomputed with project()) to position, opacity, scale,
// rotation and colour (out0..out3, the 16-float Gaussian layout). The SH basis of
// the view direction and the clamp-masked colour gradient come back too: the SH
// gradients are basis[k] * raw_grad[c] for the first `count` coefficients.
struct GaussianGrad {
  float4 out0, out1, out2, out3;
  float basis[15];
  float raw_grad[3];
};

__device__ inline void gaussian_backward(const float g[16], const Projection& p, const CameraArgs& cam,
                                         const float gr[9], const float* coeff, int count, GaussianGrad& o) {
  const float d = p.depth, d2 = d * d, d3 = d2 * d;
  const float a = p.a, b = p.b, c = p.c, det = p.det, det2 = det * det;
  // 1. conic -> 2D covariance
  const float ga = (-c * c * gr[2] + b * c * gr[3] - b * b * gr[4]) / det2;
  const float gb = (2.f * b * c * gr[2] - (det + 2.f * b * b) * gr[3] + 2.f * a * b * gr[4]) / det2;
  const float gc = (-b * b * gr[2] + a * b * gr[3] - a * a * gr[4]) / det2;
  const float G2[2][2] = {{ga, 0.5f * gb}, {0.5f * gb, gc}};
  // 2. dL/dcov3d = JWᵀ G JW, dL/dJ = 2 G JW cov3d Wᵀ
  float cov3d_grad[3][3];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++)
      cov3d_grad[r][k] = p.jw[0][r] * (G2[0][0] * p.jw[0][k] + G2[0][1] * p.jw[1][k]) +
                         p.jw[1][r] * (G2[1][0] * p.jw[0][k] + G2[1][1] * p.jw[1][k]);
  float gjw[2][3], gjwc[2][3], j_grad[2][3];
  for (int u = 0; u < 2; u++)
    for (int k = 0; k < 3; k++) gjw[u][k] = G2[u][0] * p.jw[0][k] + G2[u][1] * p.jw[1][k];
  for (int u = 0; u < 2; u++)
    for (int k = 0; k < 3; k++)
      gjwc[u][k] = gjw[u][0] * p.cov3d[0][k] + gjw[u][1] * p.cov3d[1][k] + gjw[u][2] * p.cov3d[2][k];
  const float* v = cam.view;
  for (int u = 0; u < 2; u++)
    for (int k = 0; k < 3; k++)  // (X Wᵀ)[u][k] = Σ_e X[u][e] W[k][e], W[k][e] = view[e * 4 + k]
      j_grad[u][k] = 2.f * (gjwc[u][0] * v[k] + gjwc[u][1] * v[4 + k] + gjwc[u][2] * v[8 + k]);
  // 3. J -> view-space centre, where its evaluation point was not clamped
  const float free_x = fabsf(p.ratio[0]) > p.limit[0] ? 0.f : 1.f;
  const float free_y = fabsf(p.ratio[1]) > p.limit[1] ? 0.f : 1.f;
  const float fx = p.focal[0], fy = p.focal[1];
  float t_grad[3];
  t_grad[0] = j_grad[0][2] * free_x * fx / d2;
  t_grad[1] = j_grad[1][2] * free_y * -fy / d2;
  const float depth_grad = j_grad[0][0] * (-fx / d2) + j_grad[1][1] * (fy / d2) +
                           j_grad[0][2] * (-(1.f + free_x) * fx * p.tx / d3) +
                           j_grad[1][2] * ((1.f + free_y) * fy * p.ty / d3);
  t_grad[2] = -depth_grad;
  // 4. the 2D centre through the projection
  const float ndc_x = gr[0] * 0.5f * cam.viewport[0], ndc_y = gr[1] * -0.5f * cam.viewport[1];
  const float cw = p.clip[3];
  const float clip_grad[4] = {ndc_x / cw, ndc_y / cw, 0.f, -(ndc_x * p.clip[0] + ndc_y * p.clip[1]) / (cw * cw)};
  const float* P = cam.proj;
  for (int k = 0; k < 3; k++)
    t_grad[k] += P[k * 4] * clip_grad[0] + P[k * 4 + 1] * clip_grad[1] + P[k * 4 + 2] * clip_grad[2] +
                 P[k * 4 + 3] * clip_grad[3];
  // 5. colour through the SH, masked where the forward clamp cut it
  float cp[3];
  camera_position(cam, cp);
  const float ox = g[0] - cp[0], oy = g[1] - cp[1], oz = g[2] - cp[2];
  const float dist = sqrtf(ox * ox + oy * oy + oz * oz);
  const float dx = ox / dist, dy = oy / dist, dz = oz / dist;
  float* basis = o.basis;
  sh_basis(dx, dy, dz, basis);
  float raw[3] = {g[12], g[13], g[14]};
#pragma unroll
  for (int k = 0; k < 15; k++) {  // unrolled with a guard: the basis stays in registers
    if (k < count) {
      raw[0] += basis[k] * coeff[3 * k];
      raw[1] += basis[k] * coeff[3 * k + 1];
      raw[2] += basis[k] * coeff[3 * k + 2];
    }
  }
  float* raw_grad = o.raw_grad;
  raw_grad[0] = raw[0] < 0.f ? 0.f : gr[6];
  raw_grad[1] = raw[1] < 0.f ? 0.f : gr[7];
  raw_grad[2] = raw[2] < 0.f ? 0.f : gr[8];
  float bgx[15], bgy[15], bgz[15];
  sh_basis_gradient(dx, dy, dz, bgx, bgy, bgz);
  float dgx = 0.f, dgy = 0.f, dgz = 0.f;
#pragma unroll
  for (int k = 0; k < 15; k++) {
    if (k < count) {
      const float dot = coeff[3 * k] * raw_grad[0] + coeff[3 * k + 1] * raw_grad[1] + coeff[3 * k + 2] * raw_grad[2];
      dgx += bgx[k] * dot;
      dgy += bgy[k] * dot;
      dgz += bgz[k] * dot;
    }
  }
  const float dd = dx * dgx + dy * dgy + dz * dgz;
  // 6. position: t = W p + translation
  float pos_grad[3];
  const float dir[3] = {dx, dy, dz}, dirg[3] = {dgx, dgy, dgz};
  for (int e = 0; e < 3; e++)
    pos_grad[e] = v[e * 4] * t_grad[0] + v[e * 4 + 1] * t_grad[1] + v[e * 4 + 2] * t_grad[2] +
                  (dirg[e] - dir[e] * dd) / dist;
  // 7. M = R S -> scale and rotation
  float m_grad[3][3];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++)
      m_grad[r][k] = 2.f * (cov3d_grad[r][0] * p.m[0][k] + cov3d_grad[r][1] * p.m[1][k] + cov3d_grad[r][2] * p.m[2][k]);
  float scale_grad[3], rg[3][3];
  for (int k = 0; k < 3; k++) {
    scale_grad[k] = m_grad[0][k] * p.r[0][k] + m_grad[1][k] * p.r[1][k] + m_grad[2][k] * p.r[2][k];
    for (int r = 0; r < 3; r++) rg[r][k] = m_grad[r][k] * g[4 + k];
  }
  const float qw = p.q[0], qx = p.q[1], qy = p.q[2], qz = p.q[3];
  float qg[4];
  qg[0] = 2.f * (-qz * rg[0][1] + qy * rg[0][2] + qz * rg[1][0] - qx * rg[1][2] - qy * rg[2][0] + qx * rg[2][1]);
  qg[1] = 2.f * (qy * rg[0][1] + qz * rg[0][2] + qy * rg[1][0] - 2.f * qx * rg[1][1] - qw * rg[1][2] +
                 qz * rg[2][0] + qw * rg[2][1] - 2.f * qx * rg[2][2]);
  qg[2] = 2.f * (-2.f * qy * rg[0][0] + qx * rg[0][1] + qw * rg[0][2] + qx * rg[1][0] + qz * rg[1][2] -
                 qw * rg[2][0] + qz * rg[2][1] - 2.f * qy * rg[2][2]);
  qg[3] = 2.f * (-2.f * qz * rg[0][0] - qw * rg[0][1] + qx * rg[0][2] + qw * rg[1][0] - 2.f * qz * rg[1][1] +
                 qy * rg[1][2] + qx * rg[2][0] + qy * rg[2][1]);
  const float qdot = qw * qg[0] + qx * qg[1] + qy * qg[2] + qz * qg[3];
  const float inv_l = 1.f / p.q_length;
  o.out0 = make_float4(pos_grad[0], pos_grad[1], pos_grad[2], gr[5]);
  o.out1 = make_float4(scale_grad[0], scale_grad[1], scale_grad[2], 0.f);
  o.out2 = make_float4((qg[0] - qw * qdot) * inv_l, (qg[1] - qx * qdot) * inv_l, (qg[2] - qy * qdot) * inv_l,
                       (qg[3] - qz * qdot) * inv_l);
  o.out3 = make_float4(raw_grad[0], raw_grad[1], raw_grad[2], 0.f);
}

}  // namespace dio
