/*
 * CPU oracle for the Gaussian-splatting training step, in float64.
 *
 * TEST INFRASTRUCTURE: only the tests, the smoke check and the benchmark's CPU
 * baseline load this file's library, and only as the checker or as the timed
 * CPU restatement. The product library never links or calls it.
 *
 * Each function restates one stage of the reference (paths relative to the
 * reference repository):
 *   oc_preprocess           src/shaders/preprocess.wgsl:20-87, src/checks/reference.ts:27-80
 *   oc_tile_lists           src/shaders/tiles.wgsl:23-84, src/checks/reference.ts:128-153
 *   oc_tile_lists_capped    the same with a pair capacity: tiles.wgsl:47-58 (slots past the
 *                           capacity dropped), src/tileRasterizer.ts:166,279 (the sort and
 *                           ranges over min(total, capacity) pairs)
 *   oc_rasterize            src/shaders/rasterize.wgsl:31-100, src/checks/reference.ts:156-176
 *   oc_rasterize_backward   src/shaders/rasterizeBackward.wgsl:49-163
 *   oc_preprocess_backward  src/shaders/preprocessBackward.wgsl:27-181
 *   oc_loss / oc_ssim       src/shaders/loss.wgsl:52-121, src/trainingLoss.ts:7-11,110-125,
 *                           reference/loss_reference.py:34-49, src/metrics.ts:115-167
 *   oc_activate, oc_adam_step, oc_reset_opacity, oc_accumulate_stats
 *                           src/shaders/adam.wgsl:51-134
 *   oc_densify              src/shaders/densify.wgsl:64-180, src/trainer.ts:258-337
 *   oc_train_step           src/trainer.ts:192-244 (one optimize() step)
 * Pinning of each function is recorded beside its tests.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEAR_CULL 0.2
#define BLUR 0.3
#define SH_C0 0.28209479177387814
#define SH_C1 0.4886025119029199
static const double SH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                                -1.0925484305920792, 0.5462742152789498};
static const double SH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                                0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                                -0.5900435899266435};

int oc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------- math -- */

/* R of the normalised quaternion (w, x, y, z); r[row][col]. gaussianMath.wgsl:13-25 */
static void rotation(const double* q_raw, double r[3][3], double q[4], double* length) {
  double l = sqrt(q_raw[0] * q_raw[0] + q_raw[1] * q_raw[1] + q_raw[2] * q_raw[2] + q_raw[3] * q_raw[3]);
  double w = q_raw[0] / l, x = q_raw[1] / l, y = q_raw[2] / l, z = q_raw[3] / l;
  r[0][0] = 1 - 2 * (y * y + z * z); r[0][1] = 2 * (x * y - w * z);     r[0][2] = 2 * (x * z + w * y);
  r[1][0] = 2 * (x * y + w * z);     r[1][1] = 1 - 2 * (x * x + z * z); r[1][2] = 2 * (y * z - w * x);
  r[2][0] = 2 * (x * z - w * y);     r[2][1] = 2 * (y * z + w * x);     r[2][2] = 1 - 2 * (x * x + y * y);
  if (q) { q[0] = w; q[1] = x; q[2] = y; q[3] = z; }
  if (length) *length = l;
}

/* The 15 real SH basis values of degrees 1..3 (gaussianMath.wgsl:30-46). */
static void sh_basis(const double d[3], double b[15]) {
  double x = d[0], y = d[1], z = d[2], xx = x * x, yy = y * y, zz = z * z;
  b[0] = -SH_C1 * y; b[1] = SH_C1 * z; b[2] = -SH_C1 * x;
  b[3] = SH_C2[0] * x * y; b[4] = SH_C2[1] * y * z; b[5] = SH_C2[2] * (2 * zz - xx - yy);
  b[6] = SH_C2[3] * x * z; b[7] = SH_C2[4] * (xx - yy);
  b[8] = SH_C3[0] * y * (3 * xx - yy); b[9] = SH_C3[1] * x * y * z;
  b[10] = SH_C3[2] * y * (4 * zz - xx - yy); b[11] = SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy);
  b[12] = SH_C3[4] * x * (4 * zz - xx - yy); b[13] = SH_C3[5] * z * (xx - yy);
  b[14] = SH_C3[6] * x * (xx - 3 * yy);
}

/* d(basis k)/d(x, y, z) (gaussianMath.wgsl:49-70). */
static void sh_basis_gradient(const double d[3], double g[15][3]) {
  double x = d[0], y = d[1], z = d[2], xx = x * x, yy = y * y, zz = z * z;
  memset(g, 0, sizeof(double) * 45);
  g[0][1] = -SH_C1;
  g[1][2] = SH_C1;
  g[2][0] = -SH_C1;
  g[3][0] = SH_C2[0] * y; g[3][1] = SH_C2[0] * x;
  g[4][1] = SH_C2[1] * z; g[4][2] = SH_C2[1] * y;
  g[5][0] = SH_C2[2] * -2 * x; g[5][1] = SH_C2[2] * -2 * y; g[5][2] = SH_C2[2] * 4 * z;
  g[6][0] = SH_C2[3] * z; g[6][2] = SH_C2[3] * x;
  g[7][0] = SH_C2[4] * 2 * x; g[7][1] = SH_C2[4] * -2 * y;
  g[8][0] = SH_C3[0] * 6 * x * y; g[8][1] = SH_C3[0] * (3 * xx - 3 * yy);
  g[9][0] = SH_C3[1] * y * z; g[9][1] = SH_C3[1] * x * z; g[9][2] = SH_C3[1] * x * y;
  g[10][0] = SH_C3[2] * -2 * x * y; g[10][1] = SH_C3[2] * (4 * zz - xx - 3 * yy); g[10][2] = SH_C3[2] * 8 * y * z;
  g[11][0] = SH_C3[3] * -6 * x * z; g[11][1] = SH_C3[3] * -6 * y * z; g[11][2] = SH_C3[3] * (6 * zz - 3 * xx - 3 * yy);
  g[12][0] = SH_C3[4] * (4 * zz - 3 * xx - yy); g[12][1] = SH_C3[4] * -2 * x * y; g[12][2] = SH_C3[4] * 8 * x * z;
  g[13][0] = SH_C3[5] * 2 * x * z; g[13][1] = SH_C3[5] * -2 * y * z; g[13][2] = SH_C3[5] * (xx - yy);
  g[14][0] = SH_C3[6] * (3 * xx - 3 * yy); g[14][1] = SH_C3[6] * -6 * x * y;
}

static int sh_count(int degree) { return (degree + 1) * (degree + 1) - 1; }

/* Everything the projection computes for one Gaussian, shared by forward and backward. */
typedef struct {
  double t[3], depth;
  double r[3][3], q[4], q_length;
  double m[3][3], cov3d[3][3];
  double focal[2], limit[2], ratio[2], tx, ty;
  double w[3][3], jw[2][3];
  double a, b, c, det;
  double clip[4];
  double mean[2];
  double radius;
} projection;

/* preprocess.wgsl:28-68. Returns 0 when the Gaussian is culled; apart from the near
 * cull, every intermediate is filled in either way. */
static int project(const double* g, const double* view, const double* proj, const double* vp, projection* p) {
  for (int r = 0; r < 3; r++)
    p->t[r] = view[r] * g[0] + view[4 + r] * g[1] + view[8 + r] * g[2] + view[12 + r];
  p->depth = -p->t[2];
  if (p->depth <= NEAR_CULL) return 0;
  rotation(g + 8, p->r, p->q, &p->q_length);
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) p->m[r][k] = p->r[r][k] * g[4 + k];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      p->cov3d[r][c] = p->m[r][0] * p->m[c][0] + p->m[r][1] * p->m[c][1] + p->m[r][2] * p->m[c][2];
  p->focal[0] = proj[0] * vp[0] * 0.5;
  p->focal[1] = proj[5] * vp[1] * 0.5;
  double d = p->depth;
  for (int k = 0; k < 2; k++) {
    p->limit[k] = 1.3 * vp[k] * 0.5 / p->focal[k];
    p->ratio[k] = p->t[k] / d;
  }
  double rx = p->ratio[0] < -p->limit[0] ? -p->limit[0] : (p->ratio[0] > p->limit[0] ? p->limit[0] : p->ratio[0]);
  double ry = p->ratio[1] < -p->limit[1] ? -p->limit[1] : (p->ratio[1] > p->limit[1] ? p->limit[1] : p->ratio[1]);
  p->tx = rx * d;
  p->ty = ry * d;
  double j[2][3] = {{p->focal[0] / d, 0, p->focal[0] * p->tx / (d * d)},
                    {0, -p->focal[1] / d, -p->focal[1] * p->ty / (d * d)}};
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) p->w[r][c] = view[c * 4 + r];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) p->jw[r][c] = j[r][0] * p->w[0][c] + j[r][1] * p->w[1][c] + j[r][2] * p->w[2][c];
  double tmp[2][3];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++)
      tmp[r][c] = p->jw[r][0] * p->cov3d[0][c] + p->jw[r][1] * p->cov3d[1][c] + p->jw[r][2] * p->cov3d[2][c];
  double cov[2][2];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 2; c++) cov[r][c] = tmp[r][0] * p->jw[c][0] + tmp[r][1] * p->jw[c][1] + tmp[r][2] * p->jw[c][2];
  p->a = cov[0][0] + BLUR;
  p->b = cov[0][1];
  p->c = cov[1][1] + BLUR;
  p->det = p->a * p->c - p->b * p->b;
  double mid = 0.5 * (p->a + p->c);
  double disc = mid * mid - p->det;
  double lambda = mid + sqrt(disc > 0.1 ? disc : 0.1);
  p->radius = ceil(3 * sqrt(lambda));
  for (int r = 0; r < 4; r++)
    p->clip[r] = proj[r] * p->t[0] + proj[4 + r] * p->t[1] + proj[8 + r] * p->t[2] + proj[12 + r];
  p->mean[0] = (p->clip[0] / p->clip[3] * 0.5 + 0.5) * vp[0];
  p->mean[1] = (p->clip[1] / p->clip[3] * -0.5 + 0.5) * vp[1];
  if (p->det <= 0) return 0;
  if (p->mean[0] + p->radius < 0 || p->mean[1] + p->radius < 0 || p->mean[0] - p->radius > vp[0] ||
      p->mean[1] - p->radius > vp[1])
    return 0;
  return 1;
}

static void camera_position(const double* view, double cp[3]) {
  for (int c = 0; c < 3; c++) cp[c] = -(view[c * 4] * view[12] + view[c * 4 + 1] * view[13] + view[c * 4 + 2] * view[14]);
}

void oc_preprocess(long n, const double* gaussians, const double* sh, int sh_degree, const double* view,
                   const double* proj, const double* viewport, double* splats) {
  double cp[3];
  camera_position(view, cp);
  int count = sh ? sh_count(sh_degree) : 0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    const double* g = gaussians + i * 16;
    double* s = splats + i * 12;
    memset(s, 0, sizeof(double) * 12);
    projection p;
    if (!project(g, view, proj, viewport, &p)) continue;
    double color[3] = {g[12], g[13], g[14]};
    if (count > 0) {
      double off[3] = {g[0] - cp[0], g[1] - cp[1], g[2] - cp[2]};
      double l = sqrt(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]);
      double d[3] = {off[0] / l, off[1] / l, off[2] / l}, b[15];
      sh_basis(d, b);
      for (int k = 0; k < count; k++)
        for (int ch = 0; ch < 3; ch++) color[ch] += b[k] * sh[i * 45 + k * 3 + ch];
    }
    s[0] = p.mean[0]; s[1] = p.mean[1]; s[2] = p.depth; s[3] = p.radius;
    s[4] = p.c / p.det; s[5] = -p.b / p.det; s[6] = p.a / p.det; s[7] = 0;
    for (int ch = 0; ch < 3; ch++) s[8 + ch] = color[ch] > 0 ? color[ch] : 0;
    s[11] = g[3];
  }
}

/* ---------------------------------------------------------------- tiles -- */

/* common.wgsl:37-42 with float32 edges (src/checks/reference.ts:128-136). */
static void tile_rect(const double* s, int tiles_x, int tiles_y, int rect[4]) {
  float mx = (float)s[0], my = (float)s[1], r = (float)s[3];
  float e[4] = {mx - r, my - r, mx + r, my + r};
  int lim[2] = {tiles_x, tiles_y};
  for (int k = 0; k < 4; k++) {
    double v = floor((double)e[k] / OC_TILE) + (k >= 2 ? 1 : 0);
    double hi = lim[k & 1];
    rect[k] = (int)(v < 0 ? 0 : (v > hi ? hi : v));
  }
}

typedef struct {
  uint32_t tile;
  double depth;
  uint32_t splat;
} pair_t;

static int pair_less(const void* pa, const void* pb) {
  const pair_t* a = (const pair_t*)pa;
  const pair_t* b = (const pair_t*)pb;
  if (a->tile != b->tile) return a->tile < b->tile ? -1 : 1;
  if (a->depth != b->depth) return a->depth < b->depth ? -1 : 1;
  return a->splat < b->splat ? -1 : (a->splat > b->splat);
}

/* Per-tile lists of splats front to back, ties in splat order (a stable sort of the
 * pairs in the order duplicate() writes them: tiles.wgsl:47-58). Returns the pair
 * count; at most `capacity` sorted pair splats are written. */
long oc_tile_lists(long n, const double* splats, int width, int height, uint32_t* ranges,
                   uint32_t* pair_splats, long capacity) {
  int tx = (width + OC_TILE - 1) / OC_TILE, ty = (height + OC_TILE - 1) / OC_TILE;
  long total = 0;
  for (long i = 0; i < n; i++) {
    if (splats[i * 12 + 3] == 0) continue;
    int r[4];
    tile_rect(splats + i * 12, tx, ty, r);
    total += (long)(r[2] - r[0]) * (r[3] - r[1]);
  }
  pair_t* pairs = (pair_t*)malloc(sizeof(pair_t) * (total > 0 ? total : 1));
  long k = 0;
  for (long i = 0; i < n; i++) {
    if (splats[i * 12 + 3] == 0) continue;
    int r[4];
    tile_rect(splats + i * 12, tx, ty, r);
    for (int y = r[1]; y < r[3]; y++)
      for (int x = r[0]; x < r[2]; x++) {
        pairs[k].tile = (uint32_t)(y * tx + x);
        pairs[k].depth = (double)(float)splats[i * 12 + 2];
        pairs[k].splat = (uint32_t)i;
        k++;
      }
  }
  qsort(pairs, total, sizeof(pair_t), pair_less);
  memset(ranges, 0, sizeof(uint32_t) * 2 * tx * ty);
  for (long p = 0; p < total; p++) {
    if (p < capacity) pair_splats[p] = pairs[p].splat;
    uint32_t t = pairs[p].tile;
    if (p == 0 || pairs[p - 1].tile != t) ranges[2 * t] = (uint32_t)p;
    if (p == total - 1 || pairs[p + 1].tile != t) ranges[2 * t + 1] = (uint32_t)(p + 1);
  }
  free(pairs);
  return total;
}

typedef struct {
  double depth;
  long splat;
} depth_t;

static int depth_less(const void* pa, const void* pb) {
  const depth_t* a = (const depth_t*)pa;
  const depth_t* b = (const depth_t*)pb;
  if (a->depth != b->depth) return a->depth < b->depth ? -1 : 1;
  return a->splat < b->splat ? -1 : (a->splat > b->splat);
}

/* A frame with a pair capacity (tiles.wgsl:47-58): splats take pair slots in depth
 * order (a stable sort, so ties in splat order), each splat's tiles row by row from
 * its first slot; pairs whose slot is at or past `capacity` are dropped, and the
 * kept ones are sorted and ranged as in oc_tile_lists. Returns the full pair count
 * (what the reference reads back); writes the min(total, capacity) kept pairs. */
long oc_tile_lists_capped(long n, const double* splats, int width, int height, long capacity, uint32_t* ranges,
                          uint32_t* pair_splats) {
  int tx = (width + OC_TILE - 1) / OC_TILE, ty = (height + OC_TILE - 1) / OC_TILE;
  depth_t* order = (depth_t*)malloc(sizeof(depth_t) * (n > 0 ? n : 1));
  long visible = 0, total = 0;
  for (long i = 0; i < n; i++) {
    if (splats[i * 12 + 3] == 0) continue;
    order[visible].depth = (double)(float)splats[i * 12 + 2];
    order[visible].splat = i;
    visible++;
  }
  qsort(order, visible, sizeof(depth_t), depth_less);
  for (long v = 0; v < visible; v++) {
    int r[4];
    tile_rect(splats + order[v].splat * 12, tx, ty, r);
    total += (long)(r[2] - r[0]) * (r[3] - r[1]);
  }
  long kept = total < capacity ? total : capacity;
  pair_t* pairs = (pair_t*)malloc(sizeof(pair_t) * (kept > 0 ? kept : 1));
  long slot = 0;
  for (long v = 0; v < visible && slot < kept; v++) {
    long i = order[v].splat;
    int r[4];
    tile_rect(splats + i * 12, tx, ty, r);
    for (int y = r[1]; y < r[3] && slot < kept; y++)
      for (int x = r[0]; x < r[2] && slot < kept; x++) {
        pairs[slot].tile = (uint32_t)(y * tx + x);
        pairs[slot].depth = order[v].depth;
        pairs[slot].splat = (uint32_t)i;
        slot++;
      }
  }
  qsort(pairs, kept, sizeof(pair_t), pair_less);
  memset(ranges, 0, sizeof(uint32_t) * 2 * tx * ty);
  for (long p = 0; p < kept; p++) {
    pair_splats[p] = pairs[p].splat;
    uint32_t t = pairs[p].tile;
    if (p == 0 || pairs[p - 1].tile != t) ranges[2 * t] = (uint32_t)p;
    if (p == kept - 1 || pairs[p + 1].tile != t) ranges[2 * t + 1] = (uint32_t)(p + 1);
  }
  free(pairs);
  free(order);
  return total;
}

/* ------------------------------------------------------------ rasterize -- */

static double splat_power(const double* s, double dx, double dy) {
  return -0.5 * (s[4] * dx * dx + s[6] * dy * dy) - s[5] * dx * dy;
}

/* `margin` (nullable) receives, per pixel, how close its blend came to one of the
 * reference's discontinuities: the relative distance of any examined alpha to the
 * 1/255 skip threshold or of the transmittance to the 1e-4 stop, or |power| at the
 * power > 0 skip. A float32 renderer may decide those differently within rounding. */
void oc_rasterize_margin(const double* splats, const uint32_t* ranges, const uint32_t* pair_splats, int width,
                         int height, const double* background, double* pixel_color, uint32_t* pixel_last,
                         double* margin) {
  int tx = (width + OC_TILE - 1) / OC_TILE;
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < (long)width * height; p++) {
    int x = (int)(p % width), y = (int)(p / width);
    int tile = (y / OC_TILE) * tx + x / OC_TILE;
    uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
    double T = 1, rgb[3] = {0, 0, 0}, near = 1e300;
    uint32_t last = 0;
    for (uint32_t k = start; k < end; k++) {
      const double* s = splats + (long)pair_splats[k] * 12;
      double dx = x + 0.5 - s[0], dy = y + 0.5 - s[1];
      double power = splat_power(s, dx, dy);
      double raw_alpha = s[11] * exp(power > 0 ? 0 : power);
      if (raw_alpha >= 1.0 / 255.0 && fabs(power) < near) near = fabs(power);
      if (power > 0) continue;
      double alpha = raw_alpha;
      if (alpha > 0.99) alpha = 0.99;
      double da = fabs(alpha * 255.0 - 1.0);
      if (da < near) near = da;
      if (alpha < 1.0 / 255.0) continue;
      double remaining = T * (1 - alpha);
      double dr = fabs(remaining / 0.0001 - 1.0);
      if (dr < near) near = dr;
      if (remaining < 0.0001) break;
      for (int ch = 0; ch < 3; ch++) rgb[ch] += s[8 + ch] * alpha * T;
      T = remaining;
      last = k - start + 1;
    }
    for (int ch = 0; ch < 3; ch++) pixel_color[p * 4 + ch] = rgb[ch] + T * background[ch];
    pixel_color[p * 4 + 3] = T;
    pixel_last[p] = last;
    if (margin) margin[p] = near;
  }
}

void oc_rasterize(const double* splats, const uint32_t* ranges, const uint32_t* pair_splats, int width,
                  int height, const double* background, double* pixel_color, uint32_t* pixel_last) {
  oc_rasterize_margin(splats, ranges, pair_splats, width, height, background, pixel_color, pixel_last, NULL);
}

/* The loss over each splat's footprint (no counterpart in the reference; the
 * learned densification's outcome signal): per splat, the sums over the pixels it
 * blends into of w x pixel_loss and of w, with w = alpha T its blend weight in the
 * front-to-back blend of rasterize.wgsl:64-82 (the weight its colour enters the
 * pixel with). footprint holds n x 2 doubles, overwritten. */
void oc_footprint(long n, const double* splats, const uint32_t* ranges, const uint32_t* pair_splats, int width,
                  int height, const double* pixel_loss, double* footprint) {
  memset(footprint, 0, sizeof(double) * 2 * n);
  int tx = (width + OC_TILE - 1) / OC_TILE;
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < (long)width * height; p++) {
    int x = (int)(p % width), y = (int)(p / width);
    int tile = (y / OC_TILE) * tx + x / OC_TILE;
    uint32_t start = ranges[2 * tile], end = ranges[2 * tile + 1];
    double T = 1;
    for (uint32_t k = start; k < end; k++) {
      const double* s = splats + (long)pair_splats[k] * 12;
      double dx = x + 0.5 - s[0], dy = y + 0.5 - s[1];
      double power = splat_power(s, dx, dy);
      if (power > 0) continue;
      double alpha = s[11] * exp(power);
      if (alpha > 0.99) alpha = 0.99;
      if (alpha < 1.0 / 255.0) continue;
      double remaining = T * (1 - alpha);
      if (remaining < 0.0001) break;
      double w = alpha * T;
      double* f = footprint + (long)pair_splats[k] * 2;
#pragma omp atomic
      f[0] += w * pixel_loss[p];
#pragma omp atomic
      f[1] += w;
      T = remaining;
    }
  }
}

/* rasterizeBackward.wgsl:49-163: the back-to-front replay, summed per splat. */
void oc_rasterize_backward(long n, const double* splats, const uint32_t* ranges, const uint32_t* pair_splats,
                           int width, int height, const double* background, const double* pixel_color,
                           const uint32_t* pixel_last, const double* pixel_grads, double* splat_grads) {
  memset(splat_grads, 0, sizeof(double) * 9 * n);
  int tx = (width + OC_TILE - 1) / OC_TILE;
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < (long)width * height; p++) {
    int x = (int)(p % width), y = (int)(p / width);
    int tile = (y / OC_TILE) * tx + x / OC_TILE;
    uint32_t start = ranges[2 * tile];
    double final_T = pixel_color[p * 4 + 3];
    const double* pg = pixel_grads + p * 4;
    double bg_grad = background[0] * pg[0] + background[1] * pg[1] + background[2] * pg[2];
    double T = final_T, behind[3] = {0, 0, 0}, next_alpha = 0, next_color[3] = {0, 0, 0};
    for (long k = (long)start + pixel_last[p] - 1; k >= (long)start; k--) {
      uint32_t index = pair_splats[k];
      const double* s = splats + (long)index * 12;
      double dx = x + 0.5 - s[0], dy = y + 0.5 - s[1];
      double power = splat_power(s, dx, dy);
      if (power > 0) continue;
      double G = exp(power);
      double alpha = s[11] * G;
      if (alpha > 0.99) alpha = 0.99;
      if (alpha < 1.0 / 255.0) continue;
      T /= 1 - alpha;
      for (int ch = 0; ch < 3; ch++) behind[ch] = next_alpha * next_color[ch] + (1 - next_alpha) * behind[ch];
      next_alpha = alpha;
      double dot = 0;
      for (int ch = 0; ch < 3; ch++) {
        next_color[ch] = s[8 + ch];
        dot += (s[8 + ch] - behind[ch]) * pg[ch];
      }
      double alpha_grad = T * dot - final_T / (1 - alpha) * bg_grad;
      double power_grad = alpha_grad * s[11] * G;
      double share[9];
      share[0] = power_grad * (s[4] * dx + s[5] * dy);
      share[1] = power_grad * (s[5] * dx + s[6] * dy);
      share[2] = power_grad * (-0.5 * dx * dx);
      share[3] = power_grad * (-dx * dy);
      share[4] = power_grad * (-0.5 * dy * dy);
      share[5] = alpha_grad * G;
      for (int ch = 0; ch < 3; ch++) share[6 + ch] = alpha * T * pg[ch];
      for (int v = 0; v < 9; v++) {
#pragma omp atomic
        splat_grads[(long)index * 9 + v] += share[v];
      }
    }
  }
}

/* preprocessBackward.wgsl:27-181, in float64. */
void oc_preprocess_backward(long n, const double* gaussians, const double* sh, int sh_degree,
                            const double* view, const double* proj, const double* viewport,
                            const double* splats, const double* splat_grads, double* gaussian_grads,
                            double* sh_grads) {
  double cp[3];
  camera_position(view, cp);
  int count = sh_count(sh_degree);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    double* gg = gaussian_grads + i * 16;
    double* sg = sh_grads + i * 45;
    memset(gg, 0, sizeof(double) * 16);
    memset(sg, 0, sizeof(double) * 45);
    if (splats[i * 12 + 3] == 0) continue;
    const double* g = gaussians + i * 16;
    const double* gr = splat_grads + i * 9;
    projection p;
    project(g, view, proj, viewport, &p); /* the intermediates; visibility came from the splat */
    double d = p.depth, d2 = d * d, d3 = d2 * d;
    double a = p.a, b = p.b, c = p.c, det = p.det, det2 = det * det;
    /* 1. conic -> 2D covariance */
    double ga = (-c * c * gr[2] + b * c * gr[3] - b * b * gr[4]) / det2;
    double gb = (2 * b * c * gr[2] - (det + 2 * b * b) * gr[3] + 2 * a * b * gr[4]) / det2;
    double gc = (-b * b * gr[2] + a * b * gr[3] - a * a * gr[4]) / det2;
    double G2[2][2] = {{ga, 0.5 * gb}, {0.5 * gb, gc}};
    /* 2. dL/dcov3d = JWᵀ G JW ; dL/dJ = 2 G JW cov3d Wᵀ */
    double cov3d_grad[3][3];
    for (int r = 0; r < 3; r++)
      for (int cc = 0; cc < 3; cc++) {
        double sum = 0;
        for (int u = 0; u < 2; u++)
          for (int v = 0; v < 2; v++) sum += p.jw[u][r] * G2[u][v] * p.jw[v][cc];
        cov3d_grad[r][cc] = sum;
      }
    double gjw[2][3], gjwc[2][3], j_grad[2][3];
    for (int u = 0; u < 2; u++)
      for (int cc = 0; cc < 3; cc++) gjw[u][cc] = G2[u][0] * p.jw[0][cc] + G2[u][1] * p.jw[1][cc];
    for (int u = 0; u < 2; u++)
      for (int cc = 0; cc < 3; cc++)
        gjwc[u][cc] = gjw[u][0] * p.cov3d[0][cc] + gjw[u][1] * p.cov3d[1][cc] + gjw[u][2] * p.cov3d[2][cc];
    for (int u = 0; u < 2; u++)
      for (int cc = 0; cc < 3; cc++) /* times Wᵀ: (X Wᵀ)[u][cc] = Σ_k X[u][k] W[cc][k] */
        j_grad[u][cc] = 2 * (gjwc[u][0] * p.w[cc][0] + gjwc[u][1] * p.w[cc][1] + gjwc[u][2] * p.w[cc][2]);
    /* 3. J -> view-space centre, with the clamp mask */
    double free_x = fabs(p.ratio[0]) > p.limit[0] ? 0 : 1;
    double free_y = fabs(p.ratio[1]) > p.limit[1] ? 0 : 1;
    double fx = p.focal[0], fy = p.focal[1];
    double t_grad[3];
    t_grad[0] = j_grad[0][2] * free_x * fx / d2;
    t_grad[1] = j_grad[1][2] * free_y * -fy / d2;
    t_grad[2] = 0;
    double depth_grad = j_grad[0][0] * (-fx / d2) + j_grad[1][1] * (fy / d2) +
                        j_grad[0][2] * (-(1 + free_x) * fx * p.tx / d3) +
                        j_grad[1][2] * ((1 + free_y) * fy * p.ty / d3);
    t_grad[2] -= depth_grad;
    /* 4. the 2D centre through the projection */
    double ndc_grad[2] = {gr[0] * 0.5 * viewport[0], gr[1] * -0.5 * viewport[1]};
    double cw = p.clip[3];
    double clip_grad[4] = {ndc_grad[0] / cw, ndc_grad[1] / cw, 0,
                           -(ndc_grad[0] * p.clip[0] + ndc_grad[1] * p.clip[1]) / (cw * cw)};
    for (int k = 0; k < 3; k++)
      t_grad[k] += proj[k * 4 + 0] * clip_grad[0] + proj[k * 4 + 1] * clip_grad[1] + proj[k * 4 + 2] * clip_grad[2] +
                   proj[k * 4 + 3] * clip_grad[3];
    /* 5. colour, with the clamp mask */
    double off[3] = {g[0] - cp[0], g[1] - cp[1], g[2] - cp[2]};
    double dist = sqrt(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]);
    double dir[3] = {off[0] / dist, off[1] / dist, off[2] / dist};
    double basis[15], bgrad[15][3];
    sh_basis(dir, basis);
    sh_basis_gradient(dir, bgrad);
    double raw[3] = {g[12], g[13], g[14]};
    for (int k = 0; k < count; k++)
      for (int ch = 0; ch < 3; ch++) raw[ch] += basis[k] * sh[i * 45 + k * 3 + ch];
    double raw_grad[3];
    for (int ch = 0; ch < 3; ch++) raw_grad[ch] = raw[ch] < 0 ? 0 : gr[6 + ch];
    double dir_grad[3] = {0, 0, 0};
    for (int k = 0; k < count; k++) {
      double dotk = 0;
      for (int ch = 0; ch < 3; ch++) {
        sg[k * 3 + ch] = basis[k] * raw_grad[ch];
        dotk += sh[i * 45 + k * 3 + ch] * raw_grad[ch];
      }
      for (int e = 0; e < 3; e++) dir_grad[e] += bgrad[k][e] * dotk;
    }
    double dd = dir[0] * dir_grad[0] + dir[1] * dir_grad[1] + dir[2] * dir_grad[2];
    /* 6. position */
    for (int e = 0; e < 3; e++) {
      double wt = p.w[0][e] * t_grad[0] + p.w[1][e] * t_grad[1] + p.w[2][e] * t_grad[2];
      gg[e] = wt + (dir_grad[e] - dir[e] * dd) / dist;
    }
    gg[3] = gr[5];
    /* 7. M = R S -> scale and rotation */
    double m_grad[3][3];
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++)
        m_grad[r][k] = 2 * (cov3d_grad[r][0] * p.m[0][k] + cov3d_grad[r][1] * p.m[1][k] + cov3d_grad[r][2] * p.m[2][k]);
    double rg[3][3];
    for (int k = 0; k < 3; k++) {
      gg[4 + k] = m_grad[0][k] * p.r[0][k] + m_grad[1][k] * p.r[1][k] + m_grad[2][k] * p.r[2][k];
      for (int r = 0; r < 3; r++) rg[r][k] = m_grad[r][k] * g[4 + k];
    }
    double qw = p.q[0], qx = p.q[1], qy = p.q[2], qz = p.q[3];
    double qg[4];
    qg[0] = 2 * (-qz * rg[0][1] + qy * rg[0][2] + qz * rg[1][0] - qx * rg[1][2] - qy * rg[2][0] + qx * rg[2][1]);
    qg[1] = 2 * (qy * rg[0][1] + qz * rg[0][2] + qy * rg[1][0] - 2 * qx * rg[1][1] - qw * rg[1][2] + qz * rg[2][0] +
                 qw * rg[2][1] - 2 * qx * rg[2][2]);
    qg[2] = 2 * (-2 * qy * rg[0][0] + qx * rg[0][1] + qw * rg[0][2] + qx * rg[1][0] + qz * rg[1][2] - qw * rg[2][0] +
                 qz * rg[2][1] - 2 * qy * rg[2][2]);
    qg[3] = 2 * (-2 * qz * rg[0][0] - qw * rg[0][1] + qx * rg[0][2] + qw * rg[1][0] - 2 * qz * rg[1][1] +
                 qy * rg[1][2] + qx * rg[2][0] + qy * rg[2][1]);
    double qdot = p.q[0] * qg[0] + p.q[1] * qg[1] + p.q[2] * qg[2] + p.q[3] * qg[3];
    for (int k = 0; k < 4; k++) gg[8 + k] = (qg[k] - p.q[k] * qdot) / p.q_length;
    for (int ch = 0; ch < 3; ch++) gg[12 + ch] = raw_grad[ch];
  }
}

/* ----------------------------------------------------------------- loss -- */

#define LAMBDA 0.2
#define C1 0.0001
#define C2 0.0009

static void window_weights(double w[6]) {
  double raw[6], total = 0;
  for (int k = 0; k < 6; k++) raw[k] = exp(-(double)(k * k) / (2 * 1.5 * 1.5));
  total = raw[0];
  for (int k = 1; k < 6; k++) total += 2 * raw[k];
  for (int k = 0; k < 6; k++) w[k] = raw[k] / total;
}

/* Separable 11-tap blur with zeros past the edges; `maps` planes of w*h. */
static void blur(const double* in, double* out, double* tmp, int w, int h, int maps) {
  double wt[6];
  window_weights(wt);
  long n = (long)w * h;
  for (int m = 0; m < maps; m++) {
    const double* src = in + m * n;
    double* dst = out + m * n;
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        double s = 0;
        for (int k = -5; k <= 5; k++) {
          int sx = x + k;
          if (sx >= 0 && sx < w) s += wt[abs(k)] * src[(long)y * w + sx];
        }
        tmp[(long)y * w + x] = s;
      }
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        double s = 0;
        for (int k = -5; k <= 5; k++) {
          int sy = y + k;
          if (sy >= 0 && sy < h) s += wt[abs(k)] * tmp[(long)sy * w + x];
        }
        dst[(long)y * w + x] = s;
      }
  }
}

/* loss.wgsl:52-121 per channel. Returns the loss, as TrainingLoss.read() does. */
double oc_loss(int width, int height, const double* rendered, const uint8_t* target, double* pixel_grads,
               double* pixel_loss) {
  long n = (long)width * height;
  double* mom = (double*)malloc(sizeof(double) * 5 * n);
  double* blurred = (double*)malloc(sizeof(double) * 5 * n);
  double* tmp = (double*)malloc(sizeof(double) * n);
  double* part = (double*)malloc(sizeof(double) * 3 * n);
  double* bpart = (double*)malloc(sizeof(double) * 3 * n);
  double* ssim = (double*)malloc(sizeof(double) * n);
  double scale = 1.0 / (3.0 * n);
  if (pixel_loss) memset(pixel_loss, 0, sizeof(double) * n);
  double total = 0;
  for (int ch = 0; ch < 3; ch++) {
    for (long p = 0; p < n; p++) {
      double x = rendered[p * 4 + ch], y = target[p * 4 + ch] / 255.0;
      mom[p] = x; mom[n + p] = y; mom[2 * n + p] = x * x; mom[3 * n + p] = y * y; mom[4 * n + p] = x * y;
    }
    blur(mom, blurred, tmp, width, height, 5);
    for (long p = 0; p < n; p++) {
      double mx = blurred[p], my = blurred[n + p];
      double vx = blurred[2 * n + p] - mx * mx, vy = blurred[3 * n + p] - my * my;
      double cxy = blurred[4 * n + p] - mx * my;
      double a1 = 2 * mx * my + C1, a2 = 2 * cxy + C2, b1 = mx * mx + my * my + C1, b2 = vx + vy + C2;
      double s = a1 * a2 / (b1 * b2);
      part[p] = s * (2 * my / a1 - 2 * my / a2 - 2 * mx / b1 + 2 * mx / b2);
      part[n + p] = -s / b2;
      part[2 * n + p] = 2 * s / a2;
      ssim[p] = s;
    }
    blur(part, bpart, tmp, width, height, 3);
    for (long p = 0; p < n; p++) {
      double x = rendered[p * 4 + ch], y = target[p * 4 + ch] / 255.0;
      double diff = x - y;
      double sgn = diff > 0 ? 1 : (diff < 0 ? -1 : 0);
      double g = bpart[p] + 2 * x * bpart[n + p] + y * bpart[2 * n + p];
      if (pixel_grads) {
        pixel_grads[p * 4 + ch] = ((1 - LAMBDA) * sgn - LAMBDA * g) * scale;
        pixel_grads[p * 4 + 3] = 0;
      }
      double share = ((1 - LAMBDA) * fabs(diff) - LAMBDA * ssim[p]) * scale;
      if (pixel_loss) pixel_loss[p] += share;
      total += share;
    }
  }
  free(mom); free(blurred); free(tmp); free(part); free(bpart); free(ssim);
  return total + LAMBDA;
}

/* Mean SSIM over pixels and channels of two 4-value-per-pixel images (src/metrics.ts:115-146). */
double oc_ssim(int width, int height, const double* a, const double* b) {
  long n = (long)width * height;
  double* mom = (double*)malloc(sizeof(double) * 5 * n);
  double* blurred = (double*)malloc(sizeof(double) * 5 * n);
  double* tmp = (double*)malloc(sizeof(double) * n);
  double total = 0;
  for (int ch = 0; ch < 3; ch++) {
    for (long p = 0; p < n; p++) {
      double x = a[p * 4 + ch], y = b[p * 4 + ch];
      mom[p] = x; mom[n + p] = y; mom[2 * n + p] = x * x; mom[3 * n + p] = y * y; mom[4 * n + p] = x * y;
    }
    blur(mom, blurred, tmp, width, height, 5);
    for (long p = 0; p < n; p++) {
      double mx = blurred[p], my = blurred[n + p];
      double vx = blurred[2 * n + p] - mx * mx, vy = blurred[3 * n + p] - my * my;
      double cxy = blurred[4 * n + p] - mx * my;
      total += ((2 * mx * my + C1) * (2 * cxy + C2)) / ((mx * mx + my * my + C1) * (vx + vy + C2));
    }
  }
  free(mom); free(blurred); free(tmp);
  return total / (3.0 * n);
}

/* ------------------------------------------------------------ optimizer -- */

static double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

static void activate_one(const double* params, double* out) {
  for (int k = 0; k < 3; k++) {
    out[k] = params[k];
    out[4 + k] = exp(params[4 + k]);
    out[12 + k] = 0.5 + SH_C0 * params[12 + k];
  }
  out[3] = sigmoid(params[3]);
  for (int k = 8; k < 12; k++) out[k] = params[k];
}

void oc_activate(long n, const double* params, double* gaussians) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) activate_one(params + i * 16, gaussians + i * 16);
}

void oc_accumulate_stats(long n, const double* splats, const double* splat_grads, const double* viewport,
                         double* stats) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    double radius = splats[i * 12 + 3];
    if (radius == 0) continue;
    double gx = splat_grads[i * 9] * viewport[0] * 0.5, gy = splat_grads[i * 9 + 1] * viewport[1] * 0.5;
    double* s = stats + i * 4;
    s[0] += sqrt(gx * gx + gy * gy);
    s[1] += 1;
    s[2] = s[2] > radius ? s[2] : radius;
    s[3] = 0;
  }
}

static double adam_update(double value, double grad, double* m, double lr, double bc1, double bc2) {
  m[0] = 0.9 * m[0] + (1 - 0.9) * grad;
  m[1] = 0.999 * m[1] + (1 - 0.999) * grad * grad;
  return value - lr * (m[0] / bc1) / (sqrt(m[1] / bc2) + 1e-15);
}

/* adam.wgsl:79-107. lr = (bc1, bc2, position, opacity, scale, rotation, color, sh). */
void oc_adam_step(long n, double* params, double* sh, double* moments, const double* gaussian_grads,
                  const double* sh_grads, const double* lr, double* gaussians) {
  double bc1 = lr[0], bc2 = lr[1];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    double* p = params + i * 16;
    const double* g = gaussian_grads + i * 16;
    double* m = moments + i * 61 * 2;
    for (int k = 0; k < 3; k++) {
      p[k] = adam_update(p[k], g[k], m + 2 * k, lr[2], bc1, bc2);
      double sg = g[4 + k] * exp(p[4 + k]);
      p[4 + k] = adam_update(p[4 + k], sg, m + 2 * (4 + k), lr[4], bc1, bc2);
      double cg = g[12 + k] * SH_C0;
      p[12 + k] = adam_update(p[12 + k], cg, m + 2 * (12 + k), lr[6], bc1, bc2);
    }
    double o = sigmoid(p[3]);
    p[3] = adam_update(p[3], g[3] * o * (1 - o), m + 6, lr[3], bc1, bc2);
    for (int k = 8; k < 12; k++) p[k] = adam_update(p[k], g[k], m + 2 * k, lr[5], bc1, bc2);
    for (int k = 0; k < 45; k++)
      sh[i * 45 + k] = adam_update(sh[i * 45 + k], sh_grads[i * 45 + k], m + 2 * (16 + k), lr[7], bc1, bc2);
    if (gaussians) activate_one(p, gaussians + i * 16);
  }
}

#define RESET_LOGIT (-4.59511985013459)

void oc_reset_opacity(long n, double* params, double* moments, double* gaussians) {
  for (long i = 0; i < n; i++) {
    double* p = params + i * 16;
    if (p[3] > RESET_LOGIT) p[3] = RESET_LOGIT;
    moments[i * 122 + 6] = 0;
    moments[i * 122 + 7] = 0;
    if (gaussians) activate_one(p, gaussians + i * 16);
  }
}

/* ------------------------------------------------------------- densify -- */

uint32_t oc_pcg(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

static double unit_of(uint32_t h) { return ((double)(h >> 8) + 0.5) / 16777216.0; }

static void normal3(uint32_t seed, double out[3]) {
  uint32_t h1 = oc_pcg(seed), h2 = oc_pcg(h1), h3 = oc_pcg(h2), h4 = oc_pcg(h3);
  double tau = 6.283185307179586;
  double r1 = sqrt(-2 * log(unit_of(h1))), r2 = sqrt(-2 * log(unit_of(h3)));
  out[0] = r1 * cos(tau * unit_of(h2));
  out[1] = r1 * sin(tau * unit_of(h2));
  out[2] = r2 * cos(tau * unit_of(h4));
}

enum { DROP = 0, KEEP = 1, CLONE = 2, SPLIT = 3 };

static int removed(const double* st, double opacity, double size) {
  return opacity < st[2] || (st[3] > 0 && size > st[3]);
}

static int action(const double* params, const double* stats, const double* st, long i) {
  const double* p = params + i * 16;
  const double* s = stats + i * 4;
  double grad = s[1] > 0 ? s[0] / s[1] : 0;
  double mx = p[4] > p[5] ? p[4] : p[5];
  mx = mx > p[6] ? mx : p[6];
  double size = exp(mx), opacity = sigmoid(p[3]);
  if (grad >= st[0]) {
    if (size <= st[1]) return removed(st, opacity, size) ? DROP : CLONE;
    return removed(st, opacity, size / 1.6) ? DROP : SPLIT;
  }
  return removed(st, opacity, size) ? DROP : KEEP;
}

static void copy_gaussian(long i, long o, int keep, const double* params, const double* sh, const double* moments,
                          double* np, double* nsh, double* nm) {
  memcpy(np + o * 16, params + i * 16, sizeof(double) * 16);
  memcpy(nsh + o * 45, sh + i * 45, sizeof(double) * 45);
  if (keep)
    memcpy(nm + o * 122, moments + i * 122, sizeof(double) * 122);
  else
    memset(nm + o * 122, 0, sizeof(double) * 122);
}

/* densify.wgsl:95-180 as one serial pass (the prefix sum is the running output index). */
long oc_densify(long n, const double* params, const double* sh, const double* moments, const double* stats,
                const double* st, uint32_t seed, double* np, double* nsh, double* nm, uint32_t* tally) {
  long o = 0;
  tally[0] = tally[1] = tally[2] = tally[3] = 0;
  for (long i = 0; i < n; i++) {
    int a = action(params, stats, st, i);
    tally[a]++;
    if (a == KEEP || a == CLONE) copy_gaussian(i, o, 1, params, sh, moments, np, nsh, nm);
    if (a == CLONE) {
      copy_gaussian(i, o + 1, 0, params, sh, moments, np, nsh, nm);
      np[(o + 1) * 16 + 7] = 1;
      np[(o + 1) * 16 + 15] = (double)(float)seed;
    }
    if (a == SPLIT) {
      const double* p = params + i * 16;
      double r[3][3];
      rotation(p + 8, r, NULL, NULL);
      for (int c = 0; c < 2; c++) {
        long child = o + c;
        copy_gaussian(i, child, 0, params, sh, moments, np, nsh, nm);
        double nrm[3];
        normal3(oc_pcg(seed ^ oc_pcg((uint32_t)(2 * i + c))), nrm);
        double v[3] = {exp(p[4]) * nrm[0], exp(p[5]) * nrm[1], exp(p[6]) * nrm[2]};
        for (int k = 0; k < 3; k++) {
          np[child * 16 + k] = p[k] + r[k][0] * v[0] + r[k][1] * v[1] + r[k][2] * v[2];
          np[child * 16 + 4 + k] = p[4 + k] - log(1.6);
        }
        np[child * 16 + 7] = 2;
        np[child * 16 + 15] = (double)(float)seed;
      }
    }
    o += a == DROP ? 0 : (a == KEEP ? 1 : 2);
  }
  return o;
}

/* ---------------------------------------------------------- train step -- */

/* One optimize() step of src/trainer.ts:192-244: render, loss, backward, stats and
 * (when `adam`) the Adam update with activation. Returns the loss. */
double oc_train_step(long n, double* params, double* sh, double* moments, double* stats, double* gaussians,
                     int sh_degree, const double* view, const double* proj, int width, int height,
                     const double* background, const uint8_t* target, const double* lr, int adam,
                     long* pair_count) {
  long X = (long)width * height;
  int tiles = ((width + 15) / 16) * ((height + 15) / 16);
  double vp[2] = {width, height};
  double* splats = (double*)malloc(sizeof(double) * 12 * n);
  oc_preprocess(n, gaussians, sh, sh_degree, view, proj, vp, splats);
  uint32_t* ranges = (uint32_t*)malloc(sizeof(uint32_t) * 2 * tiles);
  long total = oc_tile_lists(n, splats, width, height, ranges, NULL, 0);
  uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * (total > 0 ? total : 1));
  oc_tile_lists(n, splats, width, height, ranges, pairs, total);
  if (pair_count) *pair_count = total;
  double* color = (double*)malloc(sizeof(double) * 4 * X);
  uint32_t* last = (uint32_t*)malloc(sizeof(uint32_t) * X);
  oc_rasterize(splats, ranges, pairs, width, height, background, color, last);
  double* pgrad = (double*)malloc(sizeof(double) * 4 * X);
  double loss = oc_loss(width, height, color, target, pgrad, NULL);
  double* sgrad = (double*)malloc(sizeof(double) * 9 * n);
  oc_rasterize_backward(n, splats, ranges, pairs, width, height, background, color, last, pgrad, sgrad);
  double* ggrad = (double*)malloc(sizeof(double) * 16 * n);
  double* shgrad = (double*)malloc(sizeof(double) * 45 * n);
  oc_preprocess_backward(n, gaussians, sh, sh_degree, view, proj, vp, splats, sgrad, ggrad, shgrad);
  oc_accumulate_stats(n, splats, sgrad, vp, stats);
  if (adam) oc_adam_step(n, params, sh, moments, ggrad, shgrad, lr, gaussians);
  free(splats); free(ranges); free(pairs); free(color); free(last); free(pgrad); free(sgrad); free(ggrad);
  free(shgrad);
  return loss;
}
