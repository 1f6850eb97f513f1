/*
 * CPU oracle for the Gaussian-splatting training step. TEST INFRASTRUCTURE ONLY:
 * the tests, the smoke check and the CPU-baseline leg of the benchmark load it as
 * the checker. The product library never links or calls it.
 *
 * Everything is float64 except where the reference itself rounds to float32 on
 * purpose (tile edges, as src/checks/reference.ts:128-136 does with Math.fround).
 *
 * Layouts follow the reference's buffers (src/gaussians.ts:32-36):
 *   Gaussian  16 values: position xyz, opacity, scale xyz, pad, quaternion wxyz, rgb, pad
 *   Splat2D   12 values: mean xy, depth, radius, conic abc, pad, rgb, opacity
 *   SH rest   45 values: coefficients 1..15, rgb each, coefficient-major
 *   camera    view[16], proj[16] column-major, viewport[2]
 *   pixels    row-major; colour buffers hold 4 values per pixel (rgb + transmittance)
 */
#ifndef DIORAMA_ORACLE_H
#define DIORAMA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OC_TILE = 16 };

void oc_preprocess(long n, const double* gaussians, const double* sh, int sh_degree,
                   const double* view, const double* proj, const double* viewport,
                   double* splats);

long oc_tile_lists(long n, const double* splats, int width, int height,
                   uint32_t* ranges, uint32_t* pair_splats, long capacity);
long oc_tile_lists_capped(long n, const double* splats, int width, int height, long capacity, uint32_t* ranges,
                          uint32_t* pair_splats);

void oc_rasterize(const double* splats, const uint32_t* ranges, const uint32_t* pair_splats,
                  int width, int height, const double* background,
                  double* pixel_color, uint32_t* pixel_last);

void oc_rasterize_margin(const double* splats, const uint32_t* ranges, const uint32_t* pair_splats,
                         int width, int height, const double* background, double* pixel_color,
                         uint32_t* pixel_last, double* margin);

void oc_footprint(long n, const double* splats, const uint32_t* ranges, const uint32_t* pair_splats, int width,
                  int height, const double* pixel_loss, double* footprint);
void oc_rasterize_backward(long n, const double* splats, const uint32_t* ranges,
                           const uint32_t* pair_splats, int width, int height,
                           const double* background, const double* pixel_color,
                           const uint32_t* pixel_last, const double* pixel_grads,
                           double* splat_grads);

void oc_preprocess_backward(long n, const double* gaussians, const double* sh, int sh_degree,
                            const double* view, const double* proj, const double* viewport,
                            const double* splats, const double* splat_grads,
                            double* gaussian_grads, double* sh_grads);

double oc_loss(int width, int height, const double* rendered, const uint8_t* target,
               double* pixel_grads, double* pixel_loss);

double oc_ssim(int width, int height, const double* a, const double* b);

void oc_activate(long n, const double* params, double* gaussians);

void oc_accumulate_stats(long n, const double* splats, const double* splat_grads,
                         const double* viewport, double* stats);

void oc_adam_step(long n, double* params, double* sh, double* moments,
                  const double* gaussian_grads, const double* sh_grads,
                  const double* lr /* 8: bc1, bc2, position, opacity, scale, rotation, color, sh */,
                  double* gaussians);

void oc_reset_opacity(long n, double* params, double* moments, double* gaussians);

uint32_t oc_pcg(uint32_t v);

long oc_densify(long n, const double* params, const double* sh, const double* moments,
                const double* stats, const double* settings /* grad thr, clone max, min opacity, max size */,
                uint32_t seed, double* new_params, double* new_sh, double* new_moments,
                uint32_t* tally /* drop, keep, clone, split */);

double oc_train_step(long n, double* params, double* sh, double* moments, double* stats,
                     double* gaussians, int sh_degree, const double* view, const double* proj,
                     int width, int height, const double* background, const uint8_t* target,
                     const double* lr, int adam, long* pair_count);

/* Mean squared distance to the 3 nearest other points of pts (n x 3 float32),
 * for queries[0..m) (null: points 0..m-1), into out[m]; float32 arithmetic. */
void oc_mean_neighbor_distance2(long n, const float* pts, long m, const long* queries, float* out);

int oc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
