/* TEST INFRASTRUCTURE (the oracle): for each query point, the mean squared
 * distance to its 3 nearest other points, by brute force in float32 as
 * src/shaders/nearestNeighbors.wgsl:12-44 computes it: differences, the squared
 * length as fma(dz, dz, fma(dy, dy, dx * dx)), the three smallest kept by the
 * reference's insertion from a start of 3.4e38, their float32 mean. Only tests
 * call this. */
#include <math.h>
#include <stdint.h>

#include "oracle.h"

static inline __attribute__((always_inline)) float mean3(const float* pts, long n, long i) {
  const float qx = pts[3 * i], qy = pts[3 * i + 1], qz = pts[3 * i + 2];
  float b0 = 3.4e38f, b1 = 3.4e38f, b2 = 3.4e38f;
  for (long j = 0; j < n; j++) {
    if (j == i) continue;
    const float dx = pts[3 * j] - qx, dy = pts[3 * j + 1] - qy, dz = pts[3 * j + 2] - qz;
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    if (d2 < b2) {
      if (d2 < b1) {
        b2 = b1;
        if (d2 < b0) {
          b1 = b0;
          b0 = d2;
        } else {
          b1 = d2;
        }
      } else {
        b2 = d2;
      }
    }
  }
  return (b0 + b1 + b2) / 3.0f;
}

__attribute__((target_clones("fma", "default")))
void oc_mean_neighbor_distance2(long n, const float* pts, long m, const long* queries, float* out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (long k = 0; k < m; k++) out[k] = mean3(pts, n, queries ? queries[k] : k);
}
