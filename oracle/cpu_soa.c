/*
 * cpu_soa.c -- the stronger CPU line for the benchmark's CPU baseline: an optimised FP64
 * softened direct sum, a_i = G sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^(3/2), over
 * structure-of-arrays positions and masses.
 *
 * This is not the reference's algorithm (that one is or_bench_workpool in nb_oracle.c: AoS
 * bodies through pointers, a force loop and a detection loop per target).  It is the best plain
 * CPU formulation of the same softened sum (the potential of octree.go:314-325 without the
 * tree): the inner loop over sources has no branches and vectorises (AVX2 + FMA, four doubles
 * per instruction), the self pair contributes exactly zero (dx = dy = dz = 0 with eps > 0), and
 * OpenMP threads take contiguous target slices.  Test and benchmark infrastructure only, like
 * the rest of this directory.
 */
#include <math.h>
#include <stdint.h>

#include "nb_oracle.h"

void or_soa_soft(int64_t n, const double *restrict x, const double *restrict y,
                 const double *restrict z, const double *restrict m, double G, double eps,
                 int64_t t0, int64_t t1, double *restrict ax, double *restrict ay,
                 double *restrict az, int nthreads) {
  const double eps2 = eps * eps;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = t0; i < t1; i++) {
    const double xi = x[i], yi = y[i], zi = z[i];
    double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma omp simd reduction(+ : sx, sy, sz)
    for (int64_t j = 0; j < n; j++) {
      const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double inv = 1.0 / sqrt(r2);
      const double s = m[j] * inv * inv * inv;
      sx += dx * s;
      sy += dy * s;
      sz += dz * s;
    }
    ax[i - t0] = G * sx;
    ay[i - t0] = G * sy;
    az[i - t0] = G * sz;
  }
}
