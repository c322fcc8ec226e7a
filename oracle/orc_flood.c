/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * Triton flood step chain for FP64 (reference dtype) and FP32, plus the synthetic
 * DEM of the flood-ensemble workload.
 */
#include "orc_common.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>

#define REAL double
#define SFX _f64
#define RMAX DBL_MAX
#define SQRT sqrt
#define CBRT cbrt
#define FABS fabs
#define FMAX fmax
#define FMIN fmin
#include "orc_flood_impl.h"
#undef REAL
#undef SFX
#undef RMAX
#undef SQRT
#undef CBRT
#undef FABS
#undef FMAX
#undef FMIN

#define REAL float
#define SFX _f32
#define RMAX FLT_MAX
#define SQRT sqrtf
#define CBRT cbrtf
#define FABS fabsf
#define FMAX fmaxf
#define FMIN fminf
#include "orc_flood_impl.h"
#undef REAL
#undef SFX
#undef RMAX
#undef SQRT
#undef CBRT
#undef FABS
#undef FMAX
#undef FMIN

/* Synthetic DEM: z = sum_k A_k exp(-((x-cx_k)^2 + (y-cy_k)^2) / (2 s_k^2)) + slope*x,
 * cell centres x = (i-0.5)*dx, y = (j-0.5)*dx for interior i, j in [1, n].  Hill k
 * draws u_m = (r(seed, 4k+m) >> 11) * 2^-53, m = 0..3: A = 5 + 45 u0,
 * s = 200 + 1800 u1, cx = u2 * L, cy = u3 * L, L = nx*dx.  Ghost cells = wall. */
int orc_flood_dem_f64(double* z, int nx, int ny, double dx, uint64_t seed, int nhills, double slope, double wall)
{
    const int W = nx + 2, H = ny + 2;
    double* A = (double*)malloc(sizeof(double) * 4 * (size_t)(nhills > 0 ? nhills : 1));
    if (!A) return 1;
    const double L = (double)nx * dx;
    for (int k = 0; k < nhills; k++) {
        double u0 = orc_u53(orc_rng(seed, 4 * (uint64_t)k + 0));
        double u1 = orc_u53(orc_rng(seed, 4 * (uint64_t)k + 1));
        double u2 = orc_u53(orc_rng(seed, 4 * (uint64_t)k + 2));
        double u3 = orc_u53(orc_rng(seed, 4 * (uint64_t)k + 3));
        A[4 * k + 0] = 5.0 + 45.0 * u0;
        A[4 * k + 1] = 200.0 + 1800.0 * u1;
        A[4 * k + 2] = u2 * L;
        A[4 * k + 3] = u3 * L;
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            size_t o = (size_t)i + (size_t)W * (size_t)j;
            if (i == 0 || j == 0 || i == W - 1 || j == H - 1) {
                z[o] = wall;
                continue;
            }
            double x = ((double)i - 0.5) * dx, y = ((double)j - 0.5) * dx;
            double v = 0.0;
            for (int k = 0; k < nhills; k++) {
                double ddx = x - A[4 * k + 2], ddy = y - A[4 * k + 3], s = A[4 * k + 1];
                v = v + A[4 * k + 0] * exp(-(ddx * ddx + ddy * ddy) / (2.0 * s * s));
            }
            z[o] = v + slope * x;
        }
    free(A);
    return 0;
}
