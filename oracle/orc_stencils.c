/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * The config stencils and the reference-faithful functors, each driven the way the
 * reference runtime drives a DUAL grid with GridsToUpdate at world size 1
 * (Appfis/Runtime.hpp:325-369: OpenMP `parallel for schedule(dynamic,1)` over the
 * outermost index, i innermost; then ExchangeHelper -> Exchange on the NEXT buffer,
 * Runtime.hpp:1393-1404; then SwapHelper -> Grid::Swap, Runtime.hpp:1408-1419,
 * Grid.hpp:854-857).  After n sweeps the field is in bufs[n % 2].
 *
 * Compile with -ffp-contract=off: every expression below is evaluated in the order
 * written, one rounding per operation.
 */
#include "orc_common.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int orc_exchange(void* data, int dtype, int ndim, const int* dims, int ghost, int periodic);

#define I2(i, j, W) ((size_t)(i) + (size_t)(W) * (size_t)(j))
#define I3(i, j, k, W, H) ((size_t)(i) + (size_t)(W) * ((size_t)(j) + (size_t)(H) * (size_t)(k)))

/* ------------------------------------------------------------------------------ */
/* Heat2D, config form of HeatDiffusion/src/heat_appfis.hpp:166-174 (2D, 5-point):  */
/*   u' = u + c*((((uN + uS) + uE) + uW) - 4u),  N=(i,j-1) S=(i,j+1) E=(i+1,j) W=(i-1,j) */
/* nx, ny interior; ghost 1 (Dirichlet, held by the caller).                          */
/* ------------------------------------------------------------------------------ */

void orc_heat2d_f64(double* b0, double* b1, int nx, int ny, double c, int iters)
{
    const int W = nx + 2;
    double* buf[2] = {b0, b1};
    for (int it = 0; it < iters; it++) {
        const double* u = buf[it & 1];
        double* v = buf[(it + 1) & 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double uc = u[I2(i, j, W)];
                double s = u[I2(i, j - 1, W)] + u[I2(i, j + 1, W)];
                s = s + u[I2(i + 1, j, W)];
                s = s + u[I2(i - 1, j, W)];
                s = s - 4.0 * uc;
                v[I2(i, j, W)] = uc + c * s;
            }
    }
}

void orc_heat2d_f32(float* b0, float* b1, int nx, int ny, float c, int iters)
{
    const int W = nx + 2;
    float* buf[2] = {b0, b1};
    for (int it = 0; it < iters; it++) {
        const float* u = buf[it & 1];
        float* v = buf[(it + 1) & 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                float uc = u[I2(i, j, W)];
                float s = u[I2(i, j - 1, W)] + u[I2(i, j + 1, W)];
                s = s + u[I2(i + 1, j, W)];
                s = s + u[I2(i - 1, j, W)];
                s = s - 4.0f * uc;
                v[I2(i, j, W)] = uc + c * s;
            }
    }
}

/* Reference functor, HeatDiffusion/src/heat_appfis.hpp:166-174 (3D 7-point FTCS). */
void orc_heat3d_ref(double* b0, double* b1, int nx, int ny, int nz, double cx, double cy, double cz, int iters)
{
    const int W = nx + 2, H = ny + 2;
    double* buf[2] = {b0, b1};
    for (int it = 0; it < iters; it++) {
        const double* u = buf[it & 1];
        double* v = buf[(it + 1) & 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    double uc = u[I3(i, j, k, W, H)];
                    double val = uc + cx * (u[I3(i - 1, j, k, W, H)] + u[I3(i + 1, j, k, W, H)] - 2 * uc)
                                 + cy * (u[I3(i, j - 1, k, W, H)] + u[I3(i, j + 1, k, W, H)] - 2 * uc)
                                 + cz * (u[I3(i, j, k - 1, W, H)] + u[I3(i, j, k + 1, W, H)] - 2 * uc);
                    v[I3(i, j, k, W, H)] = val;
                }
    }
}

/* ------------------------------------------------------------------------------ */
/* GameOfLife, config form: uint8 cells, rule B3/S23, every cell written, periodic.  */
/* InitializeExecution runs ExchangeInitial on both buffers (Runtime.hpp:124-142);    */
/* every generation is followed by Exchange on the written buffer.                   */
/* ------------------------------------------------------------------------------ */

void orc_gol_u8(uint8_t* b0, uint8_t* b1, int nx, int ny, int gens)
{
    const int W = nx + 2;
    const int dims[2] = {nx + 2, ny + 2};
    uint8_t* buf[2] = {b0, b1};
    orc_exchange(b0, ORC_U8, 2, dims, 1, 1);
    orc_exchange(b1, ORC_U8, 2, dims, 1, 1);
    for (int g = 0; g < gens; g++) {
        const uint8_t* u = buf[g & 1];
        uint8_t* v = buf[(g + 1) & 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                int n = u[I2(i - 1, j - 1, W)] + u[I2(i, j - 1, W)] + u[I2(i + 1, j - 1, W)]
                        + u[I2(i - 1, j, W)] + u[I2(i + 1, j, W)]
                        + u[I2(i - 1, j + 1, W)] + u[I2(i, j + 1, W)] + u[I2(i + 1, j + 1, W)];
                v[I2(i, j, W)] = (uint8_t)(n == 3 || (n == 2 && u[I2(i, j, W)]));
            }
        orc_exchange(v, ORC_U8, 2, dims, 1, 1);
    }
}

/* Reference functor GameOfLife/src/game_appfis.hpp:121-135 (int cells): a live cell
 * dies when n < 2 or n > 4, a dead cell with n == 3 is born, and otherwise the NEXT
 * buffer is left untouched (it keeps the value from two generations ago). */
void orc_gol_ref_i32(int32_t* b0, int32_t* b1, int nx, int ny, int gens, int periodic)
{
    const int W = nx + 2;
    const int dims[2] = {nx + 2, ny + 2};
    int32_t* buf[2] = {b0, b1};
    orc_exchange(b0, ORC_I32, 2, dims, 1, periodic);
    orc_exchange(b1, ORC_I32, 2, dims, 1, periodic);
    for (int g = 0; g < gens; g++) {
        const int32_t* u = buf[g & 1];
        int32_t* v = buf[(g + 1) & 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                int n = u[I2(i - 1, j - 1, W)] + u[I2(i, j - 1, W)] + u[I2(i + 1, j - 1, W)]
                        + u[I2(i - 1, j, W)] + u[I2(i + 1, j, W)]
                        + u[I2(i - 1, j + 1, W)] + u[I2(i, j + 1, W)] + u[I2(i + 1, j + 1, W)];
                if (u[I2(i, j, W)] == 1) {
                    if (n < 2 || n > 4) v[I2(i, j, W)] = 0;
                } else {
                    if (n == 3) v[I2(i, j, W)] = 1;
                }
            }
        orc_exchange(v, ORC_I32, 2, dims, 1, periodic);
    }
}

/* ------------------------------------------------------------------------------ */
/* FDTD3d radius-4 star, config form of FDTD3d/src/fdtd3d_appfis.hpp:129-143:         */
/*   s_r = ((u[x-r] + u[x+r]) + (u[y-r] + u[y+r])) + (u[z-r] + u[z+r])                */
/*   u'  = (((c0*u + c1*s_1) + c2*s_2) + c3*s_3) + c4*s_4                            */
/* ghost 4 held at zero.  F16/BF16: storage only, arithmetic in FP32, RNE on store.   */
/* ------------------------------------------------------------------------------ */

#define FDTD_LOOP(T, LOAD, STORE, ACC, CAST)                                                        \
    do {                                                                                           \
        T* buf[2] = {(T*)b0, (T*)b1};                                                              \
        ACC c[5];                                                                                  \
        for (int r = 0; r < 5; r++) c[r] = (ACC)coeff[r];                                          \
        const size_t sy = (size_t)W, sz = (size_t)W * (size_t)H;                                   \
        for (int s = 0; s < steps; s++) {                                                          \
            const T* u = buf[s & 1];                                                               \
            T* v = buf[(s + 1) & 1];                                                               \
            _Pragma("omp parallel for schedule(dynamic, 1)")                                       \
            for (int k = 4; k < nz + 4; k++)                                                       \
                for (int j = 4; j < ny + 4; j++)                                                   \
                    for (int i = 4; i < nx + 4; i++) {                                             \
                        size_t o = I3(i, j, k, W, H);                                              \
                        ACC acc = c[0] * LOAD(u[o]);                                               \
                        for (int r = 1; r <= 4; r++) {                                             \
                            ACC sx = LOAD(u[o - r]) + LOAD(u[o + r]);                              \
                            ACC sY = LOAD(u[o - r * sy]) + LOAD(u[o + r * sy]);                    \
                            ACC sZ = LOAD(u[o - r * sz]) + LOAD(u[o + r * sz]);                    \
                            ACC sr = (sx + sY) + sZ;                                               \
                            acc = acc + c[r] * sr;                                                 \
                        }                                                                          \
                        v[o] = STORE(acc);                                                         \
                    }                                                                              \
        }                                                                                          \
    } while (0)

#define ID(x) (x)
#define LD_H(x) orc_f16_to_f32(x)
#define ST_H(x) orc_f32_to_f16(x)
#define LD_B(x) orc_bf16_to_f32(x)
#define ST_B(x) orc_f32_to_bf16(x)

int orc_fdtd3d(int dtype, void* b0, void* b1, int nx, int ny, int nz, const double* coeff, int steps)
{
    const int W = nx + 8, H = ny + 8;
    switch (dtype) {
    case ORC_F64: FDTD_LOOP(double, ID, ID, double, ID); break;
    case ORC_F32: FDTD_LOOP(float, ID, ID, float, ID); break;
    case ORC_F16: FDTD_LOOP(uint16_t, LD_H, ST_H, float, ID); break;
    case ORC_BF16: FDTD_LOOP(uint16_t, LD_B, ST_B, float, ID); break;
    default: return 1;
    }
    return 0;
}

/* Reference functor FDTD3d/src/fdtd3d_appfis.hpp:129-143: val = u; for a in 1..radius:
 * val += (x-a + x+a) + (y-a + y+a) + (z-a + z+a); val *= coeff.  ghost = radius. */
void orc_fdtd3d_ref(float* b0, float* b1, int nx, int ny, int nz, float coeff, int radius, int steps)
{
    const int R = radius, W = nx + 2 * R, H = ny + 2 * R;
    const size_t sy = (size_t)W, sz = (size_t)W * (size_t)H;
    float* buf[2] = {b0, b1};
    for (int s = 0; s < steps; s++) {
        const float* u = buf[s & 1];
        float* v = buf[(s + 1) & 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = R; k < nz + R; k++)
            for (int j = R; j < ny + R; j++)
                for (int i = R; i < nx + R; i++) {
                    size_t o = I3(i, j, k, W, H);
                    float val = u[o];
                    for (int a = 1; a <= R; a++)
                        val += (u[o - a] + u[o + a]) + (u[o - a * sy] + u[o + a * sy]) + (u[o - a * sz] + u[o + a * sz]);
                    val *= coeff;
                    v[o] = val;
                }
    }
}

/* ------------------------------------------------------------------------------ */
/* Himeno 19-point Jacobi, Himeno/src/himeno_appfis.hpp:178-198, operation order    */
/* kept; gosa = sum over the interior of ss*ss (widened to double) after each sweep. */
/* arrays: p0 p1 a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd wrk1; dims padded (shell incl.).  */
/* ------------------------------------------------------------------------------ */

#define HIMENO_LOOP(T)                                                                              \
    do {                                                                                           \
        T* P[2] = {(T*)arrays[0], (T*)arrays[1]};                                                  \
        const T *a0 = (const T*)arrays[2], *a1 = (const T*)arrays[3], *a2 = (const T*)arrays[4],   \
                *a3 = (const T*)arrays[5];                                                         \
        const T *b0 = (const T*)arrays[6], *b1 = (const T*)arrays[7], *b2 = (const T*)arrays[8];   \
        const T *c0 = (const T*)arrays[9], *c1 = (const T*)arrays[10], *c2 = (const T*)arrays[11]; \
        const T *bnd = (const T*)arrays[12], *wrk1 = (const T*)arrays[13];                         \
        const T om = (T)omega;                                                                     \
        const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;                                \
        for (int it = 0; it < iters; it++) {                                                       \
            const T* p = P[it & 1];                                                                \
            T* w2 = P[(it + 1) & 1];                                                               \
            double gosa = 0.0;                                                                     \
            _Pragma("omp parallel for schedule(dynamic, 1) reduction(+ : gosa)")                   \
            for (int k = 1; k < nz - 1; k++)                                                       \
                for (int j = 1; j < ny - 1; j++)                                                   \
                    for (int i = 1; i < nx - 1; i++) {                                             \
                        size_t o = (size_t)i + sy * (size_t)j + sz * (size_t)k;                    \
                        T s0 = a0[o] * p[o + 1] + a1[o] * p[o + sy] + a2[o] * p[o + sz]            \
                               + b0[o] * (p[o + 1 + sy] - p[o + 1 - sy] - p[o - 1 + sy] + p[o - 1 - sy]) \
                               + b1[o] * (p[o + sy + sz] - p[o - sy + sz] - p[o + sy - sz] + p[o - sy - sz]) \
                               + b2[o] * (p[o + 1 + sz] - p[o - 1 + sz] - p[o + 1 - sz] + p[o - 1 - sz]) \
                               + c0[o] * p[o - 1] + c1[o] * p[o - sy] + c2[o] * p[o - sz] + wrk1[o];  \
                        T ss = (s0 * a3[o] - p[o]) * bnd[o];                                       \
                        w2[o] = p[o] + om * ss;                                                    \
                        gosa += (double)ss * (double)ss;                                           \
                    }                                                                              \
            if (gosa_out) gosa_out[it] = gosa;                                                     \
        }                                                                                          \
    } while (0)

int orc_himeno(int dtype, void* const* arrays, int nx, int ny, int nz, double omega, int iters, double* gosa_out)
{
    if (dtype == ORC_F32) HIMENO_LOOP(float);
    else if (dtype == ORC_F64) HIMENO_LOOP(double);
    else return 1;
    return 0;
}
