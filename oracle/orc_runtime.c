/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * Runtime semantics of the APPFIS reference at MPI world size 1:
 *   seeded inputs, Grid::SetBoundary / SetBoundaryCopy, the single-rank periodic
 *   Exchange (pack -> self Sendrecv -> unpack), and Execute*Reduce.
 */
#include "orc_common.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX3(i, j, k, d) ((size_t)(i) + (size_t)(d)[0] * ((size_t)(j) + (size_t)(d)[1] * (size_t)(k)))

/* ------------------------------------------------------------------------------ */
/* seeded synthetic inputs                                                          */
/* ------------------------------------------------------------------------------ */

uint64_t orc_rng_u64(uint64_t seed, uint64_t n) { return orc_rng(seed, n); }

static void store_real(void* data, int dtype, size_t idx, double v)
{
    switch (dtype) {
    case ORC_F64: ((double*)data)[idx] = v; break;
    case ORC_F32: ((float*)data)[idx] = (float)v; break;
    case ORC_F16: ((uint16_t*)data)[idx] = orc_f32_to_f16((float)v); break;
    case ORC_BF16: ((uint16_t*)data)[idx] = orc_f32_to_bf16((float)v); break;
    case ORC_I32: ((int32_t*)data)[idx] = (int32_t)v; break;
    case ORC_U8: ((uint8_t*)data)[idx] = (uint8_t)v; break;
    case ORC_I64: ((int64_t*)data)[idx] = (int64_t)v; break;
    default: break;
    }
}

/* Interior cells get lo + (hi-lo)*u with u from r(seed, interior linear index).
 * F64 draws 53-bit u; every narrower type draws the 24-bit u (exact in float), the
 * affine map is evaluated in float and rounded to the storage type. */
int orc_fill_uniform(void* data, int dtype, int ndim, const int* dims, int ghost, uint64_t seed,
                     double lo, double hi)
{
    int d[3] = {1, 1, 1};
    for (int a = 0; a < ndim; a++) d[a] = dims[a];
    int g[3] = {0, 0, 0};
    for (int a = 0; a < ndim; a++) g[a] = ghost;
    int n[3];
    for (int a = 0; a < 3; a++) n[a] = d[a] - 2 * g[a];
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                uint64_t lin = (uint64_t)i + (uint64_t)n[0] * ((uint64_t)j + (uint64_t)n[1] * (uint64_t)k);
                uint64_t r = orc_rng(seed, lin);
                size_t idx = IDX3(i + g[0], j + g[1], k + g[2], d);
                if (dtype == ORC_F64) {
                    store_real(data, dtype, idx, lo + (hi - lo) * orc_u53(r));
                } else {
                    float v = (float)lo + ((float)hi - (float)lo) * orc_u24(r);
                    store_real(data, dtype, idx, (double)v);
                }
            }
    return 0;
}

int orc_fill_bernoulli_u8(uint8_t* data, int ndim, const int* dims, int ghost, uint64_t seed, double p)
{
    int d[3] = {1, 1, 1};
    int g[3] = {0, 0, 0};
    for (int a = 0; a < ndim; a++) { d[a] = dims[a]; g[a] = ghost; }
    int n[3];
    for (int a = 0; a < 3; a++) n[a] = d[a] - 2 * g[a];
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return 1;
    /* threshold p * 2^64 (p in [0,1]); p == 1 means always live */
    long double thr = (long double)p * 18446744073709551616.0L;
    int always = p >= 1.0;
    uint64_t t = always ? UINT64_MAX : (uint64_t)thr;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                uint64_t lin = (uint64_t)i + (uint64_t)n[0] * ((uint64_t)j + (uint64_t)n[1] * (uint64_t)k);
                uint64_t r = orc_rng(seed, lin);
                data[IDX3(i + g[0], j + g[1], k + g[2], d)] = (uint8_t)(always || r < t);
            }
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Grid::SetBoundary (reference Appfis/Grid.hpp:558-678): for g in [0,ghost) set the  */
/* g-th layer of `face` (low faces: index g; high faces: dim-ghost+g) to value.       */
/* face == -1 walks every face in order like Grid::SetBoundary(T) (Grid.hpp:530-542). */
/* ------------------------------------------------------------------------------ */

static int face_axis(int ndim, int face, int* high)
{
    *high = 0;
    if (ndim == 1) {
        if (face == ORC_LEFT) return 0;
        if (face == ORC_RIGHT) { *high = 1; return 0; }
        return -1;
    }
    if (ndim == 2) {
        switch (face) {
        case ORC_LEFT: return 0;
        case ORC_RIGHT: *high = 1; return 0;
        case ORC_UP: return 1;
        case ORC_DOWN: *high = 1; return 1;
        default: return -1;
        }
    }
    switch (face) {
    case ORC_LEFT: return 0;
    case ORC_RIGHT: *high = 1; return 0;
    case ORC_FRONT: return 1;
    case ORC_BACK: *high = 1; return 1;
    case ORC_DOWN: return 2;
    case ORC_UP: *high = 1; return 2;
    default: return -1;
    }
}

static int set_boundary_face(void* data, size_t es, int ndim, const int* dims, int ghost, int face,
                             const void* value)
{
    int high;
    int ax = face_axis(ndim, face, &high);
    if (ax < 0) return 1;
    int d[3] = {1, 1, 1};
    for (int a = 0; a < ndim; a++) d[a] = dims[a];
    unsigned char* b = (unsigned char*)data;
    for (int g = 0; g < ghost; g++) {
        int layer = high ? d[ax] - ghost + g : g;
        for (int k = 0; k < d[2]; k++)
            for (int j = 0; j < d[1]; j++)
                for (int i = 0; i < d[0]; i++) {
                    int c[3] = {i, j, k};
                    if (c[ax] != layer) continue;
                    memcpy(b + IDX3(i, j, k, d) * es, value, es);
                }
    }
    return 0;
}

int orc_set_boundary(void* data, int dtype, int ndim, const int* dims, int ghost, int face,
                     const void* value)
{
    size_t es = orc_esize(dtype);
    if (!es || ndim < 1 || ndim > 3) return 1;
    if (face >= 0) return set_boundary_face(data, es, ndim, dims, ghost, face, value);
    for (int f = 0; f < 2 * ndim; f++)
        if (set_boundary_face(data, es, ndim, dims, ghost, f, value)) return 1;
    return 0;
}

/* Grid::SetBoundaryCopy (reference Appfis/Grid.hpp:682-830): g runs from ghost-1 down
 * to 0; low faces copy layer g+1 into g; high faces copy layer dim-ghost+g-1 into
 * dim-ghost+g (so for ghost > 1 the high face reads ghost cells that are written
 * later in the same loop -- kept as the reference does it). */
static int copy_boundary_face(void* data, size_t es, int ndim, const int* dims, int ghost, int face)
{
    int high;
    int ax = face_axis(ndim, face, &high);
    if (ax < 0) return 1;
    int d[3] = {1, 1, 1};
    for (int a = 0; a < ndim; a++) d[a] = dims[a];
    unsigned char* b = (unsigned char*)data;
    for (int g = ghost - 1; g >= 0; g--) {
        int dst = high ? d[ax] - ghost + g : g;
        int src = high ? d[ax] - ghost + g - 1 : g + 1;
        for (int k = 0; k < d[2]; k++)
            for (int j = 0; j < d[1]; j++)
                for (int i = 0; i < d[0]; i++) {
                    int c[3] = {i, j, k};
                    if (c[ax] != dst) continue;
                    int s[3] = {i, j, k};
                    s[ax] = src;
                    memcpy(b + IDX3(i, j, k, d) * es, b + IDX3(s[0], s[1], s[2], d) * es, es);
                }
    }
    return 0;
}

int orc_set_boundary_copy(void* data, int dtype, int ndim, const int* dims, int ghost, int face)
{
    size_t es = orc_esize(dtype);
    if (!es || ndim < 1 || ndim > 3) return 1;
    if (face >= 0) return copy_boundary_face(data, es, ndim, dims, ghost, face);
    for (int f = 0; f < 2 * ndim; f++)
        if (copy_boundary_face(data, es, ndim, dims, ghost, f)) return 1;
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Communication::Exchange at world size 1 with periodic=true (reference             */
/* Appfis/Communication.hpp:613-678): for each axis, face1 = low, face2 = high:        */
/*   pack face1 (CopyFromArrayToBuffer: interior layers ghost+g) -> Sendrecv to self ->  */
/*   unpack into face2 ghost layers (dim-ghost+g); then pack face2 (dim-2ghost+g) ->     */
/*   unpack into face1 ghost layers (g).  Axis order X, Y, Z; each face spans the full  */
/*   padded extent of the other axes, which is what makes corners periodic.            */
/* ------------------------------------------------------------------------------ */

static void pack_layers(const unsigned char* b, unsigned char* buf, size_t es, const int* d, int ax,
                        int first_layer, int ghost)
{
    size_t n = 0;
    for (int g = 0; g < ghost; g++) {
        int layer = first_layer + g;
        for (int k = 0; k < d[2]; k++)
            for (int j = 0; j < d[1]; j++)
                for (int i = 0; i < d[0]; i++) {
                    int c[3] = {i, j, k};
                    if (c[ax] != layer) continue;
                    memcpy(buf + n * es, b + IDX3(i, j, k, d) * es, es);
                    n++;
                }
    }
}

static void unpack_layers(unsigned char* b, const unsigned char* buf, size_t es, const int* d, int ax,
                          int first_layer, int ghost)
{
    size_t n = 0;
    for (int g = 0; g < ghost; g++) {
        int layer = first_layer + g;
        for (int k = 0; k < d[2]; k++)
            for (int j = 0; j < d[1]; j++)
                for (int i = 0; i < d[0]; i++) {
                    int c[3] = {i, j, k};
                    if (c[ax] != layer) continue;
                    memcpy(b + IDX3(i, j, k, d) * es, buf + n * es, es);
                    n++;
                }
    }
}

int orc_exchange(void* data, int dtype, int ndim, const int* dims, int ghost, int periodic)
{
    size_t es = orc_esize(dtype);
    if (!es || ndim < 1 || ndim > 3) return 1;
    if (!periodic || ghost <= 0) return 0; /* neighbours are MPI_PROC_NULL: nothing moves */
    int d[3] = {1, 1, 1};
    for (int a = 0; a < ndim; a++) d[a] = dims[a];
    size_t total = (size_t)d[0] * d[1] * d[2];
    unsigned char* b = (unsigned char*)data;
    unsigned char* buf = (unsigned char*)malloc(total * es);
    if (!buf) return 1;
    for (int ax = 0; ax < ndim; ax++) {
        /* face1 -> neighbour1 (self); received into face2 */
        pack_layers(b, buf, es, d, ax, ghost, ghost);
        unpack_layers(b, buf, es, d, ax, d[ax] - ghost, ghost);
        /* face2 -> neighbour2 (self); received into face1 */
        pack_layers(b, buf, es, d, ax, d[ax] - 2 * ghost, ghost);
        unpack_layers(b, buf, es, d, ax, 0, ghost);
    }
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Execute{1,2,3}DReduceCPU (reference Appfis/Runtime.hpp:598-913): OpenMP reduction */
/* over the Config box, accumulated in T; MAX starts at numeric_limits<T>::min()       */
/* (FLT_MIN / DBL_MIN for floats: the smallest positive normal), MIN at ::max().       */
/* orc_reduce_sum_wide: SUM in int64 (integers) or long double (floats), the exact-as- */
/* possible reference value the GPU's widened sums are checked against.               */
/* ------------------------------------------------------------------------------ */

#define REDUCE_BODY(T, TMIN, TMAX)                                                              \
    do {                                                                                        \
        const T* a = (const T*)data;                                                            \
        if (op == ORC_MAX) {                                                                    \
            T m = TMIN;                                                                         \
            _Pragma("omp parallel for reduction(max : m) schedule(dynamic, 1)")                 \
            for (int k = b[4]; k < b[5]; k++)                                                   \
                for (int j = b[2]; j < b[3]; j++)                                               \
                    for (int i = b[0]; i < b[1]; i++) {                                         \
                        T v = a[IDX3(i, j, k, d)];                                              \
                        if (v > m) m = v;                                                       \
                    }                                                                           \
            *(T*)result = m;                                                                    \
        } else if (op == ORC_MIN) {                                                             \
            T m = TMAX;                                                                         \
            _Pragma("omp parallel for reduction(min : m) schedule(dynamic, 1)")                 \
            for (int k = b[4]; k < b[5]; k++)                                                   \
                for (int j = b[2]; j < b[3]; j++)                                               \
                    for (int i = b[0]; i < b[1]; i++) {                                         \
                        T v = a[IDX3(i, j, k, d)];                                              \
                        if (v < m) m = v;                                                       \
                    }                                                                           \
            *(T*)result = m;                                                                    \
        } else {                                                                                \
            T s = 0;                                                                            \
            _Pragma("omp parallel for reduction(+ : s) schedule(dynamic, 1)")                   \
            for (int k = b[4]; k < b[5]; k++)                                                   \
                for (int j = b[2]; j < b[3]; j++)                                               \
                    for (int i = b[0]; i < b[1]; i++) s += a[IDX3(i, j, k, d)];                 \
            *(T*)result = s;                                                                    \
        }                                                                                       \
    } while (0)

static void norm_box(int ndim, const int* dims, const int* box, int* d, int* b)
{
    d[0] = d[1] = d[2] = 1;
    for (int a = 0; a < ndim; a++) d[a] = dims[a];
    for (int a = 0; a < 3; a++) {
        if (a < ndim) { b[2 * a] = box[2 * a]; b[2 * a + 1] = box[2 * a + 1]; }
        else { b[2 * a] = 0; b[2 * a + 1] = 1; }
    }
}

int orc_reduce(const void* data, int dtype, int ndim, const int* dims, const int* box, int op, void* result)
{
    int d[3], b[6];
    if (ndim < 1 || ndim > 3 || op < 0 || op > 2) return 1;
    norm_box(ndim, dims, box, d, b);
    switch (dtype) {
    case ORC_F64: REDUCE_BODY(double, DBL_MIN, DBL_MAX); break;
    case ORC_F32: REDUCE_BODY(float, FLT_MIN, FLT_MAX); break;
    case ORC_I32: REDUCE_BODY(int32_t, INT_MIN, INT_MAX); break;
    case ORC_U8: REDUCE_BODY(uint8_t, 0, UCHAR_MAX); break;
    default: return 1;
    }
    return 0;
}

int orc_reduce_sum_wide(const void* data, int dtype, int ndim, const int* dims, const int* box, void* result)
{
    int d[3], b[6];
    if (ndim < 1 || ndim > 3) return 1;
    norm_box(ndim, dims, box, d, b);
    if (dtype == ORC_U8 || dtype == ORC_I32) {
        int64_t s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (int k = b[4]; k < b[5]; k++)
            for (int j = b[2]; j < b[3]; j++)
                for (int i = b[0]; i < b[1]; i++)
                    s += dtype == ORC_U8 ? (int64_t)((const uint8_t*)data)[IDX3(i, j, k, d)]
                                         : (int64_t)((const int32_t*)data)[IDX3(i, j, k, d)];
        *(int64_t*)result = s;
        return 0;
    }
    if (dtype == ORC_F32 || dtype == ORC_F64) {
        long double s = 0.0L;
        for (int k = b[4]; k < b[5]; k++)
            for (int j = b[2]; j < b[3]; j++)
                for (int i = b[0]; i < b[1]; i++)
                    s += dtype == ORC_F32 ? (long double)((const float*)data)[IDX3(i, j, k, d)]
                                          : (long double)((const double*)data)[IDX3(i, j, k, d)];
        *(double*)result = (double)s;
        return 0;
    }
    return 1;
}

/* storage conversions exposed for tests */
uint16_t orc_f32_to_f16_bits(float f) { return orc_f32_to_f16(f); }
float orc_f16_bits_to_f32(uint16_t h) { return orc_f16_to_f32(h); }
uint16_t orc_f32_to_bf16_bits(float f) { return orc_f32_to_bf16(f); }
float orc_bf16_bits_to_f32(uint16_t h) { return orc_bf16_to_f32(h); }

/* number of OpenMP threads the oracle loops use (reported with the CPU baseline) */
#ifdef _OPENMP
#include <omp.h>
#endif
int orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
