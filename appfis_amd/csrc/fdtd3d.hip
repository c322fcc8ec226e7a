// FDTD3d radius-4 25-point star (config form of the reference functor
// FDTD3d/src/fdtd3d_appfis.hpp:129-143, driven like fdtd3d_appfis.hpp:79-84 on a DUAL
// grid with a 4-cell ghost held at zero):
//
//   s_r = ((u[x-r] + u[x+r]) + (u[y-r] + u[y+r])) + (u[z-r] + u[z+r])
//   u'  = (((c0*u + c1*s_1) + c2*s_2) + c3*s_3) + c4*s_4        (no FMA contraction)
//
// 2.5D k-marching: a block owns a TX x (TY*RY) column of the x-y plane and a chunk of
// z.  Each thread keeps a 9-deep register queue along z for each of its RY rows, so
// the z-neighbours cost no extra loads; the x-y plane at the current z is staged in LDS
// with its 4-wide halo for the x and y neighbours.  Per output: one queue load, the
// plane halo share (~0.6 loads, mostly L2 hits) and one store.
//
// dtype F32 / F64: storage and arithmetic in that type.  F16 / BF16: storage only,
// arithmetic in FP32, round-to-nearest-even on store.
#include "common.hpp"

using namespace appfis;

namespace {

template <typename S> struct Cvt;
template <> struct Cvt<float> {
    using C = float;
    __device__ static float ld(float v) { return v; }
    __device__ static float st(float v) { return v; }
};
template <> struct Cvt<double> {
    using C = double;
    __device__ static double ld(double v) { return v; }
    __device__ static double st(double v) { return v; }
};
template <> struct Cvt<__half> {
    using C = float;
    __device__ static float ld(__half v) { return __half2float(v); }
    __device__ static __half st(float v) { return __float2half_rn(v); }
};
template <> struct Cvt<__hip_bfloat16> {
    using C = float;
    __device__ static float ld(__hip_bfloat16 v) { return __bfloat162float(v); }
    __device__ static __hip_bfloat16 st(float v) { return __float2bfloat16(v); }
};

constexpr int R = 4;

template <typename S, int TX, int TY, int RY, int ZC>
__global__ __launch_bounds__(TX* TY) void k_fdtd3d(const S* __restrict__ u, S* __restrict__ v, int nx, int ny, int nz,
                                                    typename Cvt<S>::C c0, typename Cvt<S>::C c1,
                                                    typename Cvt<S>::C c2, typename Cvt<S>::C c3,
                                                    typename Cvt<S>::C c4)
{
    using C = typename Cvt<S>::C;
    constexpr int PY = TY * RY;           // tile rows
    constexpr int LX = TX + 2 * R;        // LDS plane width
    constexpr int LY = PY + 2 * R;        // LDS plane height
    __shared__ C plane[LY][LX];

    const int W = nx + 2 * R, H = ny + 2 * R;
    const size_t sy = (size_t)W, sz = (size_t)W * (size_t)H;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * TX + tx;
    const int gx = blockIdx.x * TX + tx + R;         // padded x of this thread
    const int gy0 = blockIdx.y * PY + ty * RY + R;   // padded y of row 0 of this thread
    const int z0 = blockIdx.z * ZC + R;              // first padded z of the chunk
    const int z1 = min(z0 + ZC, nz + R);             // one past the last
    const bool xin = gx < nx + R;   // interior column: stores
    const bool xpad = gx < W;       // inside the padded array: loads (ghost values kept)
    const C cc[5] = {c0, c1, c2, c3, c4};

    // register queues: q[r][m] holds u at z + m - 4 for row r
    C q[RY][2 * R + 1];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const int gy = gy0 + r;
        const bool in = xpad && gy < H;
#pragma unroll
        for (int m = 0; m < 2 * R; m++) {
            const int z = z0 - R + m;
            q[r][m + 1] = in ? Cvt<S>::ld(u[(size_t)gx + sy * gy + sz * z]) : (C)0;
        }
    }

    for (int z = z0; z < z1; z++) {
        // advance the queues: q[.][8] = u at z + 4
#pragma unroll
        for (int r = 0; r < RY; r++) {
#pragma unroll
            for (int m = 0; m < 2 * R; m++) q[r][m] = q[r][m + 1];
            const int gy = gy0 + r;
            const bool in = xpad && gy < H;
            q[r][2 * R] = in ? Cvt<S>::ld(u[(size_t)gx + sy * gy + sz * (z + R)]) : (C)0;
        }
        __syncthreads();  // previous plane fully consumed
        // own cells
#pragma unroll
        for (int r = 0; r < RY; r++) plane[ty * RY + r + R][tx + R] = q[r][R];
        // halo: y-halo rows (R above, R below) x TX columns, x-halo columns (R left,
        // R right) x PY rows; clamped to the padded array
        const size_t zoff = sz * (size_t)z;
        const int bx = blockIdx.x * TX + R, by = blockIdx.y * PY + R;
        for (int p = tid; p < 2 * R * TX + 2 * R * PY; p += TX * TY) {
            int lx, ly;
            if (p < 2 * R * TX) {
                int hy = p / TX;
                lx = p - hy * TX + R;
                ly = hy < R ? hy : PY + hy;  // rows 0..R-1 and PY+R..PY+2R-1
            } else {
                int p2 = p - 2 * R * TX;
                int hx = p2 / PY;
                ly = p2 - hx * PY + R;
                lx = hx < R ? hx : TX + hx;
            }
            const int xg = bx - R + lx, yg = by - R + ly;
            C val = (C)0;
            if (xg < W && yg < H) val = Cvt<S>::ld(u[(size_t)xg + sy * yg + zoff]);
            plane[ly][lx] = val;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RY; r++) {
            const int gy = gy0 + r;
            const int ly = ty * RY + r + R, lx = tx + R;
            C acc = cc[0] * q[r][R];
#pragma unroll
            for (int k = 1; k <= R; k++) {
                C sx = plane[ly][lx - k] + plane[ly][lx + k];
                C sY = plane[ly - k][lx] + plane[ly + k][lx];
                C sZ = q[r][R - k] + q[r][R + k];
                C sr = (sx + sY) + sZ;
                acc = acc + cc[k] * sr;
            }
            if (xin && gy < ny + R) v[(size_t)gx + sy * gy + zoff] = Cvt<S>::st(acc);
        }
    }
}

constexpr int TX = 64, TY = 4, RY = 4, ZC = 64;

template <typename S>
int run(appfis_ctx* ctx, S* b0, S* b1, int nx, int ny, int nz, const double* cf, int steps)
{
    using C = typename Cvt<S>::C;
    dim3 block(TX, TY), grid(ceil_div(nx, TX), ceil_div(ny, TY * RY), ceil_div(nz, ZC));
    S* buf[2] = {b0, b1};
    for (int s = 0; s < steps; s++) {
        hipLaunchKernelGGL((k_fdtd3d<S, TX, TY, RY, ZC>), grid, block, 0, ctx->stream, buf[s & 1], buf[(s + 1) & 1], nx,
                           ny, nz, (C)cf[0], (C)cf[1], (C)cf[2], (C)cf[3], (C)cf[4]);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

}  // namespace

extern "C" int appfis_fdtd3d(appfis_ctx* ctx, int dtype, void* buf0, void* buf1, int nx, int ny, int nz,
                             const double* coeff5, int steps, int variant)
{
    APPFIS_REQUIRE(ctx && buf0 && buf1 && coeff5, "NULL argument");
    APPFIS_REQUIRE(buf0 != buf1, "the two buffers must differ");
    APPFIS_REQUIRE(nx > 0 && ny > 0 && nz > 0, "interior must be non-empty");
    APPFIS_REQUIRE((double)(nx + 8) * (ny + 8) * (nz + 8) < 2147483647.0, "grid exceeds the int index range");
    APPFIS_REQUIRE(steps >= 0, "steps must be >= 0");
    APPFIS_REQUIRE(variant == APPFIS_PLAIN || variant == APPFIS_CHOSEN, "fdtd3d: unknown variant %d", variant);
    switch (dtype) {
    case APPFIS_F32: return run<float>(ctx, (float*)buf0, (float*)buf1, nx, ny, nz, coeff5, steps);
    case APPFIS_F64: return run<double>(ctx, (double*)buf0, (double*)buf1, nx, ny, nz, coeff5, steps);
    case APPFIS_F16: return run<__half>(ctx, (__half*)buf0, (__half*)buf1, nx, ny, nz, coeff5, steps);
    case APPFIS_BF16:
        return run<__hip_bfloat16>(ctx, (__hip_bfloat16*)buf0, (__hip_bfloat16*)buf1, nx, ny, nz, coeff5, steps);
    default: return invalid("fdtd3d: dtype must be F32, F64, F16 or BF16");
    }
}
