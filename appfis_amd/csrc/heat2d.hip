// HeatDiffusion 2D 5-point explicit sweep (config form of the reference functor
// HeatDiffusion/src/heat_appfis.hpp:166-174, driven like heat_appfis.hpp:86-91 on a
// DUAL grid: sweep, then Grid::Swap).
//
//   u' = u + c*((((uN + uS) + uE) + uW) - 4u)      one rounding per operation, no FMA
//
// PLAIN : one sweep per launch, one output per thread, x along the wave.
// CHOSEN: temporal blocking.  A block loads a window of the field with a T-cell halo
//         and advances it T sweeps on chip, then writes the window's centre; the valid
//         region shrinks by one cell per sweep.  Two window forms are autotuned: in LDS
//         (k_heat2d_tb) and in registers with DPP neighbours (k_heat2d_rt).  Dirichlet
//         ghost cells are loaded and never updated, exactly as in the one-sweep form, so
//         every interior value is bit-identical to PLAIN.  HBM/MALL traffic per sweep
//         drops by ~T.
#include "common.hpp"

#include <vector>

using namespace appfis;

namespace {

// One sweep; a thread owns RY consecutive rows of one column, so the column's RY+2
// values are loaded once and its loads are all in flight together.  Block 64 x 4
// threads = a 64 x 4*RY tile; tiles are spread over the XCDs in contiguous runs.
template <typename T, int RY>
__global__ __launch_bounds__(256) void k_heat2d_plain(const T* __restrict__ u, T* __restrict__ v, int nx, int ny,
                                                      T c)
{
    const int W = nx + 2;
    const unsigned gx = gridDim.x;
    const unsigned t = xcd_tile(blockIdx.x + gx * blockIdx.y, gx * gridDim.y);
    const int i = (int)(t % gx) * 64 + threadIdx.x + 1;
    const int j0 = ((int)(t / gx) * 4 + threadIdx.y) * RY + 1;
    if (i > nx || j0 > ny) return;
    T col[RY + 2], e[RY], w[RY];
#pragma unroll
    for (int r = 0; r < RY + 2; r++) {
        const int j = j0 - 1 + r;
        col[r] = j <= ny + 1 ? u[i + (size_t)W * j] : (T)0;
    }
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const size_t o = i + (size_t)W * (j0 + r);
        e[r] = j0 + r <= ny ? u[o + 1] : (T)0;
        w[r] = j0 + r <= ny ? u[o - 1] : (T)0;
    }
#pragma unroll
    for (int r = 0; r < RY; r++) {
        if (j0 + r > ny) break;
        const T uc = col[r + 1];
        T s = col[r] + col[r + 2];
        s = s + e[r];
        s = s + w[r];
        s = s - (T)4 * uc;
        v[i + (size_t)W * (j0 + r)] = uc + c * s;
    }
}

// Temporal blocking: TX x TY output tile, TS sweeps per launch, NT threads.
// One TX x TY output tile (tile coordinates tbx, tby) of one grid, `steps` <= TS sweeps.
template <typename T, int TX, int TY, int TS, int NT>
__device__ inline void tb_tile(const T* __restrict__ u, T* __restrict__ v, int nx, int ny, T c, int steps, int tbx,
                               int tby)
{
    constexpr int RX = TX + 2 * TS, RY = TY + 2 * TS;
    constexpr int NCELL = RX * RY;
    constexpr int PER = (NCELL + NT - 1) / NT;
    __shared__ T tile[RY][RX];
    const int W = nx + 2, H = ny + 2;
    // window origin in padded coordinates (may be negative / past the array)
    const int gx0 = tbx * TX + 1 - TS;
    const int gy0 = tby * TY + 1 - TS;
    const int tid = threadIdx.x;
    for (int p = tid; p < NCELL; p += NT) {
        int ly = p / RX, lx = p - ly * RX;
        int gx = gx0 + lx, gy = gy0 + ly;
        T val = (T)0;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) val = u[(size_t)gx + (size_t)W * (size_t)gy];
        tile[ly][lx] = val;
    }
    __syncthreads();
    T nv[PER];
    for (int s = 0; s < steps; s++) {
        // cells at distance >= s+1 from the window edge can be advanced to sweep s+1
        const int lo = s + 1, hx = RX - s - 1, hy = RY - s - 1;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int p = tid + q * NT;
            int ly = p / RX, lx = p - ly * RX;
            int gx = gx0 + lx, gy = gy0 + ly;
            bool active = p < NCELL && lx >= lo && lx < hx && ly >= lo && ly < hy && gx >= 1 && gx <= nx &&
                          gy >= 1 && gy <= ny;
            if (active) {
                const T uc = tile[ly][lx];
                T sum = tile[ly - 1][lx] + tile[ly + 1][lx];
                sum = sum + tile[ly][lx + 1];
                sum = sum + tile[ly][lx - 1];
                sum = sum - (T)4 * uc;
                nv[q] = uc + c * sum;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int p = tid + q * NT;
            int ly = p / RX, lx = p - ly * RX;
            int gx = gx0 + lx, gy = gy0 + ly;
            bool active = p < NCELL && lx >= lo && lx < hx && ly >= lo && ly < hy && gx >= 1 && gx <= nx &&
                          gy >= 1 && gy <= ny;
            if (active) tile[ly][lx] = nv[q];
        }
        __syncthreads();
    }
    // write the TX x TY centre (interior cells only)
    for (int p = tid; p < TX * TY; p += NT) {
        int ly = p / TX, lx = p - ly * TX;
        int gx = tbx * TX + 1 + lx, gy = tby * TY + 1 + ly;
        if (gx <= nx && gy <= ny) v[(size_t)gx + (size_t)W * (size_t)gy] = tile[ly + TS][lx + TS];
    }
}

template <typename T, int TX, int TY, int TS, int NT>
__global__ __launch_bounds__(NT) void k_heat2d_tb(const T* __restrict__ u, T* __restrict__ v, int nx, int ny, T c,
                                                  int steps)
{
    const unsigned nbx = gridDim.x;
    const unsigned t = xcd_tile(blockIdx.x + nbx * blockIdx.y, nbx * gridDim.y);
    tb_tile<T, TX, TY, TS, NT>(u, v, nx, ny, c, steps, (int)(t % nbx), (int)(t / nbx));
}

// Ensemble: grid s owns tiles tile0 .. tile0 + tiles_x * tiles_y - 1 of one launch.
template <typename T>
struct HeatGrid {
    T* b[2];
    int nx, ny, tiles_x, tile0;
};

// The grid that owns tile `tile` of an ensemble launch: binary search over tile0 (the
// same for every thread of the block).
template <typename T>
__device__ inline int grid_of_tile(const HeatGrid<T>* __restrict__ tab, int ngrid, int tile)
{
    int lo = 0, hi = ngrid - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Register tiles: a block of NW waves holds a window of 64*WPL columns x NW*K rows, lane
// l owning columns WPL*l .. WPL*l+WPL-1 of K consecutive rows (wave w: window rows w*K ..
// w*K+K-1) in registers.  A sweep takes a cell's east / west neighbours from the lane's
// own registers, or by DPP from lanes l+1 / l-1 for its outer columns, and the north /
// south ones from the lane's own registers; only each wave's first and last rows go
// through LDS, in two slots alternating with the sweep's parity, so a sweep has one
// barrier.  Cells at the window's edge read meaningless neighbours (zero from the DPP
// shift at lanes 0 / 63, the wave's own edge rows above the first wave and below the
// last), which spoils one more ring per sweep; after TS sweeps the centre
// (64*WPL-2TS) x (NW*K-2TS) is exact.  Ghost cells keep their loaded values, so every
// interior value equals the one-sweep form's.
template <typename T, int K, int NW, int TS, int WPL>
__device__ inline void rt_tile(const T* __restrict__ u, T* __restrict__ v, int nx, int ny, T c, int steps, int tbx,
                               int tby)
{
    constexpr int TX = 64 * WPL - 2 * TS, TY = NW * K - 2 * TS;
    static_assert(TX > 0 && TY > 0, "the window must be wider and taller than its halo");
    __shared__ T edge[2][NW][2][WPL][64];
    const gmem<T>* ug = to_global(u);  // the ensemble reads u and v from its grid table
    gmem<T>* vg = to_global(v);
    const int W = nx + 2, H = ny + 2;
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    const int gx0 = tbx * TX + 1 - TS + WPL * lane;  // padded column of R[.][0]
    const int gy0 = tby * TY + 1 - TS + wv * K;      // padded row of R[0][.]
    bool col_in[WPL];
#pragma unroll
    for (int w = 0; w < WPL; w++) col_in[w] = gx0 + w >= 1 && gx0 + w <= nx;
    T R[K][WPL];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int gy = gy0 + k;
#pragma unroll
        for (int w = 0; w < WPL; w++) {
            const int gx = gx0 + w;
            R[k][w] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? ug[(size_t)gx + (size_t)W * (size_t)gy] : (T)0;
        }
    }
    // one sweep; its edge rows use LDS slot b (sweeps alternate slots)
    auto sweep = [&](const int b) {
#pragma unroll
        for (int w = 0; w < WPL; w++) {
            edge[b][wv][0][w][lane] = R[0][w];
            edge[b][wv][1][w][lane] = R[K - 1][w];
        }
        __syncthreads();
        // the row above row k before this sweep, and the row below the wave.  The block's
        // first and last waves read their own slots instead: rows outside the window only
        // reach the halo ring, which is never stored.
        const int wu = wv > 0 ? wv - 1 : wv, wd = wv < NW - 1 ? wv + 1 : wv;
        T above[WPL], dn[WPL];
#pragma unroll
        for (int w = 0; w < WPL; w++) {
            above[w] = edge[b][wu][1][w][lane];
            dn[w] = edge[b][wd][0][w][lane];
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const T east = lane_shift_z<false>(R[k][0]);       // lane l+1's first column
            const T west = lane_shift_z<true>(R[k][WPL - 1]);  // lane l-1's last column
            const int gy = gy0 + k;
            const bool row_in = gy >= 1 && gy <= ny;
            T nv[WPL];
#pragma unroll
            for (int w = 0; w < WPL; w++) {
                const T uc = R[k][w];
                T sum = above[w] + (k == K - 1 ? dn[w] : R[k + 1][w]);
                sum = sum + (w == WPL - 1 ? east : R[k][w + 1]);
                sum = sum + (w == 0 ? west : R[k][w - 1]);
                sum = sum - (T)4 * uc;
                nv[w] = (col_in[w] && row_in) ? uc + c * sum : uc;
            }
#pragma unroll
            for (int w = 0; w < WPL; w++) {
                above[w] = R[k][w];
                R[k][w] = nv[w];
            }
        }
    };
    // two sweeps per trip, so that the registers of consecutive sweeps are renamed
    // rather than copied
    for (int s = 0; s + 1 < steps; s += 2) {
        sweep(0);
        sweep(1);
    }
    if (steps & 1) sweep(0);
    // the centre: window columns TS .. TS+TX-1, rows TS .. TS+TY-1 (interior cells only)
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = wv * K + k, gy = gy0 + k;
        if (r < TS || r >= TS + TY || gy < 1 || gy > ny) continue;
#pragma unroll
        for (int w = 0; w < WPL; w++) {
            const int lx = WPL * lane + w;
            if (lx >= TS && lx < TS + TX && col_in[w]) vg[(size_t)(gx0 + w) + (size_t)W * (size_t)gy] = R[k][w];
        }
    }
}

template <typename T, int K, int NW, int TS, int WPL>
__global__ __launch_bounds__(64 * NW) void k_heat2d_rt(const T* __restrict__ u, T* __restrict__ v, int nx, int ny,
                                                       T c, int steps)
{
    constexpr int TX = 64 * WPL - 2 * TS;
    const int tiles_x = (nx + TX - 1) / TX;
    const int t = (int)xcd_tile(blockIdx.x, gridDim.x);
    const int tby = t / tiles_x;
    rt_tile<T, K, NW, TS, WPL>(u, v, nx, ny, c, steps, t - tby * tiles_x, tby);
}

template <typename T, int K, int NW, int TS, int WPL>
__global__ __launch_bounds__(64 * NW) void k_heat2d_rt_ens(const HeatGrid<T>* __restrict__ tab, int ngrid, int par,
                                                           T c, int steps)
{
    const int tile = (int)xcd_tile(blockIdx.x, gridDim.x);
    const HeatGrid<T>& G = tab[grid_of_tile(tab, ngrid, tile)];
    const int t = tile - G.tile0;
    const int tby = t / G.tiles_x;
    rt_tile<T, K, NW, TS, WPL>(G.b[par], G.b[par ^ 1], G.nx, G.ny, c, steps, t - tby * G.tiles_x, tby);
}

constexpr int PLAIN_RY = 4;

template <typename T>
int run_plain(appfis_ctx* ctx, T* b0, T* b1, int nx, int ny, T c, int iters)
{
    dim3 block(64, 4), grid(ceil_div(nx, 64), ceil_div(ny, 4 * PLAIN_RY));
    T* buf[2] = {b0, b1};
    for (int it = 0; it < iters; it++) {
        hipLaunchKernelGGL((k_heat2d_plain<T, PLAIN_RY>), grid, block, 0, ctx->stream, buf[it & 1],
                           buf[(it + 1) & 1], nx, ny, c);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

// Split `iters` sweeps into launches of at most TS sweeps such that the number of
// launches has the parity of `iters`: the field then ends in bufs[iters % 2], as
// with one sweep per launch.
inline int plan_launches(int iters, int ts, int* counts, int max_launches)
{
    int n = (iters + ts - 1) / ts;
    if ((n & 1) != (iters & 1)) n++;
    if (n > iters) n = iters;
    if (n > max_launches) return -1;
    int base = iters / n, extra = iters % n;
    for (int l = 0; l < n; l++) counts[l] = base + (l < extra ? 1 : 0);
    return n;
}

template <typename T, int TX, int TY, int TS, int NT>
int run_tb(appfis_ctx* ctx, T* b0, T* b1, int nx, int ny, T c, int iters, int depth)
{
    T* buf[2] = {b0, b1};
    dim3 grid(ceil_div(nx, TX), ceil_div(ny, TY));
    std::vector<int> counts((size_t)iters);
    int n = plan_launches(iters, depth < TS ? depth : TS, counts.data(), iters);
    if (n <= 0) return invalid("heat2d: cannot split %d sweeps", iters);
    for (int l = 0; l < n; l++) {
        hipLaunchKernelGGL((k_heat2d_tb<T, TX, TY, TS, NT>), grid, dim3(NT), 0, ctx->stream, buf[l & 1],
                           buf[(l + 1) & 1], nx, ny, c, counts[l]);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

template <typename T, int TX, int TY, int TS, int NT = 256>
int run_tb_v(appfis_ctx* ctx, void* b0, void* b1, int nx, int ny, double c, int iters, int depth)
{
    return run_tb<T, TX, TY, TS, NT>(ctx, (T*)b0, (T*)b1, nx, ny, (T)c, iters, depth);
}

template <typename T, int K, int NW, int TS, int WPL = 1>
int run_rt(appfis_ctx* ctx, void* b0, void* b1, int nx, int ny, double c, int iters, int depth)
{
    constexpr int TX = 64 * WPL - 2 * TS, TY = NW * K - 2 * TS;
    T* buf[2] = {(T*)b0, (T*)b1};
    const size_t tiles = (size_t)ceil_div(nx, TX) * ceil_div(ny, TY);
    APPFIS_REQUIRE(tiles < (1u << 31), "heat2d: too many tiles");
    std::vector<int> counts((size_t)iters);
    int n = plan_launches(iters, depth < TS ? depth : TS, counts.data(), iters);
    if (n <= 0) return invalid("heat2d: cannot split %d sweeps", iters);
    for (int l = 0; l < n; l++) {
        hipLaunchKernelGGL((k_heat2d_rt<T, K, NW, TS, WPL>), dim3((unsigned)tiles), dim3(64 * NW), 0, ctx->stream,
                           buf[l & 1], buf[(l + 1) & 1], nx, ny, (T)c, counts[l]);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

struct HeatVariant {
    const char* name;
    int (*fn)(appfis_ctx*, void*, void*, int, int, double, int, int);
    int ts;  // sweeps per launch at most
};

// temporal-blocking candidates (output tile TX x TY, TS sweeps per launch, NT threads).
// A launch costs a fixed load of its (TX+2TS) x (TY+2TS) windows plus a per-sweep part;
// larger tiles load less redundant window per output cell, and enough threads keep the
// per-sweep part short.
static const HeatVariant kHeatF64[] = {
    {"tb 32x32 ts8 nt256", run_tb_v<double, 32, 32, 8>, 8},
    {"tb 32x32 ts4 nt256", run_tb_v<double, 32, 32, 4>, 4},
    {"tb 64x32 ts8 nt256", run_tb_v<double, 64, 32, 8>, 8},
    {"tb 64x32 ts12 nt256", run_tb_v<double, 64, 32, 12>, 12},
    {"tb 64x64 ts8 nt1024", run_tb_v<double, 64, 64, 8, 1024>, 8},
    {"tb 64x64 ts12 nt1024", run_tb_v<double, 64, 64, 12, 1024>, 12},
    {"tb 64x32 ts8 nt512", run_tb_v<double, 64, 32, 8, 512>, 8},
    {"tb 64x32 ts16 nt256", run_tb_v<double, 64, 32, 16>, 16},
    // register tiles, named by output tile: 64 - 2TS columns x NW*K - 2TS rows (at 1024^2
    // each fills at most 256 blocks, one per CU)
    {"rt 48x96 ts8 k7 w16", run_rt<double, 7, 16, 8>, 8},
    {"rt 48x96 ts8 k14 w8", run_rt<double, 14, 8, 8>, 8},
    {"rt 56x88 ts4 k6 w16", run_rt<double, 6, 16, 4>, 4},
    {"rt 40x120 ts12 k9 w16", run_rt<double, 9, 16, 12>, 12},
    {"rt 32x128 ts16 k10 w16", run_rt<double, 10, 16, 16>, 16},
    // two blocks per CU at 1024^2 (one computes while the other waits at its barrier)
    {"rt 48x48 ts8 k4 w16", run_rt<double, 4, 16, 8>, 8},
    // two columns per lane: one horizontal neighbour of each cell is in the lane's
    // registers (at 1024^2 it loses to one column per lane: 0.064 against 0.060 ms)
    {"rt2 112x48 ts8 k4 w16", run_rt<double, 4, 16, 8, 2>, 8},
};
static const HeatVariant kHeatF32[] = {
    {"tb 64x32 ts8 nt256", run_tb_v<float, 64, 32, 8>, 8},
    {"tb 32x32 ts4 nt256", run_tb_v<float, 32, 32, 4>, 4},
    {"tb 64x64 ts8 nt256", run_tb_v<float, 64, 64, 8>, 8},
    {"tb 64x64 ts16 nt256", run_tb_v<float, 64, 64, 16>, 16},
    {"tb 64x64 ts8 nt1024", run_tb_v<float, 64, 64, 8, 1024>, 8},
    {"tb 64x64 ts16 nt1024", run_tb_v<float, 64, 64, 16, 1024>, 16},
    {"tb 128x64 ts16 nt1024", run_tb_v<float, 128, 64, 16, 1024>, 16},
    {"rt 48x96 ts8 k7 w16", run_rt<float, 7, 16, 8>, 8},
    {"rt 48x96 ts8 k14 w8", run_rt<float, 14, 8, 8>, 8},
    {"rt 56x88 ts4 k6 w16", run_rt<float, 6, 16, 4>, 4},
    {"rt 40x120 ts12 k9 w16", run_rt<float, 9, 16, 12>, 12},
    {"rt 32x128 ts16 k10 w16", run_rt<float, 10, 16, 16>, 16},
    {"rt 48x48 ts8 k4 w16", run_rt<float, 4, 16, 8>, 8},
    {"rt2 112x48 ts8 k4 w16", run_rt<float, 4, 16, 8, 2>, 8},
};

}  // namespace

extern "C" int appfis_heat2d(appfis_ctx* ctx, int dtype, void* buf0, void* buf1, int nx, int ny, double c, int iters,
                             int variant, int tsteps)
{
    APPFIS_REQUIRE(ctx && buf0 && buf1, "NULL argument");
    APPFIS_REQUIRE(nx > 0 && ny > 0, "interior must be non-empty (nx=%d ny=%d)", nx, ny);
    APPFIS_REQUIRE((double)(nx + 2) * (ny + 2) < 2147483647.0, "grid exceeds the int index range");
    APPFIS_REQUIRE(iters >= 0, "iters must be >= 0");
    APPFIS_REQUIRE(buf0 != buf1, "the two buffers must differ");
    if (iters == 0) return APPFIS_OK;
    if (variant == APPFIS_PLAIN) {
        if (dtype == APPFIS_F64) return run_plain<double>(ctx, (double*)buf0, (double*)buf1, nx, ny, c, iters);
        if (dtype == APPFIS_F32) return run_plain<float>(ctx, (float*)buf0, (float*)buf1, nx, ny, (float)c, iters);
        return invalid("heat2d: dtype must be F64 or F32");
    }
    if (variant == APPFIS_CHOSEN) {
        if (dtype != APPFIS_F64 && dtype != APPFIS_F32) return invalid("heat2d: dtype must be F64 or F32");
        const HeatVariant* tab = dtype == APPFIS_F64 ? kHeatF64 : kHeatF32;
        const int ntab = dtype == APPFIS_F64 ? (int)(sizeof(kHeatF64) / sizeof(kHeatF64[0]))
                                             : (int)(sizeof(kHeatF32) / sizeof(kHeatF32[0]));
        if (tsteps > 0) {
            // caller-fixed depth: the first candidate that holds that many sweeps per launch
            for (int i = 0; i < ntab; i++)
                if (tab[i].ts >= tsteps) return tab[i].fn(ctx, buf0, buf1, nx, ny, c, iters, tsteps);
            return invalid("heat2d: tsteps %d exceeds the deepest temporal blocking (16)", tsteps);
        }
        // autotuned tile shape and depth, timed once per shape on scratch buffers; every
        // candidate computes the same values in the same order
        const int key[6] = {1, dtype, nx, ny, 0, 0};
        bool cached = false;
        for (const appfis_tune_entry& e : ctx->tuned)
            if (memcmp(e.key, key, sizeof(key)) == 0) cached = true;
        void* s0 = nullptr;
        void* s1 = nullptr;
        if (!cached) {
            APPFIS_NOT_CAPTURING(ctx, "appfis_heat2d CHOSEN's first call for a shape (autotuning)");
            const size_t half = ((size_t)(nx + 2) * (ny + 2) * dtype_size(dtype) + 255) / 256 * 256;
            int st = ensure_work(ctx, 2 * half);
            if (st) return st;
            s0 = ctx->work;
            s1 = (char*)ctx->work + half;
            APPFIS_HIP_CHECK(hipMemsetAsync(ctx->work, 0, 2 * half, ctx->stream));
        }
        std::vector<const char*> names;
        for (int i = 0; i < ntab; i++) names.push_back(tab[i].name);
        int pick = 0;
        int st = tune_select(ctx, key, ntab, names.data(),
                             [&](int i) { return tab[i].fn(ctx, s0, s1, nx, ny, c, 48, tab[i].ts); }, 3, &pick);
        if (st) return st;
        return tab[pick].fn(ctx, buf0, buf1, nx, ny, c, iters, tab[pick].ts);
    }
    if (variant >= APPFIS_CANDIDATE0) {
        if (dtype != APPFIS_F64 && dtype != APPFIS_F32) return invalid("heat2d: dtype must be F64 or F32");
        const HeatVariant* tab = dtype == APPFIS_F64 ? kHeatF64 : kHeatF32;
        const int ntab = dtype == APPFIS_F64 ? (int)(sizeof(kHeatF64) / sizeof(kHeatF64[0]))
                                             : (int)(sizeof(kHeatF32) / sizeof(kHeatF32[0]));
        const int i = variant - APPFIS_CANDIDATE0;
        APPFIS_REQUIRE(i < ntab, "heat2d: candidate %d out of range (%d candidates)", i, ntab);
        return tab[i].fn(ctx, buf0, buf1, nx, ny, c, iters, tab[i].ts);
    }
    return invalid("heat2d: unknown variant %d", variant);
}

template <typename T>
using ens_kernel_t = void (*)(const HeatGrid<T>*, int, int, T, int);

// Ensemble launches with output tiles TX x TY, at most TS sweeps per launch, NT threads.
template <typename T, int TX, int TY, int TS, int NT>
int run_ensemble(appfis_ctx* ctx, ens_kernel_t<T> kern, int ngrid, const int* nx, const int* ny, void* const* buf0,
                 void* const* buf1, double c, int iters)
{
    std::vector<HeatGrid<T>> hg((size_t)ngrid);
    long long tiles = 0;
    for (int g = 0; g < ngrid; g++) {
        APPFIS_REQUIRE(nx[g] > 0 && ny[g] > 0, "heat2d ensemble: grid %d is empty", g);
        APPFIS_REQUIRE((double)(nx[g] + 2) * (ny[g] + 2) < 2147483647.0, "grid %d exceeds the int index range", g);
        APPFIS_REQUIRE(buf0[g] && buf1[g] && buf0[g] != buf1[g], "heat2d ensemble: grid %d buffers", g);
        hg[g].b[0] = (T*)buf0[g];
        hg[g].b[1] = (T*)buf1[g];
        hg[g].nx = nx[g];
        hg[g].ny = ny[g];
        hg[g].tiles_x = (int)ceil_div(nx[g], TX);
        hg[g].tile0 = (int)tiles;
        tiles += (long long)hg[g].tiles_x * ceil_div(ny[g], TY);
    }
    APPFIS_REQUIRE(tiles < (1LL << 31), "heat2d ensemble: too many tiles");
    const size_t tab_bytes = (size_t)ngrid * sizeof(HeatGrid<T>);
    int st = ensure_work(ctx, tab_bytes);
    if (st) return st;
    HeatGrid<T>* tab = (HeatGrid<T>*)ctx->work;
    APPFIS_HIP_CHECK(hipMemcpyAsync(tab, hg.data(), tab_bytes, hipMemcpyHostToDevice, ctx->stream));
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host table is a local
    std::vector<int> counts((size_t)iters);
    const int n = plan_launches(iters, TS, counts.data(), iters);
    if (n <= 0) return invalid("heat2d: cannot split %d sweeps", iters);
    // the whole launch chain is one graph, captured on the first call with these
    // arguments and replayed by later ones (it reads the grids from the table)
    const std::vector<uint64_t> key = {1, key_ptr((const void*)kern), key_ptr(tab), (uint64_t)ngrid,
                                       (uint64_t)tiles, key_bits(c), (uint64_t)iters};
    hipGraphExec_t exec = nullptr;
    st = cached_graph(ctx, key,
                      [&]() {
                          for (int l = 0; l < n; l++) {
                              hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(NT), 0, ctx->stream, tab, ngrid,
                                                 l & 1, (T)c, counts[l]);
                              APPFIS_LAUNCH_CHECK();
                          }
                          return APPFIS_OK;
                      },
                      &exec);
    if (st) return st;
    APPFIS_HIP_CHECK(hipGraphLaunch(exec, ctx->stream));
    // the table stays in use until the graph ran: wait before the workspace can be reused
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return APPFIS_OK;
}

extern "C" int appfis_heat2d_ensemble(appfis_ctx* ctx, int dtype, int ngrid, const int* nx, const int* ny,
                                      void* const* buf0, void* const* buf1, double c, int iters)
{
    APPFIS_REQUIRE(ctx && nx && ny && buf0 && buf1, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "appfis_heat2d_ensemble (it captures and replays a graph of its own)");
    APPFIS_REQUIRE(ngrid >= 1, "heat2d ensemble: need at least one grid");
    APPFIS_REQUIRE(iters >= 0, "iters must be >= 0");
    if (iters == 0) return APPFIS_OK;
    // register tiles, 48 x 96 output cells, 8 sweeps per launch, 16 waves
    if (dtype == APPFIS_F64)
        return run_ensemble<double, 48, 96, 8, 1024>(ctx, k_heat2d_rt_ens<double, 7, 16, 8, 1>, ngrid, nx, ny, buf0, buf1,
                                                     c, iters);
    if (dtype == APPFIS_F32)
        return run_ensemble<float, 48, 96, 8, 1024>(ctx, k_heat2d_rt_ens<float, 7, 16, 8, 1>, ngrid, nx, ny, buf0, buf1, c,
                                                    iters);
    return invalid("heat2d ensemble: dtype must be F64 or F32");
}

extern "C" int appfis_heat2d_candidates(int dtype, char* names, int len)
{
    if (dtype != APPFIS_F64 && dtype != APPFIS_F32) return -invalid("heat2d: dtype must be F64 or F32");
    const HeatVariant* tab = dtype == APPFIS_F64 ? kHeatF64 : kHeatF32;
    const int ntab = dtype == APPFIS_F64 ? (int)(sizeof(kHeatF64) / sizeof(kHeatF64[0]))
                                         : (int)(sizeof(kHeatF32) / sizeof(kHeatF32[0]));
    std::string all;
    for (int i = 0; i < ntab; i++) all += std::string(tab[i].name) + "\n";
    if (names && len > 0) snprintf(names, (size_t)len, "%s", all.c_str());
    return ntab;
}
