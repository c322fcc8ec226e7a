// Himeno 19-point Jacobi sweep (reference functor Himeno/src/himeno_appfis.hpp:178-198,
// driven like himeno_appfis.hpp:104-113 on a DUAL p grid) with the config's residual
//   gosa = sum over the interior of ss*ss      after every sweep.
//
// Layout: 14 dense arrays of dims (nx, ny, nz) including the 1-cell boundary shell.
//
// Kernel shape (2.5D k-march): a block owns full x rows (TX = 256 threads along x) and
// RY rows per thread, and marches a chunk of zc planes in k with the coefficients of
// the next D planes in flight (zc and D autotuned per shape: zc sets how many blocks
// share the chip and how many distinct plane regions they stream at once).  p is
// staged through an LDS ring of planes (k-1, k, k+1 with their 1-cell x/y halo), so
// each p value is read from memory about once; the 12 coefficient arrays and wrk1 are
// pure coalesced streams, read with buffer loads that the compiler can count in
// flight.  Tiles are ordered XCD-contiguously.  gosa: per-thread FP64 accumulation, wave64 DPP tree, one LDS stage, one
// partial per block, and the last block to arrive (atomic ticket) sums the partials in
// block order into gosa[it] -- deterministic, on the device, no host round trip.
//
// PLAIN (F32 or F64): the reference operation order, no FMA contraction.
// CHOSEN (defect correction, F64 problem): p = B + q with a FP64 base B and a FP32
//   correction q.  Because the sweep is affine, ss(B + q) = ss(B) + L q with
//   L q = (a3 * A q - q) * bnd, A the 18-neighbour weighted sum.  Every `outer` sweeps a
//   FP64 rebase sweep forms P = B + q, computes ss(P) exactly in FP64 (gosa from it),
//   and restarts with B = P, r = (float)ss(P), q = (float)(omega*ss(P)).  In between,
//   FP32 sweeps advance q with ss = r + L q using FP32 copies of the coefficients, so
//   the cancellation ss << p never happens in FP32.  The FP32 coefficient copy is
//   made once per call, inside the call, blocked by 64 points (Coef::inter).
#include "common.hpp"

#include <vector>

using namespace appfis;

namespace {

constexpr int TX = 256;  // threads along x (one block spans 256 interior columns)
constexpr int LW = TX + 2;  // staged row: the block's columns and their x halo

struct Coef {
    const void* a0;
    const void* a1;
    const void* a2;
    const void* a3;
    const void* b0;
    const void* b1;
    const void* b2;
    const void* c0;
    const void* c1;
    const void* c2;
    const void* bnd;
    const void* wrk1;  // PLAIN: wrk1 (the inner sweep reads r from `inter`)
    // inner sweep, MODE 2: the 11 static coefficients and r in blocks of 64 points: value
    // a (a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd r) of point p at blk_idx(p, a), so a wave's 12
    // coefficient loads fall in one 3 KB region instead of 12 arrays
    const void* inter = nullptr;
    // rebase sweep, MODE 1: the FP32 correction q added to the FP64 base while staging,
    // and the two FP32 outputs (r into slot 11 of `inter`, q' = omega * ss)
    const float* q_in = nullptr;
    float* q_out = nullptr;
};

constexpr int BLK = 12 * 64;  // floats per block of 64 points
__host__ __device__ inline size_t blk_idx(size_t p, int a) { return (p >> 6) * BLK + (size_t)a * 64 + (p & 63); }

// Block -> (x, y, z) tile of a 1-D launch: tiles x fastest, then y, then z chunks, and
// every XCD takes a contiguous run, so y-neighbour tiles (which share p rows) share an L2.
__device__ inline void tile_of(int nx, int ny, int ry, int& bx, int& by, int& bz)
{
    const int gx = (nx - 2 + TX - 1) / TX, gy = (ny - 2 + ry - 1) / ry;
    const int t = (int)xcd_tile(blockIdx.x, gridDim.x);
    bx = t % gx;
    const int r = t / gx;
    by = r % gy;
    bz = r / gy;
}

__device__ inline double block_sum_last(double v, double* partials, unsigned* ticket, double* out, double* lds)
{
    const auto add = [](double a, double b) { return a + b; };
    v = wave_reduce(v, 0.0, add);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __shared__ bool last;
    const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (lane == WAVE_TOTAL_LANE) lds[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < TX / 64; w++) s += lds[w];
        partials[bid] = s;
        __threadfence();
        unsigned prev = atomicAdd(ticket, 1u);
        last = (prev == nblk - 1);
    }
    __syncthreads();
    if (!last) return 0.0;
    __threadfence();
    double acc = 0.0;
    for (unsigned p = threadIdx.x; p < nblk; p += TX)
        acc += __hip_atomic_load(partials + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc = wave_reduce(acc, 0.0, add);
    __syncthreads();
    if (lane == WAVE_TOTAL_LANE) lds[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < TX / 64; w++) s += lds[w];
        *out = s;
        *ticket = 0u;
    }
    return 0.0;
}

// Buffer accesses: every load and store of the sweep is a buffer instruction on a
// resource of one plane (or, for the blocked coefficients, of the blocks from the plane's
// first point on), at a byte offset each thread fixes once.  Cells outside the array or
// not updated carry offset kOut = 2^31, past any resource's range, so their loads return
// 0 and their stores are dropped: no branches around memory accesses, and the compiler
// can count the loads in flight (a load under a lane branch makes it wait for all of
// them before the next plane).
constexpr uint32_t kOut = 0x80000000u;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const T* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ inline T ld(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (sizeof(T) == 4) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    } else {
        const u32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        return __hiloint2double((int)b[1], (int)b[0]);
    }
}
template <typename T>
__device__ inline void st(__amdgpu_buffer_rsrc_t r, uint32_t off, T v)
{
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 0);
    } else {
        const u32x2 b = {(unsigned)__double2loint(v), (unsigned)__double2hiint(v)};
        __builtin_amdgcn_raw_buffer_store_b64(b, r, off, 0, 0);
    }
}

// MODE 0 (PLAIN):  ss = (s0*a3 - p)*bnd with s0 including wrk1; pn = p + omega*ss
// MODE 1 (rebase sweep of the defect correction, FP64): p is the base B, and the
//                  staged plane is P = B + q (q = Coef::q_in, FP32); ss as in MODE 0
//                  from P; writes B' = P into pn, r = (float)ss into slot 11 of the
//                  blocked array Coef::inter and q' = (float)(omega*ss) into Coef::q_out
// MODE 2 (inner sweep of the defect correction, FP32): p is the correction q and the
//                  coefficients and r come from the blocked array Coef::inter;
//                  ss = r + (s0*a3 - q)*bnd with s0 excluding wrk1; qn = q + omega*ss
//
// Software pipeline along k: while plane k is computed, the p values of plane k+2 (for
// the LDS ring) and the 12 coefficient values of planes k+1 .. k+D are already in flight
// in registers (a ring of D+1 coefficient planes; the k loop is unrolled D+1 times so
// that every slot index is static).  The p ring has 4 slots, so one barrier per plane
// separates the staging of plane k+1 from its use (the slot it overwrites was last read
// two planes ago).  A thread outside the interior computes with zero coefficients, so
// its ss is 0 (it adds nothing to gosa) and its store is dropped.
template <typename T, int MODE, int D, int RY>
__global__ __launch_bounds__(TX) void k_himeno(const T* __restrict__ p, T* __restrict__ pn, Coef cf, int nx, int ny,
                                               int nz, int zc, T omega, double* partials, unsigned* ticket,
                                               double* gosa_out)
{
    static_assert(MODE == 0 || (MODE == 1 && sizeof(T) == 8) || (MODE == 2 && sizeof(T) == 4),
                  "PLAIN in F32 / F64, rebase in F64, inner sweep in F32");
    constexpr int LH = RY + 2;                    // staged rows: the block's rows and their y halo
    constexpr int NR = 4;                         // p ring slots
    constexpr int PPT = (LW * LH + TX - 1) / TX;  // staged p values per thread
    constexpr int NC = 12;                        // coefficient streams per point
    constexpr int NS = D + 1;                     // coefficient ring slots
    __shared__ T ring[NR][LH][LW];
    __shared__ double red[TX / 64];
    const int tx = threadIdx.x;
    int bx, by, bz;
    tile_of(nx, ny, RY, bx, by, bz);
    const int i = bx * TX + tx + 1;  // interior x of this thread
    const int j0 = by * RY + 1;      // first interior row of the block
    const int k0 = bz * zc + 1;
    const int k1 = min(k0 + zc, nz - 1);
    const int sy = nx, sz = nx * ny;  // < 2^31 elements (checked on the host)
    const uint32_t pbytes = (uint32_t)sz * (uint32_t)sizeof(T);  // < 2^31 (checked on the host)
    const bool xin = i < nx - 1;
    const T* arr[NC] = {(const T*)cf.a0, (const T*)cf.a1, (const T*)cf.a2, (const T*)cf.a3, (const T*)cf.b0,
                        (const T*)cf.b1, (const T*)cf.b2, (const T*)cf.c0, (const T*)cf.c1, (const T*)cf.c2,
                        (const T*)cf.bnd, (const T*)cf.wrk1};
    // staged cells of this thread: in-plane byte offsets (kOut outside the array)
    uint32_t soff[PPT];
    int slds[PPT];
    const int xb = bx * TX;
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        const int e = tx + q * TX;
        const int ly = e / LW, lx = e - ly * LW;
        const int gx = xb + lx, gy = j0 - 1 + ly;
        soff[q] = e < LW * LH && gx < nx && gy < ny ? (uint32_t)(gx + sy * gy) * (uint32_t)sizeof(T) : kOut;
        slds[q] = e < LW * LH ? e : -1;
    }
    bool rin[RY];
    int own[RY];  // in-plane element index of the thread's points
#pragma unroll
    for (int r = 0; r < RY; r++) {
        rin[r] = xin && j0 + r < ny - 1;
        own[r] = i + sy * (j0 + r);
    }
    T* ringf = &ring[0][0][0];
    auto stage_load = [&](int k, T(&v)[PPT]) {
        const auto rs = rsrc(p + (long long)sz * k, pbytes);
#pragma unroll
        for (int q = 0; q < PPT; q++) v[q] = ld<T>(rs, soff[q]);
        if constexpr (MODE == 1) {  // P = B + q, q in FP32 (half the byte offset)
            const auto rq = rsrc(cf.q_in + (long long)sz * k, pbytes / 2);
#pragma unroll
            for (int q = 0; q < PPT; q++)
                v[q] = __dadd_rn(v[q], (double)ld<float>(rq, soff[q] == kOut ? kOut : soff[q] / 2));
        }
    };
    auto stage_store = [&](int slot, const T(&v)[PPT]) {
#pragma unroll
        for (int q = 0; q < PPT; q++)
            if (slds[q] >= 0) ringf[slot * (LH * LW) + slds[q]] = v[q];
    };
    auto coef_load = [&](int k, T(&c)[RY][NC]) {
        if constexpr (MODE == 2) {
            // the blocks from the one holding the plane's first point: value a of point
            // q at ((q >> 6) - (b >> 6)) * BLK + a * 64 + (q & 63), b = the first point
            const int b = sz * k;
            const auto rs = rsrc((const float*)cf.inter + (size_t)(b >> 6) * BLK,
                                 (uint32_t)((sz >> 6) + 2) * (uint32_t)(BLK * 4));
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int q = own[r] + b;
                const uint32_t o = rin[r] ? (uint32_t)(((q >> 6) - (b >> 6)) * BLK + (q & 63)) * 4u : kOut;
#pragma unroll
                for (int a = 0; a < NC; a++) c[r][a] = ld<T>(rs, o + (uint32_t)(a * 256));
            }
        } else {
#pragma unroll
            for (int a = 0; a < NC; a++) {
                const auto rs = rsrc(arr[a] + (long long)sz * k, pbytes);
#pragma unroll
                for (int r = 0; r < RY; r++)
                    c[r][a] = ld<T>(rs, rin[r] ? (uint32_t)own[r] * (uint32_t)sizeof(T) : kOut);
            }
        }
    };
    {
        T v[PPT];
        stage_load(k0 - 1, v);
        stage_store((k0 - 1) % NR, v);
        stage_load(k0, v);
        stage_store(k0 % NR, v);
    }
    T pnext[PPT];
    stage_load(k0 + 1, pnext);  // k0 + 1 <= nz - 1: always a plane of the array
    // coefficient ring: slot s holds plane k0 + s (mod NS); planes k0 .. k0+D-1 now
    T cr[NS][RY][NC];
#pragma unroll
    for (int s = 0; s < D; s++) coef_load(min(k0 + s, nz - 2), cr[s]);
    double g = 0.0;
    for (int kb = k0; kb < k1; kb += NS) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = kb + s;
            if (k >= k1) break;
            stage_store((k + 1) % NR, pnext);
            stage_load(min(k + 2, nz - 1), pnext);
            // plane k+D into the slot plane k-1 used (past the chunk: a plane of the
            // array whose values go unused)
            coef_load(min(k + D, nz - 2), cr[(s + D) % NS]);
            __syncthreads();
            const int sm = (k - 1) % NR, s0s = k % NR, sp = (k + 1) % NR;
            const auto rw = rsrc(pn + (long long)sz * k, pbytes);
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int ly = r + 1, lx = tx + 1;
                const T* c = cr[s][r];
                const T pc = ring[s0s][ly][lx];
                T s0 = c[0] * ring[s0s][ly][lx + 1] + c[1] * ring[s0s][ly + 1][lx] + c[2] * ring[sp][ly][lx]
                       + c[4] * (ring[s0s][ly + 1][lx + 1] - ring[s0s][ly - 1][lx + 1] - ring[s0s][ly + 1][lx - 1]
                                 + ring[s0s][ly - 1][lx - 1])
                       + c[5] * (ring[sp][ly + 1][lx] - ring[sp][ly - 1][lx] - ring[sm][ly + 1][lx]
                                 + ring[sm][ly - 1][lx])
                       + c[6] * (ring[sp][ly][lx + 1] - ring[sp][ly][lx - 1] - ring[sm][ly][lx + 1]
                                 + ring[sm][ly][lx - 1])
                       + c[7] * ring[s0s][ly][lx - 1] + c[8] * ring[s0s][ly - 1][lx] + c[9] * ring[sm][ly][lx];
                T ss;
                if constexpr (MODE == 2) {
                    ss = c[11] + (s0 * c[3] - pc) * c[10];
                } else {
                    s0 = s0 + c[11];
                    ss = (s0 * c[3] - pc) * c[10];
                }
                const uint32_t so = rin[r] ? (uint32_t)own[r] * (uint32_t)sizeof(T) : kOut;
                if constexpr (MODE == 1) {
                    st<T>(rw, so, pc);
                    st<float>(rsrc(cf.q_out + (long long)sz * k, pbytes / 2), so == kOut ? kOut : so / 2,
                              (float)(omega * ss));
                    const int b = sz * k, q = own[r] + b;
                    st<float>(rsrc((const float*)cf.inter + (size_t)(b >> 6) * BLK,
                                   (uint32_t)((sz >> 6) + 2) * (uint32_t)(BLK * 4)),
                              rin[r] ? (uint32_t)(((q >> 6) - (b >> 6)) * BLK + 11 * 64 + (q & 63)) * 4u : kOut,
                              (float)ss);
                } else {
                    st<T>(rw, so, pc + omega * ss);
                }
                g += (double)ss * (double)ss;
            }
        }
    }
    block_sum_last(g, partials, ticket, gosa_out, red);
}

// The FP32 blocked coefficients of the inner sweep (see Coef::inter) from the FP64
// arrays a0..a3, b0..b2, c0..c2, bnd of cf (slot 11, r, is written by the rebase sweep);
// out holds ceil(n / 64) * BLK floats
__global__ void k_block_coefs(Coef cf, float* __restrict__ out, size_t n)
{
    const double* a[11] = {(const double*)cf.a0, (const double*)cf.a1, (const double*)cf.a2, (const double*)cf.a3,
                           (const double*)cf.b0, (const double*)cf.b1, (const double*)cf.b2, (const double*)cf.c0,
                           (const double*)cf.c1, (const double*)cf.c2, (const double*)cf.bnd};
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) {
        float* o = out + blk_idx(t, 0);
#pragma unroll
        for (int c = 0; c < 11; c++) o[c * 64] = (float)a[c][t];
    }
}

// P = B + q on the interior (final field of the defect-correction form)
__global__ void k_add_correction(const double* __restrict__ B, const float* __restrict__ q, double* __restrict__ out,
                                 size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) out[t] = __dadd_rn(B[t], (double)q[t]);
}

// 1-D launch over (x tiles) * (y tiles) * (z chunks of zc planes)
dim3 grid_of(int nx, int ny, int nz, int zc, int ry)
{
    return dim3(ceil_div(nx - 2, TX) * ceil_div(ny - 2, ry) * ceil_div(nz - 2, zc));
}

// candidate sweep shapes: planes per block (zc), coefficient prefetch depth (D) and
// rows per thread (RY).  Long chunks win on the 257x257x513 config (zc256: one 4-wave
// block per CU, 0.371 ms per FP32 sweep against 0.404 at zc128 and 0.46 at zc64, at
// any D; one 511-plane chunk of 1-row tiles is 1.5% faster again for FP64, 8% slower
// for FP32); short ones are there for shallow grids, where long chunks leave CUs
// without a block.
struct SweepShape {
    int zc, d, ry;
};
constexpr SweepShape kShapes[] = {{8, 1, 2},   {16, 1, 2},  {32, 1, 2},  {64, 1, 2},  {128, 1, 2},
                                  {128, 2, 2}, {256, 1, 2}, {256, 2, 2}, {256, 3, 2}, {256, 4, 2},
                                  {512, 1, 2}, {512, 2, 2}, {512, 2, 1}};
constexpr int NSHAPE = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
const char* const kShapeNames[NSHAPE] = {"zc8 d1",      "zc16 d1",     "zc32 d1",     "zc64 d1",     "zc128 d1",
                                         "zc128 d2",    "zc256 d1",    "zc256 d2",    "zc256 d3",    "zc256 d4",
                                         "zc512 d1",    "zc512 d2",    "zc512 d2 ry1"};
constexpr int kMinZc = 8, kMinRy = 1;

template <typename T, int MODE, int D, int RY>
void launch_one(appfis_ctx* ctx, SweepShape sh, const T* pin, T* pout, const Coef& cf, int nx, int ny, int nz,
                T omega, double* partials, unsigned* ticket, double* gosa)
{
    hipLaunchKernelGGL((k_himeno<T, MODE, D, RY>), grid_of(nx, ny, nz, sh.zc, RY), dim3(TX), 0, ctx->stream, pin,
                       pout, cf, nx, ny, nz, sh.zc, omega, partials, ticket, gosa);
}

template <typename T, int MODE>
void launch_sweep(appfis_ctx* ctx, SweepShape sh, const T* pin, T* pout, const Coef& cf, int nx, int ny, int nz,
                  T omega, double* partials, unsigned* ticket, double* gosa)
{
    auto go = [&](auto fn) { fn(ctx, sh, pin, pout, cf, nx, ny, nz, omega, partials, ticket, gosa); };
    switch (sh.ry * 10 + sh.d) {
    case 21: go(launch_one<T, MODE, 1, 2>); break;
    case 22: go(launch_one<T, MODE, 2, 2>); break;
    case 23: go(launch_one<T, MODE, 3, 2>); break;
    case 24: go(launch_one<T, MODE, 4, 2>); break;
    default: go(launch_one<T, MODE, 2, 1>); break;
    }
}

}  // namespace

extern "C" int appfis_himeno_candidates(char* names, int len)
{
    std::string all;
    for (int f = 0; f < 2; f++)
        for (int i = 0; i < NSHAPE; i++) all += std::string(f ? "dc " : "plain ") + kShapeNames[i] + "\n";
    if (names && len > 0) snprintf(names, (size_t)len, "%s", all.c_str());
    return 2 * NSHAPE;
}

extern "C" int appfis_himeno(appfis_ctx* ctx, int dtype, void* const* arrays, int nx, int ny, int nz, double omega,
                             int iters, int variant, int outer, double* gosa_host)
{
    APPFIS_REQUIRE(ctx && arrays, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "appfis_himeno (it returns gosa to the host)");
    for (int a = 0; a < 14; a++) APPFIS_REQUIRE(arrays[a], "arrays[%d] is NULL", a);
    APPFIS_REQUIRE(arrays[0] != arrays[1], "p0 and p1 must differ");
    APPFIS_REQUIRE(nx >= 3 && ny >= 3 && nz >= 3, "himeno needs at least 3 points per axis (shell included)");
    APPFIS_REQUIRE((double)nx * ny * nz < 2147483647.0, "grid exceeds the int index range");
    APPFIS_REQUIRE(iters >= 0, "iters must be >= 0");
    APPFIS_REQUIRE((double)nx * ny * (dtype == APPFIS_F32 ? 4.0 : 8.0) < 2147483648.0,
                   "a plane of the grid exceeds 2 GiB (32-bit buffer offsets)");
    if (iters == 0) return APPFIS_OK;
    const size_t nblk = grid_of(nx, ny, nz, kMinZc, kMinRy).x;  // at least the blocks any candidate launches
    // scratch: ticket | partials[nblk] | gosa[iters]
    const size_t off_part = 256, off_gosa = off_part + ((nblk * 8 + 255) / 256) * 256;
    int st = ensure_scratch(ctx, off_gosa + (size_t)iters * 8 + 256);
    if (st) return st;
    unsigned* ticket = (unsigned*)ctx->scratch;
    double* partials = (double*)((char*)ctx->scratch + off_part);
    double* gosa_dev = (double*)((char*)ctx->scratch + off_gosa);
    Coef cf{arrays[2], arrays[3], arrays[4], arrays[5], arrays[6], arrays[7], arrays[8],
            arrays[9], arrays[10], arrays[11], arrays[12], arrays[13]};

    // APPFIS_CANDIDATE0 + i: the PLAIN sweep with shape i forced (i < NSHAPE), or the
    // defect correction with inner-sweep shape i - NSHAPE forced
    int forced = -1;
    if (variant >= APPFIS_CANDIDATE0) {
        forced = variant - APPFIS_CANDIDATE0;
        APPFIS_REQUIRE(forced < 2 * NSHAPE, "himeno: candidate %d out of range (%d candidates)", forced, 2 * NSHAPE);
        variant = forced < NSHAPE ? APPFIS_PLAIN : APPFIS_CHOSEN;
        forced %= NSHAPE;
    }
    // sweep shape: timed once per (dtype, form, grid shape) -- a sweep from p0 into p1
    // (PLAIN; the first real sweep rewrites p1) or between the correction buffers (CHOSEN)
    auto tune_shape = [&](int form, const std::function<int(int)>& sweep, SweepShape* out) -> int {
        if (forced >= 0) {
            *out = kShapes[forced];
            return APPFIS_OK;
        }
        const int key[6] = {5, dtype, nx, ny, nz, form};
        int pick = 4;
        const int tst = tune_select(ctx, key, NSHAPE, kShapeNames, sweep, 2, &pick);
        *out = kShapes[pick];
        return tst;
    };
    if (variant == APPFIS_PLAIN) {
        if (dtype != APPFIS_F32 && dtype != APPFIS_F64) return invalid("himeno PLAIN: dtype must be F32 or F64");
        auto sweep = [&](SweepShape sh, int it) -> int {
            void* pin = arrays[it & 1];
            void* pout = arrays[(it + 1) & 1];
            if (dtype == APPFIS_F32)
                launch_sweep<float, 0>(ctx, sh, (const float*)pin, (float*)pout, cf, nx, ny, nz, (float)omega,
                                       partials, ticket, gosa_dev + it);
            else
                launch_sweep<double, 0>(ctx, sh, (const double*)pin, (double*)pout, cf, nx, ny, nz, omega, partials,
                                        ticket, gosa_dev + it);
            APPFIS_LAUNCH_CHECK();
            return APPFIS_OK;
        };
        SweepShape sh;
        st = tune_shape(0, [&](int c) { return sweep(kShapes[c], 0); }, &sh);
        if (st) return st;
        for (int it = 0; it < iters; it++) {
            st = sweep(sh, it);
            if (st) return st;
        }
    } else if (variant == APPFIS_CHOSEN) {
        APPFIS_REQUIRE(dtype == APPFIS_F64, "himeno CHOSEN (defect correction) takes the F64 problem");
        APPFIS_REQUIRE(((double)nx * ny / 64.0 + 2.0) * BLK * 4.0 < 2147483648.0,
                       "himeno CHOSEN: a plane of the blocked coefficients exceeds 2 GiB");
        // default rebase interval: 16 sweeps (gosa 1.3e-6 from FP64 at the 257x257x513
        // config, against 7.5e-7 at 8 and 2.6e-6 at 24; 1.08x the speed of 8)
        const int K = outer > 0 ? outer : 16;
        const size_t n = (size_t)nx * ny * nz;
        // workspace (kept by the context across calls): the blocked FP32
        // coefficients and r (48 bytes per point), q x2 (FP32) and B x2 (FP64), each
        // 256-byte aligned
        const size_t n4 = ((n * 4 + 255) / 256) * 256, n8 = ((n * 8 + 255) / 256) * 256;
        const size_t n48 = (((n + 63) / 64) * BLK * 4 + 255) / 256 * 256;  // blocked coefficients and r
        st = ensure_work(ctx, n48 + 2 * n4 + 2 * n8);
        if (st) return st;
        char* w = (char*)ctx->work;
        float* ci = (float*)w;
        float* q[2] = {(float*)(w + n48), (float*)(w + n48 + n4)};
        double* B[2] = {(double*)(w + n48 + 2 * n4), (double*)(w + n48 + 2 * n4 + n8)};
        const unsigned cb = (unsigned)std::min<size_t>(ceil_div(n, 256), 65536);
        hipLaunchKernelGGL(k_block_coefs, dim3(cb), dim3(256), 0, ctx->stream, cf, ci, n);
        APPFIS_LAUNCH_CHECK();
        // start: B = p (whole array, shell included), q = 0
        APPFIS_HIP_CHECK(hipMemcpyAsync(B[0], arrays[0], n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemcpyAsync(B[1], arrays[0], n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemsetAsync(q[0], 0, n * 4, ctx->stream));
        APPFIS_HIP_CHECK(hipMemsetAsync(q[1], 0, n * 4, ctx->stream));
        Coef c32s{};
        c32s.inter = ci;
        // sweep shape of the inner and rebase sweeps, timed as one rebase (B[0], q[0] ->
        // B[1], r, q[1]) and one inner sweep (q[0] -> q[1]): the first real sweep is a
        // rebase from B[0] and q[0] = 0, which the timing leaves as they are, and it
        // rewrites everything the timing writes
        SweepShape sh;
        Coef cr0 = cf;
        cr0.inter = ci;
        cr0.q_in = q[0];
        cr0.q_out = q[1];
        st = tune_shape(1, [&](int c) {
            launch_sweep<double, 1>(ctx, kShapes[c], B[0], B[1], cr0, nx, ny, nz, omega, partials, ticket, gosa_dev);
            APPFIS_LAUNCH_CHECK();
            launch_sweep<float, 2>(ctx, kShapes[c], q[0], q[1], c32s, nx, ny, nz, (float)omega, partials, ticket,
                                   gosa_dev);
            APPFIS_LAUNCH_CHECK();
            return APPFIS_OK;
        }, &sh);
        if (st) return st;
        int bi = 0, qi = 0;
        for (int it = 0; it < iters; it++) {
            if (it % K == 0) {
                Coef cr = cf;
                cr.inter = ci;
                cr.q_in = q[qi];
                cr.q_out = q[qi ^ 1];
                launch_sweep<double, 1>(ctx, sh, B[bi], B[bi ^ 1], cr, nx, ny, nz, omega, partials, ticket,
                                        gosa_dev + it);
                bi ^= 1;
            } else {
                launch_sweep<float, 2>(ctx, sh, q[qi], q[qi ^ 1], c32s, nx, ny, nz, (float)omega, partials, ticket,
                                       gosa_dev + it);
            }
            APPFIS_LAUNCH_CHECK();
            qi ^= 1;
        }
        // final field P = B + q (every element; the shell has q = 0) into p[iters % 2]
        hipLaunchKernelGGL(k_add_correction, dim3(cb), dim3(256), 0, ctx->stream, B[bi], q[qi],
                           (double*)arrays[iters & 1], n);
        APPFIS_LAUNCH_CHECK();
    } else {
        return invalid("himeno: unknown variant %d", variant);
    }
    if (gosa_host) {
        APPFIS_HIP_CHECK(hipMemcpyAsync(gosa_host, gosa_dev, (size_t)iters * 8, hipMemcpyDeviceToHost, ctx->stream));
        APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return APPFIS_OK;
}
