// FDTD3d radius-4 25-point star (config form of the reference functor
// FDTD3d/src/fdtd3d_appfis.hpp:129-143, driven like fdtd3d_appfis.hpp:79-84 on a DUAL
// grid with a 4-cell ghost held at zero):
//
//   s_r = ((u[x-r] + u[x+r]) + (u[y-r] + u[y+r])) + (u[z-r] + u[z+r])
//   u'  = (((c0*u + c1*s_1) + c2*s_2) + c3*s_3) + c4*s_4        (no FMA contraction)
//
// 2.5D k-march.  A block of 64 x TY threads owns a (64*VX) x (TY*RY) column of the x-y
// plane and a chunk of ZC planes.  Each thread owns VX x RY points of the plane and keeps
// a 9-deep register queue along z for each of them, so the z neighbours cost nothing;
// the y neighbours inside the thread's own RY rows come from the same registers; only
// the x neighbours and the y neighbours across thread rows are read from the plane
// staged in LDS.  The LDS plane is double-buffered and the next plane's halo and queue
// entries are prefetched into registers before the barrier, so each z step has one
// barrier and its global-load latency overlaps the previous step's arithmetic.  Every
// global load and store is a buffer instruction at a byte offset that each thread fixes
// once for the march (out-of-array slots read 0 and non-interior stores are dropped by
// the buffer range check), so the march spends no VALU work on addresses or validity.
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

// Buffer access to one z-plane of an array: the resource spans the plane's bytes, and an
// offset of 2^31 -- past any plane, since the launch checks keep the whole array under
// 2^31 elements and so a plane under 2^31 / 9 elements of at most 8 bytes -- reads 0 and
// drops a store (the hardware range check of buffer instructions).  The kernel keeps
// one fixed in-plane byte offset per access slot for the whole march, 2^31 where the
// slot is outside the array or not stored, so the accesses need no validity selects, no
// branches and no 64-bit address arithmetic.
constexpr uint32_t kOut = 0x80000000u;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename S>
__device__ inline __amdgpu_buffer_rsrc_t plane_rsrc(const S* base, int plane_bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<S*>(base), (short)0, plane_bytes, 0x00020000);
}

__device__ inline double u2d(unsigned lo, unsigned hi) { return __hiloint2double((int)hi, (int)lo); }

// one element at byte offset `off`, widened to the arithmetic type
template <typename S>
__device__ inline typename Cvt<S>::C bload(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (sizeof(S) == 2) {
        const uint16_t b = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
        S x;
        __builtin_memcpy(&x, &b, 2);
        return Cvt<S>::ld(x);
    } else if constexpr (sizeof(S) == 4) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    } else {
        const u32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        return u2d(b[0], b[1]);
    }
}

constexpr int R = 4;     // stencil radius = ghost width
constexpr int TXT = 64;  // threads along x (one wave per thread row)

// VX consecutive storage elements of a point group, moved by one buffer access (a pair
// sits at an even element offset, so its address is 4-byte aligned once the buffer is:
// the launch checks ask 4-byte aligned buffers of the 16-bit pair tilings).
template <typename S, int VX>
struct Vec {
    using C = typename Cvt<S>::C;
    __device__ static void load(__amdgpu_buffer_rsrc_t r, uint32_t off, C (&out)[VX])
    {
        if constexpr (VX == 1) {
            out[0] = bload<S>(r, off);
        } else if constexpr (sizeof(S) == 2) {
            const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
            S lo, hi;
            const uint16_t l16 = (uint16_t)(w & 0xFFFFu), h16 = (uint16_t)(w >> 16);
            __builtin_memcpy(&lo, &l16, 2);
            __builtin_memcpy(&hi, &h16, 2);
            out[0] = Cvt<S>::ld(lo);
            out[1] = Cvt<S>::ld(hi);
        } else if constexpr (sizeof(S) == 4) {
            const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
            out[0] = __uint_as_float(w[0]);
            out[1] = __uint_as_float(w[1]);
        } else {
            const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
            out[0] = u2d(w[0], w[1]);
            out[1] = u2d(w[2], w[3]);
        }
    }
    __device__ static void store(__amdgpu_buffer_rsrc_t r, uint32_t off, const C (&in)[VX])
    {
        if constexpr (sizeof(S) == 2) {
            uint16_t b[VX];
#pragma unroll
            for (int e = 0; e < VX; e++) {
                const S x = Cvt<S>::st(in[e]);
                __builtin_memcpy(&b[e], &x, 2);
            }
            if constexpr (VX == 1) __builtin_amdgcn_raw_buffer_store_b16(b[0], r, off, 0, 0);
            else __builtin_amdgcn_raw_buffer_store_b32((uint32_t)b[0] | ((uint32_t)b[1] << 16), r, off, 0, 0);
        } else if constexpr (sizeof(S) == 4) {
            if constexpr (VX == 1) {
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(in[0]), r, off, 0, 0);
            } else {
                u32x2 w;
                w[0] = __float_as_uint(in[0]);
                w[1] = __float_as_uint(in[1]);
                __builtin_amdgcn_raw_buffer_store_b64(w, r, off, 0, 0);
            }
        } else {
            if constexpr (VX == 1) {
                u32x2 w;
                w[0] = (unsigned)__double2loint(in[0]);
                w[1] = (unsigned)__double2hiint(in[0]);
                __builtin_amdgcn_raw_buffer_store_b64(w, r, off, 0, 0);
            } else {
                u32x4 w;
                w[0] = (unsigned)__double2loint(in[0]);
                w[1] = (unsigned)__double2hiint(in[0]);
                w[2] = (unsigned)__double2loint(in[1]);
                w[3] = (unsigned)__double2hiint(in[1]);
                __builtin_amdgcn_raw_buffer_store_b128(w, r, off, 0, 0);
            }
        }
    }
};

// Tile coordinates of this block: every XCD takes a contiguous run of tiles (x fastest),
// so tiles that share halos share an L2.
__device__ inline void tile_of(int& bx, int& by, int& bz)
{
    const int gx = gridDim.x, gy = gridDim.y;
    const unsigned nb = gx * gy * gridDim.z;
    const int bid = (int)xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), nb);
    bx = bid % gx;
    const int r = bid / gx;
    by = r % gy;
    bz = r / gy;
}

template <typename S, int VX, int TY, int RY, int ZC, int D, int U>
__global__ __launch_bounds__(TXT* TY) void k_fdtd3d(const S* __restrict__ u, S* __restrict__ v, int nx, int ny, int nz,
                                                     typename Cvt<S>::C c0, typename Cvt<S>::C c1,
                                                     typename Cvt<S>::C c2, typename Cvt<S>::C c3,
                                                     typename Cvt<S>::C c4)
{
    using C = typename Cvt<S>::C;
    constexpr int NT = TXT * TY;
    constexpr int PX = TXT * VX;       // tile width
    constexpr int PY = TY * RY;        // tile height
    constexpr int LX = PX + 2 * R;     // LDS plane width
    constexpr int LY = PY + 2 * R;     // LDS plane height
    constexpr int NHALO = 2 * R * PX + 2 * R * PY;
    constexpr int HPT = (NHALO + NT - 1) / NT;  // halo cells per thread
    // two plane buffers, each followed by 64 dummy cells: a thread without a halo cell
    // in a slot writes the dummy of its lane (distinct lanes, so no write conflict)
    constexpr int PLANE = LY * LX + 64;
    __shared__ C plane[2][PLANE];

    int tbx, tby, tbz;
    tile_of(tbx, tby, tbz);
    const int W = nx + 2 * R, H = ny + 2 * R;
    const long long szl = (long long)W * H;
    const int plane_bytes = (int)(szl * (long long)sizeof(S));
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * TXT + tx;
    const int bx = tbx * PX + R, by = tby * PY + R;  // padded origin of the tile
    const int z0 = tbz * ZC + R;
    const int z1 = min(z0 + ZC, nz + R);
    const C cc[5] = {c0, c1, c2, c3, c4};

    // own point group: x = gx .. gx+VX-1, rows y = gy0 + r
    const int gx = bx + tx * VX;
    const bool xload = gx < W, xstore = gx < nx + R;
    const int gy0 = by + ty * RY;
    const int own0 = gx + W * gy0;

    // halo slots of this thread: LDS position and in-plane byte offset (fixed for the
    // march; kOut outside the array, and for slots past the halo, whose load then reads
    // 0 into a dummy LDS cell)
    uint32_t hb[HPT];
    int hlds[HPT];
#pragma unroll
    for (int s = 0; s < HPT; s++) {
        const int p = tid + s * NT;
        int lx = 0, ly = 0;
        const bool huse = p < NHALO;
        if (p < 2 * R * PX) {
            const int hy = p / PX;
            lx = p - hy * PX + R;
            ly = hy < R ? hy : PY + hy;
        } else if (huse) {
            // side halos row by row: 8 consecutive slots are the R cells left and the R
            // cells right of one row, so a wave touches few cache lines
            const int p2 = p - 2 * R * PX;
            const int hy = p2 / (2 * R), hx = p2 - hy * (2 * R);
            ly = hy + R;
            lx = hx < R ? hx : PX + hx;
        }
        const int xg = bx - R + lx, yg = by - R + ly;
        hb[s] = huse && xg < W && yg < H ? (uint32_t)(xg + W * yg) * (uint32_t)sizeof(S) : kOut;
        hlds[s] = huse ? ly * LX + lx : LY * LX + tx;
    }
    // own rows: byte offsets of the loads (rows and columns of the array; outside them
    // the queue takes 0) and of the stores (interior only)
    uint32_t qb[RY], sb[RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const uint32_t off = (uint32_t)(own0 + r * W) * (uint32_t)sizeof(S);
        qb[r] = xload && gy0 + r < H ? off : kOut;
        sb[r] = xstore && gy0 + r < ny + R ? off : kOut;
    }

    // register queues.  U (the unroll of the march) a multiple of 9: plane p sits in slot
    // (p - zb + 4) mod 9 while the march is in planes zb .. zb+U-1, every slot index is a
    // compile-time constant and the queue rotates without register moves.  Otherwise
    // slot m holds plane z + m - 4 and the queue shifts by one slot per plane (moves, but
    // fewer live registers).
    constexpr int QN = 2 * R + 1;
    constexpr int UNROLL = U;
    constexpr bool ROTATE = U % QN == 0;
    static_assert(U % D == 0 && U % 2 == 0, "the unroll must be a multiple of the prefetch depth and of 2");
    C q[RY][VX][QN];
#pragma unroll
    for (int r = 0; r < RY; r++) {
#pragma unroll
        for (int m = 0; m < 2 * R; m++) {
            C t[VX];
            Vec<S, VX>::load(plane_rsrc(u + (z0 - R + m) * szl, plane_bytes), qb[r], t);
#pragma unroll
            for (int e = 0; e < VX; e++) q[r][e][ROTATE ? m : m + 1] = t[e];  // shift: before its first shift
        }
    }
    // prefetch ring: slot d holds the queue entry (plane z+R) and halo (plane z) of the
    // steps z = z0 + d (mod D)
    C qp[D][RY][VX];
    C hp[D][HPT];
    auto prefetch = [&](int z, C(&qd)[RY][VX], C(&hd)[HPT]) {
        const auto rq = plane_rsrc(u + (long long)(z + R) * szl, plane_bytes);
        const auto rh = plane_rsrc(u + (long long)z * szl, plane_bytes);
#pragma unroll
        for (int r = 0; r < RY; r++) Vec<S, VX>::load(rq, qb[r], qd[r]);
#pragma unroll
        for (int s = 0; s < HPT; s++) hd[s] = bload<S>(rh, hb[s]);
    };
#pragma unroll
    for (int d = 0; d < D; d++)
        if (z0 + d < z1) prefetch(z0 + d, qp[d], hp[d]);

    C* pl = &plane[0][0];
    for (int zb = z0; zb < z1; zb += UNROLL) {
#pragma unroll
        for (int st = 0; st < UNROLL; st++) {
            const int z = zb + st;
            if (z >= z1) break;
            const int d = st % D;
            C* buf = pl + (st & 1) * PLANE;
            // slots of planes z-4 .. z+4 at this step, m = 0 .. 8
#define QS(m) (ROTATE ? (st + (m)) % QN : (m))
            if constexpr (!ROTATE) {
#pragma unroll
                for (int r = 0; r < RY; r++)
#pragma unroll
                    for (int e = 0; e < VX; e++)
#pragma unroll
                        for (int m = 0; m < 2 * R; m++) q[r][e][m] = q[r][e][m + 1];
            }
            // incoming plane z + R
#pragma unroll
            for (int r = 0; r < RY; r++)
#pragma unroll
                for (int e = 0; e < VX; e++) q[r][e][QS(2 * R)] = qp[d][r][e];
            // stage plane z: own centres + prefetched halo
#pragma unroll
            for (int r = 0; r < RY; r++)
#pragma unroll
                for (int e = 0; e < VX; e++) buf[(ty * RY + r + R) * LX + tx * VX + e + R] = q[r][e][QS(R)];
#pragma unroll
            for (int s = 0; s < HPT; s++) buf[hlds[s]] = hp[d][s];  // unused slots hit a dummy cell
            // refill this slot with step z + D
            if (z + D < z1) prefetch(z + D, qp[d], hp[d]);
            __syncthreads();
            const auto rv = plane_rsrc(v + (long long)z * szl, plane_bytes);
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int ly = ty * RY + r + R;
                // x-window of this row, positions -R .. VX+R-1 around the lane's points: the
                // lane's own cells from its registers, the 2R outside cells from LDS
                C xw[VX + 2 * R];
#pragma unroll
                for (int j = 0; j < R; j++) {
                    xw[j] = buf[ly * LX + tx * VX + j];
                    xw[R + VX + j] = buf[ly * LX + tx * VX + R + VX + j];
                }
#pragma unroll
                for (int e = 0; e < VX; e++) xw[R + e] = q[r][e][QS(R)];
                C out[VX];
#pragma unroll
                for (int e = 0; e < VX; e++) {
                    const int lx = tx * VX + e + R;
                    C acc = cc[0] * q[r][e][QS(R)];
#pragma unroll
                    for (int k = 1; k <= R; k++) {
                        const C sx = xw[R + e - k] + xw[R + e + k];
                        const C ym = (r - k >= 0) ? q[r - k][e][QS(R)] : buf[(ly - k) * LX + lx];
                        const C yp = (r + k < RY) ? q[r + k][e][QS(R)] : buf[(ly + k) * LX + lx];
                        const C sY = ym + yp;
                        const C sZ = q[r][e][QS(R - k)] + q[r][e][QS(R + k)];
                        const C sr = (sx + sY) + sZ;
                        acc = acc + cc[k] * sr;
                    }
                    out[e] = acc;
                }
                Vec<S, VX>::store(rv, sb[r], out);
            }
#undef QS
        }
    }
}

// The same sweep with each plane tile loaded straight into LDS (FP32 storage, one x
// point per thread; the FP64 form of it ran 13% slower than FP64's two points per lane).
// The block's waves copy the whole (64+8) x (TY*RY+8) tile of a plane, halo included,
// with 16-byte buffer-to-LDS loads (`buffer_load_dwordx4 ... lds`: lane l of a load
// writes bytes 16l .. 16l+15 of its 1 KB block, and the tile's rows are contiguous in
// LDS, so chunk c of the tile is LDS bytes 16c .. 16c+15), A = 4 + D planes ahead of the
// plane being computed, into a ring of NB = A + 2 plane buffers.  No registers hold the
// tile on its way, and there is no LDS store.
//
// At step z a wave issues the loads of plane z+A, waits until its own loads of plane
// z+4 have landed, then meets the others at a plain barrier (no fence, which would wait
// for every load in flight): planes z and z+4 are then complete for every wave.  The
// wait is an explicit vmcnt.  On CDNA the vector memory counter retires loads and
// stores in issue order (as the compiler's own waitcnt placement for gfx9 assumes), so
// "at most N outstanding" means all but the last N operations are done.  A step issues
// the nload loads of one plane and then RY stores; after the loads of plane z+4 came
// the stores of that step and A-5 whole steps, then this step's loads: N =
// (A-4)*(nload+RY).  In the first A-4 steps of a chunk fewer stores have been issued,
// and N = (A-4)*nload (loads only) is used, which can only wait longer.  The wave then
// takes its own z+4 values from LDS into the register z-queue and computes plane z from
// LDS (x, y neighbours) and the queue (z neighbours), in the operation order of
// k_fdtd3d.
//
// The ring index of a plane is static in a loop unrolled NB times, so the compiler sees
// which buffer each load fills.  Between two barriers the waves are at most one step
// apart, and NB = A + 2 keeps the buffer a step's loads fill apart from the two that
// the step before reads.  Chunks past the row end read the next row (the values only
// feed points that are not stored), rows past the plane read 0 (buffer range check),
// and the padding of a buffer's last 1 KB block takes the lanes past the tile.  Launch
// checks: nx + 8 a multiple of 4, 16-byte aligned buffers.
__device__ inline void load_lds16(__amdgpu_buffer_rsrc_t r, void* lds, uint32_t off)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16, off, 0, 0, 0);
}

template <typename S, int TY, int RY, int ZC, int D>
__global__ __launch_bounds__(TXT* TY) void k_fdtd3d_dma(const S* __restrict__ u, S* __restrict__ v, int nx, int ny,
                                                         int nz, S c0, S c1, S c2, S c3, S c4)
{
    static_assert(sizeof(S) == 4, "FP32 storage");
    using C = S;
    constexpr int PY = TY * RY;
    constexpr int LX = TXT + 2 * R;               // 72 elements per tile row
    constexpr int LY = PY + 2 * R;
    constexpr int EPC = 16 / (int)sizeof(S);      // elements per 16-byte chunk
    constexpr int CPR = LX / EPC;                 // chunks per tile row
    static_assert(CPR * EPC == LX, "a tile row is whole chunks");
    constexpr int NCH = LY * CPR;                 // chunks per tile
    constexpr int NI = (NCH + 63) / 64;           // 64-lane loads per tile
    constexpr int L = (NI + TY - 1) / TY;         // loads per wave and plane, at most
    constexpr int A = 4 + D;                      // planes loaded ahead
    constexpr int NB = A + 2;                     // ring of plane buffers
    constexpr int BUF = NI * 64 * EPC;            // elements per buffer (whole 1 KB blocks)
    constexpr int QN = 2 * R + 1;
    __shared__ __align__(16) C ring[NB][BUF];

    int tbx, tby, tbz;
    tile_of(tbx, tby, tbz);
    const int W = nx + 2 * R, H = ny + 2 * R;
    const long long szl = (long long)W * H;
    const int plane_bytes = (int)(szl * (long long)sizeof(S));
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int bx = tbx * TXT + R, by = tby * PY + R;  // padded origin of the tile
    const int z0 = tbz * ZC + R;
    const int z1 = min(z0 + ZC, nz + R);
    const int zlast = z1 + R - 1;  // the last plane the chunk reads
    const C cc[5] = {c0, c1, c2, c3, c4};
    const int gx = bx + tx;
    const int gy0 = by + ty * RY;

    // this wave's tile loads: load i = ty + TY*j covers chunks 64i .. 64i+63
    uint32_t lo[L];
    int nload = 0;
#pragma unroll
    for (int j = 0; j < L; j++) {
        const int i = ty + TY * j;
        const int c = 64 * i + tx;
        lo[j] = kOut;
        if (i < NI) {
            nload = j + 1;
            if (c < NCH) {
                const int row = c / CPR, cin = c - row * CPR;
                lo[j] = (uint32_t)((bx - R + cin * EPC) + W * (by - R + row)) * (uint32_t)sizeof(S);
            }
        }
    }
    auto load_plane = [&](int p, int b) {  // tile of plane p into ring[b] (b static)
        const auto rs = plane_rsrc(u + (long long)p * szl, plane_bytes);
#pragma unroll
        for (int j = 0; j < L; j++)
            if (j < nload) load_lds16(rs, &ring[b][(ty + TY * j) * 64 * EPC], lo[j]);
    };
    // own rows: loads of the initial queue and stores (as in k_fdtd3d)
    uint32_t qb[RY], sb[RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const uint32_t off = (uint32_t)(gx + W * (gy0 + r)) * (uint32_t)sizeof(S);
        qb[r] = gx < W && gy0 + r < H ? off : kOut;
        sb[r] = gx < nx + R && gy0 + r < ny + R ? off : kOut;
    }
    // z-queue: slot m holds plane z + m - 4 at step z; before the first step, planes
    // z0-4 .. z0+3 in slots 1 .. 8
    C q[RY][QN];
#pragma unroll
    for (int r = 0; r < RY; r++)
#pragma unroll
        for (int m = 0; m < 2 * R; m++) q[r][m + 1] = bload<S>(plane_rsrc(u + (z0 - R + m) * szl, plane_bytes), qb[r]);
#pragma unroll
    for (int b = 0; b < A; b++)
        if (z0 + b <= zlast) load_plane(z0 + b, b);

    for (int zb = z0; zb < z1; zb += NB) {
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int z = zb + k;
            if (z >= z1) break;
            const bool ahead = z + A <= zlast;
            if (ahead) load_plane(z + A, (k + A) % NB);
            // this wave's loads of plane z+4 (and of every earlier plane) have landed
            if (nload > 0) {
                constexpr int N1 = (A - 4) * L, N2 = (A - 4) * (L - 1);
                constexpr int E1 = (N1 & 15) | 0x70 | 0xF00 | ((N1 >> 4) & 3) << 14;
                constexpr int E2 = (N2 & 15) | 0x70 | 0xF00 | ((N2 >> 4) & 3) << 14;
                static_assert(N1 < 64, "vmcnt range");
                constexpr int M1 = (A - 4) * (L + RY), M2 = (A - 4) * (L - 1 + RY);
                constexpr int F1 = (M1 & 15) | 0x70 | 0xF00 | ((M1 >> 4) & 3) << 14;
                constexpr int F2 = (M2 & 15) | 0x70 | 0xF00 | ((M2 >> 4) & 3) << 14;
                static_assert(M1 < 64, "vmcnt range");
                if (!ahead) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
                else if (z - z0 >= A - 4) {  // steady state: loads and stores of A-4 steps
                    if (nload == L) __builtin_amdgcn_s_waitcnt(F1);
                    else __builtin_amdgcn_s_waitcnt(F2);
                } else if (nload == L) __builtin_amdgcn_s_waitcnt(E1);  // loads only: fewer
                else __builtin_amdgcn_s_waitcnt(E2);
            }
            __builtin_amdgcn_s_barrier();
            // the wait and the barrier do not count as memory accesses for the compiler:
            // this keeps the ring reads below them (it emits no instruction)
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const C* cur = ring[k];
            const C* nxt4 = ring[(k + 4) % NB];
#pragma unroll
            for (int r = 0; r < RY; r++) {
#pragma unroll
                for (int m = 0; m < 2 * R; m++) q[r][m] = q[r][m + 1];
                q[r][2 * R] = nxt4[(ty * RY + r + R) * LX + tx + R];
            }
            const auto rv = plane_rsrc(v + (long long)z * szl, plane_bytes);
#pragma unroll
            for (int r = 0; r < RY; r++) {
                const int ly = ty * RY + r + R;
                C xw[2 * R + 1];
#pragma unroll
                for (int j = 0; j < R; j++) {
                    xw[j] = cur[ly * LX + tx + j];
                    xw[R + 1 + j] = cur[ly * LX + tx + R + 1 + j];
                }
                xw[R] = q[r][R];
                C acc = cc[0] * q[r][R];
#pragma unroll
                for (int kk = 1; kk <= R; kk++) {
                    const C sx = xw[R - kk] + xw[R + kk];
                    const C ym = (r - kk >= 0) ? q[r - kk][R] : cur[(ly - kk) * LX + tx + R];
                    const C yp = (r + kk < RY) ? q[r + kk][R] : cur[(ly + kk) * LX + tx + R];
                    const C sY = ym + yp;
                    const C sZ = q[r][R - kk] + q[r][R + kk];
                    const C sr = (sx + sY) + sZ;
                    acc = acc + cc[kk] * sr;
                }
                C out[1] = {acc};
                Vec<S, 1>::store(rv, sb[r], out);
            }
        }
    }
}

template <typename S, int TY, int RY, int ZC, int D>
int run_dma(appfis_ctx* ctx, S* b0, S* b1, int nx, int ny, int nz, const double* cf, int steps)
{
    dim3 block(TXT, TY), grid(ceil_div(nx, TXT), ceil_div(ny, TY * RY), ceil_div(nz, ZC));
    S* buf[2] = {b0, b1};
    for (int s = 0; s < steps; s++) {
        hipLaunchKernelGGL((k_fdtd3d_dma<S, TY, RY, ZC, D>), grid, block, 0, ctx->stream, buf[s & 1],
                           buf[(s + 1) & 1], nx, ny, nz, (S)cf[0], (S)cf[1], (S)cf[2], (S)cf[3], (S)cf[4]);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

template <typename S, int VX, int TY, int RY, int ZC, int D, int U>
int run(appfis_ctx* ctx, S* b0, S* b1, int nx, int ny, int nz, const double* cf, int steps)
{
    using C = typename Cvt<S>::C;
    dim3 block(TXT, TY), grid(ceil_div(nx, TXT * VX), ceil_div(ny, TY * RY), ceil_div(nz, ZC));
    S* buf[2] = {b0, b1};
    for (int s = 0; s < steps; s++) {
        hipLaunchKernelGGL((k_fdtd3d<S, VX, TY, RY, ZC, D, U>), grid, block, 0, ctx->stream, buf[s & 1],
                           buf[(s + 1) & 1], nx, ny, nz, (C)cf[0], (C)cf[1], (C)cf[2], (C)cf[3], (C)cf[4]);
        APPFIS_LAUNCH_CHECK();
    }
    return APPFIS_OK;
}

typedef int (*fdtd_fn)(appfis_ctx*, void*, void*, int, int, int, const double*, int);
struct FdtdVariant {
    const char* name;
    fdtd_fn fn;
    bool pairs;    // 16-bit storage moved in 32-bit pairs: needs 4-byte aligned buffers
    int vx;        // x points per thread: needs nx % vx == 0 (a lane stores all vx or none)
    int wmod = 0;  // tile rows loaded as 16-byte chunks: nx + 8 a multiple of wmod, 16-byte
                   // aligned buffers (0: no such needs)
    bool plain = false;  // PLAIN's fixed tiling where it fits (else row 0)
};

template <typename S, int VX, int TY, int RY, int ZC, int D, int U>
int run_v(appfis_ctx* ctx, void* b0, void* b1, int nx, int ny, int nz, const double* cf, int steps)
{
    return run<S, VX, TY, RY, ZC, D, U>(ctx, (S*)b0, (S*)b1, nx, ny, nz, cf, steps);
}

template <typename S, int TY, int RY, int ZC, int D>
int run_dma_v(appfis_ctx* ctx, void* b0, void* b1, int nx, int ny, int nz, const double* cf, int steps)
{
    return run_dma<S, TY, RY, ZC, D>(ctx, (S*)b0, (S*)b1, nx, ny, nz, cf, steps);
}

// Candidate tilings (VX x-points per thread, TY thread rows, RY rows per thread, ZC
// planes per block, D-deep prefetch, U-step unroll: 18 rotates the queue).  Index 0 of
// each table is the PLAIN default: the fastest measured tiling that fits any shape;
// PLAIN takes the row marked plain instead where the shape and buffers allow it.
static const FdtdVariant kF32[] = {
    {"vx1 ty16 ry2 zc256 d4 u4", run_v<float, 1, 16, 2, 256, 4, 4>, false, 1},
    {"vx1 ty8 ry4 zc256 d2 u2", run_v<float, 1, 8, 4, 256, 2, 2>, false, 1},
    {"vx1 ty4 ry4 zc64 d2 u2", run_v<float, 1, 4, 4, 64, 2, 2>, false, 1},
    {"vx1 ty8 ry4 zc64 d2 u2", run_v<float, 1, 8, 4, 64, 2, 2>, false, 1},
    {"vx1 ty4 ry4 zc64 d2 u18", run_v<float, 1, 4, 4, 64, 2, 18>, false, 1},
    {"vx1 ty4 ry2 zc64 d2 u18", run_v<float, 1, 4, 2, 64, 2, 18>, false, 1},
    {"vx1 ty8 ry2 zc64 d2 u18", run_v<float, 1, 8, 2, 64, 2, 18>, false, 1},
    {"vx1 ty4 ry4 zc32 d2 u2", run_v<float, 1, 4, 4, 32, 2, 2>, false, 1},
    {"vx1 ty8 ry4 zc128 d2 u2", run_v<float, 1, 8, 4, 128, 2, 2>, false, 1},
    {"vx1 ty4 ry4 zc64 d3 u6", run_v<float, 1, 4, 4, 64, 3, 6>, false, 1},
    {"vx1 ty16 ry2 zc64 d2 u2", run_v<float, 1, 16, 2, 64, 2, 2>, false, 1},
    {"vx1 ty8 ry4 zc128 d3 u6", run_v<float, 1, 8, 4, 128, 3, 6>, false, 1},
    {"vx1 ty16 ry2 zc256 d2 u2", run_v<float, 1, 16, 2, 256, 2, 2>, false, 1},
    {"vx1 ty8 ry4 zc256 d4 u4", run_v<float, 1, 8, 4, 256, 4, 4>, false, 1},
    {"vx1 ty8 ry4 zc256 d6 u6", run_v<float, 1, 8, 4, 256, 6, 6>, false, 1},
    {"vx1 ty8 ry2 zc256 d4 u4", run_v<float, 1, 8, 2, 256, 4, 4>, false, 1},
    {"vx2 ty8 ry2 zc256 d4 u4", run_v<float, 2, 8, 2, 256, 4, 4>, false, 2},
    {"vx2 ty16 ry1 zc256 d4 u4", run_v<float, 2, 16, 1, 256, 4, 4>, false, 2},
    {"vx2 ty16 ry1 zc256 d6 u6", run_v<float, 2, 16, 1, 256, 6, 6>, false, 2},
    // tiles loaded straight into LDS (k_fdtd3d_dma): "dma ty.. ry.. zc.. d.."
    {"dma ty16 ry2 zc256 d2", run_dma_v<float, 16, 2, 256, 2>, false, 1, 4, true},
    {"dma ty16 ry2 zc256 d3", run_dma_v<float, 16, 2, 256, 3>, false, 1, 4},
    {"dma ty8 ry4 zc256 d2", run_dma_v<float, 8, 4, 256, 2>, false, 1, 4},
    {"dma ty16 ry2 zc128 d2", run_dma_v<float, 16, 2, 128, 2>, false, 1, 4},
};
static const FdtdVariant kF64[] = {
    {"vx1 ty8 ry2 zc256 d4 u4", run_v<double, 1, 8, 2, 256, 4, 4>, false, 1},
    {"vx1 ty8 ry2 zc64 d2 u2", run_v<double, 1, 8, 2, 64, 2, 2>, false, 1},
    {"vx1 ty4 ry2 zc64 d2 u2", run_v<double, 1, 4, 2, 64, 2, 2>, false, 1},
    {"vx1 ty4 ry4 zc64 d2 u2", run_v<double, 1, 4, 4, 64, 2, 2>, false, 1},
    {"vx1 ty4 ry2 zc128 d2 u2", run_v<double, 1, 4, 2, 128, 2, 2>, false, 1},
    {"vx1 ty4 ry2 zc64 d2 u18", run_v<double, 1, 4, 2, 64, 2, 18>, false, 1},
    {"vx1 ty8 ry2 zc64 d2 u18", run_v<double, 1, 8, 2, 64, 2, 18>, false, 1},
    {"vx1 ty8 ry2 zc256 d2 u2", run_v<double, 1, 8, 2, 256, 2, 2>, false, 1},
    {"vx1 ty4 ry4 zc256 d2 u2", run_v<double, 1, 4, 4, 256, 2, 2>, false, 1},
    {"vx1 ty8 ry2 zc128 d2 u18", run_v<double, 1, 8, 2, 128, 2, 18>, false, 1},
    {"vx1 ty4 ry4 zc256 d4 u4", run_v<double, 1, 4, 4, 256, 4, 4>, false, 1},
    {"vx2 ty8 ry2 zc256 d2 u2", run_v<double, 2, 8, 2, 256, 2, 2>, false, 2},
    {"vx2 ty16 ry1 zc256 d2 u2", run_v<double, 2, 16, 1, 256, 2, 2>, false, 2, 0, true},
    {"vx1 ty16 ry1 zc256 d4 u4", run_v<double, 1, 16, 1, 256, 4, 4>, false, 1},
};
#define FDTD16_VARIANTS(S)                                                 \
    {"vx1 ty4 ry2 zc64 d2 u18", run_v<S, 1, 4, 2, 64, 2, 18>, false, 1},      \
    {"vx1 ty4 ry4 zc64 d2 u18", run_v<S, 1, 4, 4, 64, 2, 18>, false, 1},      \
    {"vx1 ty4 ry4 zc64 d2 u2", run_v<S, 1, 4, 4, 64, 2, 2>, false, 1},        \
    {"vx1 ty8 ry4 zc64 d2 u2", run_v<S, 1, 8, 4, 64, 2, 2>, false, 1},        \
    {"vx2 ty4 ry4 zc64 d2 u18", run_v<S, 2, 4, 4, 64, 2, 18>, true, 2},       \
    {"vx2 ty4 ry2 zc64 d2 u18", run_v<S, 2, 4, 2, 64, 2, 18>, true, 2},       \
    {"vx2 ty8 ry2 zc64 d2 u18", run_v<S, 2, 8, 2, 64, 2, 18>, true, 2},       \
    {"vx2 ty4 ry4 zc64 d2 u2", run_v<S, 2, 4, 4, 64, 2, 2>, true, 2},        \
    {"vx2 ty8 ry2 zc128 d2 u18", run_v<S, 2, 8, 2, 128, 2, 18>, true, 2, 0, true}
static const FdtdVariant kF16[] = {FDTD16_VARIANTS(__half)};
static const FdtdVariant kBF16[] = {FDTD16_VARIANTS(__hip_bfloat16)};

}  // namespace

extern "C" int appfis_fdtd3d_candidates(int dtype, char* names, int len)
{
    const FdtdVariant* tab;
    int ntab;
    switch (dtype) {
    case APPFIS_F32: tab = kF32; ntab = (int)(sizeof(kF32) / sizeof(kF32[0])); break;
    case APPFIS_F64: tab = kF64; ntab = (int)(sizeof(kF64) / sizeof(kF64[0])); break;
    case APPFIS_F16: tab = kF16; ntab = (int)(sizeof(kF16) / sizeof(kF16[0])); break;
    case APPFIS_BF16: tab = kBF16; ntab = (int)(sizeof(kBF16) / sizeof(kBF16[0])); break;
    default: return -invalid("fdtd3d: dtype must be F32, F64, F16 or BF16");
    }
    std::string all;
    for (int i = 0; i < ntab; i++) all += std::string(tab[i].name) + "\n";
    if (names && len > 0) snprintf(names, (size_t)len, "%s", all.c_str());
    return ntab;
}

extern "C" int appfis_fdtd3d(appfis_ctx* ctx, int dtype, void* buf0, void* buf1, int nx, int ny, int nz,
                             const double* coeff5, int steps, int variant)
{
    APPFIS_REQUIRE(ctx && buf0 && buf1 && coeff5, "NULL argument");
    APPFIS_REQUIRE(buf0 != buf1, "the two buffers must differ");
    APPFIS_REQUIRE(nx > 0 && ny > 0 && nz > 0, "interior must be non-empty");
    APPFIS_REQUIRE((double)(nx + 8) * (ny + 8) * (nz + 8) < 2147483647.0, "grid exceeds the int index range");
    APPFIS_REQUIRE(steps >= 0, "steps must be >= 0");
    APPFIS_REQUIRE(variant >= 0, "fdtd3d: unknown variant %d", variant);
    const FdtdVariant* tab;
    int ntab;
    switch (dtype) {
    case APPFIS_F32: tab = kF32; ntab = (int)(sizeof(kF32) / sizeof(kF32[0])); break;
    case APPFIS_F64: tab = kF64; ntab = (int)(sizeof(kF64) / sizeof(kF64[0])); break;
    case APPFIS_F16: tab = kF16; ntab = (int)(sizeof(kF16) / sizeof(kF16[0])); break;
    case APPFIS_BF16: tab = kBF16; ntab = (int)(sizeof(kBF16) / sizeof(kBF16[0])); break;
    default: return invalid("fdtd3d: dtype must be F32, F64, F16 or BF16");
    }
    const bool aligned4 = (((uintptr_t)buf0 | (uintptr_t)buf1) & 3) == 0;
    const bool aligned16 = (((uintptr_t)buf0 | (uintptr_t)buf1) & 15) == 0;
    auto fits = [&](const FdtdVariant& v) {
        return nx % v.vx == 0 && (!v.pairs || aligned4) && (v.wmod == 0 || ((nx + 8) % v.wmod == 0 && aligned16));
    };
    if (variant >= APPFIS_CANDIDATE0) {
        const int i = variant - APPFIS_CANDIDATE0;
        APPFIS_REQUIRE(i < ntab, "fdtd3d: candidate %d out of range (%d candidates)", i, ntab);
        if (tab[i].wmod)
            APPFIS_REQUIRE(fits(tab[i]), "fdtd3d: candidate %s needs nx + 8 divisible by %d and 16-byte aligned buffers",
                           tab[i].name, tab[i].wmod);
        APPFIS_REQUIRE(fits(tab[i]), "fdtd3d: candidate %s needs nx %% %d == 0%s", tab[i].name, tab[i].vx,
                       tab[i].pairs ? " and 4-byte aligned buffers" : "");
        return steps == 0 ? APPFIS_OK : tab[i].fn(ctx, buf0, buf1, nx, ny, nz, coeff5, steps);
    }
    if (steps == 0) return APPFIS_OK;
    if (variant == APPFIS_PLAIN) {
        // a fixed tiling: the table's row marked plain where the shape and the buffers
        // admit it, else row 0 -- one launch per sweep either way
        int i0 = 0;
        for (int i = 0; i < ntab; i++)
            if (tab[i].plain && fits(tab[i])) {
                i0 = i;
                break;
            }
        return tab[i0].fn(ctx, buf0, buf1, nx, ny, nz, coeff5, steps);
    }
    // CHOSEN: the fastest tiling for this shape, timed once on scratch buffers.  Every
    // candidate evaluates the same expression in the same order: results are identical.
    std::vector<int> cand;
    std::vector<const char*> names;
    for (int i = 0; i < ntab; i++)
        if (fits(tab[i])) {
            cand.push_back(i);
            names.push_back(tab[i].name);
        }
    const int key[6] = {3, dtype, nx, ny, nz, aligned16 ? 2 : aligned4 ? 1 : 0};
    int pick = 0;
    const size_t bytes = (size_t)(nx + 8) * (ny + 8) * (nz + 8) * dtype_size(dtype);
    int st = APPFIS_OK;
    bool cached = false;
    for (const appfis_tune_entry& e : ctx->tuned)
        if (memcmp(e.key, key, sizeof(key)) == 0) cached = true;
    void* s0 = nullptr;
    void* s1 = nullptr;
    if (!cached) {
        APPFIS_NOT_CAPTURING(ctx, "appfis_fdtd3d CHOSEN's first call for a shape (autotuning)");
        const size_t half = (bytes + 255) / 256 * 256;
        st = ensure_work(ctx, 2 * half);
        if (st) return st;
        s0 = ctx->work;
        s1 = (char*)ctx->work + half;
        APPFIS_HIP_CHECK(hipMemsetAsync(ctx->work, 0, 2 * half, ctx->stream));
    }
    st = tune_select(ctx, key, (int)cand.size(), names.data(),
                     [&](int i) { return tab[cand[i]].fn(ctx, s0, s1, nx, ny, nz, coeff5, 2); }, 3, &pick);
    if (st) return st;
    return tab[cand[pick]].fn(ctx, buf0, buf1, nx, ny, nz, coeff5, steps);
}
