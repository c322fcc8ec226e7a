// GameOfLife CHOSEN form: the same B3/S23 generations on a bit-packed copy of the field.
//
// The call packs the interior of the uint8 field into 32-bit words (bit b of word w of
// row y = cell (32w + b, y), interior coordinates), advances the packed field (many
// generations per launch), and its last launch computes the last generation while
// unpacking the last two back into the two uint8 buffers, rings exchanged: on return both buffers hold exactly what the PLAIN form
// leaves (generation `gens` in buf[gens % 2], generation gens-1 in the other).  The
// packed layout needs nx % 32 == 0 so that the periodic wrap falls on word boundaries;
// appfis_gol runs the byte kernel for other widths.  Cells must hold 0 or 1.
//
// One generation, per word, with 32 cells evaluated at once by bitwise full adders:
//   h(row)  = W + C + E                (2-bit horizontal sum of each row; W/E are the
//                                       word shifted by one cell, funnel-shifted in from
//                                       the neighbouring words, wrapping at the row end)
//   S       = h(y-1) + h(y) + h(y+1)   (3x3 sum including the cell itself, 0..9)
//   alive'  = (S == 3) | (alive & S == 4)
// A thread owns one word column over K consecutive rows, so each row's horizontal sum
// is computed once and used by the three output rows that need it.
#include "common.hpp"

using namespace appfis;

namespace golbits {

// 3-input bitwise functions, one v_bitop3_b32 each (gfx950).  Both are symmetric in
// their inputs, so the lookup table does not depend on the operand order.  Written as
// separate instructions on purpose: sharing a ^ b between the two lets the compiler
// fall back to two-input operations.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);  // a ^ b ^ c
}
__device__ __forceinline__ uint32_t maj(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);  // at least two of a, b, c
}

// B3/S23 from the adders of the three rows' 2-bit horizontal sums: with s0 = xor3 and
// c0 = maj of the low bit-planes and v1 = xor3, c1 = maj of the high ones, the 3x3 sum
// is S = s0 + 2(c0 + v1) + 4c1, so
//   S == 3  <=>  s0 & P,   P = (c0 ^ v1) & ~c1
//   S == 4  <=>  ~s0 & Q,  Q = ~(c0 ^ v1) & (c0 ^ c1)
// and alive' = (s0 & P) | (alive & ~s0 & Q): four three-input operations (truth-table
// index a*4 + b*2 + c for bitop3(a, b, c)), checked against S == 3 | (alive & S == 4)
// for every triple of 2-bit sums.
__device__ __forceinline__ uint32_t life_rule(uint32_t s0, uint32_t c0, uint32_t v1, uint32_t c1, uint32_t alive)
{
    const uint32_t p = __builtin_amdgcn_bitop3_b32(c0, v1, c1, 0x14);
    const uint32_t q = __builtin_amdgcn_bitop3_b32(c0, v1, c1, 0x42);
    const uint32_t t = __builtin_amdgcn_bitop3_b32(alive, s0, q, 0x20);
    return __builtin_amdgcn_bitop3_b32(s0, p, t, 0xEA);
}

// cells x = 32w .. 32w+31 of interior row y (uint8, 0/1) -> one word
__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ u, uint32_t* __restrict__ P, int nx,
                                              int ny, int NW)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)NW * ny) return;
    const int y = (int)(t / NW), w = (int)(t - (size_t)y * NW);
    const uint8_t* row = u + (size_t)(y + 1) * (nx + 2) + 1 + 32 * w;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t x = ((uint32_t)row[4 * k] | ((uint32_t)row[4 * k + 1] << 8) |
                            ((uint32_t)row[4 * k + 2] << 16) | ((uint32_t)row[4 * k + 3] << 24)) &
                           0x01010101u;
        // byte j of x (0/1) lands on bit 24 + j: the other partial products fall on
        // distinct bits below 24 or above 31, so nothing carries into bits 24..27
        word |= ((x * 0x01020408u) >> 24) << (4 * k);
    }
    P[t] = word;
}

template <int K>
__global__ __launch_bounds__(256) void k_step(const uint32_t* __restrict__ P, uint32_t* __restrict__ Q, int NW, int ny,
                                              unsigned nitems)
{
    // consecutive blocks share rows at their band edges: keep them on one XCD
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= nitems) return;
    const int band = (int)(t / (unsigned)NW), w = (int)(t - (unsigned)band * NW);
    const int wl = w == 0 ? NW - 1 : w - 1;
    const int wr = w == NW - 1 ? 0 : w + 1;
    const int y0 = band * K;
    uint32_t h0[K + 2], h1[K + 2], cen[K + 2];
#pragma unroll
    for (int r = 0; r < K + 2; r++) {
        int y = y0 - 1 + r;  // periodic rows; y < ny + K, so the loop ends (once if K <= ny)
        if (y < 0) y += ny;
        while (y >= ny) y -= ny;
        const uint32_t* row = P + (size_t)y * NW;
        const uint32_t c = row[w], l = row[wl], rr = row[wr];
        const uint32_t W = __builtin_amdgcn_alignbit(c, l, 31);  // cell x-1 at bit x
        const uint32_t E = __builtin_amdgcn_alignbit(rr, c, 1);  // cell x+1 at bit x
        h0[r] = xor3(W, c, E);
        h1[r] = maj(W, c, E);
        cen[r] = c;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int y = y0 + k;
        if (y >= ny) break;
        // S = h[k] + h[k+1] + h[k+2] through the carry-save adders of the two bit-planes
        const uint32_t s0 = xor3(h0[k], h0[k + 1], h0[k + 2]);
        const uint32_t c0 = maj(h0[k], h0[k + 1], h0[k + 2]);
        const uint32_t u1 = xor3(h1[k], h1[k + 1], h1[k + 2]);
        const uint32_t c1 = maj(h1[k], h1[k + 1], h1[k + 2]);
        Q[(size_t)y * NW + w] = life_rule(s0, c0, u1, c1, cen[k + 1]);
    }
}

constexpr int KROWS = 8;

// One padded uint8 row from a packed word: the 32 cells of word w, and the ring cell
// that the periodic exchange (X first) copies from it -- ghost column nx+1 holds cell 0,
// ghost column 0 holds cell nx-1.
__device__ __forceinline__ void put_word(uint8_t* __restrict__ u, size_t W, int prow, int w, int NW, uint32_t word)
{
    uint8_t* row = u + (size_t)prow * W;
#pragma unroll
    for (int b = 0; b < 32; b++) row[1 + 32 * w + b] = (uint8_t)((word >> b) & 1u);
    if (w == 0) row[W - 1] = (uint8_t)(word & 1u);
    if (w == NW - 1) row[0] = (uint8_t)(word >> 31);
}

// The last generation of a call fused with the unpacking: from the packed generation
// n-1 in P it computes generation n (as k_step does) and writes generation n into `un`
// and generation n-1 into `uprev`, both with their periodic rings -- what one k_step,
// unpacking both generations and exchanging both rings would leave, in one pass.  The exchange's Y step copies
// padded row ny (ghost columns included) into ghost row 0 and padded row 1 into ghost
// row ny+1, so the thread of interior row ny-1 also writes row 0 and that of row 0 also
// writes row ny+1.  Reads only P: `uprev` may be the buffer the call's input came from.
template <int K>
__global__ __launch_bounds__(256) void k_step_unpack(const uint32_t* __restrict__ P, uint8_t* __restrict__ un,
                                                     uint8_t* __restrict__ uprev, int nx, int NW, int ny,
                                                     unsigned nitems)
{
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= nitems) return;
    const int band = (int)(t / (unsigned)NW), w = (int)(t - (unsigned)band * NW);
    const int wl = w == 0 ? NW - 1 : w - 1;
    const int wr = w == NW - 1 ? 0 : w + 1;
    const int y0 = band * K;
    const size_t W = (size_t)nx + 2;
    uint32_t h0[K + 2], h1[K + 2], cen[K + 2];
#pragma unroll
    for (int r = 0; r < K + 2; r++) {
        int y = y0 - 1 + r;
        if (y < 0) y += ny;
        while (y >= ny) y -= ny;
        const uint32_t* row = P + (size_t)y * NW;
        const uint32_t c = row[w], l = row[wl], rr = row[wr];
        const uint32_t Wc = __builtin_amdgcn_alignbit(c, l, 31);
        const uint32_t Ec = __builtin_amdgcn_alignbit(rr, c, 1);
        h0[r] = xor3(Wc, c, Ec);
        h1[r] = maj(Wc, c, Ec);
        cen[r] = c;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int y = y0 + k;
        if (y >= ny) break;
        const uint32_t s0 = xor3(h0[k], h0[k + 1], h0[k + 2]);
        const uint32_t c0 = maj(h0[k], h0[k + 1], h0[k + 2]);
        const uint32_t u1 = xor3(h1[k], h1[k + 1], h1[k + 2]);
        const uint32_t c1 = maj(h1[k], h1[k + 1], h1[k + 2]);
        const uint32_t next = life_rule(s0, c0, u1, c1, cen[k + 1]);
        put_word(un, W, y + 1, w, NW, next);
        put_word(uprev, W, y + 1, w, NW, cen[k + 1]);
        if (y == ny - 1) {
            put_word(un, W, 0, w, NW, next);
            put_word(uprev, W, 0, w, NW, cen[k + 1]);
        }
        if (y == 0) {
            put_word(un, W, ny + 1, w, NW, next);
            put_word(uprev, W, ny + 1, w, NW, cen[k + 1]);
        }
    }
}

// Temporal blocking: G generations per launch.  A block of NWAVES waves spans whole
// packed rows: NW = 64*WX words, lane l of wave (wx, wy) owns word 64*wx + l, and wave
// row wy holds K consecutive rows in registers, so the block's window is WY*K rows of
// the full periodic row width.  Its centre OUT = WY*K - 2G rows are exact after G
// generations (each generation invalidates one more window row at each end); blocks
// are spaced OUT rows apart.  Per generation:
//   A  each wave publishes its lane-0 and lane-63 words (all K rows) to LDS; barrier
//   B  horizontal sums of its K rows: the neighbour words come from lanes l-1 / l+1 by
//      DPP wave shifts, and across wave edges (lanes 0 / 63) from the LDS words of the
//      waves left / right (periodic in wx); the first and last rows' sums go to LDS;
//      barrier
//   C  vertical combine: the sums of the rows above / below the wave's band come from
//      the waves above / below (zero outside the window), then the rule.
template <int K, int G, int NWAVES>
__global__ __launch_bounds__(64 * NWAVES) void k_step_tb(const uint32_t* __restrict__ P, uint32_t* __restrict__ Q,
                                                          int NW, int ny, int WX, int OUT, int gens)
{
    __shared__ uint32_t edge_lo[NWAVES][K], edge_hi[NWAVES][K];  // lane-0 / lane-63 words
    __shared__ uint32_t sum_top[2][NWAVES][64], sum_bot[2][NWAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wx = wv % WX, wy = wv / WX;
    const int WY = NWAVES / WX;
    const int w = wx * 64 + lane;
    const int left = wy * WX + (wx == 0 ? WX - 1 : wx - 1);
    const int right = wy * WX + (wx == WX - 1 ? 0 : wx + 1);
    // bands in XCD-contiguous order: neighbouring bands (whose rows form each other's halo)
    // run on one XCD and share its L2
    const int band = (int)xcd_tile(blockIdx.x, gridDim.x);
    const int y0 = band * OUT - G + wy * K;  // global row of register row 0
    // periodic row of register row 0 (the window may start up to G rows below 0 and end
    // past ny); then consecutive rows step by one word row, wrapping once at ny when
    // K <= ny (the general loop otherwise).  32-bit word offsets: NW * ny < 2^31.
    int yw = y0 % ny;
    if (yw < 0) yw += ny;
    uint32_t R[K];
    if (K <= ny) {
        int y = yw, off = yw * NW + w;
#pragma unroll
        for (int k = 0; k < K; k++) {
            R[k] = P[off];
            ++y;
            off += NW;
            if (y == ny) {
                y = 0;
                off = w;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            int y = yw + k;
            while (y >= ny) y -= ny;
            R[k] = P[y * NW + w];
        }
    }
    for (int g = 0; g < gens; g++) {
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < K; k++) edge_lo[wv][k] = R[k];
        if (lane == 63)
#pragma unroll
            for (int k = 0; k < K; k++) edge_hi[wv][k] = R[k];
        __syncthreads();
        // the neighbouring waves' edge words, read by every lane (one address per wave:
        // LDS broadcast) so that no read waits behind a lane mask
        uint32_t eh[K], el[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            eh[k] = edge_hi[left][k];
            el[k] = edge_lo[right][k];
        }
        uint32_t h0[K], h1[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            // lane l receives lane l-1 (wave_shr:1) / lane l+1 (wave_shl:1); lanes 0 / 63
            // have no source lane and keep the `old` operand: the neighbouring wave's word
            const uint32_t l = (uint32_t)__builtin_amdgcn_update_dpp((int)eh[k], (int)R[k], 0x138, 0xF, 0xF, false);
            const uint32_t r = (uint32_t)__builtin_amdgcn_update_dpp((int)el[k], (int)R[k], 0x130, 0xF, 0xF, false);
            const uint32_t c = R[k];
            const uint32_t Wc = __builtin_amdgcn_alignbit(c, l, 31);
            const uint32_t Ec = __builtin_amdgcn_alignbit(r, c, 1);
            h0[k] = xor3(Wc, c, Ec);
            h1[k] = maj(Wc, c, Ec);
        }
        sum_top[0][wv][lane] = h0[0];
        sum_top[1][wv][lane] = h1[0];
        sum_bot[0][wv][lane] = h0[K - 1];
        sum_bot[1][wv][lane] = h1[K - 1];
        __syncthreads();
        uint32_t a0 = 0u, a1 = 0u, z0 = 0u, z1 = 0u;
        if (wy > 0) {
            a0 = sum_bot[0][wv - WX][lane];
            a1 = sum_bot[1][wv - WX][lane];
        }
        if (wy < WY - 1) {
            z0 = sum_top[0][wv + WX][lane];
            z1 = sum_top[1][wv + WX][lane];
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t u0 = k == 0 ? a0 : h0[k - 1], u1 = k == 0 ? a1 : h1[k - 1];
            const uint32_t d0 = k == K - 1 ? z0 : h0[k + 1], d1 = k == K - 1 ? z1 : h1[k + 1];
            const uint32_t s0 = xor3(u0, h0[k], d0);
            const uint32_t c0 = maj(u0, h0[k], d0);
            const uint32_t v1 = xor3(u1, h1[k], d1);
            const uint32_t c1 = maj(u1, h1[k], d1);
            R[k] = life_rule(s0, c0, v1, c1, R[k]);
        }
    }
    // output rows: window rows G .. G+OUT-1 of the block, global rows band*OUT + r
    const int ybase = band * OUT - G + wy * K;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = wy * K + k - G;  // output row within the block's OUT rows
        const int y = ybase + k;
        if (r >= 0 && r < OUT && y < ny) Q[y * NW + w] = R[k];
    }
}

// One barrier per generation.  Each wave publishes its raw edge words (lanes 0 / 63,
// all K rows) and its first and last rows into the LDS buffers of the generation's
// parity; after the barrier every wave computes the horizontal sums of its K rows and
// of the two rows next to its band (those of the waves above / below, with their edge
// words) itself.  The buffers alternate by generation, so a wave still reading
// generation g's buffers is never overwritten: the next writes to them come after the
// barrier of generation g+1, which every wave reaches only after finishing g.
template <int K, int G, int NWAVES>
__global__ __launch_bounds__(64 * NWAVES) void k_step_tb1(const uint32_t* __restrict__ P, uint32_t* __restrict__ Q,
                                                           int NW, int ny, int WX, int OUT, int gens)
{
    __shared__ uint32_t edge_lo[2][NWAVES][K], edge_hi[2][NWAVES][K];  // lane-0 / lane-63 words
    __shared__ uint32_t row_top[2][NWAVES][64], row_bot[2][NWAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wx = wv % WX, wy = wv / WX;
    const int WY = NWAVES / WX;
    const int w = wx * 64 + lane;
    const int left = wy * WX + (wx == 0 ? WX - 1 : wx - 1);
    const int right = wy * WX + (wx == WX - 1 ? 0 : wx + 1);
    const int band = (int)xcd_tile(blockIdx.x, gridDim.x);  // XCD-contiguous band order
    const int y0 = band * OUT - G + wy * K;
    int yw = y0 % ny;
    if (yw < 0) yw += ny;
    uint32_t R[K];
    if (K <= ny) {
        int y = yw, off = yw * NW + w;
#pragma unroll
        for (int k = 0; k < K; k++) {
            R[k] = P[off];
            ++y;
            off += NW;
            if (y == ny) {
                y = 0;
                off = w;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            int y = yw + k;
            while (y >= ny) y -= ny;
            R[k] = P[y * NW + w];
        }
    }
    for (int g = 0; g < gens; g++) {
        const int b = g & 1;
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < K; k++) edge_lo[b][wv][k] = R[k];
        if (lane == 63)
#pragma unroll
            for (int k = 0; k < K; k++) edge_hi[b][wv][k] = R[k];
        row_top[b][wv][lane] = R[0];
        row_bot[b][wv][lane] = R[K - 1];
        __syncthreads();
        // horizontal sums of rows -1 .. K (rows -1 / K: the neighbouring waves' band
        // edges, zero outside the window)
        uint32_t h0[K + 2], h1[K + 2];
        auto hsum = [&](uint32_t c, uint32_t lo_edge, uint32_t hi_edge, uint32_t& s0, uint32_t& s1) {
            const uint32_t l = (uint32_t)__builtin_amdgcn_update_dpp((int)lo_edge, (int)c, 0x138, 0xF, 0xF, false);
            const uint32_t r = (uint32_t)__builtin_amdgcn_update_dpp((int)hi_edge, (int)c, 0x130, 0xF, 0xF, false);
            const uint32_t Wc = __builtin_amdgcn_alignbit(c, l, 31);
            const uint32_t Ec = __builtin_amdgcn_alignbit(r, c, 1);
            s0 = xor3(Wc, c, Ec);
            s1 = maj(Wc, c, Ec);
        };
#pragma unroll
        for (int k = 0; k < K; k++) hsum(R[k], edge_hi[b][left][k], edge_lo[b][right][k], h0[k + 1], h1[k + 1]);
        if (wy > 0) {
            hsum(row_bot[b][wv - WX][lane], edge_hi[b][left - WX][K - 1], edge_lo[b][right - WX][K - 1], h0[0],
                 h1[0]);
        } else {
            h0[0] = 0u;
            h1[0] = 0u;
        }
        if (wy < WY - 1) {
            hsum(row_top[b][wv + WX][lane], edge_hi[b][left + WX][0], edge_lo[b][right + WX][0], h0[K + 1],
                 h1[K + 1]);
        } else {
            h0[K + 1] = 0u;
            h1[K + 1] = 0u;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t s0 = xor3(h0[k], h0[k + 1], h0[k + 2]);
            const uint32_t c0 = maj(h0[k], h0[k + 1], h0[k + 2]);
            const uint32_t v1 = xor3(h1[k], h1[k + 1], h1[k + 2]);
            const uint32_t c1 = maj(h1[k], h1[k + 1], h1[k + 2]);
            R[k] = life_rule(s0, c0, v1, c1, R[k]);
        }
    }
    const int ybase = band * OUT - G + wy * K;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = wy * K + k - G;
        const int y = ybase + k;
        if (r >= 0 && r < OUT && y < ny) Q[y * NW + w] = R[k];
    }
}

// Full-row waves: each wave holds K rows of the whole packed width, NW = 64 * WPL words,
// lane l holding words WPL*l .. WPL*l + WPL - 1 of each row.  A word's horizontal
// neighbours are then the lane's own registers, except at the lane's two ends, which
// come from the adjacent lanes by DPP wave rotates (wave_ror / wave_rol), and the
// rotation is the periodic wrap of the row: no LDS for the horizontal direction, and
// 16-byte loads and stores.  The block's NWAVES waves stack bands of K rows.  Each
// generation a wave publishes the horizontal sums (both bit-planes) of its first and
// last rows into LDS slots double-buffered by generation parity; after the one barrier
// it reads the sums of the rows just above and below its band (zero outside the
// block's window).  A wave still reading generation g's slots is never overwritten:
// the next writes to them come after the barrier of generation g+1, which every wave
// reaches only after finishing g.  G generations per launch, OUT = NWAVES*K - 2G
// output rows per block.  Window row r holds a valid state after generation g (counted
// from 0) only for g + 1 <= r <= NWAVES*K - 2 - g: the waves at the window's ends skip
// the rule once all their rows are past that range, and their horizontal sums once
// their rows were already past it before the generation (nothing valid reads them then;
// they still meet every barrier).
template <int WPL>
struct WordVec;
template <> struct WordVec<1> {
    using V = uint32_t;
    __device__ static void get(const V& v, uint32_t (&w)[1]) { w[0] = v; }
    __device__ static V put(const uint32_t (&w)[1]) { return w[0]; }
};
template <> struct WordVec<2> {
    using V = uint2;
    __device__ static void get(const V& v, uint32_t (&w)[2]) { w[0] = v.x; w[1] = v.y; }
    __device__ static V put(const uint32_t (&w)[2]) { return make_uint2(w[0], w[1]); }
};
template <> struct WordVec<4> {
    using V = uint4;
    __device__ static void get(const V& v, uint32_t (&w)[4]) { w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w; }
    __device__ static V put(const uint32_t (&w)[4]) { return make_uint4(w[0], w[1], w[2], w[3]); }
};

// The generation loop of the full-row form: `gens` generations of the K rows in R
// (window rows wv*K .. wv*K+K-1 of the block); with KEEP, Rp receives the rows as they
// were before the last generation.
template <int WPL, int K, int G, int NWAVES, bool KEEP>
__device__ __forceinline__ void tw_generations(uint32_t (&R)[K][WPL], uint32_t (&Rp)[K][WPL],
                                               uint32_t (*hs)[NWAVES][2][2][64 * WPL], int lane, int wv, int gens)
{
    using VW = WordVec<WPL>;
    using V = typename VW::V;
    const int top = wv * K, bot = wv * K + K - 1, last = NWAVES * K - 1;  // window rows of the wave
    for (int g = 0; g < gens; g++) {
        const int b = g & 1;
        const bool sums = bot >= g && top <= last - g;        // some row valid before generation g
        const bool rule = bot >= g + 1 && top <= last - 1 - g;  // some row valid after it
        if constexpr (KEEP) {
            if (g == gens - 1) {
#pragma unroll
                for (int k = 0; k < K; k++)
#pragma unroll
                    for (int w = 0; w < WPL; w++) Rp[k][w] = R[k][w];
            }
        }
        uint32_t h0[K + 2][WPL], h1[K + 2][WPL];
        if (sums) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                // the words left of the lane's first and right of its last (periodic row;
                // a rotation has a source for every lane)
                const uint32_t lft = (uint32_t)__builtin_amdgcn_mov_dpp((int)R[k][WPL - 1], 0x13C, 0xF, 0xF, true);
                const uint32_t rgt = (uint32_t)__builtin_amdgcn_mov_dpp((int)R[k][0], 0x134, 0xF, 0xF, true);
#pragma unroll
                for (int w = 0; w < WPL; w++) {
                    const uint32_t c = R[k][w];
                    const uint32_t l = w == 0 ? lft : R[k][w - 1];
                    const uint32_t r = w == WPL - 1 ? rgt : R[k][w + 1];
                    const uint32_t Wc = __builtin_amdgcn_alignbit(c, l, 31);
                    const uint32_t Ec = __builtin_amdgcn_alignbit(r, c, 1);
                    h0[k + 1][w] = xor3(Wc, c, Ec);
                    h1[k + 1][w] = maj(Wc, c, Ec);
                }
            }
            reinterpret_cast<V*>(hs[b][wv][0][0])[lane] = VW::put(h0[1]);
            reinterpret_cast<V*>(hs[b][wv][0][1])[lane] = VW::put(h1[1]);
            reinterpret_cast<V*>(hs[b][wv][1][0])[lane] = VW::put(h0[K]);
            reinterpret_cast<V*>(hs[b][wv][1][1])[lane] = VW::put(h1[K]);
        }
        __syncthreads();
        if (!rule) continue;
        {
            // rows -1 and K: whole-vector reads from the neighbouring waves' slots.  The
            // block's first and last waves read their own slots instead: rows outside the
            // window only reach the halo rows, which are never stored.
            const int up = wv > 0 ? wv - 1 : wv, dn = wv < NWAVES - 1 ? wv + 1 : wv;
            VW::get(reinterpret_cast<const V*>(hs[b][up][1][0])[lane], h0[0]);
            VW::get(reinterpret_cast<const V*>(hs[b][up][1][1])[lane], h1[0]);
            VW::get(reinterpret_cast<const V*>(hs[b][dn][0][0])[lane], h0[K + 1]);
            VW::get(reinterpret_cast<const V*>(hs[b][dn][0][1])[lane], h1[K + 1]);
        }
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int w = 0; w < WPL; w++) {
                const uint32_t s0 = xor3(h0[k][w], h0[k + 1][w], h0[k + 2][w]);
                const uint32_t c0 = maj(h0[k][w], h0[k + 1][w], h0[k + 2][w]);
                const uint32_t v1 = xor3(h1[k][w], h1[k + 1][w], h1[k + 2][w]);
                const uint32_t c1 = maj(h1[k][w], h1[k + 1][w], h1[k + 2][w]);
                R[k][w] = life_rule(s0, c0, v1, c1, R[k][w]);
            }
    }
}

template <int WPL, int K, int G, int NWAVES>
__global__ __launch_bounds__(64 * NWAVES) void k_step_tw(const uint32_t* __restrict__ P, uint32_t* __restrict__ Q,
                                                          int NW, int ny, int WX, int OUT, int gens)
{
    using VW = WordVec<WPL>;
    using V = typename VW::V;
    (void)WX;
    // [parity][wave][first / last row][bit-plane][word of the row]
    __shared__ uint32_t hs[2][NWAVES][2][2][64 * WPL];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int band = (int)xcd_tile(blockIdx.x, gridDim.x);  // XCD-contiguous band order
    const int y0 = band * OUT - G + wv * K;  // global row of register row 0
    int yw = y0 % ny;
    if (yw < 0) yw += ny;
    uint32_t R[K][WPL], Rp[K][WPL];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int y = yw + k;  // periodic rows (the window may wrap more than once when ny < K)
        while (y >= ny) y -= ny;
        VW::get(reinterpret_cast<const V*>(P + (size_t)y * NW)[lane], R[k]);
    }
    tw_generations<WPL, K, G, NWAVES, false>(R, Rp, hs, lane, wv, gens);
    // output rows: window rows G .. G+OUT-1 of the block, global rows band*OUT + r
    const int ybase = band * OUT - G + wv * K;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = wv * K + k - G;
        const int y = ybase + k;
        if (r >= 0 && r < OUT && y < ny) reinterpret_cast<V*>(Q + (size_t)y * NW)[lane] = VW::put(R[k]);
    }
}

// 16 cells of a padded uint8 row from 16 bits (bit j = cell j), as one 16-byte store
// (the compiler merges the byte stores), plus the ghost cell the periodic exchange
// copies from it: chunk 0 holds cell 0, which ghost column nx+1 takes; the last chunk
// holds cell nx-1, which ghost column 0 takes.
// 16 bytes at any address, allowed to alias the uint8 field it reads and writes
typedef unsigned int u32x4_u __attribute__((ext_vector_type(4), aligned(1), may_alias));
__device__ __forceinline__ void put_chunk(uint8_t* __restrict__ row, size_t W, int c, int nchunks, uint32_t h)
{
    u32x4_u d;
#pragma unroll
    for (int q = 0; q < 4; q++) d[q] = (((h >> (4 * q)) & 0xFu) * 0x00204081u) & 0x01010101u;  // bit j -> byte j
    *reinterpret_cast<u32x4_u*>(row + 1 + 16 * c) = d;
    if (c == 0) row[W - 1] = (uint8_t)(h & 1u);
    if (c == nchunks - 1) row[0] = (uint8_t)((h >> 15) & 1u);
}

// The first and last launches of a call, fused with the uint8 conversions of k_pack and
// k_step_unpack.  IN_U8: the window comes from the padded uint8 field u (interior cells
// (x, y) at (y+1)*(nx+2) + 1 + x, 0 or 1) and is packed in registers; else from P.
// OUT_U8: the output rows go out as uint8 -- generation n (the launch's last) into un,
// generation n-1 into uprev, with the periodic ring as the exchange leaves it (the
// block holding interior row ny-1 also writes ghost row 0, the one holding row 0 ghost
// row ny+1) -- else into Q, packed.  The uint8 rows move in 16-byte chunks, chunk
// 64p + l of a row by lane l in pass p, so that each access instruction covers 1 KB of
// the row; the chunks' 16-bit halves of the lanes' words pass through a per-wave row in
// LDS (each lane owns WPL consecutive words).  A launch that reads u must not write it:
// the host pairs IN_U8 and OUT_U8 only across two launches.
template <int WPL, int K, int G, int NWAVES, bool IN_U8, bool OUT_U8>
__global__ __launch_bounds__(64 * NWAVES) void k_step_tw_io(const uint8_t* __restrict__ u, const uint32_t* __restrict__ P,
                                                             uint32_t* __restrict__ Q, uint8_t* __restrict__ un,
                                                             uint8_t* __restrict__ uprev, int nx, int ny, int OUT,
                                                             int gens)
{
    using VW = WordVec<WPL>;
    using V = typename VW::V;
    constexpr int NW = 64 * WPL;  // words of a row
    constexpr int NC = 2 * NW;    // 16-cell chunks of a row
    const size_t W = (size_t)nx + 2;
    __shared__ uint32_t hs[2][NWAVES][2][2][NW];
    __shared__ uint32_t stage[NWAVES][NW];  // one row of the wave, as words
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // chunk c = half c of the row (little-endian); may_alias: the same LDS is also read
    // and written as words
    typedef uint16_t __attribute__((may_alias)) half_t;
    half_t* st16 = reinterpret_cast<half_t*>(stage[wv]);
    const int band = (int)xcd_tile(blockIdx.x, gridDim.x);
    const int y0 = band * OUT - G + wv * K;
    int yw = y0 % ny;
    if (yw < 0) yw += ny;
    uint32_t R[K][WPL], Rp[K][WPL];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int y = yw + k;
        while (y >= ny) y -= ny;
        if constexpr (IN_U8) {
            const uint8_t* row = u + (size_t)(y + 1) * W + 1;
#pragma unroll
            for (int p = 0; p < NC / 64; p++) {
                const int c = 64 * p + lane;
                const u32x4_u b = *reinterpret_cast<const u32x4_u*>(row + 16 * c);
                uint32_t h = 0;
#pragma unroll
                for (int q = 0; q < 4; q++)
                    h |= (((b[q] & 0x01010101u) * 0x01020408u) >> 24) << (4 * q);  // byte j -> bit j, as in k_pack
                st16[c] = (uint16_t)h;
            }
#pragma unroll
            for (int w = 0; w < WPL; w++) R[k][w] = stage[wv][WPL * lane + w];
        } else {
            VW::get(reinterpret_cast<const V*>(P + (size_t)y * NW)[lane], R[k]);
        }
    }
    tw_generations<WPL, K, G, NWAVES, OUT_U8>(R, Rp, hs, lane, wv, gens);
    const int ybase = band * OUT - G + wv * K;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int r = wv * K + k - G;
        const int y = ybase + k;
        if (r < 0 || r >= OUT || y >= ny) continue;
        if constexpr (OUT_U8) {
#pragma unroll
            for (int f = 0; f < 2; f++) {
                uint8_t* const dst = f == 0 ? un : uprev;
#pragma unroll
                for (int w = 0; w < WPL; w++) stage[wv][WPL * lane + w] = f == 0 ? R[k][w] : Rp[k][w];
#pragma unroll
                for (int p = 0; p < NC / 64; p++) {
                    const int c = 64 * p + lane;
                    const uint32_t h = st16[c];
                    put_chunk(dst + (size_t)(y + 1) * W, W, c, NC, h);
                    if (y == ny - 1) put_chunk(dst, W, c, NC, h);
                    if (y == 0) put_chunk(dst + (size_t)(ny + 1) * W, W, c, NC, h);
                }
            }
        } else {
            reinterpret_cast<V*>(Q + (size_t)y * NW)[lane] = VW::put(R[k]);
        }
    }
}

typedef void (*tw_io_kernel)(const uint8_t*, const uint32_t*, uint32_t*, uint8_t*, uint8_t*, int, int, int, int);
struct TbVariant {
    const char* name;
    void (*kernel)(const uint32_t*, uint32_t*, int, int, int, int, int);
    int k, g, nwaves;
    int wpl;  // full-row form: words per lane (the wave spans the row); 0: WX waves per row
    // full-row form: the first launch of a call reading the uint8 field, and the last one
    // writing both uint8 buffers (k_step_tw_io); null for the other forms
    tw_io_kernel kin = nullptr, kout = nullptr;
};
#define TW_IO(WPL, K, G, NWAVES) k_step_tw_io<WPL, K, G, NWAVES, true, false>, k_step_tw_io<WPL, K, G, NWAVES, false, true>

static const TbVariant kTb[] = {
    {"tb k16 g8 w16", k_step_tb<16, 8, 16>, 16, 8, 16, 0}, {"tb k10 g4 w16", k_step_tb<10, 4, 16>, 10, 4, 16, 0},
    {"tb k20 g4 w8", k_step_tb<20, 4, 8>, 20, 4, 8, 0},    {"tb k12 g4 w8", k_step_tb<12, 4, 8>, 12, 4, 8, 0},
    {"tb k24 g8 w16", k_step_tb<24, 8, 16>, 24, 8, 16, 0}, {"tb k8 g4 w16", k_step_tb<8, 4, 16>, 8, 4, 16, 0},
    {"tb k12 g4 w16", k_step_tb<12, 4, 16>, 12, 4, 16, 0}, {"tb k32 g8 w8", k_step_tb<32, 8, 8>, 32, 8, 8, 0},
    {"tb k16 g4 w8", k_step_tb<16, 4, 8>, 16, 4, 8, 0},    {"tb k11 g5 w16", k_step_tb<11, 5, 16>, 11, 5, 16, 0},
    {"tb k13 g6 w16", k_step_tb<13, 6, 16>, 13, 6, 16, 0},
    {"tb1 k10 g4 w16", k_step_tb1<10, 4, 16>, 10, 4, 16, 0}, {"tb1 k11 g5 w16", k_step_tb1<11, 5, 16>, 11, 5, 16, 0},
    {"tb1 k13 g6 w16", k_step_tb1<13, 6, 16>, 13, 6, 16, 0}, {"tb1 k16 g8 w16", k_step_tb1<16, 8, 16>, 16, 8, 16, 0},
    {"tb1 k20 g4 w8", k_step_tb1<20, 4, 8>, 20, 4, 8, 0},    {"tb1 k12 g4 w16", k_step_tb1<12, 4, 16>, 12, 4, 16, 0},
    // 4K - 2G = 32 output rows per block like k10 g4 (256 blocks fill 256 CUs at 8192
    // rows), with more generations per launch to spread the launch's fixed load / store
    {"tb1 k11 g6 w16", k_step_tb1<11, 6, 16>, 11, 6, 16, 0}, {"tb1 k12 g8 w16", k_step_tb1<12, 8, 16>, 12, 8, 16, 0},
    // full-row waves (nx = 2048 * WPL), also 32 output rows per block at 8192 rows
    {"tw4 k3 g8 w16", k_step_tw<4, 3, 8, 16>, 3, 8, 16, 4, TW_IO(4, 3, 8, 16)},
    {"tw4 k4 g8 w12", k_step_tw<4, 4, 8, 12>, 4, 8, 12, 4, TW_IO(4, 4, 8, 12)},
    {"tw2 k3 g8 w16", k_step_tw<2, 3, 8, 16>, 3, 8, 16, 2, TW_IO(2, 3, 8, 16)},
    {"tw1 k3 g8 w16", k_step_tw<1, 3, 8, 16>, 3, 8, 16, 1, TW_IO(1, 3, 8, 16)},
};

}  // namespace golbits

namespace {

// `gens` generations from P[*cur] with the one-generation kernel
int run_step(appfis_ctx* ctx, uint32_t* const P[2], int* cur, int NW, int ny, int gens)
{
    const int nbands = (ny + golbits::KROWS - 1) / golbits::KROWS;
    const unsigned nitems = (unsigned)NW * (unsigned)nbands;
    for (int g = 0; g < gens; g++) {
        hipLaunchKernelGGL(golbits::k_step<golbits::KROWS>, dim3(ceil_div(nitems, 256)), dim3(256), 0, ctx->stream,
                           P[*cur], P[*cur ^ 1], NW, ny, nitems);
        APPFIS_LAUNCH_CHECK();
        *cur ^= 1;
    }
    return APPFIS_OK;
}

// rows of a block's window: the full-row form stacks NWAVES bands; the other forms
// spread each row over WX = NW / 64 waves
int tb_rows(const golbits::TbVariant& v, int NW)
{
    return v.wpl > 0 ? v.nwaves * v.k : (v.nwaves / (NW / 64)) * v.k;
}

bool tb_fits(const golbits::TbVariant& v, int NW)
{
    if (v.wpl > 0) return NW == 64 * v.wpl && v.nwaves * v.k > 2 * v.g;
    if (NW % 64) return false;
    const int WX = NW / 64;
    return WX <= v.nwaves && v.nwaves % WX == 0 && (v.nwaves / WX) * v.k > 2 * v.g;
}

// `gens` generations from P[*cur] with temporal-blocking variant v (G per launch)
int run_tb(appfis_ctx* ctx, const golbits::TbVariant& v, uint32_t* const P[2], int* cur, int NW, int ny, int gens)
{
    const int WX = NW / 64;
    const int OUT = tb_rows(v, NW) - 2 * v.g;
    const unsigned blocks = ceil_div((size_t)ny, (size_t)OUT);
    for (int done = 0; done < gens;) {
        const int n = gens - done < v.g ? gens - done : v.g;
        hipLaunchKernelGGL(v.kernel, dim3(blocks), dim3(64 * v.nwaves), 0, ctx->stream, P[*cur], P[*cur ^ 1], NW, ny,
                           WX, OUT, n);
        APPFIS_LAUNCH_CHECK();
        *cur ^= 1;
        done += n;
    }
    return APPFIS_OK;
}

}  // namespace

// Runs `gens` >= 1 generations of the nx x ny periodic field (nx % 32 == 0) held in
// buf0; see the file comment for what the two buffers hold on return.  When the packed
// row width is a multiple of 64 words (and at most 1024), the generations run
// temporally blocked (tile height and depth autotuned per shape).  With a full-row
// variant and two or more generations, all of them run in its launches (at least two),
// the first reading the uint8 field and the last writing both uint8 buffers; otherwise
// k_pack packs the field, generations 1 .. gens-1 run blocked, and the last one runs
// fused with the unpacking (k_step_unpack), so that generation gens-1 is still at hand.
int gol_bits_run(appfis_ctx* ctx, uint8_t* b0, uint8_t* b1, int nx, int ny, int gens, int force)
{
    APPFIS_REQUIRE(nx % 32 == 0, "gol: the bit-packed form needs nx %% 32 == 0 (nx = %d)", nx);
    const int NW = nx / 32;
    const int ntb = (int)(sizeof(golbits::kTb) / sizeof(golbits::kTb[0]));
    if (force >= 0) {
        APPFIS_REQUIRE(force < ntb, "gol: candidate %d out of range (%d candidates)", force, ntb);
        APPFIS_REQUIRE(tb_fits(golbits::kTb[force], NW),
                       "gol: candidate %s needs nx that its block of waves spans (nx = %d)",
                       golbits::kTb[force].name, nx);
    }
    const size_t words = (size_t)NW * ny;
    const size_t half = (words * 4 + 255) / 256 * 256;
    int st = ensure_work(ctx, 2 * half);
    if (st) return st;
    uint32_t* const P[2] = {(uint32_t*)ctx->work, (uint32_t*)((char*)ctx->work + half)};
    // autotune the blocked variant on the two packed buffers before they hold the field
    // (timing does not depend on the cell values)
    std::vector<int> cand;
    std::vector<const char*> names;
    for (int i = 0; i < ntb; i++)
        if (tb_fits(golbits::kTb[i], NW)) {
            cand.push_back(i);
            names.push_back(golbits::kTb[i].name);
        }
    int pick = force;
    if (force < 0 && !cand.empty() && gens > 1) {
        const int key[6] = {4, APPFIS_U8, nx, ny, 0, 0};
        int c = 0;
        st = tune_select(ctx, key, (int)cand.size(), names.data(),
                         [&](int i) {
                             int tmp = 0;
                             return run_tb(ctx, golbits::kTb[cand[i]], P, &tmp, NW, ny, 48);
                         },
                         4, &c);
        if (st) return st;
        pick = cand[c];
    }
    uint8_t* buf[2] = {b0, b1};
    if (pick >= 0 && golbits::kTb[pick].kin && gens >= 2) {
        // full-row form over two or more launches: the first packs the uint8 field as it
        // loads its window, the last writes generations gens and gens-1 back as uint8
        // (rings included) -- after the first has finished reading b0, which it may
        // overwrite.  Up to G generations: gens-1 in the first launch, 1 in the last;
        // beyond: G per launch, the rest in the last.
        const golbits::TbVariant& v = golbits::kTb[pick];
        const int OUT = tb_rows(v, NW) - 2 * v.g;
        const dim3 grid(ceil_div((size_t)ny, (size_t)OUT)), block(64 * v.nwaves);
        const int L = gens <= v.g ? 2 : (gens + v.g - 1) / v.g;
        const int first = gens <= v.g ? gens - 1 : v.g;
        const int last = gens - first - (L - 2) * v.g;
        st = ktime_begin(ctx);
        if (st) return st;
        hipLaunchKernelGGL(v.kin, grid, block, 0, ctx->stream, b0, nullptr, P[0], nullptr, nullptr, nx, ny, OUT,
                           first);
        APPFIS_LAUNCH_CHECK();
        int cur = 0;
        for (int l = 1; l < L - 1; l++) {
            hipLaunchKernelGGL(v.kernel, grid, block, 0, ctx->stream, P[cur], P[cur ^ 1], NW, ny, NW / 64, OUT, v.g);
            APPFIS_LAUNCH_CHECK();
            cur ^= 1;
        }
        hipLaunchKernelGGL(v.kout, grid, block, 0, ctx->stream, nullptr, P[cur], nullptr, buf[gens & 1],
                           buf[(gens & 1) ^ 1], nx, ny, OUT, last);
        APPFIS_LAUNCH_CHECK();
        return ktime_end(ctx, L, gens);
    }
    const unsigned pblocks = ceil_div(words, 256);
    hipLaunchKernelGGL(golbits::k_pack, dim3(pblocks), dim3(256), 0, ctx->stream, b0, P[0], nx, ny, NW);
    APPFIS_LAUNCH_CHECK();
    int cur = 0;
    st = ktime_begin(ctx);
    if (st) return st;
    int launches = gens - 1;
    if (pick >= 0 && gens > 1) {
        st = run_tb(ctx, golbits::kTb[pick], P, &cur, NW, ny, gens - 1);
        launches = (gens - 1 + golbits::kTb[pick].g - 1) / golbits::kTb[pick].g;
    } else {
        st = run_step(ctx, P, &cur, NW, ny, gens - 1);
    }
    if (st) return st;
    st = ktime_end(ctx, launches, gens - 1);
    if (st) return st;
    // generation gens-1 is in P[cur]: the last generation runs fused with the unpacking
    // of generations gens (into buf[gens % 2]) and gens-1 (the other buffer), rings
    // exchanged
    const int nbands = (ny + golbits::KROWS - 1) / golbits::KROWS;
    const unsigned nitems = (unsigned)NW * (unsigned)nbands;
    hipLaunchKernelGGL(golbits::k_step_unpack<golbits::KROWS>, dim3(ceil_div(nitems, 256)), dim3(256), 0,
                       ctx->stream, P[cur], buf[gens & 1], buf[(gens & 1) ^ 1], nx, NW, ny, nitems);
    APPFIS_LAUNCH_CHECK();
    return APPFIS_OK;
}

extern "C" int appfis_gol_candidates(char* names, int len)
{
    const int ntb = (int)(sizeof(golbits::kTb) / sizeof(golbits::kTb[0]));
    std::string all;
    for (int i = 0; i < ntb; i++) all += std::string(golbits::kTb[i].name) + "\n";
    if (names && len > 0) snprintf(names, (size_t)len, "%s", all.c_str());
    return ntb;
}
