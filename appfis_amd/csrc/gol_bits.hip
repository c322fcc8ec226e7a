// GameOfLife CHOSEN form: the same B3/S23 generations on a bit-packed copy of the field.
//
// The call packs the interior of the uint8 field into 32-bit words (bit b of word w of
// row y = cell (32w + b, y), interior coordinates), advances the packed field one
// generation per launch, and unpacks the last two generations back into the two uint8
// buffers, ring exchanged: on return both buffers hold exactly what the PLAIN form
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

__device__ __forceinline__ uint32_t maj(uint32_t a, uint32_t b, uint32_t c)
{
    const uint32_t m = a ^ b;
    return (m & c) | (~m & a);  // a == b ? a : c
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

// one word -> cells x = 32w .. 32w+31 of interior row y
__global__ __launch_bounds__(256) void k_unpack(const uint32_t* __restrict__ P, uint8_t* __restrict__ u, int nx,
                                                int ny, int NW)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)NW * ny) return;
    const int y = (int)(t / NW), w = (int)(t - (size_t)y * NW);
    uint8_t* row = u + (size_t)(y + 1) * (nx + 2) + 1 + 32 * w;
    const uint32_t word = P[t];
#pragma unroll
    for (int b = 0; b < 32; b++) row[b] = (uint8_t)((word >> b) & 1u);
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
        h0[r] = W ^ c ^ E;
        h1[r] = maj(W, c, E);
        cen[r] = c;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int y = y0 + k;
        if (y >= ny) break;
        // S = h[k] + h[k+1] + h[k+2]: bits s0 (1), s1 (2), s2 (4); the 8s bit is implied
        const uint32_t s0 = h0[k] ^ h0[k + 1] ^ h0[k + 2];
        const uint32_t c0 = maj(h0[k], h0[k + 1], h0[k + 2]);
        const uint32_t u1 = h1[k] ^ h1[k + 1] ^ h1[k + 2];
        const uint32_t c1 = maj(h1[k], h1[k + 1], h1[k + 2]);
        const uint32_t s1 = u1 ^ c0;
        const uint32_t s2 = c1 ^ (u1 & c0);
        // S <= 9, so S == 3 <=> s0 s1 ~s2 and S == 4 <=> ~s0 ~s1 s2
        const uint32_t eq3 = s0 & s1 & ~s2;
        const uint32_t eq4 = s2 & ~(s0 | s1);
        Q[(size_t)y * NW + w] = eq3 | (cen[k + 1] & eq4);
    }
}

constexpr int KROWS = 8;

}  // namespace golbits

// Runs `gens` >= 1 generations of the nx x ny periodic field (nx % 32 == 0) held in
// buf0; see the file comment for what the two buffers hold on return.
int gol_bits_run(appfis_ctx* ctx, uint8_t* b0, uint8_t* b1, int nx, int ny, int gens, int tsteps)
{
    (void)tsteps;
    const int NW = nx / 32;
    const size_t words = (size_t)NW * ny;
    const size_t half = (words * 4 + 255) / 256 * 256;
    int st = ensure_work(ctx, 2 * half);
    if (st) return st;
    uint32_t* P[2] = {(uint32_t*)ctx->work, (uint32_t*)((char*)ctx->work + half)};
    const unsigned pblocks = ceil_div(words, 256);
    hipLaunchKernelGGL(golbits::k_pack, dim3(pblocks), dim3(256), 0, ctx->stream, b0, P[0], nx, ny, NW);
    APPFIS_LAUNCH_CHECK();
    const int nbands = (ny + golbits::KROWS - 1) / golbits::KROWS;
    const unsigned nitems = (unsigned)NW * (unsigned)nbands;
    const unsigned sblocks = ceil_div(nitems, 256);
    for (int g = 0; g < gens; g++) {
        hipLaunchKernelGGL(golbits::k_step<golbits::KROWS>, dim3(sblocks), dim3(256), 0, ctx->stream, P[g & 1],
                           P[(g + 1) & 1], NW, ny, nitems);
        APPFIS_LAUNCH_CHECK();
    }
    // generation gens -> buf[gens % 2], generation gens-1 -> the other buffer, rings exchanged
    uint8_t* buf[2] = {b0, b1};
    const int dims[2] = {nx + 2, ny + 2};
    for (int back = 0; back < 2; back++) {
        const int g = gens - back;
        uint8_t* dst = buf[g & 1];
        if (g > 0) {
            hipLaunchKernelGGL(golbits::k_unpack, dim3(pblocks), dim3(256), 0, ctx->stream, P[g & 1], dst, nx, ny,
                               NW);
            APPFIS_LAUNCH_CHECK();
        }
        st = appfis_exchange(ctx, dst, APPFIS_U8, 2, dims, 1, 1);
        if (st) return st;
    }
    return APPFIS_OK;
}
