// GameOfLife 2D, uint8 cells, periodic, rule B3/S23 -- config form of the reference
// functor GameOfLife/src/game_appfis.hpp:121-135 driven on a periodic DUAL grid, where
// every generation is followed by the single-rank Exchange of the written buffer
// (Communication.hpp:613-678, X then Y over the full extents, so corners wrap).
//
// PLAIN: one generation per launch.  Each lane owns one CHUNK-byte (16) aligned chunk
//   of the padded buffer viewed as a linear byte array (rows are W = nx+2 bytes, so
//   chunks do not align with rows).  For a chunk whose cells lie in interior rows the
//   three input rows are p-W, p, p+W: the centre row is the aligned chunk itself, the
//   rows above and below are funnel-shifted out of aligned 16-byte loads by the
//   compile-time byte offset W mod 16 (one kernel instance per residue).  Neighbour
//   counts are SWAR byte sums (column sums, then +/-1 byte shifts with v_alignbyte),
//   the B3/S23 rule is a zero-byte test, and the chunk is stored with 16-byte stores.
//   Ghost cells in a chunk (x = 0 or W-1, and the two ghost rows) take the wrapped
//   value new(xw, yw), recomputed from the current buffer whose ring is valid, so the
//   written buffer leaves with its ring already exchanged: no second launch.
//
// CHOSEN: gol_bits.hip (bit-packed copy of the field, 32 cells per word) when nx % 32 == 0,
//   this byte kernel otherwise.
#include "common.hpp"

using namespace appfis;

namespace gol {

__device__ __forceinline__ uint32_t abyte(uint32_t hi, uint32_t lo, uint32_t s)
{
    return __builtin_amdgcn_alignbyte(hi, lo, s);
}

// CB bytes starting `off` bytes (0..15, compile time) into the CB + 16 bytes a[], plus
// the byte before (left) and after (right) that window.
template <int OFF, int CB>
__device__ __forceinline__ void window(const uint32_t (&a)[CB / 4 + 4], uint32_t (&w)[CB / 4], uint32_t& left,
                                       uint32_t& right)
{
    constexpr int D = OFF >> 2, S = OFF & 3;
#pragma unroll
    for (int d = 0; d < CB / 4; d++) {
        if constexpr (S == 0) w[d] = a[D + d];
        else w[d] = abyte(a[D + d + 1], a[D + d], S);
    }
    constexpr int L = OFF - 1, R = OFF + CB;  // byte positions in a[]
    left = (a[L >> 2] >> ((L & 3) * 8)) & 0xFFu;
    right = (a[R >> 2] >> ((R & 3) * 8)) & 0xFFu;
}

// Row of a byte offset p < 2^31 of rows of W >= 2 bytes, by multiply-high: with
// 2^(l-1) < W <= 2^l and m = floor(2^(31+l) / W) + 1 (< 2^32),
// floor(p*m / 2^(31+l)) = floor(p / W), because the excess p*(m - 2^(31+l)/W)/2^(31+l)
// is below 2^31/2^(31+l) <= 1/W.
struct RowDiv {
    uint32_t m;
    int sh;  // l - 1
    __device__ __forceinline__ int row(int p) const { return (int)(__umulhi((uint32_t)p, m) >> sh); }
};

inline RowDiv make_rowdiv(int W)
{
    int l = 0;
    while ((1LL << l) < W) l++;
    RowDiv r;
    r.m = (uint32_t)((1ULL << (31 + l)) / (uint64_t)W + 1ULL);
    r.sh = l - 1;
    return r;
}

__device__ __forceinline__ uint32_t zero_bytes(uint32_t t)
{
    // bytes of t are <= 0x0B: (t + 0x7F) sets bit 7 exactly for non-zero bytes
    return (~(t + 0x7F7F7F7Fu) & 0x80808080u) >> 7;
}

__device__ __forceinline__ uint8_t rule(int n, uint8_t alive) { return (uint8_t)(n == 3 || (n == 2 && alive)); }

// New value of cell (x, y) of the padded buffer, ghost cells mapped to the wrapped
// interior cell; neighbours read from the current buffer (valid ring).
__device__ __noinline__ uint8_t cell_new(const uint8_t* __restrict__ u, int W, int H, int x, int y)
{
    const int nx = W - 2, ny = H - 2;
    int xw = x == 0 ? nx : (x == W - 1 ? 1 : x);
    int yw = y == 0 ? ny : (y == H - 1 ? 1 : y);
    const uint8_t* r0 = u + (size_t)(yw - 1) * W;
    const uint8_t* r1 = r0 + W;
    const uint8_t* r2 = r1 + W;
    int n = r0[xw - 1] + r0[xw] + r0[xw + 1] + r1[xw - 1] + r1[xw + 1] + r2[xw - 1] + r2[xw] + r2[xw + 1];
    return rule(n, r1[xw]);
}

template <int DB, int CB>
__global__ __launch_bounds__(256) void k_gol_plain(const uint8_t* __restrict__ u, uint8_t* __restrict__ v, int W,
                                                   int H, int nchunks, int nbytes, RowDiv rd)
{
    constexpr int ND = CB / 4;   // dwords of a chunk
    constexpr int NA = ND + 4;   // dwords of the aligned span holding a neighbour row's window
    // above-row window starts at p - W: offset (16 - DB) % 16 in its aligned span;
    // below-row window starts at p + W: offset DB.
    constexpr int DA = (16 - DB) & 15;
    // nbytes < 2^31 (checked on the host): 32-bit offsets throughout
    // blocks two apart read each other's rows: keep them on one XCD (shared L2)
    const int c = (int)(xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
    const int lane = threadIdx.x & 63;
    const int p0 = c * CB;
    const bool valid = c < nchunks;
    const bool full = valid && p0 + CB <= nbytes;

    // rows of the first and last cell of the chunk, and whether the fast path applies
    // (all CB cells in interior rows and every load in bounds) -- before any load, so
    // that all loads of the lane are in flight together
    const int y0 = rd.row(p0);
    const int y1 = rd.row(p0 + CB - 1);
    const int below_end = ((p0 + W) & ~15) + 4 * NA;
    const bool fast = full && y0 >= 1 && y1 <= H - 2 && p0 - W >= 16 && below_end <= nbytes;
    // centre chunk (every valid full lane loads it; shared with neighbours below) and,
    // on the fast path, the aligned spans holding the rows above and below
    uint32_t m[ND], a8[NA], b8[NA];
#pragma unroll
    for (int d = 0; d < ND; d++) m[d] = 0u;
    uint32_t mlb = 0u, mrb = 0u;
    if (full) {
#pragma unroll
        for (int q = 0; q < ND / 4; q++) {
            const uint4 t = *(const uint4*)(u + p0 + 16 * q);
            m[4 * q] = t.x; m[4 * q + 1] = t.y; m[4 * q + 2] = t.z; m[4 * q + 3] = t.w;
        }
    }
    if (fast) {
        const int pa = (p0 - W) & ~15, pb = (p0 + W) & ~15;
#pragma unroll
        for (int q = 0; q < NA / 4; q++) {
            const uint4 ta = *(const uint4*)(u + pa + 16 * q);
            a8[4 * q] = ta.x; a8[4 * q + 1] = ta.y; a8[4 * q + 2] = ta.z; a8[4 * q + 3] = ta.w;
            const uint4 tb = *(const uint4*)(u + pb + 16 * q);
            b8[4 * q] = tb.x; b8[4 * q + 1] = tb.y; b8[4 * q + 2] = tb.z; b8[4 * q + 3] = tb.w;
        }
        // the centre row's outer neighbours for lanes whose adjacent lane does not hold them
        if (lane == 0 || c == 0) mlb = u[p0 - 1];
        if (lane == 63 || c + 1 >= nchunks || p0 + 2 * CB > nbytes) mrb = u[p0 + CB];
    }
    // left / right neighbour bytes of the centre row from the adjacent lanes
    const uint32_t prevl = __shfl_up(m[ND - 1], 1, 64);
    const uint32_t next0 = __shfl_down(m[0], 1, 64);

    if (fast) {
        const uint32_t ml = (lane == 0 || c == 0) ? mlb : (prevl >> 24);
        const uint32_t mr = (lane == 63 || c + 1 >= nchunks || p0 + 2 * CB > nbytes) ? mrb : (next0 & 0xFFu);
        uint32_t up[ND], dn[ND], ul, ur, dl, dr;
        if constexpr (DA == 0) {
            // W multiple of 16: the window is the aligned span's start; edge bytes come
            // from the neighbouring bytes (loaded directly: in bounds by the fast test)
#pragma unroll
            for (int d = 0; d < ND; d++) { up[d] = a8[d]; dn[d] = b8[d]; }
            ul = u[p0 - W - 1];
            ur = a8[ND] & 0xFFu;
            dl = u[p0 + W - 1];
            dr = b8[ND] & 0xFFu;
        } else {
            window<DA, CB>(a8, up, ul, ur);
            window<DB, CB>(b8, dn, dl, dr);
        }
        // column sums (bytes <= 3), then x-1 / x+1 shifted copies
        uint32_t s[ND];
#pragma unroll
        for (int d = 0; d < ND; d++) s[d] = up[d] + m[d] + dn[d];
        const uint32_t sl = ul + ml + dl, sr = ur + mr + dr;
        uint32_t out[ND];
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const uint32_t lo = d == 0 ? (sl << 24) : s[d - 1];
            const uint32_t hi = d == ND - 1 ? sr : s[d + 1];
            const uint32_t left = abyte(s[d], lo, 3);
            const uint32_t right = abyte(hi, s[d], 1);
            const uint32_t n = left + s[d] + right - m[d];
            // alive' = (n == 3) | (alive & n == 2) = ((n | alive) == 3) for n <= 8, alive <= 1
            out[d] = zero_bytes((n | m[d]) ^ 0x03030303u);
        }
        // patch ghost cells (x == 0 or x == W-1) inside the chunk
        const int x0 = p0 - y0 * W;
        if (x0 + CB - 1 >= W - 1 || x0 == 0) {
#pragma unroll
            for (int b = 0; b < CB; b++) {
                const int p = p0 + b;
                const int y = (unsigned)p / (unsigned)W, x = p - y * W;
                if (x == 0 || x == W - 1) {
                    const uint32_t nvv = cell_new(u, W, H, x, y);
                    out[b >> 2] = (out[b >> 2] & ~(0xFFu << ((b & 3) * 8))) | (nvv << ((b & 3) * 8));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < ND / 4; q++) {
            uint4 o;
            o.x = out[4 * q]; o.y = out[4 * q + 1]; o.z = out[4 * q + 2]; o.w = out[4 * q + 3];
            *(uint4*)(v + p0 + 16 * q) = o;
        }
    } else if (valid) {
        // slow path: chunk touches a ghost row, the buffer end, or an edge window
        for (int b = 0; b < CB; b++) {
            const int p = p0 + b;
            if (p >= nbytes) break;
            const int y = (unsigned)p / (unsigned)W, x = p - y * W;
            v[p] = cell_new(u, W, H, x, y);
        }
    }
}

constexpr int CHUNK = 16;  // bytes (cells) per lane (32 measured slower: fewer lanes in flight)

typedef void (*gol_kernel_t)(const uint8_t*, uint8_t*, int, int, int, int, RowDiv);

}  // namespace gol

int gol_bits_run(appfis_ctx* ctx, uint8_t* b0, uint8_t* b1, int nx, int ny, int gens, int force);

extern "C" int appfis_gol(appfis_ctx* ctx, uint8_t* buf0, uint8_t* buf1, int nx, int ny, int gens, int variant,
                          int tsteps)
{
    (void)tsteps;  // reserved: the blocked kernels' depth comes from the tuned candidate
    APPFIS_REQUIRE(ctx && buf0 && buf1, "NULL argument");
    APPFIS_REQUIRE(buf0 != buf1, "the two buffers must differ");
    APPFIS_REQUIRE(nx >= 2 && ny >= 2, "periodic GoL needs an interior of at least 2x2 (got %dx%d)", nx, ny);
    APPFIS_REQUIRE((double)(nx + 2) * (ny + 2) < 2147483647.0, "grid exceeds the int index range");
    APPFIS_REQUIRE(gens >= 0, "gens must be >= 0");
    APPFIS_REQUIRE(((uintptr_t)buf0 & 15) == 0 && ((uintptr_t)buf1 & 15) == 0, "GoL buffers must be 16-byte aligned");
    if (gens == 0) return APPFIS_OK;
    APPFIS_REQUIRE(variant >= 0, "gol: unknown variant %d", variant);
    if (variant >= APPFIS_CANDIDATE0)  // a forced blocked candidate of the bit-packed form
        return gol_bits_run(ctx, buf0, buf1, nx, ny, gens, variant - APPFIS_CANDIDATE0);
    // the bit-packed form needs the periodic wrap on word boundaries; other widths run
    // the byte kernel, which gives the same bytes
    if (variant == APPFIS_CHOSEN && nx % 32 == 0) return gol_bits_run(ctx, buf0, buf1, nx, ny, gens, -1);
    const int W = nx + 2, H = ny + 2;
    const int nbytes = W * H;
    const int nchunks = (nbytes + gol::CHUNK - 1) / gol::CHUNK;
    static constexpr gol::gol_kernel_t table[16] = {
        gol::k_gol_plain<0, gol::CHUNK>,  gol::k_gol_plain<1, gol::CHUNK>,  gol::k_gol_plain<2, gol::CHUNK>,
        gol::k_gol_plain<3, gol::CHUNK>,  gol::k_gol_plain<4, gol::CHUNK>,  gol::k_gol_plain<5, gol::CHUNK>,
        gol::k_gol_plain<6, gol::CHUNK>,  gol::k_gol_plain<7, gol::CHUNK>,  gol::k_gol_plain<8, gol::CHUNK>,
        gol::k_gol_plain<9, gol::CHUNK>,  gol::k_gol_plain<10, gol::CHUNK>, gol::k_gol_plain<11, gol::CHUNK>,
        gol::k_gol_plain<12, gol::CHUNK>, gol::k_gol_plain<13, gol::CHUNK>, gol::k_gol_plain<14, gol::CHUNK>,
        gol::k_gol_plain<15, gol::CHUNK>};
    gol::gol_kernel_t k = table[W & 15];
    // the ring of the input buffer must be valid (reference: ExchangeInitial on both
    // buffers in InitializeExecution, Runtime.hpp:124-142)
    const int dims[2] = {W, H};
    int st = appfis_exchange(ctx, buf0, APPFIS_U8, 2, dims, 1, 1);
    if (st) return st;
    uint8_t* buf[2] = {buf0, buf1};
    const unsigned blocks = ceil_div((size_t)nchunks, 256);
    const gol::RowDiv rd = gol::make_rowdiv(W);
    st = ktime_begin(ctx);
    if (st) return st;
    for (int g = 0; g < gens; g++) {
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, ctx->stream, buf[g & 1], buf[(g + 1) & 1], W, H, nchunks,
                           nbytes, rd);
        APPFIS_LAUNCH_CHECK();
    }
    return ktime_end(ctx, gens, gens);
}
