// Shared internals of libappfis_hip.so (gfx950 / CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/appfis_hip.h"

// one autotuning decision: kernel id, dtype and shape -> index of the chosen variant
struct appfis_tune_entry {
    int key[6];
    int choice;
    std::string report;
};

struct appfis_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr;
    hipEvent_t ev_stop = nullptr;
    // device scratch reused across calls (reductions, flood scalars); grown on demand
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // large device workspace of the chosen forms (mixed-precision copies, flood
    // intermediates); grown on demand, kept until the context is destroyed
    void* work = nullptr;
    size_t work_bytes = 0;
    // pinned host staging for small results
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    int num_cu = 0;
    std::vector<appfis_tune_entry> tuned;  // autotuner cache
    std::string tune_report;               // text of the most recent decision
    // instantiated launch chains of the ensemble and flood loops, keyed by every value
    // their kernel arguments hold (the tables they read are uploaded by each call); a
    // later call with the same key replays the graph instead of capturing it again
    struct graph_entry {
        std::vector<uint64_t> key;
        hipGraphExec_t exec;
    };
    std::vector<graph_entry> graphs;
    // appfis_kernel_timing: events around the generation kernel's launches of the most
    // recent appfis_gol call, with their launch and generation counts
    bool ktime = false;
    hipEvent_t kt0 = nullptr, kt1 = nullptr;
    int kt_launches = -1, kt_units = 0;
};

namespace appfis {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
int invalid(const char* fmt, ...);

// APPFIS_ERR_UNSUPPORTED (with a message naming `what`) when ctx's stream is being
// captured into a graph: for the steps that would allocate, free, synchronise or begin a
// capture of their own, none of which a capture allows.
int refuse_if_capturing(appfis_ctx* ctx, const char* what);

// Ensure ctx->scratch holds at least `bytes`; returns status.
int ensure_scratch(appfis_ctx* ctx, size_t bytes);
int ensure_work(appfis_ctx* ctx, size_t bytes);
int ensure_pinned(appfis_ctx* ctx, size_t bytes);

// The launch chain of `key` as an instantiated graph: the cached one, else `enqueue`
// (which only launches kernels on ctx->stream) captured, instantiated and cached.  The
// cache holds a few graphs per context; the oldest is destroyed, after the stream has
// drained, to make room, and all of them when the workspace their arguments point into
// is reallocated.
int cached_graph(appfis_ctx* ctx, std::vector<uint64_t> key, const std::function<int()>& enqueue,
                 hipGraphExec_t* exec);
// Destroy every cached graph (the stream must have drained).
int clear_graphs(appfis_ctx* ctx);

// appfis_kernel_timing: the events before the first and after the last timed launch of
// a call (no-ops unless enabled, and while the stream is being captured).
int ktime_begin(appfis_ctx* ctx);
int ktime_end(appfis_ctx* ctx, int launches, int units);

inline uint64_t key_bits(double x)
{
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return u;
}
inline uint64_t key_ptr(const void* p) { return (uint64_t)(uintptr_t)p; }

inline size_t dtype_size(int dtype)
{
    switch (dtype) {
    case APPFIS_U8: return 1;
    case APPFIS_I32: return 4;
    case APPFIS_F32: return 4;
    case APPFIS_F64: return 8;
    case APPFIS_F16: return 2;
    case APPFIS_BF16: return 2;
    case APPFIS_I64: return 8;
    default: return 0;
    }
}

// Autotuner: look up / run / record the variant choice for a (kernel, dtype, shape) key.
// run_variant(i) must enqueue the timed work of candidate i on ctx->stream (on scratch
// buffers, never on the caller's data) and return a status; each candidate is run once
// untimed, then `reps` times under hipEvents; the fastest mean wins.
int tune_select(appfis_ctx* ctx, const int key[6], int ncand, const char* const* names,
                const std::function<int(int)>& run_variant, int reps, int* choice);

// Looks up (once per device, with `device` current) the device address of the fixed
// scratch array behind appfis_reduce_device_noalloc; called by appfis_ctx_create.
int noalloc_scratch_setup(int device);

// Validates ndim in 1..3, dims > 0, returns padded extents in d[3] (missing axes = 1).
int check_shape(int ndim, const int* dims, int* d);

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Workgroup b of a launch is placed on XCD b % 8 (round-robin over gfx950's 8 XCDs, each
// with its own L2).  xcd_tile maps b to a tile index such that every XCD takes one
// contiguous run of the nb tiles: tiles that read each other's halo rows then share an
// L2 instead of fetching the same lines from HBM once per XCD.  Bijective on [0, nb).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb)
{
    const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Wave64 reductions through DPP (the cross-lane moves are operand modifiers of VALU
// moves; nothing goes through the LDS crossbar).  The tree is fixed: quad swaps (lane
// xor 1, xor 2), the half-row mirror (i <-> 7-i) and the row mirror (i <-> 15-i), after
// which every lane holds its 16-lane row's total; then row_bcast:15 adds row 0 into
// row 1 and row 2 into row 3, and row_bcast:31 adds row 1 into row 3.  The wave total
// ends in lane 63 (WAVE_TOTAL_LANE).  Lanes that a move does not write receive the
// identity, so the other lanes hold partial values.  Every lane of the wave must be
// active.
constexpr int WAVE_TOTAL_LANE = 63;

template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T dpp_move(T identity, T v)
{
    static_assert(sizeof(T) % 4 == 0, "DPP moves 32-bit words");
    constexpr int NW = (int)(sizeof(T) / 4);
    int vi[NW], oi[NW], ri[NW];
    __builtin_memcpy(vi, &v, sizeof(T));
    __builtin_memcpy(oi, &identity, sizeof(T));
#pragma unroll
    for (int w = 0; w < NW; w++) ri[w] = __builtin_amdgcn_update_dpp(oi[w], vi[w], CTRL, ROWS, 0xF, false);
    T r;
    __builtin_memcpy(&r, ri, sizeof(T));
    return r;
}

template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, T identity, Op op)
{
    v = op(v, dpp_move<0xB1, 0xF>(identity, v));   // quad_perm [1,0,3,2]
    v = op(v, dpp_move<0x4E, 0xF>(identity, v));   // quad_perm [2,3,0,1]
    v = op(v, dpp_move<0x141, 0xF>(identity, v));  // row_half_mirror
    v = op(v, dpp_move<0x140, 0xF>(identity, v));  // row_mirror
    v = op(v, dpp_move<0x142, 0xA>(identity, v));  // row_bcast:15 into rows 1 and 3
    v = op(v, dpp_move<0x143, 0xC>(identity, v));  // row_bcast:31 into rows 2 and 3
    return v;
}

// The value of lane WAVE_TOTAL_LANE, in every lane.
template <typename T>
__device__ __forceinline__ T wave_total(T v)
{
    static_assert(sizeof(T) % 4 == 0, "readlane moves 32-bit words");
    constexpr int NW = (int)(sizeof(T) / 4);
    int vi[NW];
    __builtin_memcpy(vi, &v, sizeof(T));
#pragma unroll
    for (int w = 0; w < NW; w++) vi[w] = __builtin_amdgcn_readlane(vi[w], WAVE_TOTAL_LANE);
    T r;
    __builtin_memcpy(&r, vi, sizeof(T));
    return r;
}

// Lane shifts across the wave (DPP wave_shr:1 / wave_shl:1): from_left gives lane l the
// value of lane l-1, from_right the value of lane l+1 (lanes 0 / 63 keep their own).
__device__ __forceinline__ int from_left(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int from_right(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false); }
template <bool LEFT>
__device__ __forceinline__ float lane_shift(float v)
{
    const int b = __float_as_int(v);
    return __int_as_float(LEFT ? from_left(b) : from_right(b));
}
template <bool LEFT>
__device__ __forceinline__ double lane_shift(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const unsigned l2 = (unsigned)(LEFT ? from_left(lo) : from_right(lo));
    const unsigned h2 = (unsigned)(LEFT ? from_left(hi) : from_right(hi));
    return __longlong_as_double((long long)(((unsigned long long)h2 << 32) | l2));
}

// As lane_shift, but lanes 0 / 63 (no source lane) receive zero: one DPP move per 32-bit
// word, with no copy of the old value.  For values whose edge lanes are discarded.
template <bool LEFT>
__device__ __forceinline__ float lane_shift_z(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), LEFT ? 0x138 : 0x130, 0xF, 0xF, true));
}
template <bool LEFT>
__device__ __forceinline__ double lane_shift_z(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned l2 = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, LEFT ? 0x138 : 0x130, 0xF, 0xF, true);
    const unsigned h2 = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), LEFT ? 0x138 : 0x130, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)h2 << 32) | l2));
}

// Global memory.  A pointer that a kernel reads from a table in memory (rather than from
// its arguments) is generic to the compiler, which then emits flat loads and stores; flat
// accesses count against both the vector-memory and the LDS/scalar wait counters, so
// every wait for one of them waits for all.  to_global states the address space.
template <typename T>
using gmem = T __attribute__((address_space(1)));
template <typename T>
__device__ __forceinline__ gmem<T>* to_global(T* p)
{
    return (gmem<T>*)p;
}

// splitmix64 (Steele, Lea, Flood 2014): r(seed, n) = mix(seed + (n+1)*golden)
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rng(uint64_t seed, uint64_t n)
{
    return mix64(seed + (n + 1) * 0x9E3779B97F4A7C15ULL);
}

}  // namespace appfis

#define APPFIS_HIP_CHECK(expr)                                                      \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) return appfis::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define APPFIS_LAUNCH_CHECK()                                                       \
    do {                                                                            \
        hipError_t _e = hipGetLastError();                                          \
        if (_e != hipSuccess) return appfis::hip_fail(_e, "kernel launch", __FILE__, __LINE__); \
    } while (0)

#define APPFIS_NOT_CAPTURING(ctx, what)                                             \
    do {                                                                            \
        int _st = appfis::refuse_if_capturing((ctx), (what));                      \
        if (_st) return _st;                                                        \
    } while (0)

#define APPFIS_REQUIRE(cond, ...)                                                   \
    do {                                                                            \
        if (!(cond)) return appfis::invalid(__VA_ARGS__);                           \
    } while (0)
