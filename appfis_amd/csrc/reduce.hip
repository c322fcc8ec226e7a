// Box reductions (SUM / MIN / MAX) in a single pass.
//
// Replaces Execute{1,2,3}DReduceGPU + Kernel*DReduce* + KernelReduce*Loop of the
// reference (Appfis/Runtime_CUDA.hpp:554-740, 1249-1666), whose first pass scans the
// whole padded array with %,/ index math, then loops log_256 passes and copies one
// scalar back.  Here: grid-stride over the Config box only, per-thread accumulation
// in a wide type, a wave64 DPP reduction, one LDS stage per block, per-block
// partials, and the last block to arrive (atomic ticket) combines the partials in a
// fixed order -- so the result is deterministic and never leaves the device unless
// the host asked for it.
//
// Semantics kept from the reference CPU path (Appfis/Runtime.hpp:598-913): MAX
// starts at numeric_limits<T>::min(), MIN at numeric_limits<T>::max().  SUM is
// widened (int64 for integer types, compensated FP64 for floats) instead of being
// accumulated in T; the T-typed result is the widened sum rounded once.
#include "common.hpp"

#include <cfloat>
#include <climits>
#include <mutex>

using namespace appfis;

namespace {

constexpr int RB = 256;        // threads per block
constexpr int MAX_BLOCKS = 2048;

struct Box {
    int d0, d1;             // padded extents along x, y
    int s0, s1, s2;         // box start
    int n0, n1, n2;         // box extents
};

__device__ inline size_t box_offset(const Box& b, uint64_t t)
{
    int i = (int)(t % (uint64_t)b.n0);
    uint64_t r = t / (uint64_t)b.n0;
    int j = (int)(r % (uint64_t)b.n1);
    int k = (int)(r / (uint64_t)b.n1);
    return (size_t)(i + b.s0) + (size_t)b.d0 * ((size_t)(j + b.s1) + (size_t)b.d1 * (size_t)(k + b.s2));
}

// --- compensated FP64 pair (Neumaier / TwoSum) ---------------------------------------
struct KSum {
    double s, c;
};
__device__ inline void two_sum_add(KSum& a, double x)
{
    double t = __dadd_rn(a.s, x);
    double bp = __dsub_rn(t, a.s);
    double err = __dadd_rn(__dsub_rn(a.s, __dsub_rn(t, bp)), __dsub_rn(x, bp));
    a.s = t;
    a.c = __dadd_rn(a.c, err);
}
__device__ inline KSum ksum_merge(KSum a, KSum b)
{
    two_sum_add(a, b.s);
    a.c = __dadd_rn(a.c, b.c);
    return a;
}

// wave reductions (DPP tree of common.hpp; the total is in lane WAVE_TOTAL_LANE)
template <typename T>
__device__ inline T wave_max(T v, T identity)
{
    return wave_reduce(v, identity, [](T a, T b) { return b > a ? b : a; });
}
template <typename T>
__device__ inline T wave_min(T v, T identity)
{
    return wave_reduce(v, identity, [](T a, T b) { return b < a ? b : a; });
}
__device__ inline long long wave_sum_i64(long long v)
{
    return wave_reduce(v, 0LL, [](long long a, long long b) { return a + b; });
}
__device__ inline KSum wave_sum_k(KSum v)
{
    return wave_reduce(v, KSum{0.0, 0.0}, [](KSum a, KSum b) { return ksum_merge(a, b); });
}

template <typename T> struct Lim;
template <> struct Lim<uint8_t> { static constexpr uint8_t lo = 0, hi = UCHAR_MAX; };
template <> struct Lim<int32_t> { static constexpr int32_t lo = INT_MIN, hi = INT_MAX; };
template <> struct Lim<float> { static constexpr float lo = FLT_MIN, hi = FLT_MAX; };
template <> struct Lim<double> { static constexpr double lo = DBL_MIN, hi = DBL_MAX; };

template <typename T, int OP> struct Ops;
// MAX/MIN of uint8 carry an int accumulator so every partial is a 4- or 8-byte word.
template <typename T> struct Wide { using type = T; };
template <> struct Wide<uint8_t> { using type = int; };

template <typename T> struct Ops<T, APPFIS_MAX> {
    using A = typename Wide<T>::type;
    __device__ static A init() { return Lim<T>::lo; }
    __device__ static void add(A& a, T v) { a = (A)v > a ? (A)v : a; }
    __device__ static A merge(A a, A b) { return b > a ? b : a; }
    __device__ static A wave(A a) { return wave_max(a, init()); }
};
template <typename T> struct Ops<T, APPFIS_MIN> {
    using A = typename Wide<T>::type;
    __device__ static A init() { return Lim<T>::hi; }
    __device__ static void add(A& a, T v) { a = (A)v < a ? (A)v : a; }
    __device__ static A merge(A a, A b) { return b < a ? b : a; }
    __device__ static A wave(A a) { return wave_min(a, init()); }
};
template <typename T> struct OpsSumInt {
    using A = long long;
    __device__ static A init() { return 0; }
    __device__ static void add(A& a, T v) { a += (long long)v; }
    __device__ static A merge(A a, A b) { return a + b; }
    __device__ static A wave(A a) { return wave_sum_i64(a); }
};
template <typename T> struct OpsSumFp {
    using A = KSum;
    __device__ static A init() { return KSum{0.0, 0.0}; }
    __device__ static void add(A& a, T v) { two_sum_add(a, (double)v); }
    __device__ static A merge(A a, A b) { return ksum_merge(a, b); }
    __device__ static A wave(A a) { return wave_sum_k(a); }
};
template <> struct Ops<uint8_t, APPFIS_SUM> : OpsSumInt<uint8_t> {};
template <> struct Ops<int32_t, APPFIS_SUM> : OpsSumInt<int32_t> {};
template <> struct Ops<float, APPFIS_SUM> : OpsSumFp<float> {};
template <> struct Ops<double, APPFIS_SUM> : OpsSumFp<double> {};

// Output conversion: MAX/MIN write T; SUM writes the wide result (int64 or double).
template <typename A> __device__ inline void write_wide(void* out, A a);
template <> __device__ inline void write_wide<long long>(void* out, long long a) { *(long long*)out = a; }
template <> __device__ inline void write_wide<KSum>(void* out, KSum a) { *(double*)out = __dadd_rn(a.s, a.c); }

// Partials written by other blocks are read past the (non-coherent) L1.
template <typename A> __device__ inline A load_l2(const A* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <> __device__ inline KSum load_l2<KSum>(const KSum* p)
{
    KSum k;
    k.s = __hip_atomic_load(&p->s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    k.c = __hip_atomic_load(&p->c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return k;
}

template <typename T, int OP>
__global__ __launch_bounds__(RB) void k_reduce(const T* __restrict__ data, Box b, uint64_t n,
                                               typename Ops<T, OP>::A* partials, unsigned* ticket, void* out)
{
    using O = Ops<T, OP>;
    using A = typename O::A;
    __shared__ A lds[RB / 64];
    __shared__ bool is_last;
    A acc = O::init();
    for (uint64_t t = (uint64_t)blockIdx.x * RB + threadIdx.x; t < n; t += (uint64_t)gridDim.x * RB)
        O::add(acc, data[box_offset(b, t)]);
    acc = O::wave(acc);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == WAVE_TOTAL_LANE) lds[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        A blk = lds[0];
        for (int w = 1; w < RB / 64; w++) blk = O::merge(blk, lds[w]);
        partials[blockIdx.x] = blk;
        __threadfence();
        unsigned prev = atomicAdd(ticket, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    // last block: combine partials in a fixed order (thread t takes t, t+RB, ...)
    A v = O::init();
    for (unsigned p = threadIdx.x; p < gridDim.x; p += RB) {
        v = O::merge(v, load_l2(partials + p));
    }
    v = O::wave(v);
    if (lane == WAVE_TOTAL_LANE) lds[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        A r = lds[0];
        for (int w = 1; w < RB / 64; w++) r = O::merge(r, lds[w]);
        if constexpr (OP == APPFIS_SUM) write_wide(out, r);
        else *(T*)out = (T)r;
        *ticket = 0u;  // ready for the next reduction on this context
    }
}

// Layout of a reduction's scratch: [0,256) ticket, [256, 256 + MAX_BLOCKS*16) partials
// (one 16-byte slot per block at most: KSum is the widest accumulator).
constexpr size_t SCRATCH_PARTIALS = 256;
constexpr size_t SCRATCH_BYTES = SCRATCH_PARTIALS + (size_t)MAX_BLOCKS * 16;
static_assert(SCRATCH_BYTES == APPFIS_NOALLOC_REDUCE_SCRATCH_BYTES, "header and kernel disagree on the scratch size");

unsigned reduce_blocks(const appfis_ctx* ctx, uint64_t n)
{
    unsigned blocks = ceil_div(n, RB);
    unsigned cap = (unsigned)(ctx->num_cu > 0 ? ctx->num_cu * 4 : 1024);
    if (cap > MAX_BLOCKS) cap = MAX_BLOCKS;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return blocks;
}

// scratch: SCRATCH_BYTES of device memory whose ticket word is zero at launch (the last
// block resets it, so consecutive reductions on one scratch need no clearing)
template <typename T, int OP>
int launch_reduce(appfis_ctx* ctx, const void* data, const Box& b, uint64_t n, void* dev_out, void* scratch)
{
    using A = typename Ops<T, OP>::A;
    const unsigned blocks = reduce_blocks(ctx, n);
    if (SCRATCH_PARTIALS + (size_t)blocks * sizeof(A) > SCRATCH_BYTES)
        return invalid("reduce: %u partials exceed the %zu-byte scratch", blocks, SCRATCH_BYTES);
    unsigned* ticket = (unsigned*)scratch;
    A* partials = (A*)((char*)scratch + SCRATCH_PARTIALS);
    hipLaunchKernelGGL((k_reduce<T, OP>), dim3(blocks), dim3(RB), 0, ctx->stream, (const T*)data, b, n, partials,
                       ticket, dev_out);
    APPFIS_LAUNCH_CHECK();
    return APPFIS_OK;
}

template <typename T>
int dispatch_op(appfis_ctx* ctx, const void* data, const Box& b, uint64_t n, int op, void* dev_out, void* scratch)
{
    switch (op) {
    case APPFIS_MAX: return launch_reduce<T, APPFIS_MAX>(ctx, data, b, n, dev_out, scratch);
    case APPFIS_MIN: return launch_reduce<T, APPFIS_MIN>(ctx, data, b, n, dev_out, scratch);
    case APPFIS_SUM: return launch_reduce<T, APPFIS_SUM>(ctx, data, b, n, dev_out, scratch);
    default: return invalid("unknown reduction op %d", op);
    }
}

int make_box(int ndim, const int* dims, const int* box, Box* b, uint64_t* n)
{
    int d[3];
    int st = check_shape(ndim, dims, d);
    if (st) return st;
    APPFIS_REQUIRE(box, "box is NULL");
    int s[3] = {0, 0, 0}, e[3] = {1, 1, 1};
    for (int a = 0; a < ndim; a++) {
        s[a] = box[2 * a];
        e[a] = box[2 * a + 1];
        APPFIS_REQUIRE(s[a] >= 0 && e[a] <= d[a], "box [%d,%d) outside dims[%d] = %d", s[a], e[a], a, d[a]);
    }
    b->d0 = d[0];
    b->d1 = d[1];
    b->s0 = s[0];
    b->s1 = s[1];
    b->s2 = s[2];
    b->n0 = e[0] > s[0] ? e[0] - s[0] : 0;
    b->n1 = e[1] > s[1] ? e[1] - s[1] : 0;
    b->n2 = e[2] > s[2] ? e[2] - s[2] : 0;
    *n = (uint64_t)b->n0 * b->n1 * b->n2;
    return APPFIS_OK;
}

int reduce_to_device(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims, const int* box, int op,
                     void* dev_out, void* scratch)
{
    Box b;
    uint64_t n;
    int st = make_box(ndim, dims, box, &b, &n);
    if (st) return st;
    if (n == 0) b.n0 = b.n1 = b.n2 = 0;  // empty box: kernel writes the identity
    switch (dtype) {
    case APPFIS_U8: return dispatch_op<uint8_t>(ctx, data, b, n, op, dev_out, scratch);
    case APPFIS_I32: return dispatch_op<int32_t>(ctx, data, b, n, op, dev_out, scratch);
    case APPFIS_F32: return dispatch_op<float>(ctx, data, b, n, op, dev_out, scratch);
    case APPFIS_F64: return dispatch_op<double>(ctx, data, b, n, op, dev_out, scratch);
    default: return invalid("reduce: unsupported dtype %d", dtype);
    }
}

// The context scratch of the allocating entry points: SCRATCH_BYTES, then 256 bytes for
// the result that the host copies back.
int ctx_reduce_scratch(appfis_ctx* ctx, void** scratch, void** result)
{
    int st = ensure_scratch(ctx, SCRATCH_BYTES + 256);
    if (st) return st;
    *scratch = ctx->scratch;
    if (result) *result = (char*)ctx->scratch + SCRATCH_BYTES;
    return APPFIS_OK;
}

}  // namespace

// ---- fixed scratch of appfis_reduce_device_noalloc --------------------------------
// One copy per GPU, zero at module load and held until the process ends.  Host code
// never passes the symbol itself to a kernel (on the host it only names the variable):
// it passes the device address that hipGetSymbolAddress returned for the current GPU.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

__global__ void k_clear_ticket(unsigned* ticket) { *ticket = 0u; }

}  // namespace

namespace appfis {

int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return invalid("device %d: the fixed reduction scratch supports devices 0..%d", device, MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return APPFIS_OK;
    int cur = -1;
    APPFIS_HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) return invalid("noalloc_scratch_setup: device %d is not current (%d is)", device, cur);
    void* ptr = nullptr;
    APPFIS_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return invalid("hipGetSymbolAddress returned NULL for the reduction scratch");
    g_scratch_ptr[device] = ptr;
    return APPFIS_OK;
}

}  // namespace appfis

namespace {

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

int appfis_reduce_device(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims, const int* box,
                         int op, void* dev_result)
{
    APPFIS_REQUIRE(ctx && data && dev_result, "NULL argument");
    void* scr = nullptr;
    int st = ctx_reduce_scratch(ctx, &scr, nullptr);
    if (st) return st;
    return reduce_to_device(ctx, data, dtype, ndim, dims, box, op, dev_result, scr);
}

int appfis_reduce_device_noalloc(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims,
                                 const int* box, int op, void* dev_result)
{
    APPFIS_REQUIRE(ctx && data && dev_result, "NULL argument");
    void* scr = noalloc_scratch(ctx->device);
    APPFIS_REQUIRE(scr, "device %d has no fixed reduction scratch (context not created through appfis_ctx_create)",
                   ctx->device);
    // the ticket must start at zero: a one-thread kernel clears it (a kernel, not a
    // memset node, so that the call can be captured into a graph)
    hipLaunchKernelGGL(k_clear_ticket, dim3(1), dim3(1), 0, ctx->stream, (unsigned*)scr);
    APPFIS_LAUNCH_CHECK();
    return reduce_to_device(ctx, data, dtype, ndim, dims, box, op, dev_result, scr);
}

int appfis_reduce_sum_wide(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims, const int* box,
                           void* result)
{
    APPFIS_REQUIRE(ctx && data && result, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "a reduction with a host result (use appfis_reduce_device)");
    void *scr = nullptr, *dev_out = nullptr;
    int st = ctx_reduce_scratch(ctx, &scr, &dev_out);
    if (st) return st;
    st = reduce_to_device(ctx, data, dtype, ndim, dims, box, APPFIS_SUM, dev_out, scr);
    if (st) return st;
    APPFIS_HIP_CHECK(hipMemcpyAsync(result, dev_out, 8, hipMemcpyDeviceToHost, ctx->stream));
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return APPFIS_OK;
}

int appfis_reduce(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims, const int* box, int op,
                  void* result)
{
    APPFIS_REQUIRE(ctx && data && result, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "a reduction with a host result (use appfis_reduce_device)");
    void *scr = nullptr, *dev_out = nullptr;
    int st = ctx_reduce_scratch(ctx, &scr, &dev_out);
    if (st) return st;
    st = reduce_to_device(ctx, data, dtype, ndim, dims, box, op, dev_out, scr);
    if (st) return st;
    unsigned char wide[8];
    APPFIS_HIP_CHECK(hipMemcpyAsync(wide, dev_out, 8, hipMemcpyDeviceToHost, ctx->stream));
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    size_t es = dtype_size(dtype);
    if (op != APPFIS_SUM) {
        memcpy(result, wide, es);
        return APPFIS_OK;
    }
    // SUM: round the widened result to T once
    if (dtype == APPFIS_U8 || dtype == APPFIS_I32) {
        long long v;
        memcpy(&v, wide, 8);
        if (dtype == APPFIS_U8) *(uint8_t*)result = (uint8_t)v;
        else *(int32_t*)result = (int32_t)v;
    } else {
        double v;
        memcpy(&v, wide, 8);
        if (dtype == APPFIS_F32) *(float*)result = (float)v;
        else *(double*)result = v;
    }
    return APPFIS_OK;
}

}  // extern "C"
