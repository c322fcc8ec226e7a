// Context, memory, timing, graph capture and seeded inputs.
//
// Replaces the device/stream part of Topology::Initialize (reference
// Appfis/Topology.hpp:124-151: cudaSetDevice(rank % n) + one stream) and the
// allocation/copy part of Grid::CreateSubGridsGPU / CopySubGridsGPUToCPU
// (Appfis/Grid.hpp:403-479).
#include "common.hpp"

#include <cstdarg>
#include <new>

namespace appfis {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? APPFIS_ERR_NOMEM : APPFIS_ERR_HIP;
}

int invalid(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return APPFIS_ERR_INVALID;
}

int refuse_if_capturing(appfis_ctx* ctx, const char* what)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    APPFIS_HIP_CHECK(hipStreamIsCapturing(ctx->stream, &cs));
    if (cs == hipStreamCaptureStatusNone) return APPFIS_OK;
    set_error("%s cannot run while the context's stream is captured into a graph (it would allocate, free, "
              "synchronise or capture on its own); call it once before the capture, or use a capture-safe entry point",
              what);
    return APPFIS_ERR_UNSUPPORTED;
}

int ensure_scratch(appfis_ctx* ctx, size_t bytes)
{
    if (ctx->scratch_bytes >= bytes) return APPFIS_OK;
    APPFIS_NOT_CAPTURING(ctx, "growing the context scratch");
    if (ctx->scratch) {
        APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        APPFIS_HIP_CHECK(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    APPFIS_HIP_CHECK(hipMalloc(&ctx->scratch, want));
    APPFIS_HIP_CHECK(hipMemsetAsync(ctx->scratch, 0, want, ctx->stream));
    ctx->scratch_bytes = want;
    return APPFIS_OK;
}

int ensure_work(appfis_ctx* ctx, size_t bytes)
{
    if (ctx->work_bytes >= bytes) return APPFIS_OK;
    APPFIS_NOT_CAPTURING(ctx, "growing the context workspace");
    if (ctx->work) {
        APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        // the cached graphs hold pointers into the old workspace
        int st = clear_graphs(ctx);
        if (st) return st;
        APPFIS_HIP_CHECK(hipFree(ctx->work));
        ctx->work = nullptr;
        ctx->work_bytes = 0;
    }
    APPFIS_HIP_CHECK(hipMalloc(&ctx->work, bytes));
    ctx->work_bytes = bytes;
    return APPFIS_OK;
}

int ensure_pinned(appfis_ctx* ctx, size_t bytes)
{
    if (ctx->pinned_bytes >= bytes) return APPFIS_OK;
    APPFIS_NOT_CAPTURING(ctx, "growing the pinned staging buffer");
    if (ctx->pinned) {
        APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        APPFIS_HIP_CHECK(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    APPFIS_HIP_CHECK(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
    ctx->pinned_bytes = want;
    return APPFIS_OK;
}

static bool capturing(appfis_ctx* ctx)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int ktime_begin(appfis_ctx* ctx)
{
    if (!ctx->ktime || capturing(ctx)) return APPFIS_OK;
    ctx->kt_launches = -1;  // until ktime_end: a call that fails in between leaves no time
    APPFIS_HIP_CHECK(hipEventRecord(ctx->kt0, ctx->stream));
    return APPFIS_OK;
}

int ktime_end(appfis_ctx* ctx, int launches, int units)
{
    if (!ctx->ktime || capturing(ctx)) return APPFIS_OK;
    APPFIS_HIP_CHECK(hipEventRecord(ctx->kt1, ctx->stream));
    ctx->kt_launches = launches;
    ctx->kt_units = units;
    return APPFIS_OK;
}

constexpr size_t kMaxGraphs = 8;

int clear_graphs(appfis_ctx* ctx)
{
    hipError_t first = hipSuccess;
    for (auto& g : ctx->graphs) {
        hipError_t e = hipGraphExecDestroy(g.exec);
        if (first == hipSuccess) first = e;
    }
    ctx->graphs.clear();
    if (first != hipSuccess) return hip_fail(first, "graph cache teardown", __FILE__, __LINE__);
    return APPFIS_OK;
}

int cached_graph(appfis_ctx* ctx, std::vector<uint64_t> key, const std::function<int()>& enqueue,
                 hipGraphExec_t* exec)
{
    for (auto& g : ctx->graphs)
        if (g.key == key) {
            *exec = g.exec;
            return APPFIS_OK;
        }
    APPFIS_NOT_CAPTURING(ctx, "capturing a launch chain");
    APPFIS_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != APPFIS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "graph capture", __FILE__, __LINE__);
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    hipError_t e2 = hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "graph instantiate", __FILE__, __LINE__);
    if (e2 != hipSuccess) {
        (void)hipGraphExecDestroy(ge);
        return hip_fail(e2, "graph destroy", __FILE__, __LINE__);
    }
    if (ctx->graphs.size() >= kMaxGraphs) {
        // the oldest may still be running: drain the stream before destroying it
        hipError_t es = hipStreamSynchronize(ctx->stream);
        hipError_t ed = hipGraphExecDestroy(ctx->graphs.front().exec);
        ctx->graphs.erase(ctx->graphs.begin());
        if (es == hipSuccess) es = ed;
        if (es != hipSuccess) {
            (void)hipGraphExecDestroy(ge);
            return hip_fail(es, "graph cache eviction", __FILE__, __LINE__);
        }
    }
    ctx->graphs.push_back({std::move(key), ge});
    *exec = ge;
    return APPFIS_OK;
}

int tune_select(appfis_ctx* ctx, const int key[6], int ncand, const char* const* names,
                const std::function<int(int)>& run_variant, int reps, int* choice)
{
    for (const appfis_tune_entry& e : ctx->tuned)
        if (memcmp(e.key, key, sizeof(e.key)) == 0) {
            *choice = e.choice;
            return APPFIS_OK;
        }
    APPFIS_NOT_CAPTURING(ctx, "autotuning (the first CHOSEN call for a shape)");
    std::string rep = "autotune kernel " + std::to_string(key[0]) + " dtype " + std::to_string(key[1]) + " shape " +
                      std::to_string(key[2]) + "x" + std::to_string(key[3]) + "x" + std::to_string(key[4]) + ":";
    int best = 0;
    float best_ms = 0.f;
    // events of its own: the caller's timer (ctx->ev_start/ev_stop) may be running
    hipEvent_t e0 = nullptr, e1 = nullptr;
    APPFIS_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(hipErrorOutOfMemory, "hipEventCreate", __FILE__, __LINE__);
    }
    struct Guard {
        hipEvent_t a, b;
        ~Guard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1};
    for (int i = 0; i < ncand; i++) {
        int st = run_variant(i);  // warm-up (code load, caches)
        if (st) return st;
        APPFIS_HIP_CHECK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; r++) {
            st = run_variant(i);
            if (st) return st;
        }
        APPFIS_HIP_CHECK(hipEventRecord(e1, ctx->stream));
        APPFIS_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        APPFIS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        char line[160];
        snprintf(line, sizeof(line), " [%s %.4f ms]", names[i], ms);
        rep += line;
        if (i == 0 || ms < best_ms) {
            best = i;
            best_ms = ms;
        }
    }
    rep += std::string(" -> ") + names[best];
    appfis_tune_entry e;
    memcpy(e.key, key, sizeof(e.key));
    e.choice = best;
    e.report = rep;
    ctx->tuned.push_back(e);
    ctx->tune_report = rep;
    *choice = best;
    return APPFIS_OK;
}

int check_shape(int ndim, const int* dims, int* d)
{
    if (ndim < 1 || ndim > 3) return invalid("ndim must be 1, 2 or 3 (got %d)", ndim);
    if (!dims) return invalid("dims is NULL");
    d[0] = d[1] = d[2] = 1;
    for (int a = 0; a < ndim; a++) {
        if (dims[a] <= 0) return invalid("dims[%d] = %d must be positive", a, dims[a]);
        d[a] = dims[a];
    }
    if ((double)d[0] * d[1] * d[2] > 2147483647.0)
        return invalid("array of %dx%dx%d exceeds the int index range of APPFIS::Array", d[0], d[1], d[2]);
    return APPFIS_OK;
}

}  // namespace appfis

using namespace appfis;

extern "C" {

int appfis_abi_version(void) { return APPFIS_ABI_VERSION; }

const char* appfis_status_string(int status)
{
    switch (status) {
    case APPFIS_OK: return "ok";
    case APPFIS_ERR_INVALID: return "invalid argument";
    case APPFIS_ERR_HIP: return "HIP runtime error";
    case APPFIS_ERR_UNSUPPORTED: return "unsupported";
    case APPFIS_ERR_NOMEM: return "out of device memory";
    default: return "unknown status";
    }
}

const char* appfis_last_error(void) { return g_err; }

int appfis_tune_report(appfis_ctx* ctx, char* buf, int len)
{
    APPFIS_REQUIRE(ctx && buf && len > 0, "bad argument");
    std::string all;
    for (const appfis_tune_entry& e : ctx->tuned) all += e.report + "\n";
    snprintf(buf, (size_t)len, "%s", all.c_str());
    return APPFIS_OK;
}

int appfis_device_count(int* count)
{
    APPFIS_REQUIRE(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return APPFIS_OK;
}

int appfis_ctx_create(int device, appfis_ctx** out)
{
    APPFIS_REQUIRE(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    APPFIS_HIP_CHECK(hipGetDeviceCount(&n));
    APPFIS_REQUIRE(n > 0, "no HIP device visible");
    APPFIS_REQUIRE(device >= 0, "device index %d is negative", device);
    device = device % n;  // Topology.hpp:124-149 selects rank % deviceCount
    APPFIS_HIP_CHECK(hipSetDevice(device));
    appfis_ctx* c = new (std::nothrow) appfis_ctx();
    APPFIS_REQUIRE(c, "host allocation failed");
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // timer events are timing markers only: no system-scope fence (cache writeback and
    // invalidate) when they are recorded, which would add to the interval they measure
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_start, hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_stop, hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        appfis_ctx_destroy(c);
        return hip_fail(e, "context creation", __FILE__, __LINE__);
    }
    int st = noalloc_scratch_setup(device);
    if (st) {
        appfis_ctx_destroy(c);
        return st;
    }
    *out = c;
    return APPFIS_OK;
}

int appfis_ctx_destroy(appfis_ctx* ctx)
{
    if (!ctx) return APPFIS_OK;
    // best-effort teardown: report the first failure but release everything
    hipError_t first = hipSuccess;
    auto note = [&](hipError_t e) {
        if (first == hipSuccess && e != hipSuccess) first = e;
    };
    note(hipSetDevice(ctx->device));
    if (ctx->stream) note(hipStreamSynchronize(ctx->stream));
    for (auto& g : ctx->graphs) note(hipGraphExecDestroy(g.exec));
    ctx->graphs.clear();
    if (ctx->scratch) note(hipFree(ctx->scratch));
    if (ctx->work) note(hipFree(ctx->work));
    if (ctx->pinned) note(hipHostFree(ctx->pinned));
    if (ctx->ev_start) note(hipEventDestroy(ctx->ev_start));
    if (ctx->ev_stop) note(hipEventDestroy(ctx->ev_stop));
    if (ctx->kt0) note(hipEventDestroy(ctx->kt0));
    if (ctx->kt1) note(hipEventDestroy(ctx->kt1));
    if (ctx->stream) note(hipStreamDestroy(ctx->stream));
    delete ctx;
    if (first != hipSuccess) return hip_fail(first, "context teardown", __FILE__, __LINE__);
    return APPFIS_OK;
}

int appfis_ctx_device(appfis_ctx* ctx, int* device)
{
    APPFIS_REQUIRE(ctx && device, "NULL argument");
    *device = ctx->device;
    return APPFIS_OK;
}

void* appfis_ctx_stream(appfis_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int appfis_sync(appfis_ctx* ctx)
{
    APPFIS_REQUIRE(ctx, "ctx is NULL");
    APPFIS_NOT_CAPTURING(ctx, "appfis_sync");
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return APPFIS_OK;
}

int appfis_device_name(appfis_ctx* ctx, char* buf, int len)
{
    APPFIS_REQUIRE(ctx && buf && len > 0, "bad argument");
    hipDeviceProp_t p;
    APPFIS_HIP_CHECK(hipGetDeviceProperties(&p, ctx->device));
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return APPFIS_OK;
}

int appfis_malloc(appfis_ctx* ctx, size_t bytes, void** out)
{
    APPFIS_REQUIRE(ctx && out, "NULL argument");
    *out = nullptr;
    APPFIS_NOT_CAPTURING(ctx, "appfis_malloc");
    if (bytes == 0) bytes = 1;
    APPFIS_HIP_CHECK(hipSetDevice(ctx->device));
    APPFIS_HIP_CHECK(hipMalloc(out, bytes));
    APPFIS_HIP_CHECK(hipMemsetAsync(*out, 0, bytes, ctx->stream));
    return APPFIS_OK;
}

int appfis_free(appfis_ctx* ctx, void* ptr)
{
    APPFIS_REQUIRE(ctx, "ctx is NULL");
    if (!ptr) return APPFIS_OK;
    APPFIS_NOT_CAPTURING(ctx, "appfis_free");
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    APPFIS_HIP_CHECK(hipFree(ptr));
    return APPFIS_OK;
}

int appfis_memcpy_h2d(appfis_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    APPFIS_REQUIRE(ctx && dst && src, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "appfis_memcpy_h2d");
    APPFIS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return APPFIS_OK;
}

int appfis_memcpy_d2h(appfis_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    APPFIS_REQUIRE(ctx && dst && src, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "appfis_memcpy_d2h");
    APPFIS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return APPFIS_OK;
}

// 16-byte copy, one element per thread: each byte is read once and written once
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int appfis_memcpy_d2d(appfis_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    APPFIS_REQUIRE(ctx && dst && src, "NULL argument");
    size_t head = 0;
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) {
        const size_t n = bytes / 16;
        head = n * 16;
        if (n) {
            hipLaunchKernelGGL(k_copy16, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (const uint4*)src,
                               (uint4*)dst, n);
            APPFIS_LAUNCH_CHECK();
        }
    }
    if (bytes > head)
        APPFIS_HIP_CHECK(hipMemcpyAsync((char*)dst + head, (const char*)src + head, bytes - head,
                                        hipMemcpyDeviceToDevice, ctx->stream));
    return APPFIS_OK;
}

int appfis_memset(appfis_ctx* ctx, void* dst, int value, size_t bytes)
{
    APPFIS_REQUIRE(ctx && dst, "NULL argument");
    // a memset node in a captured graph is not reliable on replay: fill with a kernel
    // (appfis_fill_value) inside a capture
    APPFIS_NOT_CAPTURING(ctx, "appfis_memset");
    APPFIS_HIP_CHECK(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return APPFIS_OK;
}

int appfis_timer_start(appfis_ctx* ctx)
{
    APPFIS_REQUIRE(ctx, "ctx is NULL");
    APPFIS_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
    return APPFIS_OK;
}

int appfis_timer_stop(appfis_ctx* ctx, float* ms)
{
    APPFIS_REQUIRE(ctx && ms, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "appfis_timer_stop");
    APPFIS_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));
    APPFIS_HIP_CHECK(hipEventSynchronize(ctx->ev_stop));
    APPFIS_HIP_CHECK(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
    return APPFIS_OK;
}

int appfis_kernel_timing(appfis_ctx* ctx, int on)
{
    APPFIS_REQUIRE(ctx, "ctx is NULL");
    if (on && !ctx->kt0) {
        APPFIS_NOT_CAPTURING(ctx, "appfis_kernel_timing");
        // timing markers only: no system-scope fence (cache writeback and invalidate)
        // when they are recorded, so they do not slow the work they bracket
        APPFIS_HIP_CHECK(hipEventCreateWithFlags(&ctx->kt0, hipEventDisableSystemFence));
        APPFIS_HIP_CHECK(hipEventCreateWithFlags(&ctx->kt1, hipEventDisableSystemFence));
    }
    ctx->ktime = on != 0;
    ctx->kt_launches = -1;
    return APPFIS_OK;
}

int appfis_kernel_time(appfis_ctx* ctx, double* ms, int* launches, int* generations)
{
    APPFIS_REQUIRE(ctx && ms && launches && generations, "NULL argument");
    APPFIS_REQUIRE(ctx->kt_launches >= 0, "appfis_kernel_time: no call was timed since appfis_kernel_timing");
    APPFIS_NOT_CAPTURING(ctx, "appfis_kernel_time");
    float t = 0.0f;
    APPFIS_HIP_CHECK(hipEventSynchronize(ctx->kt1));
    APPFIS_HIP_CHECK(hipEventElapsedTime(&t, ctx->kt0, ctx->kt1));
    *ms = t;
    *launches = ctx->kt_launches;
    *generations = ctx->kt_units;
    return APPFIS_OK;
}

int appfis_graph_begin(appfis_ctx* ctx)
{
    APPFIS_REQUIRE(ctx, "ctx is NULL");
    APPFIS_NOT_CAPTURING(ctx, "appfis_graph_begin (a capture is already active)");
    APPFIS_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    return APPFIS_OK;
}

int appfis_graph_end(appfis_ctx* ctx, void** graph_exec)
{
    APPFIS_REQUIRE(ctx && graph_exec, "NULL argument");
    hipGraph_t g = nullptr;
    APPFIS_HIP_CHECK(hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipError_t e2 = hipGraphDestroy(g);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    *graph_exec = (void*)ge;
    return APPFIS_OK;
}

int appfis_graph_launch(appfis_ctx* ctx, void* graph_exec, int times)
{
    APPFIS_REQUIRE(ctx && graph_exec, "NULL argument");
    for (int t = 0; t < times; t++) APPFIS_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, ctx->stream));
    return APPFIS_OK;
}

int appfis_graph_destroy(appfis_ctx* ctx, void* graph_exec)
{
    APPFIS_REQUIRE(ctx, "ctx is NULL");
    if (graph_exec) APPFIS_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return APPFIS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Seeded inputs: one thread per interior cell, x fastest; writes only the interior.
// ---------------------------------------------------------------------------------

template <typename T>
__device__ inline T cvt_store(float v);
template <> __device__ inline float cvt_store<float>(float v) { return v; }
template <> __device__ inline __half cvt_store<__half>(float v) { return __float2half_rn(v); }
template <> __device__ inline __hip_bfloat16 cvt_store<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

template <typename T>
__global__ void k_fill_uniform_narrow(T* data, int d0, int d1, int n0, int n1, int n2, int g0, int g1, int g2,
                                      uint64_t seed, float lo, float hi)
{
    uint64_t lin = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = (uint64_t)n0 * n1 * n2;
    if (lin >= total) return;
    int i = (int)(lin % (uint64_t)n0);
    uint64_t r2 = lin / (uint64_t)n0;
    int j = (int)(r2 % (uint64_t)n1);
    int k = (int)(r2 / (uint64_t)n1);
    float u = (float)(rng(seed, lin) >> 40) * (1.0f / 16777216.0f);
    float v = __fadd_rn(lo, __fmul_rn(__fsub_rn(hi, lo), u));
    size_t o = (size_t)(i + g0) + (size_t)d0 * ((size_t)(j + g1) + (size_t)d1 * (size_t)(k + g2));
    data[o] = cvt_store<T>(v);
}

__global__ void k_fill_uniform_f64(double* data, int d0, int d1, int n0, int n1, int n2, int g0, int g1, int g2,
                                   uint64_t seed, double lo, double hi)
{
    uint64_t lin = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = (uint64_t)n0 * n1 * n2;
    if (lin >= total) return;
    int i = (int)(lin % (uint64_t)n0);
    uint64_t r2 = lin / (uint64_t)n0;
    int j = (int)(r2 % (uint64_t)n1);
    int k = (int)(r2 / (uint64_t)n1);
    double u = (double)(rng(seed, lin) >> 11) * (1.0 / 9007199254740992.0);
    double v = __dadd_rn(lo, __dmul_rn(__dsub_rn(hi, lo), u));
    size_t o = (size_t)(i + g0) + (size_t)d0 * ((size_t)(j + g1) + (size_t)d1 * (size_t)(k + g2));
    data[o] = v;
}

__global__ void k_fill_bernoulli(uint8_t* data, int d0, int d1, int n0, int n1, int n2, int g0, int g1, int g2,
                                 uint64_t seed, uint64_t thr, int always)
{
    uint64_t lin = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = (uint64_t)n0 * n1 * n2;
    if (lin >= total) return;
    int i = (int)(lin % (uint64_t)n0);
    uint64_t r2 = lin / (uint64_t)n0;
    int j = (int)(r2 % (uint64_t)n1);
    int k = (int)(r2 / (uint64_t)n1);
    uint64_t r = rng(seed, lin);
    size_t o = (size_t)(i + g0) + (size_t)d0 * ((size_t)(j + g1) + (size_t)d1 * (size_t)(k + g2));
    data[o] = (uint8_t)(always || r < thr);
}

template <typename T>
__global__ void k_fill_value(T* data, size_t count, T v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) data[i] = v;
}

extern "C" {

int appfis_fill_uniform(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost, uint64_t seed,
                        double lo, double hi)
{
    APPFIS_REQUIRE(ctx && data, "NULL argument");
    int d[3];
    int st = check_shape(ndim, dims, d);
    if (st) return st;
    APPFIS_REQUIRE(ghost >= 0, "ghost must be >= 0");
    int g[3] = {0, 0, 0};
    for (int a = 0; a < ndim; a++) g[a] = ghost;
    int n[3];
    for (int a = 0; a < 3; a++) {
        n[a] = d[a] - 2 * g[a];
        APPFIS_REQUIRE(n[a] > 0, "interior extent along axis %d is empty", a);
    }
    uint64_t total = (uint64_t)n[0] * n[1] * n[2];
    unsigned blocks = ceil_div(total, 256);
    switch (dtype) {
    case APPFIS_F64:
        hipLaunchKernelGGL(k_fill_uniform_f64, dim3(blocks), dim3(256), 0, ctx->stream, (double*)data, d[0], d[1], n[0],
                           n[1], n[2], g[0], g[1], g[2], seed, lo, hi);
        break;
    case APPFIS_F32:
        hipLaunchKernelGGL(k_fill_uniform_narrow<float>, dim3(blocks), dim3(256), 0, ctx->stream, (float*)data, d[0],
                           d[1], n[0], n[1], n[2], g[0], g[1], g[2], seed, (float)lo, (float)hi);
        break;
    case APPFIS_F16:
        hipLaunchKernelGGL(k_fill_uniform_narrow<__half>, dim3(blocks), dim3(256), 0, ctx->stream, (__half*)data, d[0],
                           d[1], n[0], n[1], n[2], g[0], g[1], g[2], seed, (float)lo, (float)hi);
        break;
    case APPFIS_BF16:
        hipLaunchKernelGGL(k_fill_uniform_narrow<__hip_bfloat16>, dim3(blocks), dim3(256), 0, ctx->stream,
                           (__hip_bfloat16*)data, d[0], d[1], n[0], n[1], n[2], g[0], g[1], g[2], seed, (float)lo,
                           (float)hi);
        break;
    default: return invalid("fill_uniform: unsupported dtype %d", dtype);
    }
    APPFIS_LAUNCH_CHECK();
    return APPFIS_OK;
}

int appfis_fill_bernoulli_u8(appfis_ctx* ctx, uint8_t* data, int ndim, const int* dims, int ghost, uint64_t seed,
                             double p)
{
    APPFIS_REQUIRE(ctx && data, "NULL argument");
    int d[3];
    int st = check_shape(ndim, dims, d);
    if (st) return st;
    APPFIS_REQUIRE(ghost >= 0, "ghost must be >= 0");
    APPFIS_REQUIRE(p >= 0.0 && p <= 1.0, "p must lie in [0,1]");
    int g[3] = {0, 0, 0};
    for (int a = 0; a < ndim; a++) g[a] = ghost;
    int n[3];
    for (int a = 0; a < 3; a++) {
        n[a] = d[a] - 2 * g[a];
        APPFIS_REQUIRE(n[a] > 0, "interior extent along axis %d is empty", a);
    }
    int always = p >= 1.0;
    uint64_t thr = always ? UINT64_MAX : (uint64_t)((long double)p * 18446744073709551616.0L);
    uint64_t total = (uint64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_fill_bernoulli, dim3(ceil_div(total, 256)), dim3(256), 0, ctx->stream, data, d[0], d[1],
                       n[0], n[1], n[2], g[0], g[1], g[2], seed, thr, always);
    APPFIS_LAUNCH_CHECK();
    return APPFIS_OK;
}

int appfis_fill_value(appfis_ctx* ctx, void* data, int dtype, size_t count, double value)
{
    APPFIS_REQUIRE(ctx && data, "NULL argument");
    if (count == 0) return APPFIS_OK;
    unsigned blocks = ceil_div(count, 256);
    if (blocks > 65536) blocks = 65536;
    switch (dtype) {
    case APPFIS_F64:
        hipLaunchKernelGGL(k_fill_value<double>, dim3(blocks), dim3(256), 0, ctx->stream, (double*)data, count, value);
        break;
    case APPFIS_F32:
        hipLaunchKernelGGL(k_fill_value<float>, dim3(blocks), dim3(256), 0, ctx->stream, (float*)data, count,
                           (float)value);
        break;
    case APPFIS_I32:
        hipLaunchKernelGGL(k_fill_value<int32_t>, dim3(blocks), dim3(256), 0, ctx->stream, (int32_t*)data, count,
                           (int32_t)value);
        break;
    case APPFIS_U8:
        hipLaunchKernelGGL(k_fill_value<uint8_t>, dim3(blocks), dim3(256), 0, ctx->stream, (uint8_t*)data, count,
                           (uint8_t)value);
        break;
    case APPFIS_I64:
        hipLaunchKernelGGL(k_fill_value<int64_t>, dim3(blocks), dim3(256), 0, ctx->stream, (int64_t*)data, count,
                           (int64_t)value);
        break;
    default: return invalid("fill_value: unsupported dtype %d", dtype);
    }
    APPFIS_LAUNCH_CHECK();
    return APPFIS_OK;
}

}  // extern "C"
