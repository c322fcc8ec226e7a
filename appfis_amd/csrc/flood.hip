// C entry points of the Triton flood step chain and the FP64 form (kernels and driver in
// flood_impl.hpp; the FP32 form is instantiated in flood_f32.hip).  Built without FMA
// contraction: the FP64 form keeps the functors' operation order.
#include "flood_impl.hpp"

extern "C" {

int appfis_flood_run(appfis_ctx* ctx, int dtype, void* h, void* qx, void* qy, const void* z, const void* n2, int nx,
                     int ny, const appfis_flood_params* params, int steps, int variant, double* time_host,
                     double* volume_host)
{
    APPFIS_REQUIRE(ctx && params, "NULL argument");
    APPFIS_REQUIRE(steps >= 0, "steps must be >= 0");
    APPFIS_REQUIRE(variant == APPFIS_PLAIN || variant == APPFIS_CHOSEN, "flood: unknown variant %d", variant);
    void* hh[1] = {h};
    void* qq[1] = {qx};
    void* rr[1] = {qy};
    const void* zz[1] = {z};
    const void* nn[1] = {n2};
    const bool graph = variant == APPFIS_CHOSEN;
    if (dtype == APPFIS_F64)
        return run_flood<double>(ctx, 1, &nx, &ny, hh, qq, rr, zz, nn, params, steps, graph, time_host, volume_host,
                                 nullptr);
    if (dtype == APPFIS_F32)
        return flood_run_f32(ctx, 1, &nx, &ny, hh, qq, rr, zz, nn, params, steps, graph, time_host, volume_host,
                             nullptr, nullptr, nullptr);
    return invalid("flood: dtype must be F64 or F32");
}

int appfis_flood_run_src(appfis_ctx* ctx, int dtype, void* h, void* qx, void* qy, const void* z, const void* n2,
                         int nx, int ny, const appfis_flood_params* params, const appfis_flood_sources* sources,
                         int steps, int variant, double* time_host, double* volume_host, int* steps_done)
{
    APPFIS_REQUIRE(ctx && params, "NULL argument");
    APPFIS_REQUIRE(steps >= 0, "steps must be >= 0");
    APPFIS_REQUIRE(variant == APPFIS_PLAIN || variant == APPFIS_CHOSEN, "flood: unknown variant %d", variant);
    APPFIS_REQUIRE(!sources || sources->count >= 0, "sources: negative count");
    void* hh[1] = {h};
    void* qq[1] = {qx};
    void* rr[1] = {qy};
    const void* zz[1] = {z};
    const void* nn[1] = {n2};
    const bool graph = variant == APPFIS_CHOSEN;
    if (dtype == APPFIS_F64)
        return run_flood<double>(ctx, 1, &nx, &ny, hh, qq, rr, zz, nn, params, steps, graph, time_host, volume_host,
                                 nullptr, sources, steps_done);
    if (dtype == APPFIS_F32)
        return flood_run_f32(ctx, 1, &nx, &ny, hh, qq, rr, zz, nn, params, steps, graph, time_host, volume_host,
                             nullptr, sources, steps_done);
    return invalid("flood: dtype must be F64 or F32");
}

int appfis_flood_ensemble_run(appfis_ctx* ctx, int dtype, int nscen, const int* nx, const int* ny, void* const* h,
                              void* const* qx, void* const* qy, const void* const* z, const void* const* n2,
                              const appfis_flood_params* params, int steps, double* time_host, double* volume_host)
{
    APPFIS_REQUIRE(ctx && params && nx && ny && h && qx && qy && z && n2, "NULL argument");
    APPFIS_REQUIRE(nscen >= 1, "nscen must be >= 1");
    APPFIS_REQUIRE(steps >= 0, "steps must be >= 0");
    if (dtype == APPFIS_F64)
        return run_flood<double>(ctx, nscen, nx, ny, h, qx, qy, z, n2, params, steps, true, time_host, volume_host,
                                 nullptr);
    if (dtype == APPFIS_F32)
        return flood_run_f32(ctx, nscen, nx, ny, h, qx, qy, z, n2, params, steps, true, time_host, volume_host,
                             nullptr, nullptr, nullptr);
    return invalid("flood: dtype must be F64 or F32");
}

int appfis_flood_dem(appfis_ctx* ctx, int dtype, void* z, int nx, int ny, double dx, uint64_t seed, int nhills,
                     double slope, double wall)
{
    APPFIS_REQUIRE(ctx && z, "NULL argument");
    APPFIS_NOT_CAPTURING(ctx, "appfis_flood_dem");
    APPFIS_REQUIRE(nx >= 1 && ny >= 1 && nhills >= 0 && nhills <= 4096, "bad DEM arguments");
    std::vector<double> hills((size_t)(nhills > 0 ? nhills : 1) * 4);
    const double L = (double)nx * dx;
    for (int k = 0; k < nhills; k++) {
        double u[4];
        for (int m = 0; m < 4; m++) u[m] = (double)(rng(seed, 4 * (uint64_t)k + m) >> 11) * (1.0 / 9007199254740992.0);
        hills[4 * k + 0] = 5.0 + 45.0 * u[0];
        hills[4 * k + 1] = 200.0 + 1800.0 * u[1];
        hills[4 * k + 2] = u[2] * L;
        hills[4 * k + 3] = u[3] * L;
    }
    int st = ensure_scratch(ctx, 4096 + hills.size() * 8);
    if (st) return st;
    double* dh = (double*)((char*)ctx->scratch + 4096);
    APPFIS_HIP_CHECK(hipMemcpyAsync(dh, hills.data(), hills.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    const size_t n = (size_t)(nx + 2) * (ny + 2);
    if (dtype == APPFIS_F64)
        hipLaunchKernelGGL(k_dem<double>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (double*)z, nx, ny, dx, dh,
                           nhills, slope, wall);
    else if (dtype == APPFIS_F32)
        hipLaunchKernelGGL(k_dem<float>, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, (float*)z, nx, ny, dx, dh,
                           nhills, slope, wall);
    else
        return invalid("flood_dem: dtype must be F64 or F32");
    APPFIS_LAUNCH_CHECK();
    // the host vector dies on return: make the upload complete first
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return APPFIS_OK;
}

}  // extern "C"
