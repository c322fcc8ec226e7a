// Triton flood step chain.
#include "common.hpp"

extern "C" {
int appfis_flood_run(appfis_ctx* ctx, int dtype, void* h, void* qx, void* qy, const void* z, const void* n2, int nx,
                     int ny, const appfis_flood_params* params, int steps, int variant, double* time_host,
                     double* volume_host)
{
    appfis::set_error("flood: not built yet");
    return APPFIS_ERR_UNSUPPORTED;
}
int appfis_flood_ensemble_run(appfis_ctx* ctx, int dtype, int nscen, const int* nx, const int* ny, void* const* h,
                              void* const* qx, void* const* qy, const void* const* z, const void* const* n2,
                              const appfis_flood_params* params, int steps, double* time_host, double* volume_host)
{
    appfis::set_error("flood: not built yet");
    return APPFIS_ERR_UNSUPPORTED;
}
int appfis_flood_dem(appfis_ctx* ctx, int dtype, void* z, int nx, int ny, double dx, uint64_t seed, int nhills,
                     double slope, double wall)
{
    appfis::set_error("flood: not built yet");
    return APPFIS_ERR_UNSUPPORTED;
}
}
