// The FP32 bulk form of the flood step chain (kernels and driver in flood_impl.hpp),
// in a translation unit of its own so that it can be compiled with multiply-adds
// contracted into FMAs (-ffp-contract=fast in the Makefile): its results are held to a
// stated tolerance against FP64, not to the functors' rounding, and the contraction
// takes 3% off its step time.  The C entry points (flood.hip) call it for APPFIS_F32.
#include "flood_impl.hpp"

namespace appfis {

int flood_run_f32(appfis_ctx* ctx, int nscen, const int* nx, const int* ny, void* const* h, void* const* qx,
                  void* const* qy, const void* const* z, const void* const* n2, const appfis_flood_params* prm,
                  int steps, bool graph, double* time_host, double* volume_host, double* dt_host,
                  const appfis_flood_sources* src, int* steps_done)
{
    return run_flood<float>(ctx, nscen, nx, ny, h, qx, qy, z, n2, prm, steps, graph, time_host, volume_host, dt_host,
                            src, steps_done);
}

}  // namespace appfis
