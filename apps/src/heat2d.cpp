// HeatDiffusion 2D config workload on the APPFIS API: the config stencil written as a
// user functor (generic Execute2D path), the hand-written kernel (APPFIS::kernels::Heat2D)
// and the batched ensemble (APPFIS::kernels::Heat2DEnsemble, the same grid among two
// others of different sizes) on identical inputs; prints one JSON line with the
// throughputs and whether the three fields are bit-identical.
//   heat2d [-dimx=1024] [-dimy=1024] [-iteration=500]
#include <cstdio>
#include <cstring>
#include <vector>

#include "Appfis/Appfis.hpp"

namespace APPFIS
{
    PARALLEL void STENCIL(const int i, const int j, double c, Array<double>* cur, Array<double>* next);
}

PARALLEL void APPFIS::STENCIL(const int i, const int j, double c, Array<double>* cur, Array<double>* next)
{
    const double u = cur->Get(i, j);
    double s = cur->Get(i, j - 1) + cur->Get(i, j + 1);
    s = s + cur->Get(i + 1, j);
    s = s + cur->Get(i - 1, j);
    s = s - 4.0 * u;
    next->Set(i, j, u + c * s);
}

int main(int argc, char* argv[])
{
    int nx = 1024, ny = 1024, iters = 500;
    if (APPFIS::CheckArgument(argc, argv, "dimx")) nx = APPFIS::GetArgument(argc, argv, "dimx");
    if (APPFIS::CheckArgument(argc, argv, "dimy")) ny = APPFIS::GetArgument(argc, argv, "dimy");
    if (APPFIS::CheckArgument(argc, argv, "iteration")) iters = APPFIS::GetArgument(argc, argv, "iteration");
    APPFIS::Attribute at;
    at.ghostLayer = 1;
    at.periodic = false;
    APPFIS::Initialize(argc, argv, APPFIS::DIM_2D, APPFIS::DIM_2D, at);
    APPFIS::Grid<double> ga(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL), gb(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL);
    APPFIS::Grid<double> gc(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL), gd(APPFIS::DIM_2D, nx / 2 + 3, ny + 5, APPFIS::DUAL),
        ge(APPFIS::DIM_2D, 37, 29, APPFIS::DUAL);
    // a grid of the same shape for one untimed call, so the hand kernel's one-time
    // autotuning for this shape stays outside the timed region
    APPFIS::Grid<double> gw(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL);
    gw.SetBoundary(0.0);
    ga.SetBoundary(0.0);
    gb.SetBoundary(0.0);
    gc.SetBoundary(0.0);
    gd.SetBoundary(0.0);
    ge.SetBoundary(0.0);
    APPFIS::Config<double> config(ga);
    config.GridsToUpdate(ga);
    APPFIS::InitializeExecution(ga, gb, gc, gd, ge, gw);
    const int dims[2] = {nx + 2, ny + 2};
    APPFIS::detail::Check(appfis_fill_uniform(APPFIS::GetContext(), gc.GetSubGridsGPUData(), APPFIS_F64, 2, dims, 1,
                                              1001, 0.0, 1.0), "fill");
    APPFIS::detail::Check(appfis_fill_uniform(APPFIS::GetContext(), ga.GetSubGridsGPUData(), APPFIS_F64, 2, dims, 1,
                                              1001, 0.0, 1.0), "fill");
    APPFIS::detail::Check(appfis_fill_uniform(APPFIS::GetContext(), gb.GetSubGridsGPUData(), APPFIS_F64, 2, dims, 1,
                                              1001, 0.0, 1.0), "fill");
    APPFIS::kernels::Heat2D(gw, 0.1, 2, APPFIS_CHOSEN);
    APPFIS::Timer timer;
    timer.Start("GENERIC");
    for (int it = 0; it < iters; it++) APPFIS::Execute2D(config, 0.1, ga.GetSubGrid(), ga.GetSubGridNext());
    timer.Stop("GENERIC");
    timer.Start("KERNEL");
    APPFIS::kernels::Heat2D(gb, 0.1, iters, APPFIS_CHOSEN);
    timer.Stop("KERNEL");
    std::vector<APPFIS::Grid<double>*> ensemble = {&gd, &gc, &ge};
    timer.Start("ENSEMBLE");
    APPFIS::kernels::Heat2DEnsemble(ensemble, 0.1, iters);
    timer.Stop("ENSEMBLE");
    double sa = 0, sb = 0;
    APPFIS::Execute2DReduce(config, sa, APPFIS::SUM, ga);
    APPFIS::Execute2DReduce(config, sb, APPFIS::SUM, gb);
    APPFIS::FinalizeExecution(ga, gb, gc, gd, ge, gw);
    const size_t n = (size_t)(nx + 2) * (ny + 2);
    const bool same = memcmp(ga.GetSubGrid()->GetData(), gb.GetSubGrid()->GetData(), n * sizeof(double)) == 0 &&
                      memcmp(ga.GetSubGrid()->GetData(), gc.GetSubGrid()->GetData(), n * sizeof(double)) == 0;
    printf("{\"app\": \"heat2d\", \"nx\": %d, \"ny\": %d, \"iterations\": %d, \"checksum\": %.17g, "
           "\"generic_glups\": %.3f, \"kernel_glups\": %.3f, \"bit_identical\": %s, \"checksum_kernel\": %.17g}\n",
           nx, ny, iters, sa, timer.CalculateBLUPS("GENERIC", nx * ny, iters),
           timer.CalculateBLUPS("KERNEL", nx * ny, iters), same ? "true" : "false", sb);
    APPFIS::Finalize();
    return same ? 0 : 1;
}
