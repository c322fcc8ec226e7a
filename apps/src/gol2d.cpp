// GameOfLife config workload (uint8 cells, periodic, B3/S23, every cell written) on
// the APPFIS API: user functor through the generic Execute2D path (periodic exchange
// on the device after every generation) against the hand-written kernel
// (APPFIS::kernels::GameOfLife).  Then the result goes through the single-rank
// Gather / Scatter of a uint8 grid (a type the reference's MPI path does not take,
// Communication.hpp:67-83) and a binary file round trip (Array::SaveBinaryFile /
// LoadBinaryFile): the gathered full grid must be the subgrid's interior, and a fresh
// grid loaded from the file and scattered must hold it again.  Prints one JSON line
// (live counts, GLUP/s, identities); -output also writes the gathered field to
// output/gol.bin (uint8, row-major, nx x ny).
//   gol2d [-dimx=8192] [-dimy=8192] [-iteration=1000] [-chosen] [-output]
#include <cstdio>
#include <cstring>
#include <string>

#include "Appfis/Appfis.hpp"

namespace APPFIS
{
    PARALLEL void STENCIL(const int i, const int j, Array<uint8_t>* cur, Array<uint8_t>* next);
}

PARALLEL void APPFIS::STENCIL(const int i, const int j, Array<uint8_t>* cur, Array<uint8_t>* next)
{
    const int n = cur->Get(i - 1, j - 1) + cur->Get(i, j - 1) + cur->Get(i + 1, j - 1) + cur->Get(i - 1, j) +
                  cur->Get(i + 1, j) + cur->Get(i - 1, j + 1) + cur->Get(i, j + 1) + cur->Get(i + 1, j + 1);
    next->Set(i, j, (uint8_t)(n == 3 || (n == 2 && cur->Get(i, j))));
}

int main(int argc, char* argv[])
{
    int nx = 8192, ny = 8192, gens = 1000;
    if (APPFIS::CheckArgument(argc, argv, "dimx")) nx = APPFIS::GetArgument(argc, argv, "dimx");
    if (APPFIS::CheckArgument(argc, argv, "dimy")) ny = APPFIS::GetArgument(argc, argv, "dimy");
    if (APPFIS::CheckArgument(argc, argv, "iteration")) gens = APPFIS::GetArgument(argc, argv, "iteration");
    const int variant = APPFIS::CheckArgument(argc, argv, "chosen") ? APPFIS_CHOSEN : APPFIS_PLAIN;
    APPFIS::Attribute at;
    at.ghostLayer = 1;
    at.periodic = true;
    APPFIS::Initialize(argc, argv, APPFIS::DIM_2D, APPFIS::DIM_2D, at);
    APPFIS::Grid<uint8_t> ga(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL), gb(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL);
    // a grid of the same shape for one untimed call, so the hand kernel's one-time
    // autotuning for this shape stays outside the timed region
    APPFIS::Grid<uint8_t> gw(APPFIS::DIM_2D, nx, ny, APPFIS::DUAL);
    APPFIS::Config<uint8_t> config(ga);
    config.GridsToUpdate(ga);
    APPFIS::InitializeExecution(ga, gb, gw);
    const int dims[2] = {nx + 2, ny + 2};
    for (auto* g : {&ga, &gb})
        APPFIS::detail::Check(appfis_fill_bernoulli_u8(APPFIS::GetContext(), g->GetSubGridsGPUData(), 2, dims, 1,
                                                       1002, 0.25), "fill");
    // periodic ring of the seeded buffer (the reference exchanges at InitializeExecution)
    APPFIS::detail::Check(appfis_exchange(APPFIS::GetContext(), ga.GetSubGridsGPUData(), APPFIS_U8, 2, dims, 1, 1),
                          "exchange");
    APPFIS::kernels::GameOfLife(gw, 2, variant);
    APPFIS::Timer timer;
    timer.Start("GENERIC");
    for (int g = 0; g < gens; g++) APPFIS::Execute2D(config, ga.GetSubGrid(), ga.GetSubGridNext());
    timer.Stop("GENERIC");
    timer.Start("KERNEL");
    APPFIS::kernels::GameOfLife(gb, gens, variant);
    timer.Stop("KERNEL");
    int64_t la = 0, lb = 0;
    const int box[4] = {1, nx + 1, 1, ny + 1};
    APPFIS::detail::Check(appfis_reduce_sum_wide(APPFIS::GetContext(), ga.GetSubGridsGPUData(), APPFIS_U8, 2, dims,
                                                 box, &la), "live count");
    APPFIS::detail::Check(appfis_reduce_sum_wide(APPFIS::GetContext(), gb.GetSubGridsGPUData(), APPFIS_U8, 2, dims,
                                                 box, &lb), "live count");
    APPFIS::FinalizeExecution(ga, gb, gw);
    const size_t n = (size_t)(nx + 2) * (ny + 2);
    const bool same = memcmp(ga.GetSubGrid()->GetData(), gb.GetSubGrid()->GetData(), n) == 0;
    // uint8 Gather of the result, a binary file round trip, and a Scatter into a new grid
    ga.AllocateFullGrid();
    APPFIS::Gather(ga.GetFullGrid(), ga.GetSubGrid());
    bool sg = true;
    const uint8_t* sub = ga.GetSubGrid()->GetData();
    const uint8_t* full = ga.GetFullGrid()->GetData();
    for (int j = 0; j < ny && sg; j++) sg = memcmp(full + (size_t)j * nx, sub + (size_t)(j + 1) * (nx + 2) + 1, nx) == 0;
    const std::string path = APPFIS::CheckArgument(argc, argv, "output") ? "./output/gol.bin" : "./gol_roundtrip.bin";
    ga.GetFullGrid()->SaveBinaryFile(path);
    APPFIS::Grid<uint8_t> gc(APPFIS::DIM_2D, nx, ny, APPFIS::SINGLE);
    gc.AllocateFullGrid();
    gc.GetFullGrid()->LoadBinaryFile(path);
    APPFIS::Scatter(gc.GetFullGrid(), gc.GetSubGrid());
    const uint8_t* sc = gc.GetSubGrid()->GetData();
    for (int j = 1; j <= ny && sg; j++)
        sg = memcmp(sc + (size_t)j * (nx + 2) + 1, sub + (size_t)j * (nx + 2) + 1, nx) == 0;
    if (!APPFIS::CheckArgument(argc, argv, "output")) std::remove(path.c_str());
    printf("{\"app\": \"gol2d\", \"nx\": %d, \"ny\": %d, \"generations\": %d, \"live_count\": %lld, "
           "\"live_count_kernel\": %lld, \"generic_glups\": %.3f, \"kernel_glups\": %.3f, \"kernel_form\": \"%s\", "
           "\"bit_identical\": %s, \"scatter_gather_u8\": %s}\n",
           nx, ny, gens, (long long)la, (long long)lb, timer.CalculateBLUPS("GENERIC", nx * ny, gens),
           timer.CalculateBLUPS("KERNEL", nx * ny, gens), variant == APPFIS_PLAIN ? "plain" : "chosen",
           same ? "true" : "false", sg ? "true" : "false");
    APPFIS::Finalize();
    return same && sg ? 0 : 1;
}
