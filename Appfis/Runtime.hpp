/*
 * Runtime.hpp -- execution lifecycle of the APPFIS API (reference
 * Appfis/Runtime.hpp:23-170): InitializeExecution / FinalizeExecution and, through
 * Runtime_HIP.hpp, Execute{1,2,3}D, Execute{1,2,3}DReduce and HaloExchange.
 */
#ifndef APPFIS_RUNTIME_HPP
#define APPFIS_RUNTIME_HPP

#include "Common.hpp"
#include "Communication.hpp"
#include "Config.hpp"
#include "Grid.hpp"
#include "Runtime_HIP.hpp"

namespace APPFIS
{
    // Uploads the grid (current and, for DUAL, next buffer) to the device and switches
    // GetSubGrid() to the device views.  The initial halo exchange of both buffers
    // (Runtime.hpp:124-142) runs on the device after the upload.
    template <typename T>
    void InitializeExecution(Grid<T>& grid)
    {
        EXECUTION_FLAG = true;
        grid.CreateSubGridsGPU();
        if (GetGhostLayer() > 0 && grid.GetGridType() != FLAT && IsPeriodic()) {
            Array<T>* h = grid.GetSubGridCPU();
            detail::Check(appfis_exchange(GetContext(), grid.GetSubGridsGPUData(), detail::DtypeOf<T>(),
                                          h->GetDimCount(), h->GetDims(), GetGhostLayer(), 1),
                          "initial halo exchange");
            if (grid.GetGridType() == DUAL)
                detail::Check(appfis_exchange(GetContext(), grid.GetSubGridsNextGPUData(), detail::DtypeOf<T>(),
                                              h->GetDimCount(), h->GetDims(), GetGhostLayer(), 1),
                              "initial halo exchange");
        }
    }

    template <typename T, typename... TArgs>
    void InitializeExecution(Grid<T>& grid, TArgs&&... args)
    {
        InitializeExecution(grid);
        InitializeExecution(args...);
    }

    // Waits for the device and downloads the CURRENT buffer (Grid.hpp:471-479).
    template <typename T>
    void FinalizeExecution(Grid<T>& grid)
    {
        EXECUTION_FLAG = false;
        detail::SyncDevice();
        grid.CopySubGridsGPUToCPU();
    }

    template <typename T, typename... TArgs>
    void FinalizeExecution(Grid<T>& grid, TArgs&&... args)
    {
        FinalizeExecution(grid);
        FinalizeExecution(args...);
    }
}  // namespace APPFIS

#endif
