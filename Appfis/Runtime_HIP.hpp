/*
 * Runtime_HIP.hpp -- device side of the APPFIS runtime on MI355X (replaces the
 * reference Appfis/Runtime_CUDA.hpp; a new design, not a translation).
 *
 * User stencils are compile-time functors, so the generic launch stays a template in
 * the application's translation unit (built by hipcc for gfx950), with x along the
 * 64-wide wavefront so consecutive lanes touch consecutive addresses.  In 2D a block of
 * 64 x 4 threads covers 64 x 16 points (each thread 4 rows, 4 apart); in 3D a block of
 * 64 x 4 threads covers 64 x 4 points on 8 consecutive planes (each thread marches them
 * in order, so the z-neighbours a functor reads are still in its CU's L1); in 1D one
 * thread per point.  Blocks are taken in XCD-contiguous order: workgroup b runs on XCD
 * b % 8, and the remap gives each XCD one contiguous run of tiles, so tiles that read
 * each other's halo points share an L2.  The points of one Execute are independent
 * (the functor reads the current buffers and writes the next ones), so the order in
 * which they are computed does not change a result.
 * Everything else goes through the C ABI (include/appfis_hip.h): allocation and copies,
 * the single-rank halo exchange, boundary fill and the single-pass reductions.
 *
 * Differences from the reference GPU launcher (Runtime_CUDA.hpp:231-551, 554-740):
 *   - no host synchronisation after every Execute: kernels, exchanges and swaps are
 *     queued on the one stream; the host waits only when a value must come back
 *     (Execute*Reduce, FinalizeExecution, Timer::Stop);
 *   - the periodic exchange is an on-device copy per axis, not pack -> D2H -> MPI ->
 *     H2D -> unpack;
 *   - reductions scan only the Config box and never use the grid's NEXT buffer as
 *     scratch (the reference overwrites it, Runtime_CUDA.hpp:561-562, 624-625).
 */
#ifndef APPFIS_RUNTIME_HIP_HPP
#define APPFIS_RUNTIME_HIP_HPP

#include <iostream>

#include "Common.hpp"
#include "Config.hpp"
#include "Grid.hpp"

namespace APPFIS
{
    namespace hip_detail
    {
        template <typename... TArgs>
        __global__ __launch_bounds__(256) void Kernel1D(int sx, int ex, TArgs... args)
        {
            const int i = sx + (int)(blockIdx.x * blockDim.x + threadIdx.x);
            if (i < ex) STENCIL(i, args...);
        }

        constexpr int ROWS_2D = 4;    // rows per thread in 2D (block tile 64 x 16)
        constexpr int PLANES_3D = 8;  // planes per thread in 3D (block tile 64 x 4 x 8)

        // Tile of flat block b of nb: every XCD (b % 8) takes one contiguous run.
        __device__ __forceinline__ unsigned XcdTile(unsigned b, unsigned nb)
        {
            const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
            return x * q + (x < r ? x : r) + (b >> 3);
        }

        template <typename... TArgs>
        __global__ __launch_bounds__(256) void Kernel2D(int sx, int ex, int sy, int ey, int tiles_x, TArgs... args)
        {
            const int t = (int)XcdTile(blockIdx.x, gridDim.x);
            const int ty = t / tiles_x, tx = t - ty * tiles_x;
            const int i = sx + tx * 64 + (int)threadIdx.x;
            if (i >= ex) return;
            const int j0 = sy + ty * (4 * ROWS_2D) + (int)threadIdx.y;
#pragma unroll
            for (int r = 0; r < ROWS_2D; r++) {
                const int j = j0 + 4 * r;
                if (j < ey) STENCIL(i, j, args...);
            }
        }

        template <typename... TArgs>
        __global__ __launch_bounds__(256) void Kernel3D(int sx, int ex, int sy, int ey, int sz, int ez, int tiles_x,
                                                        int tiles_y, TArgs... args)
        {
            const int t = (int)XcdTile(blockIdx.x, gridDim.x);
            const int txy = t % (tiles_x * tiles_y), tz = t / (tiles_x * tiles_y);
            const int ty = txy / tiles_x, tx = txy - ty * tiles_x;
            const int i = sx + tx * 64 + (int)threadIdx.x;
            const int j = sy + ty * 4 + (int)threadIdx.y;
            if (i >= ex || j >= ey) return;
            const int k0 = sz + tz * PLANES_3D;
            const int k1 = k0 + PLANES_3D < ez ? k0 + PLANES_3D : ez;
            for (int k = k0; k < k1; k++) STENCIL(i, j, k, args...);
        }

        inline void CheckLaunch(const char* what)
        {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                std::cerr << "APPFIS: " << what << " launch failed: " << hipGetErrorString(e) << std::endl;
                exit(EXIT_FAILURE);
            }
        }

        inline void RequireExecution(const char* what)
        {
            if (!EXECUTION_FLAG) {
                std::cerr << "APPFIS: " << what << " called outside InitializeExecution/FinalizeExecution" << std::endl;
                exit(EXIT_FAILURE);
            }
        }

        inline unsigned Blocks(int n, int per) { return (unsigned)((n + per - 1) / per); }

        // exchange of the NEXT device buffer, then swap (ExchangeHelper + SwapHelper of
        // the reference, Runtime.hpp:1393-1419)
        template <typename T>
        void ExchangeNextAndSwap(Grid<T>* g)
        {
            if (g->GetGridType() != FLAT && GetGhostLayer() > 0 && IsPeriodic() && g->GetSubGridsNextGPUData()) {
                Array<T>* host = g->GetSubGridNextCPU();
                detail::Check(appfis_exchange(GetContext(), g->GetSubGridsNextGPUData(), detail::DtypeOf<T>(),
                                              host->GetDimCount(), host->GetDims(), GetGhostLayer(), 1),
                              "device halo exchange");
            }
            g->Swap();
        }

        template <typename T>
        void ExchangeAndSwap(Config<T>& config)
        {
            for (Grid<T>* g : config.updatedGrids) ExchangeNextAndSwap(g);
        }

        template <typename T>
        void Reduce(Config<T>& config, T& var, REDUCTION rType, Grid<T>& grid, int nd)
        {
            RequireExecution("Execute*Reduce");
            static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value ||
                              std::is_same<T, int32_t>::value || std::is_same<T, uint8_t>::value,
                          "reductions take double, float, int or uint8 grids");
            Array<T>* host = grid.GetSubGridCPU();
            int box[6] = {config.GetStartX(), config.GetEndX(), 0, 1, 0, 1};
            if (nd > 1) {
                box[2] = config.GetStartY();
                box[3] = config.GetEndY();
            }
            if (nd > 2) {
                box[4] = config.GetStartZ();
                box[5] = config.GetEndZ();
            }
            const int op = rType == MAX ? APPFIS_MAX : rType == MIN ? APPFIS_MIN : APPFIS_SUM;
            T local{};
            detail::Check(appfis_reduce(GetContext(), grid.GetSubGridsGPUData(), detail::DtypeOf<T>(),
                                        host->GetDimCount(), host->GetDims(), box, op, &local),
                          "device reduction");
            var = AllReduce(local, rType);
        }
    }  // namespace hip_detail

    template <typename T, typename... TArgs>
    void Execute1D(Config<T>& config, TArgs... args)
    {
        hip_detail::RequireExecution("Execute1D");
        const int sx = config.GetStartX(), ex = config.GetEndX();
        if (ex > sx) {
            hipLaunchKernelGGL(hip_detail::Kernel1D<TArgs...>, dim3(hip_detail::Blocks(ex - sx, 256)), dim3(256), 0,
                               GetHipStream(), sx, ex, args...);
            hip_detail::CheckLaunch("Execute1D");
        }
        hip_detail::ExchangeAndSwap(config);
    }

    template <typename T, typename... TArgs>
    void Execute2D(Config<T>& config, TArgs... args)
    {
        hip_detail::RequireExecution("Execute2D");
        const int sx = config.GetStartX(), ex = config.GetEndX();
        const int sy = config.GetStartY(), ey = config.GetEndY();
        if (ex > sx && ey > sy) {
            const unsigned tx = hip_detail::Blocks(ex - sx, 64), ty = hip_detail::Blocks(ey - sy, 4 * hip_detail::ROWS_2D);
            hipLaunchKernelGGL(hip_detail::Kernel2D<TArgs...>, dim3(tx * ty), dim3(64, 4), 0, GetHipStream(), sx, ex, sy,
                               ey, (int)tx, args...);
            hip_detail::CheckLaunch("Execute2D");
        }
        hip_detail::ExchangeAndSwap(config);
    }

    template <typename T, typename... TArgs>
    void Execute3D(Config<T>& config, TArgs... args)
    {
        hip_detail::RequireExecution("Execute3D");
        const int sx = config.GetStartX(), ex = config.GetEndX();
        const int sy = config.GetStartY(), ey = config.GetEndY();
        const int sz = config.GetStartZ(), ez = config.GetEndZ();
        if (ex > sx && ey > sy && ez > sz) {
            const unsigned tx = hip_detail::Blocks(ex - sx, 64), ty = hip_detail::Blocks(ey - sy, 4),
                           tz = hip_detail::Blocks(ez - sz, hip_detail::PLANES_3D);
            hipLaunchKernelGGL(hip_detail::Kernel3D<TArgs...>, dim3(tx * ty * tz), dim3(64, 4, 1), 0, GetHipStream(), sx,
                               ex, sy, ey, sz, ez, (int)tx, (int)ty, args...);
            hip_detail::CheckLaunch("Execute3D");
        }
        hip_detail::ExchangeAndSwap(config);
    }

    template <typename T>
    void Execute1DReduce(Config<T>& config, T& var, REDUCTION rType, Grid<T>& grid)
    {
        hip_detail::Reduce(config, var, rType, grid, 1);
    }

    template <typename T>
    void Execute2DReduce(Config<T>& config, T& var, REDUCTION rType, Grid<T>& grid)
    {
        hip_detail::Reduce(config, var, rType, grid, 2);
    }

    template <typename T>
    void Execute3DReduce(Config<T>& config, T& var, REDUCTION rType, Grid<T>& grid)
    {
        hip_detail::Reduce(config, var, rType, grid, 3);
    }

    // HaloExchange: exchange the NEXT buffer, then swap (reference Runtime.hpp:175-191)
    template <typename T>
    void HaloExchange(Grid<T>* grid)
    {
        hip_detail::RequireExecution("HaloExchange");
        hip_detail::ExchangeNextAndSwap(grid);
    }
}  // namespace APPFIS

#endif
