/*
 * Kernels_HIP.hpp -- the hand-written MI355X kernels of the config stencils, applied
 * to APPFIS grids in place of a generic Execute loop.
 *
 * Each call runs `n` sweeps on the grid's device buffers (current -> next -> ...) and
 * leaves the grid swapped exactly as n calls of Execute with GridsToUpdate(grid) would:
 * after an odd count the former NEXT buffer is current.  The grids must be between
 * InitializeExecution and FinalizeExecution.  Variants: APPFIS_PLAIN (one sweep per
 * launch, reference operation order) or APPFIS_CHOSEN (temporal blocking / packed /
 * mixed precision; see include/appfis_hip.h).
 */
#ifndef APPFIS_KERNELS_HIP_HPP
#define APPFIS_KERNELS_HIP_HPP

#include <cstdint>
#include <vector>

#include "Common.hpp"
#include "Grid.hpp"

namespace APPFIS
{
    namespace kernels
    {
        namespace detail_k
        {
            template <typename T>
            void RequireDual(Grid<T>& g, const char* what)
            {
                if (!EXECUTION_FLAG || g.GetGridType() != DUAL || !g.HasDeviceMirror()) {
                    std::cerr << "APPFIS::kernels::" << what << " needs a DUAL grid inside InitializeExecution"
                              << std::endl;
                    exit(EXIT_FAILURE);
                }
            }
            template <typename T>
            void SwapParity(Grid<T>& g, int n)
            {
                if (n & 1) g.Swap();
            }
        }  // namespace detail_k

        // HeatDiffusion 2D 5-point (ghost 1): u' = u + c*(uN + uS + uE + uW - 4u)
        template <typename T>
        void Heat2D(Grid<T>& g, double c, int iters, int variant = APPFIS_CHOSEN, int tsteps = 0)
        {
            detail_k::RequireDual(g, "Heat2D");
            Array<T>* h = g.GetSubGridCPU();
            const int nx = h->GetDim(X) - 2, ny = h->GetDim(Y) - 2;
            APPFIS::detail::Check(appfis_heat2d(GetContext(), APPFIS::detail::DtypeOf<T>(), g.GetSubGridsGPUData(),
                                                g.GetSubGridsNextGPUData(), nx, ny, c, iters, variant, tsteps),
                                  "appfis_heat2d");
            detail_k::SwapParity(g, iters);
        }

        // Many HeatDiffusion 2D grids of any sizes advanced together (one launch per chunk
        // of sweeps over a tile work-list spanning every grid, replayed as a hipGraph);
        // values identical to Heat2D on each grid.  All grids of one element type.
        template <typename T>
        void Heat2DEnsemble(std::vector<Grid<T>*>& grids, double c, int iters)
        {
            std::vector<int> nx, ny;
            std::vector<void*> b0, b1;
            for (Grid<T>* g : grids) {
                detail_k::RequireDual(*g, "Heat2DEnsemble");
                Array<T>* h = g->GetSubGridCPU();
                nx.push_back(h->GetDim(X) - 2);
                ny.push_back(h->GetDim(Y) - 2);
                b0.push_back(g->GetSubGridsGPUData());
                b1.push_back(g->GetSubGridsNextGPUData());
            }
            if (grids.empty()) return;
            APPFIS::detail::Check(appfis_heat2d_ensemble(GetContext(), APPFIS::detail::DtypeOf<T>(),
                                                         (int)grids.size(), nx.data(), ny.data(), b0.data(),
                                                         b1.data(), c, iters),
                                  "appfis_heat2d_ensemble");
            for (Grid<T>* g : grids) detail_k::SwapParity(*g, iters);
        }

        // GameOfLife B3/S23 on a periodic uint8 grid (ghost 1)
        inline void GameOfLife(Grid<uint8_t>& g, int gens, int variant = APPFIS_CHOSEN, int tsteps = 0)
        {
            detail_k::RequireDual(g, "GameOfLife");
            Array<uint8_t>* h = g.GetSubGridCPU();
            const int nx = h->GetDim(X) - 2, ny = h->GetDim(Y) - 2;
            APPFIS::detail::Check(appfis_gol(GetContext(), g.GetSubGridsGPUData(), g.GetSubGridsNextGPUData(), nx, ny,
                                             gens, variant, tsteps),
                                  "appfis_gol");
            detail_k::SwapParity(g, gens);
        }

        // FDTD3d radius-4 star (ghost 4): u' = c0 u + sum_r c_r (6 neighbours at r)
        template <typename T>
        void FDTD3d(Grid<T>& g, const double coeff5[5], int steps, int variant = APPFIS_CHOSEN)
        {
            detail_k::RequireDual(g, "FDTD3d");
            Array<T>* h = g.GetSubGridCPU();
            const int nx = h->GetDim(X) - 8, ny = h->GetDim(Y) - 8, nz = h->GetDim(Z) - 8;
            APPFIS::detail::Check(appfis_fdtd3d(GetContext(), APPFIS::detail::DtypeOf<T>(), g.GetSubGridsGPUData(),
                                                g.GetSubGridsNextGPUData(), nx, ny, nz, coeff5, steps, variant),
                                  "appfis_fdtd3d");
            detail_k::SwapParity(g, steps);
        }

        // Himeno 19-point Jacobi with gosa per sweep (p DUAL, the 12 coefficient grids
        // SINGLE, all with the boundary shell as ghost 1)
        template <typename T>
        std::vector<double> Himeno(Grid<T>& p, Grid<T>& a0, Grid<T>& a1, Grid<T>& a2, Grid<T>& a3, Grid<T>& b0,
                                   Grid<T>& b1, Grid<T>& b2, Grid<T>& c0, Grid<T>& c1, Grid<T>& c2, Grid<T>& bnd,
                                   Grid<T>& wrk1, double omega, int iters, int variant = APPFIS_PLAIN)
        {
            detail_k::RequireDual(p, "Himeno");
            Array<T>* h = p.GetSubGridCPU();
            void* arrays[14] = {p.GetSubGridsGPUData(), p.GetSubGridsNextGPUData(), a0.GetSubGridsGPUData(),
                                a1.GetSubGridsGPUData(), a2.GetSubGridsGPUData(), a3.GetSubGridsGPUData(),
                                b0.GetSubGridsGPUData(), b1.GetSubGridsGPUData(), b2.GetSubGridsGPUData(),
                                c0.GetSubGridsGPUData(), c1.GetSubGridsGPUData(), c2.GetSubGridsGPUData(),
                                bnd.GetSubGridsGPUData(), wrk1.GetSubGridsGPUData()};
            std::vector<double> gosa((size_t)(iters > 0 ? iters : 1));
            APPFIS::detail::Check(appfis_himeno(GetContext(), APPFIS::detail::DtypeOf<T>(), arrays, h->GetDim(X),
                                                h->GetDim(Y), h->GetDim(Z), omega, iters, variant, 0, gosa.data()),
                                  "appfis_himeno");
            gosa.resize((size_t)(iters > 0 ? iters : 0));
            detail_k::SwapParity(p, iters);
            return gosa;
        }

        // Triton shallow-water flood step chain (Roe flux, implicit Manning friction, wet /
        // dry fix; triton_appfis.hpp:316-386 of the reference) on SINGLE grids h, qx, qy
        // (updated in place), dem z (wall height in the ghost ring) and n2 (Manning n^2),
        // ghost 1, with optional hydrograph point sources.  Runs while the simulated time
        // is below params.end_time (> 0), at most max_steps steps; returns the steps run.
        template <typename T>
        int Flood(Grid<T>& h, Grid<T>& qx, Grid<T>& qy, Grid<T>& z, Grid<T>& n2, const appfis_flood_params& params,
                  const appfis_flood_sources* sources, int max_steps, int variant = APPFIS_CHOSEN)
        {
            Grid<T>* gs[5] = {&h, &qx, &qy, &z, &n2};
            for (Grid<T>* g : gs)
                if (!EXECUTION_FLAG || !g->HasDeviceMirror()) {
                    std::cerr << "APPFIS::kernels::Flood needs its grids inside InitializeExecution" << std::endl;
                    exit(EXIT_FAILURE);
                }
            Array<T>* a = h.GetSubGridCPU();
            const int nx = a->GetDim(X) - 2, ny = a->GetDim(Y) - 2;
            int done = 0;
            APPFIS::detail::Check(appfis_flood_run_src(GetContext(), APPFIS::detail::DtypeOf<T>(),
                                                       h.GetSubGridsGPUData(), qx.GetSubGridsGPUData(),
                                                       qy.GetSubGridsGPUData(), z.GetSubGridsGPUData(),
                                                       n2.GetSubGridsGPUData(), nx, ny, &params, sources, max_steps,
                                                       variant, nullptr, nullptr, &done),
                                  "appfis_flood_run_src");
            return done;
        }

        // One scenario of FloodEnsemble: its SINGLE grids, as Flood takes them.
        template <typename T>
        struct FloodScenario {
            Grid<T>* h;
            Grid<T>* qx;
            Grid<T>* qy;
            Grid<T>* z;
            Grid<T>* n2;
        };

        // Many flood scenarios of any sizes advanced together: every phase of a step is
        // one launch over a strip work-list spanning all scenarios, each scenario with its
        // own dt, and the steps run as replayed 20-step hipGraph chunks.  Each scenario's
        // values are the ones Flood gives it alone (no point sources).  Runs `steps`
        // steps, or fewer for a scenario whose simulated time reaches params.end_time
        // (when > 0) first.
        template <typename T>
        void FloodEnsemble(std::vector<FloodScenario<T>>& scen, const appfis_flood_params& params, int steps)
        {
            std::vector<int> nx, ny;
            std::vector<void*> h, qx, qy;
            std::vector<const void*> z, n2;
            for (FloodScenario<T>& s : scen) {
                Grid<T>* gs[5] = {s.h, s.qx, s.qy, s.z, s.n2};
                for (Grid<T>* g : gs)
                    if (!g || !EXECUTION_FLAG || !g->HasDeviceMirror()) {
                        std::cerr << "APPFIS::kernels::FloodEnsemble needs its grids inside InitializeExecution"
                                  << std::endl;
                        exit(EXIT_FAILURE);
                    }
                Array<T>* a = s.h->GetSubGridCPU();
                nx.push_back(a->GetDim(X) - 2);
                ny.push_back(a->GetDim(Y) - 2);
                h.push_back(s.h->GetSubGridsGPUData());
                qx.push_back(s.qx->GetSubGridsGPUData());
                qy.push_back(s.qy->GetSubGridsGPUData());
                z.push_back(s.z->GetSubGridsGPUData());
                n2.push_back(s.n2->GetSubGridsGPUData());
            }
            if (scen.empty()) return;
            APPFIS::detail::Check(appfis_flood_ensemble_run(GetContext(), APPFIS::detail::DtypeOf<T>(),
                                                            (int)scen.size(), nx.data(), ny.data(), h.data(),
                                                            qx.data(), qy.data(), z.data(), n2.data(), &params,
                                                            steps, nullptr, nullptr),
                                  "appfis_flood_ensemble_run");
        }
    }  // namespace kernels
}  // namespace APPFIS

#endif
