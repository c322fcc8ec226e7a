/*
 * Communication.hpp -- data movement between ranks, at MPI world size 1
 * (API of the reference Appfis/Communication.hpp:24-46).
 *
 * One rank owns the whole domain, so
 *   Scatter / Gather   copy the full array into / out of the interior of the local
 *                      subgrid (any element type; the reference's MPI subarray path
 *                      only takes int/float/double, Communication.hpp:67-83);
 *   Exchange           is the periodic self-exchange of the host buffers (X, then Y,
 *                      then Z, each face over the full extent of the other axes), and
 *                      nothing at all on a non-periodic axis (no neighbour);
 *   AllReduce          returns its argument;
 *   ProcessSync        has nothing to wait for.
 * The device-side exchange used during execution is appfis_exchange (Runtime.hpp).
 */
#ifndef APPFIS_COMMUNICATION_HPP
#define APPFIS_COMMUNICATION_HPP

#include <cstring>
#include <iostream>

#include "Array.hpp"
#include "Common.hpp"
#include "HaloBuffer.hpp"

namespace APPFIS
{
    namespace detail
    {
        template <typename T>
        void CopyInterior(Array<T>* global, Array<T>* local, bool toLocal)
        {
            const int n = global->GetDimCount();
            const int g = GetGhostLayer();
            const int gx = global->GetDim(X), gy = n > 1 ? global->GetDim(Y) : 1, gz = n > 2 ? global->GetDim(Z) : 1;
            const int lx = local->GetDim(X), ly = n > 1 ? local->GetDim(Y) : 1;
            const int ox = g, oy = n > 1 ? g : 0, oz = n > 2 ? g : 0;
            T* G = global->GetData();
            T* L = local->GetData();
            for (int k = 0; k < gz; k++)
                for (int j = 0; j < gy; j++) {
                    T* grow = G + (size_t)gx * ((size_t)j + (size_t)gy * k);
                    T* lrow = L + (size_t)ox + (size_t)lx * ((size_t)(j + oy) + (size_t)ly * (k + oz));
                    if (toLocal) memcpy(lrow, grow, sizeof(T) * gx);
                    else memcpy(grow, lrow, sizeof(T) * gx);
                }
        }
    }  // namespace detail

    template <typename T>
    void Scatter(Array<T>* global, Array<T>* local)
    {
        if (global->GetDimCount() != GetDimCount()) {
            std::cerr << "Partition failed. Data dimension and partition are different." << std::endl;
            exit(EXIT_FAILURE);
        }
        if (global->GetData() == nullptr || local->GetData() == nullptr) return;
        detail::CopyInterior(global, local, true);
    }

    template <typename T>
    void Gather(Array<T>* global, Array<T>* local)
    {
        if (global->GetDimCount() != GetDimCount()) {
            std::cerr << "Gather failed. Data dimension and partition are different." << std::endl;
            exit(EXIT_FAILURE);
        }
        if (local->GetData() == nullptr) return;
        if (global->GetData() == nullptr) global->AllocateMemory();
        detail::CopyInterior(global, local, false);
    }

    // host-side periodic self-exchange (pack face1 -> unpack face2, pack face2 -> unpack
    // face1 per axis, as Communication.hpp:613-678 with Runtime.hpp:917-1173 at size 1)
    template <typename T>
    void Exchange(Array<T>* arr, HaloBuffer<T>* buffer)
    {
        (void)buffer;
        if (!IsPeriodic()) return;
        const int n = arr->GetDimCount();
        const int g = GetGhostLayer();
        const int d[3] = {arr->GetDim(X), n > 1 ? arr->GetDim(Y) : 1, n > 2 ? arr->GetDim(Z) : 1};
        T* a = arr->GetData();
        auto at = [&](int i, int j, int k) -> T& { return a[(size_t)i + (size_t)d[0] * ((size_t)j + (size_t)d[1] * k)]; };
        for (int ax = 0; ax < n; ax++) {
            for (int l = 0; l < g; l++) {
                for (int k = 0; k < d[2]; k++)
                    for (int j = 0; j < d[1]; j++)
                        for (int i = 0; i < d[0]; i++) {
                            const int c[3] = {i, j, k};
                            if (c[ax] != l) continue;
                            int lo[3] = {i, j, k}, hi[3] = {i, j, k}, src_lo[3] = {i, j, k}, src_hi[3] = {i, j, k};
                            lo[ax] = l;                         // low ghost layer
                            hi[ax] = d[ax] - g + l;             // high ghost layer
                            src_lo[ax] = d[ax] - 2 * g + l;     // last interior layers
                            src_hi[ax] = g + l;                 // first interior layers
                            at(hi[0], hi[1], hi[2]) = at(src_hi[0], src_hi[1], src_hi[2]);
                            at(lo[0], lo[1], lo[2]) = at(src_lo[0], src_lo[1], src_lo[2]);
                        }
            }
        }
    }

    template <typename T>
    T AllReduce(T local, REDUCTION rType)
    {
        (void)rType;
        return local;
    }

    template <typename T>
    void ExchangeInitial(Array<T>* arr, HaloBuffer<T>* buffer)
    {
        Exchange(arr, buffer);
    }

    inline void ProcessSync() {}
}  // namespace APPFIS

#endif
