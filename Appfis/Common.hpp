/*
 * Common.hpp -- free functions over the topology (reference Appfis/Common.hpp:19-226).
 */
#ifndef APPFIS_COMMON_HPP
#define APPFIS_COMMON_HPP

#include "Topology.hpp"

namespace APPFIS
{
    inline void Initialize(int argc, char* argv[], int dataDim, int partitionDim, Attribute attribute = ATTRIBUTE)
    {
        Topology::GetInstance().Initialize(argc, argv, dataDim, partitionDim, attribute);
    }

    inline void Finalize() { Topology::GetInstance().Finalize(); }

    inline int GetRank() { return Topology::GetInstance()._rank; }
    inline int GetMasterRank() { return Topology::GetInstance()._master; }
    inline int GetRankCount() { return Topology::GetInstance()._nRank; }
    inline int GetNeighbourCount() { return Topology::GetInstance()._nNeighbour; }
    inline int GetDimCount() { return Topology::GetInstance()._nDim; }
    inline int GetPartitionDim() { return Topology::GetInstance()._pDim; }
    inline int GetGhostLayer() { return Topology::GetInstance()._ghostLayer; }
    inline bool IsMaster() { return Topology::GetInstance()._rank == Topology::GetInstance()._master; }
    inline bool IsOverlap() { return Topology::GetInstance()._overlap; }
    inline bool IsPeriodic() { return Topology::GetInstance()._periodic; }
    inline int* GetBlocks() { return Topology::GetInstance()._blocks; }
    inline int* GetCoords() { return Topology::GetInstance()._coords; }
    inline int GetNeighboursRank(int face) { return Topology::GetInstance()._neighboursRanks[face]; }
    inline Communicator& GetCommunicator() { return Topology::GetInstance()._comm; }

    // the single HIP stream of this process (the reference's GetCudaStream)
    inline hipStream_t& GetHipStream() { return Topology::GetInstance()._stream; }

    // the C ABI context behind the stream (include/appfis_hip.h)
    inline appfis_ctx* GetContext()
    {
        Topology& t = Topology::GetInstance();
        if (!t._ctx) {
            std::cerr << "APPFIS::Initialize must be called before any device work." << std::endl;
            exit(EXIT_FAILURE);
        }
        return t._ctx;
    }

    inline void SetMasterRank(int master)
    {
        Topology& t = Topology::GetInstance();
        if (master < 0 || master >= t._nRank)
            std::cerr << "Setting master rank to " << master << " has failed. Default value 0." << std::endl;
        else
            t._master = master;
    }

    // local subgrid extents: 2*ghost + full/blocks (+1 on the low coordinates that
    // take the remainder), with the reference's fatal checks (Common.hpp:197-226)
    inline void GetLocalPartitionDims(int nDim, int* coords, int* fullDataDims, int* subDataDims)
    {
        int* blocks = GetBlocks();
        const int g = GetGhostLayer();
        for (int i = 0; i < nDim; i++) {
            if (fullDataDims[i] < blocks[i]) {
                std::cerr << "Number of process in a dimension: " << blocks[i]
                          << " is larger than data dimension: " << fullDataDims[i] << std::endl;
                exit(EXIT_FAILURE);
            }
            if (fullDataDims[i] / blocks[i] < 2 * g) {
                std::cerr << "Subgrid Data dimension should be at least: " << (2 * g)
                          << ". Found: " << (fullDataDims[i] / blocks[i]) << std::endl;
                exit(EXIT_FAILURE);
            }
        }
        for (int i = 0; i < nDim; i++) {
            subDataDims[i] = 2 * g + fullDataDims[i] / blocks[i];
            const int rem = fullDataDims[i] % blocks[i];
            if (coords[i] < rem && rem > 0) subDataDims[i]++;
        }
    }

    inline void GetLocalPartitionDims(int nDim, int* fullDataDims, int* subDataDims)
    {
        if (nDim != GetDimCount()) {
            std::cerr << "Dimension not matched." << std::endl;
            return;
        }
        GetLocalPartitionDims(nDim, GetCoords(), fullDataDims, subDataDims);
    }

    namespace detail
    {
        inline void SyncDevice()
        {
            Topology& t = Topology::GetInstance();
            if (t._ctx) Check(appfis_sync(t._ctx), "device synchronisation");
        }
    }  // namespace detail
}  // namespace APPFIS

#endif
