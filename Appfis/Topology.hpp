/*
 * Topology.hpp -- process topology singleton of the APPFIS API at MPI world size 1
 * (reference Appfis/Topology.hpp:104-222).
 *
 * One process drives one MI355X: Initialize picks the device (rank % device count,
 * Topology.hpp:124-149) and creates the library context that owns the single HIP
 * stream every kernel runs on (Topology.hpp:151).  The Cartesian decomposition of
 * the reference collapses to one block per axis: coords 0, neighbours = self on a
 * periodic axis and "no neighbour" (PROC_NULL) otherwise, exactly what
 * MPI_Cart_create/MPI_Cart_shift return on one rank.
 */
#ifndef APPFIS_TOPOLOGY_HPP
#define APPFIS_TOPOLOGY_HPP

#include <cstdlib>
#include <iostream>
#include <string>

#include "Attribute.hpp"
#include "Env.hpp"
#include "appfis_hip.h"

namespace APPFIS
{
    constexpr int PROC_NULL = -1;  // "no neighbour" (MPI_PROC_NULL of the reference)

    // single-rank communicator handle (the reference returns an MPI_Comm)
    struct Communicator {
        int rank = 0;
        int size = 1;
    };

    namespace detail
    {
        // C ABI status -> the reference's fatal-misuse convention (stderr + exit)
        inline void Check(int status, const char* what)
        {
            if (status != APPFIS_OK) {
                std::cerr << "APPFIS: " << what << " failed: " << appfis_status_string(status) << ": "
                          << appfis_last_error() << std::endl;
                exit(EXIT_FAILURE);
            }
        }
    }  // namespace detail

    class Topology
    {
    public:
        int _rank = 0;
        int _master = 0;
        int _nRank = 1;
        int _nNeighbour = 0;
        int _nDim = 0;
        int _pDim = 0;
        int _ghostLayer = 1;
        bool _overlap = false;
        bool _periodic = false;
        int _blocks[3] = {1, 1, 1};
        int _periods[3] = {0, 0, 0};
        int _coords[3] = {0, 0, 0};
        int _neighboursRanks[6] = {PROC_NULL, PROC_NULL, PROC_NULL, PROC_NULL, PROC_NULL, PROC_NULL};
        Communicator _comm;
        appfis_ctx* _ctx = nullptr;
        hipStream_t _stream = nullptr;
        bool _initialized = false;

        static Topology& GetInstance()
        {
            static Topology instance;
            return instance;
        }
        Topology(Topology const&) = delete;
        void operator=(Topology const&) = delete;

        ~Topology()
        {
            if (_ctx) {
                appfis_ctx_destroy(_ctx);
                _ctx = nullptr;
            }
        }

        void Initialize(int argc, char* argv[], int dataDim, int partitionDim, Attribute attribute = ATTRIBUTE)
        {
            (void)argc;
            (void)argv;
            if (partitionDim > dataDim) {
                std::cerr << "Partition dimension can not be larger than data dimension." << std::endl;
                exit(EXIT_FAILURE);
            }
            if (dataDim < DIM_1D || dataDim > DIM_3D) {
                std::cerr << "Data dimension must be 1, 2 or 3." << std::endl;
                exit(EXIT_FAILURE);
            }
            _nDim = dataDim;
            _pDim = partitionDim;
            _ghostLayer = attribute.ghostLayer;
            _overlap = attribute.overlap;
            _periodic = attribute.periodic;
            _nNeighbour = 2 * _nDim;
            for (int i = 0; i < 3; i++) {
                _blocks[i] = 1;
                _coords[i] = 0;
                _periods[i] = i < _nDim ? (int)_periodic : 0;
            }
            for (int f = 0; f < 6; f++) _neighboursRanks[f] = _periodic && f < _nNeighbour ? _rank : PROC_NULL;
            if (_rank == _master) {
                std::string s = "Process Dims:";
                for (int i = 0; i < _nDim; i++) s += " " + std::to_string(_blocks[i]);
                std::cout << s << std::endl;
            }
            if (!_ctx) detail::Check(appfis_ctx_create(_rank, &_ctx), "device context creation");
            _stream = (hipStream_t)appfis_ctx_stream(_ctx);
            _initialized = true;
        }

        void Finalize()
        {
            if (_ctx) detail::Check(appfis_sync(_ctx), "device synchronisation");
        }

    private:
        Topology() {}
    };
}  // namespace APPFIS

#endif
