/*
 * HaloBuffer.hpp -- per-face halo buffers of a grid (API of the reference
 * Appfis/HaloBuffer.hpp:22-53).  Face sizes follow HaloBuffer.hpp:185-228: 1D faces hold
 * one cell, 2D faces dims[Y] (LEFT/RIGHT) and dims[X] (UP/DOWN), 3D faces dims[Y]*dims[Z]
 * (LEFT/RIGHT), dims[X]*dims[Y] (UP/DOWN) and dims[X]*dims[Z] (FRONT/BACK), each times the
 * ghost width.
 *
 * At world size 1 the device path never stages halos through these buffers: the
 * periodic self-exchange is one on-device copy per axis (appfis_exchange).  They are
 * kept, in plain host memory, for the host-side Exchange of the setup phase and for
 * source compatibility.
 */
#ifndef APPFIS_HALOBUFFER_HPP
#define APPFIS_HALOBUFFER_HPP

#include <iostream>
#include <vector>

#include "Common.hpp"
#include "Env.hpp"

namespace APPFIS
{
    template <class T>
    class HaloBuffer
    {
    private:
        int _nFace = 0;
        std::vector<int> _haloSizes;
        std::vector<std::vector<T>> _send;
        std::vector<std::vector<T>> _recv;

    public:
        HaloBuffer() {}
        ~HaloBuffer() { Reset(); }
        HaloBuffer(const HaloBuffer<T>& o) = default;
        HaloBuffer<T>& operator=(const HaloBuffer<T>& o) = default;

        void Set(int face, int index, T val) { _send[face][index] = val; }
        T Get(int face, int index) const { return _recv[face][index]; }

        void Create(int nDim, int* dims)
        {
            Reset();
            _nFace = 2 * nDim;
            _haloSizes.assign(_nFace, 0);
            if (_nFace == 2) {
                _haloSizes[LEFT] = _haloSizes[RIGHT] = 1;
            } else if (_nFace == 4) {
                _haloSizes[LEFT] = _haloSizes[RIGHT] = dims[Y];
                _haloSizes[UP] = _haloSizes[DOWN] = dims[X];
            } else if (_nFace == 6) {
                _haloSizes[LEFT] = _haloSizes[RIGHT] = dims[Y] * dims[Z];
                _haloSizes[UP] = _haloSizes[DOWN] = dims[X] * dims[Y];
                _haloSizes[FRONT] = _haloSizes[BACK] = dims[X] * dims[Z];
            } else {
                std::cerr << "Dimension not supported" << std::endl;
                _nFace = 0;
                return;
            }
            const int g = GetGhostLayer();
            if (g > 0)
                for (int f = 0; f < _nFace; f++) _haloSizes[f] *= g;
            _send.assign(_nFace, std::vector<T>());
            _recv.assign(_nFace, std::vector<T>());
            for (int f = 0; f < _nFace; f++) {
                _send[f].assign(_haloSizes[f], T());
                _recv[f].assign(_haloSizes[f], T());
            }
        }

        void Reset()
        {
            _nFace = 0;
            _haloSizes.clear();
            _send.clear();
            _recv.clear();
        }

        int GetHaloSize(int face) const { return _haloSizes[face]; }
        int* GetHaloSizes() { return _haloSizes.data(); }
        int GetFaceCount() const { return _nFace; }
        T* GetSendBuffer(int face) { return _send[face].data(); }
        T* GetRecvBuffer(int face) { return _recv[face].data(); }
    };
}  // namespace APPFIS

#endif
