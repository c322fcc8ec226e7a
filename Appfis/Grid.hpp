/*
 * Grid.hpp -- the APPFIS grid (API of the reference Appfis/Grid.hpp:23-98): a full
 * array on the master, one (SINGLE/FLAT) or two (DUAL, ping-pong) local subgrids with a
 * ghost frame, a halo buffer, and the device mirror used during execution.
 *
 * Device mirror (reference Grid.hpp:403-479): InitializeExecution uploads every local
 * buffer into device memory from the C ABI and writes, next to it, a device image of an
 * Array view of that buffer; GetSubGrid()/GetSubGridNext() return those images while
 * EXECUTION_FLAG is set.  FinalizeExecution downloads the current buffer only, as the
 * reference does.
 */
#ifndef APPFIS_GRID_HPP
#define APPFIS_GRID_HPP

#include <cstdint>
#include <cstring>
#include <iostream>
#include <type_traits>
#include <utility>
#include <vector>

#include "Array.hpp"
#include "Common.hpp"
#include "Env.hpp"
#include "HaloBuffer.hpp"

namespace APPFIS
{
    namespace detail
    {
        // C ABI element type of T: exact for the reduction types, by size otherwise
        // (boundary fill and halo copies only move bytes)
        template <typename T>
        constexpr int DtypeOf()
        {
            if (std::is_same<T, double>::value) return APPFIS_F64;
            if (std::is_same<T, float>::value) return APPFIS_F32;
            if (std::is_same<T, int32_t>::value) return APPFIS_I32;
            if (std::is_same<T, uint8_t>::value) return APPFIS_U8;
            if (std::is_same<T, int64_t>::value) return APPFIS_I64;
            return sizeof(T) == 1 ? APPFIS_U8 : sizeof(T) == 2 ? APPFIS_F16 : sizeof(T) == 4 ? APPFIS_I32
                                                                                           : APPFIS_I64;
        }
        template <typename T>
        constexpr bool BytesOnlyType()
        {
            return sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8;
        }
    }  // namespace detail

    template <class T>
    class Grid
    {
    private:
        int _nDim = 0;
        int _currIndex = 0;
        int _nextIndex = 0;
        Array<T> _full;
        std::vector<Array<T>> _holder;
        HaloBuffer<T> _haloBuffer;
        GRID_TYPE _gridType = SINGLE;
        Array<T>* _dev[2] = {nullptr, nullptr};  // device images of the Array views
        T* _devData[2] = {nullptr, nullptr};

    public:
        Grid() {}
        ~Grid() { FreeDevice(); }

        Grid(int nDim, int dimX, GRID_TYPE gridType) { Create(nDim, dimX, gridType); }
        Grid(int nDim, int dimX, int dimY, GRID_TYPE gridType) { Create(nDim, dimX, dimY, gridType); }
        Grid(int nDim, int dimX, int dimY, int dimZ, GRID_TYPE gridType) { Create(nDim, dimX, dimY, dimZ, gridType); }
        Grid(int nDim, int* dims, GRID_TYPE gridType) { Create(nDim, dims, gridType); }

        Grid(const Grid<T>& grid)
        {
            std::cout << "Grid copy" << std::endl;
            CopyHost(grid);
        }

        Grid& operator=(const Grid<T>& grid)
        {
            std::cout << "Grid assign" << std::endl;
            if (this != &grid) {
                FreeDevice();
                Reset();
                CopyHost(grid);
            }
            return *this;
        }

        void Create(int nDim, int dimX, GRID_TYPE gridType)
        {
            int d[1] = {dimX};
            Create(nDim, d, gridType);
        }
        void Create(int nDim, int dimX, int dimY, GRID_TYPE gridType)
        {
            int d[2] = {dimX, dimY};
            Create(nDim, d, gridType);
        }
        void Create(int nDim, int dimX, int dimY, int dimZ, GRID_TYPE gridType)
        {
            int d[3] = {dimX, dimY, dimZ};
            Create(nDim, d, gridType);
        }
        void Create(int nDim, int* dims, GRID_TYPE gridType)
        {
            _nDim = nDim;
            _gridType = gridType;
            _full.Resize(nDim, dims);
            CreateSubGrids();
        }

        int GetDimCount() const { return _nDim; }
        Array<T>* GetFullGrid() { return &_full; }

        Array<T>* GetSubGrid()
        {
            if (EXECUTION_FLAG && _dev[_currIndex]) return _dev[_currIndex];
            return &_holder.at(_currIndex);
        }
        Array<T>* GetSubGridNext()
        {
            if (EXECUTION_FLAG && _dev[_nextIndex]) return _dev[_nextIndex];
            return &_holder.at(_nextIndex);
        }
        Array<T>* GetSubGridCPU() { return &_holder.at(_currIndex); }
        Array<T>* GetSubGridNextCPU() { return &_holder.at(_nextIndex); }

        GRID_TYPE GetGridType() const { return _gridType; }
        HaloBuffer<T>* GetBuffer() { return &_haloBuffer; }

        void AllocateFullGrid() { _full.AllocateMemory(); }

        // local extents per Common.hpp (FLAT: the full extents), zero-filled buffers
        void CreateSubGrids()
        {
            _holder = std::vector<Array<T>>(2);
            int sub[3] = {0, 0, 0};
            if (_gridType == FLAT) {
                for (int i = 0; i < _nDim; i++) sub[i] = _full.GetDims()[i];
            } else {
                GetLocalPartitionDims(_nDim, _full.GetDims(), sub);
            }
            _currIndex = 0;
            _nextIndex = 0;
            _holder.at(0).Resize(_nDim, sub);
            _holder.at(0).AllocateMemory();
            if (_gridType == DUAL) {
                _nextIndex = 1;
                _holder.at(1).Resize(_nDim, sub);
                _holder.at(1).AllocateMemory();
            }
            _haloBuffer.Create(_nDim, sub);
        }

        // ---- boundary fill (reference Grid.hpp:530-830) ------------------------------
        // On the host buffers before execution; on the device buffers during execution.
        void SetBoundary(T value)
        {
            for (int f = 0; f < _haloBuffer.GetFaceCount(); f++) SetBoundary(f, value);
        }

        void SetBoundary(int face, T value)
        {
            if (DeviceActive()) {
                DeviceSetBoundary(_currIndex, face, &value);
                if (_gridType == DUAL) DeviceSetBoundary(_nextIndex, face, &value);
                return;
            }
            SetBoundary(&_holder.at(_currIndex), face, value);
            if (_gridType == DUAL) SetBoundary(&_holder.at(_nextIndex), face, value);
        }

        void SetBoundary(Array<T>* arr, int face, T value)
        {
            const int g = GetGhostLayer();
            const int n = arr->GetDimCount();
            const int dx = arr->GetDim(X), dy = n > 1 ? arr->GetDim(Y) : 1, dz = n > 2 ? arr->GetDim(Z) : 1;
            int ax, high;
            if (!FaceAxis(n, face, &ax, &high)) return;
            const int ext[3] = {dx, dy, dz};
            for (int l = 0; l < g; l++) {
                const int layer = high ? ext[ax] - g + l : l;
                ForFace(ext, ax, layer, [&](int i, int j, int k) { Put(arr, n, i, j, k, value); });
            }
        }

        void SetBoundaryCopy()
        {
            for (int f = 0; f < _haloBuffer.GetFaceCount(); f++) SetBoundaryCopy(f);
        }

        void SetBoundaryCopy(int face)
        {
            if (DeviceActive()) {
                DeviceSetBoundaryCopy(_currIndex, face);
                if (_gridType == DUAL) DeviceSetBoundaryCopy(_nextIndex, face);
                return;
            }
            SetBoundaryCopy(&_holder.at(_currIndex), face);
            if (_gridType == DUAL) SetBoundaryCopy(&_holder.at(_nextIndex), face);
        }

        // g walks ghost-1 .. 0; low faces copy layer g+1, high faces layer dim-ghost+g-1
        void SetBoundaryCopy(Array<T>* arr, int face)
        {
            const int g = GetGhostLayer();
            const int n = arr->GetDimCount();
            const int dx = arr->GetDim(X), dy = n > 1 ? arr->GetDim(Y) : 1, dz = n > 2 ? arr->GetDim(Z) : 1;
            int ax, high;
            if (!FaceAxis(n, face, &ax, &high)) return;
            const int ext[3] = {dx, dy, dz};
            for (int l = g - 1; l >= 0; l--) {
                const int dst = high ? ext[ax] - g + l : l;
                const int src = high ? ext[ax] - g + l - 1 : l + 1;
                ForFace(ext, ax, dst, [&](int i, int j, int k) {
                    int s[3] = {i, j, k};
                    s[ax] = src;
                    Put(arr, n, i, j, k, Take(arr, n, s[0], s[1], s[2]));
                });
            }
        }

        void Reset()
        {
            _nDim = 0;
            _full.Reset();
            for (auto& h : _holder) h.Reset();
            _haloBuffer.Reset();
            _gridType = SINGLE;
            _currIndex = 0;
            _nextIndex = 0;
        }

        void Swap() { std::swap(_currIndex, _nextIndex); }

        // ---- device mirror ---------------------------------------------------------
        void CreateSubGridsGPU()
        {
            appfis_ctx* ctx = GetContext();
            const int nbuf = _gridType == DUAL ? 2 : 1;
            for (int b = 0; b < nbuf; b++) {
                const int idx = b == 0 ? _currIndex : _nextIndex;
                Array<T>& h = _holder.at(idx);
                const size_t bytes = (size_t)h.GetSize() * sizeof(T);
                if (!_devData[idx]) {
                    void* p = nullptr;
                    detail::Check(appfis_malloc(ctx, bytes, &p), "device allocation of a subgrid");
                    _devData[idx] = (T*)p;
                    detail::Check(appfis_malloc(ctx, sizeof(Array<T>), &p), "device allocation of an Array view");
                    _dev[idx] = (Array<T>*)p;
                }
                if (bytes) detail::Check(appfis_memcpy_h2d(ctx, _devData[idx], h.GetData(), bytes), "upload");
                Array<T> img;
                img._nDim = h._nDim;
                img._size = h._size;
                for (int a = 0; a < 3; a++) img._dimsStore[a] = h._dimsStore[a];
                img._data = _devData[idx];
                img._owns = false;
                const ptrdiff_t dims_off = (const char*)img._dimsStore - (const char*)&img;
                img._dims = (int*)((char*)_dev[idx] + dims_off);
                detail::Check(appfis_memcpy_h2d(ctx, _dev[idx], &img, sizeof(Array<T>)), "upload of an Array view");
                img._data = nullptr;
                img._dims = img._dimsStore;
            }
        }

        void CopySubGridsGPUToCPU()
        {
            if (!_devData[_currIndex]) return;
            Array<T>& h = _holder.at(_currIndex);
            const size_t bytes = (size_t)h.GetSize() * sizeof(T);
            if (bytes) detail::Check(appfis_memcpy_d2h(GetContext(), h.GetData(), _devData[_currIndex], bytes),
                                     "download");
        }

        T* GetSubGridsGPUData() { return _devData[_currIndex]; }
        T* GetSubGridsNextGPUData() { return _devData[_nextIndex]; }
        bool HasDeviceMirror() const { return _devData[0] != nullptr; }

    private:
        bool DeviceActive() const { return EXECUTION_FLAG && _devData[_currIndex] != nullptr; }

        void DeviceSetBoundary(int idx, int face, const T* value)
        {
            static_assert(detail::BytesOnlyType<T>(), "element size must be 1, 2, 4 or 8 bytes");
            detail::Check(appfis_set_boundary(GetContext(), _devData[idx], detail::DtypeOf<T>(), _nDim,
                                              _holder.at(idx).GetDims(), GetGhostLayer(), face, value),
                          "device boundary fill");
        }
        void DeviceSetBoundaryCopy(int idx, int face)
        {
            static_assert(detail::BytesOnlyType<T>(), "element size must be 1, 2, 4 or 8 bytes");
            detail::Check(appfis_set_boundary_copy(GetContext(), _devData[idx], detail::DtypeOf<T>(), _nDim,
                                                   _holder.at(idx).GetDims(), GetGhostLayer(), face),
                          "device boundary copy");
        }

        void FreeDevice()
        {
            Topology& t = Topology::GetInstance();
            for (int b = 0; b < 2; b++) {
                if (t._ctx) {
                    if (_devData[b]) appfis_free(t._ctx, _devData[b]);
                    if (_dev[b]) appfis_free(t._ctx, _dev[b]);
                }
                _devData[b] = nullptr;
                _dev[b] = nullptr;
            }
        }

        void CopyHost(const Grid<T>& g)
        {
            _nDim = g._nDim;
            _gridType = g._gridType;
            _full = g._full;
            _holder = std::vector<Array<T>>(2);
            _currIndex = g._currIndex;
            _nextIndex = g._nextIndex;
            if (g._holder.size() == 2) {
                _holder.at(_currIndex) = g._holder.at(_currIndex);
                _holder.at(_nextIndex) = g._holder.at(_nextIndex);
            }
            _haloBuffer = g._haloBuffer;
        }

        // face -> (axis, high side) after APPFIS::DIRECTION (3D: UP/DOWN are z)
        static bool FaceAxis(int n, int face, int* ax, int* high)
        {
            *high = 0;
            if (n == 1) {
                if (face > RIGHT) return false;
                *ax = X;
                *high = face == RIGHT;
                return true;
            }
            if (n == 2) {
                if (face > DOWN) return false;
                *ax = face <= RIGHT ? X : Y;
                *high = face == RIGHT || face == DOWN;
                return true;
            }
            switch (face) {
            case LEFT: *ax = X; return true;
            case RIGHT: *ax = X; *high = 1; return true;
            case FRONT: *ax = Y; return true;
            case BACK: *ax = Y; *high = 1; return true;
            case DOWN: *ax = Z; return true;
            case UP: *ax = Z; *high = 1; return true;
            default: return false;
            }
        }

        template <typename F>
        static void ForFace(const int* ext, int ax, int layer, F f)
        {
            int lo[3] = {0, 0, 0}, hi[3] = {ext[0], ext[1], ext[2]};
            lo[ax] = layer;
            hi[ax] = layer + 1;
            for (int k = lo[2]; k < hi[2]; k++)
                for (int j = lo[1]; j < hi[1]; j++)
                    for (int i = lo[0]; i < hi[0]; i++) f(i, j, k);
        }

        static void Put(Array<T>* a, int n, int i, int j, int k, T v)
        {
            if (n == 1) a->Set(i, v);
            else if (n == 2) a->Set(i, j, v);
            else a->Set(i, j, k, v);
        }
        static T Take(Array<T>* a, int n, int i, int j, int k)
        {
            return n == 1 ? a->Get(i) : n == 2 ? a->Get(i, j) : a->Get(i, j, k);
        }
    };
}  // namespace APPFIS

#endif
