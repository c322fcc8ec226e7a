/*
 * Array.hpp -- dense APPFIS array (API of the reference Appfis/Array.hpp:22-78).
 *
 * Row-major, x fastest, int sizes and indices: idx = i + dimX*j + dimX*dimY*k
 * (Array.hpp:579-590).  Get/Set run on host and device.
 *
 * Device views: during execution the runtime hands user stencils `Array<T>*` that
 * point to a device-resident image of the array (built by Grid): the image is a
 * byte copy of a non-owning Array whose _data is the device buffer and whose _dims
 * points at the image's own inline extents, so an access costs the data pointer and
 * two extents read from one cache line plus the element itself.
 */
#ifndef APPFIS_ARRAY_HPP
#define APPFIS_ARRAY_HPP

#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "Env.hpp"
#include "Utils.hpp"

namespace APPFIS
{
    template <class T>
    class Grid;

    template <class T>
    class Array
    {
    private:
        int _nDim = 0;
        int _size = 0;
        int* _dims = _dimsStore;  // always points at the inline extents of this object
        T* _data = nullptr;
        int _dimsStore[3] = {0, 0, 0};
        bool _owns = true;  // false for views (device images)

        friend class Grid<T>;

    public:
        Array() {}
        ~Array() { Reset(); }

        Array(int nDim, int dimX)
        {
            int d[1] = {dimX};
            Shape(nDim, d);
        }
        Array(int nDim, int dimX, int dimY)
        {
            int d[2] = {dimX, dimY};
            Shape(nDim, d);
        }
        Array(int nDim, int dimX, int dimY, int dimZ)
        {
            int d[3] = {dimX, dimY, dimZ};
            Shape(nDim, d);
        }
        Array(int nDim, int* dims) { Shape(nDim, dims); }

        Array(const Array<T>& arr)
        {
            std::cout << "Array copy" << std::endl;
            CopyFrom(arr);
        }

        Array& operator=(const Array<T>& arr)
        {
            std::cout << "Array assign" << std::endl;
            if (this != &arr) {
                Reset();
                CopyFrom(arr);
            }
            return *this;
        }

        // Host and device forms of the element access.  On the device the elements are
        // a device allocation (global memory) and the extents are this object's inline
        // copy; saying so lets the compiler emit global loads and stores, where a pointer
        // read from memory would otherwise give flat ones.
        __host__ void Set(int i, T val) { _data[i] = val; }
        __host__ void Set(int i, int j, T val) { _data[Get1DIndex(i, j)] = val; }
        __host__ void Set(int i, int j, int k, T val) { _data[Get1DIndex(i, j, k)] = val; }
        __host__ T Get(int i) { return _data[i]; }
        __host__ T Get(int i, int j) { return _data[Get1DIndex(i, j)]; }
        __host__ T Get(int i, int j, int k) { return _data[Get1DIndex(i, j, k)]; }
        __device__ void Set(int i, T val) { Elems()[i] = val; }
        __device__ void Set(int i, int j, T val) { Elems()[DevIndex(i, j)] = val; }
        __device__ void Set(int i, int j, int k, T val) { Elems()[DevIndex(i, j, k)] = val; }
        __device__ T Get(int i) { return Elems()[i]; }
        __device__ T Get(int i, int j) { return Elems()[DevIndex(i, j)]; }
        __device__ T Get(int i, int j, int k) { return Elems()[DevIndex(i, j, k)]; }

        int GetDimCount() const { return _nDim; }
        int GetSize() const { return _size; }
        int GetDim(int dim) const { return _dims[dim]; }
        int* GetDims() const { return _dims; }

        T* GetData() { return _data; }

        void SetData(T* data)
        {
            if (_data == nullptr && _size > 0) AllocateMemory();
            for (int i = 0; i < _size; i++) _data[i] = data[i];
        }

        void FillData(T value)
        {
            if (_data == nullptr && _size > 0) AllocateMemory();
            for (int i = 0; i < _size; i++) _data[i] = value;
        }

        void Square()
        {
            if (_data == nullptr && _size > 0) AllocateMemory();
            for (int i = 0; i < _size; i++) _data[i] = _data[i] * _data[i];
        }

        void Reset()
        {
            if (_data != nullptr && _owns) delete[] _data;
            _data = nullptr;
            _nDim = 0;
            _size = 0;
            _dims = _dimsStore;
            _dimsStore[0] = _dimsStore[1] = _dimsStore[2] = 0;
            _owns = true;
        }

        // zero-initialised, like the reference (Array.hpp:339-342)
        void AllocateMemory()
        {
            if (_data != nullptr && _owns) delete[] _data;
            _data = new T[_size]();
            _owns = true;
        }

        void Resize(int nDim, int* dims)
        {
            Reset();
            Shape(nDim, dims);
        }

        // whitespace-separated values, the first `headerSize` lines skipped
        // (reference Array.hpp:377-417; a last line without a newline is not read)
        void LoadAsciiFile(const std::string& filePath, int headerSize = 0)
        {
            std::ifstream in(filePath.c_str());
            if (!in.is_open()) {
                std::cerr << "Can not read " << filePath << std::endl;
                return;
            }
            if (_data == nullptr && _size > 0) AllocateMemory();
            int index = 0, lineNum = 0;
            std::string line;
            std::getline(in, line);
            while (!in.eof()) {
                lineNum++;
                if (lineNum > headerSize) {
                    std::vector<std::string> vals = Split(line, ' ');
                    for (auto it = vals.begin(); it != vals.end() && index < _size; ++it, ++index)
                        _data[index] = static_cast<T>(atof(it->c_str()));
                }
                std::getline(in, line);
            }
        }

        // raw values after `headerSize` elements of T (reference Array.hpp:420-443)
        void LoadBinaryFile(const std::string& filePath, int headerSize = 0)
        {
            std::ifstream in(filePath.c_str(), std::ios::binary);
            if (!in.is_open()) {
                std::cerr << "Can not read " << filePath << std::endl;
                return;
            }
            if (_data == nullptr && _size > 0) AllocateMemory();
            if (headerSize > 0) in.seekg((std::streamoff)headerSize * sizeof(T), std::ios::beg);
            in.read((char*)_data, (std::streamsize)sizeof(T) * _size);
        }

        // fixed, 6 decimals, one row per line (reference Array.hpp:446-505)
        void SaveAsciiFile(const std::string& filePath)
        {
            CreateDir(GetParentDir(filePath.c_str()));
            std::ofstream out(filePath.c_str());
            if (!out.is_open() || _data == nullptr) {
                std::cerr << "Can not write to " << filePath << std::endl;
                return;
            }
            out.precision(6);
            out << std::fixed;
            const int nx = GetDim(X);
            const int ny = _nDim > 1 ? GetDim(Y) : 1;
            const int nz = _nDim > 2 ? GetDim(Z) : 1;
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++) {
                    for (int i = 0; i < nx; i++) {
                        out << +_data[i + nx * j + nx * ny * k];
                        if (i < nx - 1) out << " ";
                    }
                    if (_nDim > 1) out << "\n";
                }
        }

        void SaveBinaryFile(const std::string& filePath)
        {
            CreateDir(GetParentDir(filePath.c_str()));
            std::ofstream out(filePath.c_str(), std::ios::binary);
            if (!out.is_open() || _data == nullptr) {
                std::cerr << "Can not write to " << filePath << std::endl;
                return;
            }
            out.write((const char*)_data, (std::streamsize)_size * sizeof(T));
        }

        std::string ToString()
        {
            std::string s;
            if (_data == nullptr) return s;
            const int nx = GetDim(X);
            const int ny = _nDim > 1 ? GetDim(Y) : 1;
            const int nz = _nDim > 2 ? GetDim(Z) : 1;
            bool indent = false;
            for (int k = 0; k < nz; k++) {
                for (int j = 0; j < ny; j++) {
                    if (_nDim == DIM_3D && indent) s += "    ";
                    for (int i = 0; i < nx; i++) s += " " + std::to_string(_data[i + nx * j + nx * ny * k]);
                    if (_nDim > 1) s += "\n";
                }
                indent = !indent;
            }
            return s;
        }

        T** DataAddress() { return &_data; }
        int** DimsAddress() { return &_dims; }

    private:
        void Shape(int nDim, const int* dims)
        {
            _nDim = nDim;
            _size = 1;
            for (int i = 0; i < nDim; i++) {
                _dimsStore[i] = dims[i];
                _size *= dims[i];
            }
            _dims = _dimsStore;
        }

        void CopyFrom(const Array<T>& arr)
        {
            if (arr._nDim == 0) return;
            Shape(arr._nDim, arr._dims);
            if (arr._data != nullptr) {
                AllocateMemory();
                for (int i = 0; i < _size; i++) _data[i] = arr._data[i];
            }
        }

        EXEC_SPACE int Get1DIndex(int i, int j) { return i + _dims[X] * j; }
        EXEC_SPACE int Get1DIndex(int i, int j, int k) { return i + _dims[X] * j + _dims[X] * _dims[Y] * k; }
        using GlobalT = T __attribute__((address_space(1)));
        __device__ GlobalT* Elems() const { return (GlobalT*)_data; }
        __device__ int DevIndex(int i, int j) const { return i + _dimsStore[X] * j; }
        __device__ int DevIndex(int i, int j, int k) const
        {
            return i + _dimsStore[X] * j + _dimsStore[X] * _dimsStore[Y] * k;
        }
    };
}  // namespace APPFIS

#endif
