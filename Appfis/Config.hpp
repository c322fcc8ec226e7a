/*
 * Config.hpp -- iteration box of an Execute call and the grids it updates
 * (API of the reference Appfis/Config.hpp:20-188).  Boxes are half-open [start, end)
 * per axis; the default box of a grid is [ghost, dim - ghost); the skip constructors
 * give [skipStart, dim - skipEnd).
 */
#ifndef APPFIS_CONFIG_HPP
#define APPFIS_CONFIG_HPP

#include <vector>

#include "Grid.hpp"

namespace APPFIS
{
    template <class T>
    class Config
    {
    private:
        int _dimX = 0, _dimY = 0, _dimZ = 0;
        int _startDimX = 0, _endDimX = 0;
        int _startDimY = 0, _endDimY = 0;
        int _startDimZ = 0, _endDimZ = 0;

    public:
        std::vector<Grid<T>*> updatedGrids;

        Config() {}
        ~Config() {}

        Config(Grid<T>& g)
        {
            Array<T>* arr = g.GetSubGridCPU();
            const int n = arr->GetDimCount();
            const int gh = GetGhostLayer();
            SetX(arr->GetDim(X), gh, gh);
            if (n > 1) SetY(arr->GetDim(Y), gh, gh);
            if (n > 2) SetZ(arr->GetDim(Z), gh, gh);
        }
        Config(Grid<T>& g, int skipStartX, int skipEndX) { SetX(g.GetSubGridCPU()->GetDim(X), skipStartX, skipEndX); }
        Config(Grid<T>& g, int skipStartX, int skipEndX, int skipStartY, int skipEndY)
        {
            SetX(g.GetSubGridCPU()->GetDim(X), skipStartX, skipEndX);
            SetY(g.GetSubGridCPU()->GetDim(Y), skipStartY, skipEndY);
        }
        Config(Grid<T>& g, int skipStartX, int skipEndX, int skipStartY, int skipEndY, int skipStartZ, int skipEndZ)
        {
            SetX(g.GetSubGridCPU()->GetDim(X), skipStartX, skipEndX);
            SetY(g.GetSubGridCPU()->GetDim(Y), skipStartY, skipEndY);
            SetZ(g.GetSubGridCPU()->GetDim(Z), skipStartZ, skipEndZ);
        }
        Config(int dimX, int skipStartX, int skipEndX) { SetX(dimX, skipStartX, skipEndX); }
        Config(int dimX, int skipStartX, int skipEndX, int dimY, int skipStartY, int skipEndY)
        {
            SetX(dimX, skipStartX, skipEndX);
            SetY(dimY, skipStartY, skipEndY);
        }
        Config(int dimX, int skipStartX, int skipEndX, int dimY, int skipStartY, int skipEndY, int dimZ,
               int skipStartZ, int skipEndZ)
        {
            SetX(dimX, skipStartX, skipEndX);
            SetY(dimY, skipStartY, skipEndY);
            SetZ(dimZ, skipStartZ, skipEndZ);
        }

        int GetDimX() const { return _dimX; }
        int GetDimY() const { return _dimY; }
        int GetDimZ() const { return _dimZ; }
        int GetStartX() const { return _startDimX; }
        int GetStartY() const { return _startDimY; }
        int GetStartZ() const { return _startDimZ; }
        int GetEndX() const { return _endDimX; }
        int GetEndY() const { return _endDimY; }
        int GetEndZ() const { return _endDimZ; }

        bool IsUpdated() { return !updatedGrids.empty(); }

        void GridsToUpdate(Grid<T>& grid) { updatedGrids.push_back(&grid); }

        template <typename... TArgs>
        void GridsToUpdate(Grid<T>& grid, TArgs&&... args)
        {
            updatedGrids.push_back(&grid);
            GridsToUpdate(args...);
        }

    private:
        void SetX(int d, int s, int e)
        {
            _dimX = d;
            _startDimX = s;
            _endDimX = d - e;
        }
        void SetY(int d, int s, int e)
        {
            _dimY = d;
            _startDimY = s;
            _endDimY = d - e;
        }
        void SetZ(int d, int s, int e)
        {
            _dimZ = d;
            _startDimZ = s;
            _endDimZ = d - e;
        }
    };
}  // namespace APPFIS

#endif
