// cellSpace.h -- host raster matrix of the gpro interface (gpro/cellSpace.h
// of the reference: CellSpace<elemType> :117-236).  A CellSpace owns one
// contiguous row-major block; operator[](iRow) returns the row pointer so
// that user code written as cellSpace[iRow][iCol] keeps working.
#pragma once

#include <algorithm>
#include <cstring>
#include <vector>

#include "basicTypes.h"

namespace GPRO {

template <class elemType>
class CellSpace {
  public:
    CellSpace() {}
    explicit CellSpace(const SpaceDims &dims) { initMem(dims); }
    CellSpace(const SpaceDims &dims, const elemType &initVal) { initMem(dims), initVals(initVal); }
    CellSpace(int nRows, int nCols) { initMem(SpaceDims(nRows, nCols)); }
    CellSpace(int nRows, int nCols, const elemType &initVal) { initMem(SpaceDims(nRows, nCols)), initVals(initVal); }

    bool initMem(const SpaceDims &dims)
    {
        if (dims.nRows() < 0 || dims.nCols() < 0) {
            cerr << "CellSpace::initMem: negative dimensions (" << dims << ")" << endl;
            return false;
        }
        _dims = dims;
        _data.assign((size_t)dims.size(), elemType());
        return true;
    }
    bool initVals(const elemType &v)
    {
        std::fill(_data.begin(), _data.end(), v);
        return true;
    }
    void clear()
    {
        _data.clear();
        _data.shrink_to_fit();
        _dims = SpaceDims();
    }

    bool empty() const { return _data.empty(); }
    int nRows() const { return _dims.nRows(); }
    int nCols() const { return _dims.nCols(); }
    const SpaceDims &dims() const { return _dims; }
    long long size() const { return _dims.size(); }
    template <class T2>
    bool sameDims(const CellSpace<T2> &rhs) const { return _dims == rhs.dims(); }

    elemType *operator[](int iRow) { return _data.data() + (size_t)iRow * nCols(); }
    const elemType *operator[](int iRow) const { return _data.data() + (size_t)iRow * nCols(); }
    elemType &at(const CellCoord &c) { return (*this)[c.iRow()][c.iCol()]; }
    const elemType &at(const CellCoord &c) const { return (*this)[c.iRow()][c.iCol()]; }

    bool validCoord(const CellCoord &c, bool warning = true) const
    {
        bool ok = c.iRow() >= 0 && c.iRow() < nRows() && c.iCol() >= 0 && c.iCol() < nCols();
        if (!ok && warning) cerr << "CellSpace: coordinate (" << c << ") outside (" << _dims << ")" << endl;
        return ok;
    }
    bool validCoord(int iRow, int iCol, bool warning = true) const { return validCoord(CellCoord(iRow, iCol), warning); }

    // contiguous storage (the reference's public _matrix member)
    elemType *data() { return _data.data(); }
    const elemType *data() const { return _data.data(); }

    bool updateCellAs(const CellCoord &c, const elemType &v, bool warning = true)
    {
        if (!validCoord(c, warning)) return false;
        at(c) = v;
        return true;
    }
    long long count(const elemType &v) const { return (long long)std::count(_data.begin(), _data.end(), v); }

  private:
    SpaceDims _dims;
    std::vector<elemType> _data;
};

}  // namespace GPRO
