// basicTypes.h -- value types of the gpro interface (gpro/basicTypes.h of
// the reference: DomDcmpType/ProgramType enums :51-93, SpaceDims :184-254,
// CellCoord :261-369, CoordBR :376-515), re-declared for the MI355X build.
#pragma once

#include <iostream>
#include <string>
#include <vector>

#define DEFAULT = default
#define OVERRIDE override

namespace GPRO {

using std::cerr;
using std::cout;
using std::endl;
using std::string;
using std::vector;

enum DomDcmpType { NON_DCMP = 0, ROWWISE_DCMP, COLWISE_DCMP, BLOCK_DCMP };
enum DomDcmpMethod { SPACE_DIM = 0, SPACE_WEIGHT, COMPUTE_WEIGHT };

// The reference reserves CUDA_Type for a GPU path and leaves it empty
// (rasterOperator.h:347-348); on this build HIP_Type names the GPU path.
// Operators that have a device implementation use it under every type.
enum ProgramType { MPI_Type = 0, MPI_OpenMP_Type, CUDA_Type, Serial_Type, HIP_Type = CUDA_Type };

enum MeshDir {
    NORTH_DIR = 0, NORTHEAST_DIR, EAST_DIR, SOUTHEAST_DIR, SOUTH_DIR, SOUTHWEST_DIR, WEST_DIR, NORTHWEST_DIR
};

class SpaceDims {
  public:
    SpaceDims() : _nRows(0), _nCols(0) {}
    SpaceDims(int nRows, int nCols) : _nRows(nRows), _nCols(nCols) {}
    int nRows() const { return _nRows; }
    int nCols() const { return _nCols; }
    void nRows(int n) { _nRows = n; }
    void nCols(int n) { _nCols = n; }
    long long size() const { return (long long)_nRows * _nCols; }
    bool isNone() const { return _nRows <= 0 || _nCols <= 0; }
    bool operator==(const SpaceDims &o) const { return _nRows == o._nRows && _nCols == o._nCols; }
    bool operator!=(const SpaceDims &o) const { return !(*this == o); }

  private:
    int _nRows, _nCols;
};

inline std::ostream &operator<<(std::ostream &os, const SpaceDims &d)
{
    return os << d.nRows() << " " << d.nCols();
}

class CellCoord {
  public:
    CellCoord() : _iRow(0), _iCol(0) {}
    CellCoord(int iRow, int iCol) : _iRow(iRow), _iCol(iCol) {}
    int iRow() const { return _iRow; }
    int iCol() const { return _iCol; }
    void iRow(int r) { _iRow = r; }
    void iCol(int c) { _iCol = c; }
    bool operator==(const CellCoord &o) const { return _iRow == o._iRow && _iCol == o._iCol; }
    bool operator!=(const CellCoord &o) const { return !(*this == o); }
    CellCoord operator+(const CellCoord &o) const { return CellCoord(_iRow + o._iRow, _iCol + o._iCol); }

  private:
    int _iRow, _iCol;
};

inline std::ostream &operator<<(std::ostream &os, const CellCoord &c) { return os << c.iRow() << " " << c.iCol(); }

// Bounding rectangle [nwCorner, seCorner], both inclusive.
class CoordBR {
  public:
    CoordBR() {}
    CoordBR(const CellCoord &nw, const CellCoord &se) : _nw(nw), _se(se) {}
    const CellCoord &nwCorner() const { return _nw; }
    const CellCoord &seCorner() const { return _se; }
    void nwCorner(const CellCoord &c) { _nw = c; }
    void seCorner(const CellCoord &c) { _se = c; }
    void nwCorner(int r, int c) { _nw = CellCoord(r, c); }
    void seCorner(int r, int c) { _se = CellCoord(r, c); }
    int minIRow() const { return _nw.iRow(); }
    int minICol() const { return _nw.iCol(); }
    int maxIRow() const { return _se.iRow(); }
    int maxICol() const { return _se.iCol(); }
    int nRows() const { return _se.iRow() - _nw.iRow() + 1; }
    int nCols() const { return _se.iCol() - _nw.iCol() + 1; }
    long long size() const { return (long long)nRows() * nCols(); }
    bool valid(const SpaceDims &d) const
    {
        return minIRow() >= 0 && minICol() >= 0 && maxIRow() < d.nRows() && maxICol() < d.nCols() &&
               minIRow() <= maxIRow() && minICol() <= maxICol();
    }
    bool ifContain(const CellCoord &c) const
    {
        return c.iRow() >= minIRow() && c.iRow() <= maxIRow() && c.iCol() >= minICol() && c.iCol() <= maxICol();
    }

  private:
    CellCoord _nw, _se;
};

inline std::ostream &operator<<(std::ostream &os, const CoordBR &b) { return os << b.nwCorner() << "; " << b.seCorner(); }

typedef std::vector<int> IntVect;

}  // namespace GPRO
