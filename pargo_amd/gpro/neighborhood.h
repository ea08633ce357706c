// neighborhood.h -- analysis window of the gpro interface
// (gpro/neighborhood.h of the reference: Neighborhood<elemType> :31-98,
// calcWorkBR :424-468, the .nbr reader operator>> :627-664).
//
// A .nbr file holds a count followed by `count` triples "iRow iCol weight";
// a count of 100 or more instead means a square window of side
// floor(sqrt(count)) with zero weights.  The window's bounding rectangle
// decides the work rectangle of every layer that carries it: the cells
// whose whole window lies inside the raster.
#pragma once

#include <algorithm>
#include <cmath>
#include <fstream>
#include <istream>

#include "basicTypes.h"

namespace GPRO {

template <class elemType>
class Neighborhood {
  public:
    Neighborhood() {}
    explicit Neighborhood(bool isLocal)
    {
        if (isLocal) fromVect(vector<CellCoord>(1, CellCoord(0, 0)), 1.0);
    }
    Neighborhood(const vector<CellCoord> &coords, double weight = 1.0) { fromVect(coords, weight); }
    Neighborhood(const vector<CellCoord> &coords, const vector<double> &weights) { fromVects(coords, weights); }
    template <class T2>
    explicit Neighborhood(const Neighborhood<T2> &rhs) { fromVects(rhs.coords(), rhs.weights()); }

    bool fromVect(const vector<CellCoord> &coords, double weight = 1.0)
    {
        return fromVects(coords, vector<double>(coords.size(), weight));
    }
    bool fromVects(const vector<CellCoord> &coords, const vector<double> &weights)
    {
        clear();
        if (coords.size() != weights.size()) {
            cerr << "Neighborhood: " << coords.size() << " coordinates but " << weights.size() << " weights" << endl;
            return false;
        }
        // a window must be non-empty, hold its centre and list each offset
        // once (neighborhood.h:112-139)
        if (coords.empty()) {
            cerr << "Neighborhood: the coordinate vector is empty" << endl;
            return false;
        }
        if (std::find(coords.begin(), coords.end(), CellCoord(0, 0)) == coords.end()) {
            cerr << "Neighborhood: the coordinate vector has to have a coordinate [0, 0]" << endl;
            return false;
        }
        for (size_t i = 0; i < coords.size(); i++) {
            if (std::find(_coords.begin(), _coords.end(), coords[i]) != _coords.end()) {
                cerr << "Neighborhood: a duplicated coordinate [" << coords[i] << "] has been found" << endl;
                clear();
                return false;
            }
            _coords.push_back(coords[i]);
            _weights.push_back(weights[i]);
        }
        int r0 = _coords[0].iRow(), r1 = r0, c0 = _coords[0].iCol(), c1 = c0;
        for (const CellCoord &c : _coords) {
            r0 = std::min(r0, c.iRow()), r1 = std::max(r1, c.iRow());
            c0 = std::min(c0, c.iCol()), c1 = std::max(c1, c.iCol());
        }
        _mbr = CoordBR(CellCoord(r0, c0), CellCoord(r1, c1));
        return true;
    }
    void clear()
    {
        _coords.clear();
        _weights.clear();
        _mbr = CoordBR();
    }

    bool empty() const { return _coords.empty(); }
    int size() const { return (int)_coords.size(); }
    const vector<CellCoord> &coords() const { return _coords; }
    const vector<double> &weights() const { return _weights; }
    const CellCoord &operator[](int i) const { return _coords.at(i); }
    double weight(int i) const { return _weights.at(i); }
    const CoordBR &getMBR() const { return _mbr; }
    int minIRow() const { return _mbr.minIRow(); }
    int minICol() const { return _mbr.minICol(); }
    int maxIRow() const { return _mbr.maxIRow(); }
    int maxICol() const { return _mbr.maxICol(); }
    int nRows() const { return empty() ? 0 : _mbr.nRows(); }
    int nCols() const { return empty() ? 0 : _mbr.nCols(); }

    // True when the window is exactly the 3x3 Moore window (any order).
    bool isMoore3x3() const
    {
        if (size() != 9) return false;
        return minIRow() == -1 && maxIRow() == 1 && minICol() == -1 && maxICol() == 1;
    }

    // Work rectangle of a raster of `dims` under this window.
    bool calcWorkBR(CoordBR &workBR, const SpaceDims &dims) const
    {
        if (empty()) {
            cerr << "Neighborhood::calcWorkBR: empty neighborhood" << endl;
            return false;
        }
        if (dims.isNone()) {
            cerr << "Neighborhood::calcWorkBR: empty dimensions" << endl;
            return false;
        }
        if (dims.nRows() < nRows() || dims.nCols() < nCols()) {
            cerr << "Neighborhood::calcWorkBR: dimensions (" << dims << ") smaller than the neighborhood ("
                 << nRows() << " " << nCols() << ")" << endl;
            return false;
        }
        workBR.nwCorner(-minIRow(), -minICol());
        workBR.seCorner(dims.nRows() - maxIRow() - 1, dims.nCols() - maxICol() - 1);
        if (!workBR.valid(dims)) {
            cerr << "Neighborhood::calcWorkBR: invalid work rectangle (" << workBR << ")" << endl;
            return false;
        }
        if ((maxIRow() > 0 && workBR.nRows() < maxIRow()) || (minIRow() < 0 && workBR.nRows() < -minIRow()) ||
            (maxICol() > 0 && workBR.nCols() < maxICol()) || (minICol() < 0 && workBR.nCols() < -minICol())) {
            cerr << "Neighborhood::calcWorkBR: work rectangle (" << workBR << ") too small for the neighborhood"
                 << endl;
            return false;
        }
        return true;
    }

  private:
    vector<CellCoord> _coords;
    vector<double> _weights;
    CoordBR _mbr;
};

// .nbr text reader
template <class elemType>
std::istream &operator>>(std::istream &is, Neighborhood<elemType> &nbr)
{
    int n = 0;
    vector<CellCoord> coords;
    vector<double> weights;
    if (!(is >> n)) {
        nbr.clear();
        return is;
    }
    if (n >= 100) {
        const int half = (int)std::sqrt((double)n) / 2;
        for (int r = -half; r <= half; r++)
            for (int c = -half; c <= half; c++) coords.push_back(CellCoord(r, c)), weights.push_back(0.0);
    } else {
        for (int i = 0; i < n; i++) {
            int r, c;
            double w;
            if (!(is >> r >> c >> w)) break;
            coords.push_back(CellCoord(r, c)), weights.push_back(w);
        }
    }
    nbr.fromVects(coords, weights);
    return is;
}

}  // namespace GPRO
