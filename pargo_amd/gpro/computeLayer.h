// computeLayer.h -- coarse compute-load grid and load-balanced row bands
// (gpro/computeLayer.h of the reference: ComputeLayer<elemType> :41-125,
// _init :208-260, getCompuLoad :378-443, readComputeLoadFile :445-454,
// writeComputeIntensityFile :456-513).
//
// The grid holds one estimated load per block of granularity x granularity
// data cells.  getCompuLoad cuts the data rows into bands of near-equal
// summed load (pargo_load_row_bands) and returns the band of `rank`; with
// one process per GPU and one GPU the band is the whole work rectangle.
#pragma once

#include "../../include/pargo.h"
#include "rasterLayer.h"

namespace GPRO {

inline bool ifDecomposeBySpace(const string &arg)
{
    if (arg == "space") return true;
    if (arg == "compute") return false;
    cerr << "-dcmp must be space or compute (got " << arg << "); using space" << endl;
    return true;
}

template <class elemType>
class ComputeLayer : public RasterLayer<elemType> {
  public:
    ComputeLayer() : RasterLayer<elemType>("computeLayer") {}
    explicit ComputeLayer(const string &name) : RasterLayer<elemType>(name) {}

    void addRasterLayerSerial(RasterLayer<elemType> *dataLayer) { _vDataLayers.push_back(dataLayer); }
    void addRasterLayer(RasterLayer<elemType> *dataLayer) { _vDataLayers.push_back(dataLayer); }
    const vector<RasterLayer<elemType> *> &dataLayers() const { return _vDataLayers; }
    int getComputeGrain() const { return _comptGrain; }

    // Coarse grid over the first data layer's work rectangle: ceil(rows / g)
    // x ceil(cols / g) cells of size cellSize * g, geotransform shifted to
    // the work rectangle's corner, every load 0.
    bool init(const char *neighborFile, int comptGrain)
    {
        if (_vDataLayers.empty() || !_vDataLayers[0]->metaData() || comptGrain < 1) return false;
        if (neighborFile) {
            if (!this->readNeighborhood(neighborFile)) return false;
        } else {
            this->newLocalNbrhood();
        }
        const MetaData &rhs = *_vDataLayers[0]->metaData();
        _comptGrain = comptGrain;
        MetaData m;
        m.cellSize = rhs.cellSize * comptGrain;
        m.row = rhs._localworkBR.nRows() / comptGrain + (rhs._localworkBR.nRows() % comptGrain ? 1 : 0);
        m.column = rhs._localworkBR.nCols() / comptGrain + (rhs._localworkBR.nCols() % comptGrain ? 1 : 0);
        m.format = rhs.format;
        m.projection = rhs.projection;
        m.geoKeys = rhs.geoKeys;
        m.noData = rhs.noData;
        m._domDcmpType = NON_DCMP;
        m._glbDims = SpaceDims(m.row, m.column);
        m._localdims = m._glbDims;
        m._MBR = CoordBR(CellCoord(0, 0), CellCoord(m.row - 1, m.column - 1));
        if (!this->nbrhood()->calcWorkBR(m._localworkBR, m._glbDims)) return false;
        m.dataType = dataTypeOf<elemType>();
        for (int i = 0; i < 6; i++) m.pTransform[i] = rhs.pTransform[i];
        m.pTransform[0] += rhs._localworkBR.minICol() * rhs.cellSize;
        m.pTransform[3] -= rhs._localworkBR.minIRow() * rhs.cellSize;
        m.pTransform[1] *= comptGrain;
        m.pTransform[5] *= comptGrain;
        this->setMetaData(m);
        this->newCellSpace(m._localdims, (elemType)0);
        return true;
    }

    // Every band's data rows (getCompuLoad's cut, computeLayer.h:396-428):
    // load-grid cuts scaled by dataRows / loadRows and shifted to the data
    // work rectangle; the last band ends at the work rectangle's last row.
    bool getCompuLoadBands(int nSubSpcs, vector<CoordBR> &bands)
    {
        bands.clear();
        if (_vDataLayers.empty() || !this->cellSpace() || !this->metaData()) return false;
        const MetaData &dm = *_vDataLayers[0]->metaData();
        CoordBR glbWorkBR;
        const Neighborhood<elemType> *dn = _vDataLayers[0]->nbrhood();
        if (dn && !dn->empty()) {
            if (!dn->calcWorkBR(glbWorkBR, dm._glbDims)) return false;
        } else {
            glbWorkBR = CoordBR(CellCoord(0, 0), CellCoord(dm._glbDims.nRows() - 1, dm._glbDims.nCols() - 1));
        }
        const int loadRows = this->metaData()->_glbDims.nRows(), loadCols = this->metaData()->_glbDims.nCols();
        if (nSubSpcs < 1 || nSubSpcs > loadRows) {
            cerr << "ComputeLayer::getCompuLoad: invalid number of SubSpaces (" << nSubSpcs << ") for " << loadRows
                 << " load rows" << endl;
            return false;
        }
        std::vector<double> load((size_t)loadRows * loadCols);
        const elemType *src = this->cellSpace()->data();
        for (size_t i = 0; i < load.size(); i++) load[i] = (double)src[i];
        std::vector<int> first(nSubSpcs + 1);
        if (pargo_load_row_bands(load.data(), loadRows, loadCols, this->metaData()->noData, nSubSpcs,
                                 first.data()) != PARGO_OK)
            return false;
        const double convertGrain = dm._glbDims.nRows() / (double)loadRows;
        int subEnd = glbWorkBR.minIRow() - 1;
        for (int i = 0; i < nSubSpcs - 1; ++i) {
            const int subBegin = (int)(first[i] * convertGrain + glbWorkBR.minIRow());
            subEnd = (int)(first[i + 1] * convertGrain + glbWorkBR.minIRow() - 1);
            bands.push_back(CoordBR(CellCoord(subBegin, glbWorkBR.minICol()), CellCoord(subEnd, glbWorkBR.maxICol())));
        }
        bands.push_back(
            CoordBR(CellCoord(subEnd + 1, glbWorkBR.minICol()), CellCoord(glbWorkBR.maxIRow(), glbWorkBR.maxICol())));
        return true;
    }

    bool getCompuLoad(DomDcmpType dcmpType, int nSubSpcs, CoordBR &subWorkBR, int rank = 0)
    {
        if (dcmpType != ROWWISE_DCMP) {
            cerr << "compute layer decomposition: " << dcmpType << " not supported" << endl;
            return false;
        }
        vector<CoordBR> bands;
        if (!getCompuLoadBands(nSubSpcs, bands) || rank < 0 || rank >= (int)bands.size()) return false;
        subWorkBR = bands[rank];
        return true;
    }

    bool readComputeLoadFile(const char *loadFile) { return this->readFile(loadFile, NON_DCMP); }
    bool writeComputeIntensityFile(const char *outputfile) { return this->writeFile(outputfile); }

    vector<RasterLayer<elemType> *> _vDataLayers;

  private:
    int _comptGrain = 1;
};

}  // namespace GPRO
