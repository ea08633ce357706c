// idwOperator.h -- inverse distance weighting of point samples onto a grid
// (apps/spatial/idw/idwOperator.h of the reference: SamplePoint :16-22,
// extent_info :24-29, SampleBlock :31-33, class IDWOperator :35-116).
//
// Same interface.  creatSampleBlocks() buckets the samples into blocks on the
// host, as the reference does, and uploads them once (pargo_idw_prepare);
// Run() computes every cell of the work rectangle on the GPU
// (pargo_idw_run) with the reference's ring search, K-set and weights.
#pragma once

#include <string>
#include <vector>

#include "../../gpro/rasterOperator.h"

using namespace GPRO;

#define NODATA_DEFINE -9999

struct pargo_idw_plan;

struct SamplePoint {
    double x = 0, y = 0, value = 0;
};

struct extent_info {
    double minX = 0, maxX = 0, minY = 0, maxY = 0;
};

struct SampleBlock {
    vector<SamplePoint> samplePoints;
};

// Host copy of a mask layer laid out for the device lookup: the kernel maps
// row r of the grid to min(trunc(r * R / rows), R - 1) with R the mask's
// global rows, while the reference caps at the last row it holds locally
// (rowAtOtherLayer, gpro/rasterLayer.h:1289-1294).  Rows past the local
// ones repeat the last local row, which gives the same lookup.
bool idwMaskForDevice(RasterLayer<int> *mask, std::vector<int> &buf, int &rows, int &cols);

class IDWOperator : public RasterOperator<double> {
  public:
    IDWOperator() {}
    IDWOperator(float cellsize, int nbrPoints, int idw_power, double bufferSize, double blockSize)
        : _cellSize(cellsize), _nbrPoints(nbrPoints), _idw_power(idw_power), _idw_buffer(bufferSize),
          _blockSize(blockSize)
    {
    }
    ~IDWOperator() override;

    double getBlockSize() { return _blockSize; }
    // Number of records of a point shapefile; *pSpatialRefWkt receives the
    // text of its .prj (owned by this operator).
    int readSampleNums(const char *filename, char **pSpatialRefWkt);
    // Reads x, y and DBF field fieldIdx of every point, and the grid extent
    // (idwOperator.cpp:63-149).
    bool readSamples(const char *filename, int fieldIdx, char **pSpatialRefWkt, vector<SamplePoint> &samples);
    // Buckets samples into blockSize blocks in file order and uploads them.
    void creatSampleBlocks(vector<SamplePoint> &samples);
    const vector<SampleBlock> *getSampleBlocks() { return &_pSampleBlocks; }
    void idwLayer(RasterLayer<double> &layerD, char **pSpatialRefWkt, DomDcmpType dcmpType = ROWWISE_DCMP);
    void idwLayer(RasterLayer<double> &layerD, char **pSpatialRefWkt, CoordBR &subWorkBR);
    // The whole grid as one band (idwOperator.cpp:277-326).
    void idwLayerSerial(RasterLayer<double> &layerD, char **pSpatialRefWkt) { idwLayer(layerD, pSpatialRefWkt, NON_DCMP); }
    void maskLayer(RasterLayer<int> &layerD) { _pMaskLayer = &layerD; }
    void maskLayer(RasterLayer<int> &layerD, CoordBR &) { _pMaskLayer = &layerD; }
    bool isTermination() override { return flag; }
    // There is no per-cell host rule: the operator runs on the GPU only.
    bool Operator(const CellCoord &coord, bool operFlag) override;
    bool WorkDevice(const CoordBR &workBR, bool &ok) override;

    int getBlockRowIndexByCoord(double y) { return (int)((_glb_extent.maxY - y) / _blockSize); }
    int getBlockColIndexByCoord(double x) { return (int)((x - _glb_extent.minX) / _blockSize); }
    int getBlockRowIndexByCellIndex(int iRow, double cellSize)
    {
        return std::min(int((iRow + 0.5) * cellSize / _blockSize), _blockRows - 1);
    }
    int getBlockColIndexByCellIndex(int iCol, double cellSize)
    {
        return std::min(int((iCol + 0.5) * cellSize / _blockSize), _blockCols - 1);
    }
    double getYByCellIndex(int iRow, double cellSize) { return double(_glb_extent.maxY - (iRow + 0.5) * cellSize); }
    double getXByCellIndex(int iCol, double cellSize) { return double((iCol + 0.5) * cellSize + _glb_extent.minX); }
    int getBlockCols() { return _blockCols; }
    int getBlockRows() { return _blockRows; }
    float getCellSize() { return _cellSize; }
    double getGlobalMinX() { return _glb_extent.minX; }
    double getGlobalMaxX() { return _glb_extent.maxX; }
    double getGlobalMinY() { return _glb_extent.minY; }
    double getGlobalMaxY() { return _glb_extent.maxY; }
    RasterLayer<int> *getMaskLayer() { return _pMaskLayer; }
    int getNbrPoints() { return _nbrPoints; }
    int getIdwPower() { return _idw_power; }
    double getIdwBuffer() { return _idw_buffer; }
    // The device copy of the sample blocks (null before creatSampleBlocks).
    pargo_idw_plan *plan() { return _plan; }
    void initIdwLayerGlobalInfo(RasterLayer<double> &layerD, char **pSpatialRefWkt);
    // Number of distance evaluations of the last Run().
    unsigned long long lastCandidates() const { return _candidates; }

  private:
    bool setLayerMeta(RasterLayer<double> &layerD, char **pSpatialRefWkt);

    float _cellSize = 0;
    int _nRows = 0, _nCols = 0;
    double _noData = NODATA_DEFINE;
    bool flag = true;
    unsigned long long _candidates = 0;

  protected:
    int _nbrPoints = 0;
    int _idw_power = 0;
    double _idw_buffer = 0;
    extent_info _glb_extent;
    int _sample_nums = 0;
    double _blockSize = 0;
    vector<SampleBlock> _pSampleBlocks;
    int _blockRows = 0;
    int _blockCols = 0;
    RasterLayer<double> *_pIDWLayer = nullptr;
    RasterLayer<int> *_pMaskLayer = nullptr;
    std::string _wkt;
    pargo_idw_plan *_plan = nullptr;
};
