// transformation.h -- from the data domain to the compute domain
// (gpro/transformation.h of the reference: class Transformation :34-79,
// Operator :155-197, run :199-228).
//
// Same interface.  The reference's default Operator() is the intensity
// ratio: each compute-layer cell sums validLoad over the valid data cells of
// its granule and nodataLoad over the NoData ones, over every data layer of
// the ComputeLayer.  run() computes that for the whole compute layer on the
// GPU (pargo_intensity_load, one call per data layer, accumulating in
// order).  A subclass with its own estimate overrides run(), as
// IdwTransformation does.
#pragma once

#include <cmath>
#include <vector>

#include "../../include/pargo.h"
#include "application.h"
#include "computeLayer.h"

namespace GPRO {

template <class elemType>
class Transformation {
  public:
    Transformation() {}
    explicit Transformation(ComputeLayer<elemType> *pLayer) : _pComptLayer(pLayer) { Configure(pLayer, false); }
    Transformation(elemType nodataLoad, elemType validLoad, ComputeLayer<elemType> *pLayer)
        : _nodataLoad(nodataLoad), _validLoad(validLoad), _pComptLayer(pLayer)
    {
        Configure(pLayer, false);
    }
    virtual ~Transformation() {}

    virtual bool isTermination(bool isIter = false) { return isIter; }
    bool Configure(ComputeLayer<elemType> *pLayer, bool isCommunication)
    {
        (void)isCommunication;
        _pComptLayer = pLayer;
        return pLayer != nullptr;
    }
    bool paramInit()
    {
        if (!_pComptLayer || _pComptLayer->_vDataLayers.empty()) {
            cerr << "Datalayers used for calculating compute layer should not be null." << endl;
            return false;
        }
        const MetaData *m = _pComptLayer->_vDataLayers[0]->metaData();
        if (!m) return false;
        _noData = m->noData;
        _computGrain = _pComptLayer->getComputeGrain();
        _dataWorkBR = m->_localworkBR;
        return true;
    }

    // The intensity ratio over the whole compute layer.
    virtual bool run()
    {
        if (_nodataLoad < 0 || _validLoad < 0) {
            cerr << "The load specified cannot be negative." << endl;
            return false;
        }
        if (!paramInit() || !_pComptLayer->cellSpace()) return false;
        CellSpace<elemType> &computL = *_pComptLayer->cellSpace();
        pargo_ctx *ctx = Application::device();
        if (!ctx) return false;
        const RasterLayer<elemType> &first = *_pComptLayer->_vDataLayers[0];
        const int rows = first.cellSpace()->nRows(), cols = first.cellSpace()->nCols();
        const size_t n = (size_t)rows * cols;
        int lr = 0, lc = 0;
        pargo_grid g{rows, cols, 0, 0.0, _noData};
        if (!deviceOk(pargo_intensity_load(ctx, &g, nullptr, _dataWorkBR.minIRow(), _dataWorkBR.minICol(),
                                           _dataWorkBR.nRows(), _dataWorkBR.nCols(), (int)_computGrain,
                                           (double)_nodataLoad, (double)_validLoad, 0, nullptr, &lr, &lc),
                      "pargo_intensity_load"))
            return false;
        if (computL.nRows() != lr || computL.nCols() != lc) {
            cerr << "[ERROR] Transformation: the compute layer is " << computL.nRows() << " x " << computL.nCols()
                 << ", the granularity gives " << lr << " x " << lc << " (ComputeLayer::init not called?)" << endl;
            return false;
        }
        void *d_data = nullptr, *d_load = nullptr;
        bool good = deviceOk(pargo_alloc(ctx, n * sizeof(double), &d_data), "pargo_alloc(data)") &&
                    deviceOk(pargo_alloc(ctx, (size_t)lr * lc * sizeof(double), &d_load), "pargo_alloc(load)");
        std::vector<double> host(good ? n : 0), load(good ? (size_t)lr * lc : 0);
        for (size_t L = 0; good && L < _pComptLayer->_vDataLayers.size(); L++) {
            const RasterLayer<elemType> &layer = *_pComptLayer->_vDataLayers[L];
            if (!layer.cellSpace() || layer.cellSpace()->nRows() != rows || layer.cellSpace()->nCols() != cols) {
                cerr << "[ERROR] Transformation: data layer " << layer.name() << " differs in size" << endl;
                good = false;
                break;
            }
            const elemType *src = layer.cellSpace()->data();
            for (size_t i = 0; i < n; i++) host[i] = (double)src[i];
            pargo_grid gl{rows, cols, 0, 0.0, layer.metaData() ? layer.metaData()->noData : _noData};
            good = deviceOk(pargo_h2d(ctx, d_data, host.data(), n * sizeof(double)), "pargo_h2d(data)") &&
                   deviceOk(pargo_intensity_load(ctx, &gl, (const double *)d_data, _dataWorkBR.minIRow(),
                                                 _dataWorkBR.minICol(), _dataWorkBR.nRows(), _dataWorkBR.nCols(),
                                                 (int)_computGrain, (double)_nodataLoad, (double)_validLoad,
                                                 L > 0, (double *)d_load, &lr, &lc),
                            "pargo_intensity_load") &&
                   deviceOk(pargo_sync(ctx), "pargo_sync");
        }
        good = good && deviceOk(pargo_d2h(ctx, load.data(), d_load, load.size() * sizeof(double)), "pargo_d2h(load)") &&
               deviceOk(pargo_sync(ctx), "pargo_sync");
        for (void *p : {d_data, d_load})
            if (p) pargo_free(ctx, p);
        if (!good) return false;
        elemType *dst = computL.data();
        for (size_t i = 0; i < load.size(); i++) dst[i] = (elemType)load[i];
        return true;
    }

  protected:
    elemType _nodataLoad = 0;  ///< intensity value for NoData grid
    elemType _validLoad = 0;   ///< intensity value for valid grid
    double _noData = 0;
    double _computGrain = 1;
    CoordBR _dataWorkBR;
    ComputeLayer<elemType> *_pComptLayer = nullptr;
    int Termination = 1;
};

}  // namespace GPRO
