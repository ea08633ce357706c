// idwTransformation.cpp -- see idwTransformation.h.
#include "idwTransformation.h"

#include <climits>
#include <vector>

#include "../../../include/pargo.h"

bool IdwTransformation::run()
{
    if (!_pComptLayer || !_idwOperator || !_pComptLayer->cellSpace() || !_idwOperator->plan()) {
        cerr << "[ERROR] IdwTransformation: the compute layer must be initialised and the samples blocked" << endl;
        return false;
    }
    if (_pComptLayer->dataLayers().empty()) {
        cerr << "Datalayers used for calculating compute layer should not be null." << endl;
        return false;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    CellSpace<double> &cl = *_pComptLayer->cellSpace();
    const double nodata = _pComptLayer->dataLayers()[0]->metaData()->noData;
    int rows = 0, cols = 0;
    const int g = _pComptLayer->getComputeGrain();
    if (!deviceOk(pargo_idw_load(ctx, _idwOperator->plan(), _idwOperator->getNbrPoints(), g, nullptr, 0, 0, 0, nodata,
                                 nullptr, &rows, &cols),
                  "pargo_idw_load"))
        return false;
    if (rows != cl.nRows() || cols != cl.nCols()) {
        cerr << "[ERROR] IdwTransformation: compute layer is " << cl.nRows() << " x " << cl.nCols() << ", expected "
             << rows << " x " << cols << " at granularity " << g << endl;
        return false;
    }
    void *d_load = nullptr, *d_mask = nullptr;
    int mrows = 0, mcols = 0;
    double mnodata = 0;
    bool good = deviceOk(pargo_alloc(ctx, sizeof(double) * (size_t)rows * cols, &d_load), "pargo_alloc(load)");
    RasterLayer<int> *mask = _idwOperator->getMaskLayer();
    std::vector<int> mbuf;
    if (good && mask) {
        mnodata = mask->metaData() ? mask->metaData()->noData : 0;
        good = idwMaskForDevice(mask, mbuf, mrows, mcols) &&
               deviceOk(pargo_alloc(ctx, mbuf.size() * sizeof(int), &d_mask), "pargo_alloc(mask)") &&
               deviceOk(pargo_h2d(ctx, d_mask, mbuf.data(), mbuf.size() * sizeof(int)), "pargo_h2d(mask)");
    }
    good = good && deviceOk(pargo_idw_load(ctx, _idwOperator->plan(), _idwOperator->getNbrPoints(), g,
                                           (const int32_t *)d_mask, mrows, mcols, mnodata, nodata, (double *)d_load,
                                           &rows, &cols),
                            "pargo_idw_load");
    good = good && deviceOk(pargo_d2h(ctx, cl.data(), d_load, sizeof(double) * (size_t)rows * cols), "pargo_d2h(load)") &&
           deviceOk(pargo_sync(ctx), "pargo_sync");
    for (void *p : {d_load, d_mask})
        if (p) pargo_free(ctx, p);
    return good;
}
