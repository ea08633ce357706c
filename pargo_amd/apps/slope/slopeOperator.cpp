// slopeOperator.cpp -- see slopeOperator.h.
#include "slopeOperator.h"

#include "../../gpro/focalDevice.h"
#include "../../../include/pargo.h"

SlopeAlgor GetSlopeAlgorithm(const string &arg)
{
    static const char *names[] = {"FD", "FFD", "MD", "SD", "SFD", "TFD", "TFDW"};
    for (int i = 0; i < 7; i++)
        if (StringMatch(arg, names[i])) return (SlopeAlgor)i;
    return FD;
}

void SlopeOperator::demLayer(RasterLayer<double> &layerD)
{
    _pDEMLayer = &layerD;
    _pDEMNbrhood = layerD.nbrhood();
    cellSize = layerD._pMetaData ? layerD._pMetaData->cellSize : 0;
    noData = layerD._pMetaData ? layerD._pMetaData->noData : -9999.;
    Configure(_pDEMLayer, false);
}

void SlopeOperator::slopeLayer(RasterLayer<double> &layerD)
{
    _pSlopeLayer = &layerD;
    Configure(_pSlopeLayer, false);
}

bool SlopeOperator::isTermination()
{
    num--;
    return num > 0;
}

bool SlopeOperator::Operator(const CellCoord &, bool)
{
    cerr << "[ERROR] SlopeOperator runs on the GPU only (WorkDevice)" << endl;
    return false;
}

bool SlopeOperator::WorkDevice(const CoordBR &workBR, bool &ok)
{
    // the enum order equals the library's method codes
    static_assert(PARGO_SLOPE_FD == FD && PARGO_SLOPE_TFDW == TFDW, "method codes");
    ok = false;
    if (!_pDEMLayer || !_pSlopeLayer) {
        cerr << "[ERROR] SlopeOperator: demLayer() and slopeLayer() must be set before Run()" << endl;
        return true;
    }
    if (!focalWindowIsMoore(_pDEMNbrhood, "SlopeOperator")) return true;
    ok = focalOnDevice(*_pDEMLayer, workBR, _pSlopeLayer, (int)calcAlgor, nullptr, 256);
    return true;
}
