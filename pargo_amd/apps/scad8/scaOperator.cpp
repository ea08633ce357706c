// scaOperator.cpp -- see scaOperator.h.
#include "scaOperator.h"

#include "../../../include/pargo.h"
#include "../../gpro/focalDevice.h"
#include "../../gpro/hydroDevice.h"

void SCAOperator::d8Layer(RasterLayer<double> &layerD)
{
    _pD8Layer = &layerD;
    _pD8Nbrhood = layerD.nbrhood();
    _cellSize = layerD._pMetaData ? layerD._pMetaData->cellSize : 0;
    _noData = layerD._pMetaData ? layerD._pMetaData->noData : 0;
    Configure(_pD8Layer, false);
}

void SCAOperator::scaLayer(RasterLayer<double> &layerD)
{
    _pSCALayer = &layerD;
    Configure(_pSCALayer, true);
}

bool SCAOperator::isTermination()
{
    return true;
}

bool SCAOperator::Operator(const CellCoord &, bool)
{
    cerr << "[ERROR] SCAOperator runs on the GPU only (WorkDevice)" << endl;
    return false;
}

bool SCAOperator::WorkDevice(const CoordBR &workBR, bool &ok)
{
    ok = false;
    if (!_pD8Layer || !_pSCALayer) {
        cerr << "[ERROR] SCAOperator: d8Layer() and scaLayer() must be set before Run()" << endl;
        return true;
    }
    if (!focalWindowIsMoore(_pD8Nbrhood, "SCAOperator")) return true;
    ok = scaOnDevice(*_pD8Layer, workBR, *_pSCALayer, _converge ? PARGO_SCA_CONVERGE : PARGO_SCA_REFERENCE);
    return true;
}
