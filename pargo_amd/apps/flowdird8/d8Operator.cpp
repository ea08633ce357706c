// d8Operator.cpp -- see d8Operator.h.
#include "d8Operator.h"

#include "../../gpro/focalDevice.h"

void D8Operator::demLayer(RasterLayer<double> &layerD)
{
    _pDEMLayer = &layerD;
    _pDEMNbrhood = layerD.nbrhood();
    cellSize = layerD._pMetaData ? layerD._pMetaData->cellSize : 0;
    noData = layerD._pMetaData ? layerD._pMetaData->noData : 0;
    Configure(_pDEMLayer, false);
}

void D8Operator::d8Layer(RasterLayer<double> &layerD)
{
    _pD8Layer = &layerD;
    Configure(_pD8Layer, false);
}

bool D8Operator::isTermination()
{
    num--;
    return num > 0;
}

bool D8Operator::Operator(const CellCoord &, bool)
{
    cerr << "[ERROR] D8Operator runs on the GPU only (WorkDevice)" << endl;
    return false;
}

bool D8Operator::WorkDevice(const CoordBR &workBR, bool &ok)
{
    ok = false;
    if (!_pDEMLayer || !_pD8Layer) {
        cerr << "[ERROR] D8Operator: demLayer() and d8Layer() must be set before Run()" << endl;
        return true;
    }
    if (dirType != 8 && dirType != 256) {
        cerr << "[ERROR] D8Operator: dirType must be 8 or 256, got " << dirType << endl;
        return true;
    }
    if (!focalWindowIsMoore(_pDEMNbrhood, "D8Operator")) return true;
    ok = focalOnDevice(*_pDEMLayer, workBR, nullptr, 0, _pD8Layer, dirType);
    return true;
}
