// pitRemoveOperator.cpp -- see pitRemoveOperator.h.
#include "pitRemoveOperator.h"

#include "../../gpro/focalDevice.h"
#include "../../gpro/hydroDevice.h"

void PitRemoveOperator::demLayer(RasterLayer<double> &layerD)
{
    _pDEMLayer = &layerD;
    _pDEMNbrhood = layerD.nbrhood();
    cellSize = layerD._pMetaData ? layerD._pMetaData->cellSize : 0;
    noData = layerD._pMetaData ? layerD._pMetaData->noData : 0;
    if (layerD.cellSpace()) {
        _xSize = layerD.cellSpace()->nRows();
        _ySize = layerD.cellSpace()->nCols();
    }
    Configure(_pDEMLayer, false);
}

void PitRemoveOperator::wdemLayer(RasterLayer<double> &layerD)
{
    _pwDEMLayer = &layerD;
    Configure(_pwDEMLayer, true);
}

bool PitRemoveOperator::isTermination()
{
    return flag;
}

bool PitRemoveOperator::Operator(const CellCoord &, bool)
{
    cerr << "[ERROR] PitRemoveOperator runs on the GPU only (WorkDevice)" << endl;
    return false;
}

bool PitRemoveOperator::WorkDevice(const CoordBR &workBR, bool &ok)
{
    ok = false;
    if (!_pDEMLayer || !_pwDEMLayer) {
        cerr << "[ERROR] PitRemoveOperator: demLayer() and wdemLayer() must be set before Run()" << endl;
        return true;
    }
    if (!focalWindowIsMoore(_pDEMNbrhood, "PitRemoveOperator")) return true;
    ok = fillOnDevice(*_pDEMLayer, workBR, *_pwDEMLayer);
    return true;
}
