// pitRemoveOperator.h -- depression filling (apps/hydrology/pitremove/
// pitRemoveOperator.h of the reference, class PitRemoveOperator :87-113).
//
// Same interface (demLayer, wdemLayer, Run); WorkDevice() fills the whole
// work rectangle through pargo_pit_remove, which gives the values the
// reference's sweeps settle at.
#pragma once

#include "../../gpro/rasterOperator.h"

using namespace GPRO;

class PitRemoveOperator : public RasterOperator<double> {
  public:
    PitRemoveOperator() {}
    ~PitRemoveOperator() override {}
    void demLayer(RasterLayer<double> &layerD);
    void wdemLayer(RasterLayer<double> &layerD);
    bool isTermination() override;
    // There is no per-cell host rule: the operator runs on the GPU only.
    bool Operator(const CellCoord &coord, bool operFlag) override;
    bool WorkDevice(const CoordBR &workBR, bool &ok) override;

  public:
    double cellSize = 0;
    double noData = 0;
    int _xSize = 0;  // rows
    int _ySize = 0;  // columns
    bool flag = true;
    RasterLayer<double> *_pDEMLayer = nullptr;
    RasterLayer<double> *_pwDEMLayer = nullptr;
    Neighborhood<double> *_pDEMNbrhood = nullptr;
};
