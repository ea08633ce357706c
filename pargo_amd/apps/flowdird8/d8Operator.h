// d8Operator.h -- D8 flow direction (apps/hydrology/flowdird8/d8Operator.h
// of the reference, class D8Operator :13-40).
//
// Same interface; WorkDevice() runs the whole work rectangle through
// pargo_slope_d8 (D8 only).  Output values follow the reference: the ESRI
// bit-sum of tied steepest directions (dirType 256) or the median tied
// window index 0..8 (dirType 8), -1 for flat cells and pits, NoData for
// windows holding NoData and for the border.
#pragma once

#include "../../gpro/rasterOperator.h"

using namespace GPRO;

class D8Operator : public RasterOperator<double> {
  public:
    D8Operator() {}
    explicit D8Operator(int type) : dirType(type) {}
    ~D8Operator() override {}
    void demLayer(RasterLayer<double> &layerD);
    void d8Layer(RasterLayer<double> &layerD);
    bool isTermination() override;
    // There is no per-cell host rule: the operator runs on the GPU only.
    bool Operator(const CellCoord &coord, bool operFlag) override;
    bool WorkDevice(const CoordBR &workBR, bool &ok) override;

  protected:
    double cellSize = 0;
    double noData = 0;
    int dirType = 0;
    int num = 0;
    RasterLayer<double> *_pDEMLayer = nullptr;
    RasterLayer<double> *_pD8Layer = nullptr;
    Neighborhood<double> *_pDEMNbrhood = nullptr;
};
