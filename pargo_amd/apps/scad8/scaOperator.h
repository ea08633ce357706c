// scaOperator.h -- D8 contributing area (apps/hydrology/scad8/scaOperator.h
// of the reference, class SCAOperator :13-41).
//
// Same interface (d8Layer, scaLayer, Run); WorkDevice() runs the whole work
// rectangle through pargo_d8_classify + pargo_sca_d8.  The result is the
// reference's: one accumulation sweep in row-major order (its loop stops
// there, scaOperator.cpp:126), areas in cells, NoData on NoData cells and
// on the border.  setConverge(true) repeats the sweep until it settles
// (full D8 flow accumulation), which the reference does not offer.
#pragma once

#include "../../gpro/rasterOperator.h"

using namespace GPRO;

class SCAOperator : public RasterOperator<double> {
  public:
    SCAOperator() {}
    ~SCAOperator() override {}
    void d8Layer(RasterLayer<double> &layerD);
    void scaLayer(RasterLayer<double> &layerD);
    void setConverge(bool converge) { _converge = converge; }
    bool isTermination() override;
    // There is no per-cell host rule: the operator runs on the GPU only.
    bool Operator(const CellCoord &coord, bool operFlag) override;
    bool WorkDevice(const CoordBR &workBR, bool &ok) override;

  protected:
    double _cellSize = 0;
    double _noData = 0;
    bool _converge = false;
    RasterLayer<double> *_pD8Layer = nullptr;
    RasterLayer<double> *_pSCALayer = nullptr;
    Neighborhood<double> *_pD8Nbrhood = nullptr;
};
