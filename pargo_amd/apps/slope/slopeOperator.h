// slopeOperator.h -- slope of a DEM (apps/morphology/slope/slopeOperator.h
// of the reference: enum SlopeAlgor :20-28, class SlopeOperator :30-55).
//
// Same interface; the per-cell Operator() of the reference is replaced by
// WorkDevice(), which runs the whole work rectangle through pargo_slope_d8
// (slope only).  Output is tan(slope) in the DEM's NoData convention.
#pragma once

#include "../../gpro/rasterOperator.h"

using namespace GPRO;

enum SlopeAlgor {
    FD,   ///< third-order finite difference weighted by 1/d^2 (Horn)
    FFD,  ///< frame finite difference
    MD,   ///< maximum downslope
    SD,   ///< simple difference
    SFD,  ///< second-order finite difference
    TFD,  ///< third-order finite difference
    TFDW  ///< third-order finite difference weighted by 1/d
};

// Method name to SlopeAlgor, case-insensitive; unknown names give FD
// (slopeOperator.cpp:3-26).
SlopeAlgor GetSlopeAlgorithm(const string &arg);

class SlopeOperator : public RasterOperator<double> {
  public:
    SlopeOperator() {}
    ~SlopeOperator() override {}
    void demLayer(RasterLayer<double> &layerD);
    void slopeLayer(RasterLayer<double> &layerD);
    void calcAlgorithm(SlopeAlgor algor) { calcAlgor = algor; }
    bool isTermination() override;
    // There is no per-cell host rule: the operator runs on the GPU only.
    bool Operator(const CellCoord &coord, bool operFlag) override;
    bool WorkDevice(const CoordBR &workBR, bool &ok) override;

  protected:
    double cellSize = 0;
    double noData = -9999.;
    int num = 0;
    SlopeAlgor calcAlgor = FD;
    RasterLayer<double> *_pDEMLayer = nullptr;
    RasterLayer<double> *_pSlopeLayer = nullptr;
    Neighborhood<double> *_pDEMNbrhood = nullptr;
};
