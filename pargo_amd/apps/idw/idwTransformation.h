// idwTransformation.h -- IDW compute-load estimate on a ComputeLayer
// (apps/spatial/idw/idwTransformation.h of the reference, class
// IdwTransformation :26-34; Operator in idwTransformation.cpp:3-62).
// run() fills every cell of the compute layer on the GPU (pargo_idw_load).
#pragma once

#include "../../gpro/computeLayer.h"
#include "idwOperator.h"

class IdwTransformation {
  public:
    IdwTransformation(ComputeLayer<double> *comptLayer, IDWOperator *idwOperator)
        : _pComptLayer(comptLayer), _idwOperator(idwOperator)
    {
    }
    bool run();

  private:
    ComputeLayer<double> *_pComptLayer;
    IDWOperator *_idwOperator;
};
