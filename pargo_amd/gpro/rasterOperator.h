// rasterOperator.h -- operator base class of the gpro interface
// (gpro/rasterOperator.h of the reference: RasterOperator<elemType> :38-88,
// comptLayer :110-116, Configure :126-146, Work :147-353, Run :355-361).
//
// The reference's Work() visits every cell of the work rectangle and calls
// the virtual Operator(coord) on each, repeating the sweep until the
// operator sets Termination; its CUDA_Type branch is empty.  Here Work()
// first offers the whole rectangle to WorkDevice(), which an operator with a
// GPU implementation (slope, D8 flow direction, IDW) overrides to run one
// device call; those operators never fall back to the per-cell loop.  The
// per-cell loop remains for user-written operators that have no device code.
#pragma once

#include "application.h"
#include "rasterLayer.h"

namespace GPRO {

template <class elemType>
class RasterOperator {
  public:
    RasterOperator() {}
    virtual ~RasterOperator() {}

    // Per-cell rule of a user-written operator.
    virtual bool Operator(const CellCoord &, bool operFlag) { return operFlag; }
    virtual bool isTermination() { return false; }

    // Whole-rectangle device implementation.  Returns false when the
    // operator has none; otherwise sets `ok` to the outcome.
    virtual bool WorkDevice(const CoordBR &workBR, bool &ok) { return (void)workBR, (void)ok, false; }

    bool Work(const CoordBR *pWorkBR)
    {
        if (!pWorkBR) {
            cerr << "[ERROR] RasterOperator::Work: no work rectangle (Configure was not called)" << endl;
            return false;
        }
        bool ok = true;
        if (WorkDevice(*pWorkBR, ok)) return ok;
        // per-cell sweep, repeated while an operator clears Termination
        do {
            Termination = 1;
            if (Application::_programType == MPI_OpenMP_Type) {
                int failed = 0;
#pragma omp parallel for reduction(| : failed)
                for (int r = pWorkBR->minIRow(); r <= pWorkBR->maxIRow(); r++)
                    for (int c = pWorkBR->minICol(); c <= pWorkBR->maxICol(); c++)
                        if (!Operator(CellCoord(r, c), true)) failed = 1;
                if (failed) ok = false;
            } else {
                for (int r = pWorkBR->minIRow(); r <= pWorkBR->maxIRow() && ok; r++)
                    for (int c = pWorkBR->minICol(); c <= pWorkBR->maxICol(); c++)
                        if (!Operator(CellCoord(r, c), true)) {
                            cout << "Operator is not successes!" << endl;
                            ok = false;
                            break;
                        }
            }
        } while (ok && !Termination);
        return ok;
    }

    // Takes the work rectangle and decomposition of `pLayer`
    // (rasterOperator.h:126-146); communication is a no-op with one process.
    bool Configure(RasterLayer<elemType> *pLayer, bool isCommunication)
    {
        if (!pLayer || !pLayer->_pMetaData) {
            cerr << "[ERROR] RasterOperator::Configure: layer without metadata" << endl;
            return false;
        }
        _pWorkBR = &pLayer->_pMetaData->_localworkBR;
        _domDcmpType = pLayer->_pMetaData->_domDcmpType;
        if (isCommunication) {
            CommVec.push_back(pLayer);
            commFlag = true;
        }
        return true;
    }

    bool Run() { return Work(_pWorkBR); }

    void comptLayer(RasterLayer<elemType> &layerD)
    {
        _pComptLayer = &layerD;
        _pComptLayer->cellSpace()->initVals(0);
    }

  protected:
    DomDcmpType _domDcmpType = NON_DCMP;

  public:
    RasterLayer<elemType> *_pComptLayer = nullptr;
    vector<RasterLayer<elemType> *> CommVec;
    CoordBR *_pWorkBR = nullptr;
    bool commFlag = false;
    int Termination = 1;
    double computeTimeExceptLastCell = 0;
    double operatorReduceTime = 0;
    bool _writePreExpLoad = false;
};

}  // namespace GPRO
