// hydroDevice.h -- runs libpargo's hydrology kernels on RasterLayers.
//
// scaOnDevice copies the D8 layer (float64 window-index codes, the
// flowdird8 -dirtype 8 output) to device memory, converts it to flow
// classes with the reference's own comparisons, accumulates, and copies the
// areas of the work rectangle back (the other cells keep the SCA layer's
// NoData from copyLayerInfo).
#pragma once

#include "rasterLayer.h"

namespace GPRO {

// mode is PARGO_SCA_REFERENCE or PARGO_SCA_CONVERGE.
bool scaOnDevice(const RasterLayer<double> &d8, const CoordBR &workBR, RasterLayer<double> &sca, int mode);

}  // namespace GPRO
