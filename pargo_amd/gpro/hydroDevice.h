// hydroDevice.h -- runs libpargo's hydrology kernels on RasterLayers.
//
// scaOnDevice copies the D8 layer (float64 window-index codes, the
// flowdird8 -dirtype 8 output) to device memory, converts it to flow
// classes with the reference's own comparisons, accumulates, and copies the
// areas of the work rectangle back (the other cells keep the SCA layer's
// NoData from copyLayerInfo).  fillOnDevice does the same for depression
// filling.
#pragma once

#include "rasterLayer.h"

namespace GPRO {

// mode is PARGO_SCA_REFERENCE or PARGO_SCA_CONVERGE.
bool scaOnDevice(const RasterLayer<double> &d8, const CoordBR &workBR, RasterLayer<double> &sca, int mode);

// Depression filling of the DEM layer into wdem (pargo_pit_remove): the DEM
// goes to the device as float32 (a value float32 cannot hold is an error),
// the filled values of the work rectangle come back as doubles.
bool fillOnDevice(const RasterLayer<double> &dem, const CoordBR &workBR, RasterLayer<double> &wdem);

}  // namespace GPRO
