// focalDevice.h -- runs the 3x3 focal kernels of libpargo on RasterLayers.
//
// The DEM layer is copied to device memory as float32 (exact for float32 and
// 16-bit integer GeoTIFFs; a value that float32 cannot hold is reported as
// an error rather than rounded), the fused slope/D8 kernel runs once over the
// raster, and the results are copied back into the work rectangle of the
// output layers in the reference's encoding: NoData windows get the layer's
// NoData, flat D8 cells get -1 (apps/hydrology/flowdird8/d8Operator.cpp:73-76).
#pragma once

#include "rasterLayer.h"

namespace GPRO {

// True when `nbr` makes the reference's operators read a 3x3 window
// ((int)sqrt(size) / 2 == 1, slopeOperator.cpp:54) and its bounding
// rectangle is [-1, 1] x [-1, 1], so that the work rectangle is the
// interior.  Otherwise prints why and returns false.
bool focalWindowIsMoore(const Neighborhood<double> *nbr, const char *op);

// slope_method is a PARGO_SLOPE_* value; dir_type is 8 or 256.  Either
// output layer may be null.
bool focalOnDevice(const RasterLayer<double> &dem, const CoordBR &workBR, RasterLayer<double> *slope,
                   int slope_method, RasterLayer<double> *d8, int dir_type);

}  // namespace GPRO
