// metaData.h -- description of a raster layer (gpro/metaData.h of the
// reference: MetaData :45-75).  The fields keep the reference's names and
// meaning.  GDAL's GDALDataType is replaced by DataType, and the projection
// is carried in two forms: `projection` (WKT text, e.g. from a shapefile's
// .prj) and `geoKeys` (the GeoTIFF key directory of an input raster, copied
// verbatim to outputs).  This build runs one process per GPU with the whole
// raster resident, so myrank is 0, processor_number is 1 and the actual and
// virtual rectangles coincide.
#pragma once

#include <string>
#include <vector>

#include "basicTypes.h"

namespace GPRO {

enum DataType { DT_Unknown = 0, DT_Byte, DT_Int32, DT_Float32, DT_Float64 };

template <class T>
inline DataType dataTypeOf() { return DT_Float64; }
template <>
inline DataType dataTypeOf<float>() { return DT_Float32; }
template <>
inline DataType dataTypeOf<int>() { return DT_Int32; }
template <>
inline DataType dataTypeOf<unsigned char>() { return DT_Byte; }

// GeoTIFF georeferencing keys (GeoKeyDirectory, GeoDoubleParams,
// GeoAsciiParams) as read from an input file.
struct GeoKeys {
    std::vector<unsigned short> keys;
    std::vector<double> doubles;
    std::string ascii;
    bool empty() const { return keys.empty(); }
};

class MetaData {
  public:
    string format = "GTiff";  ///< output format; only GTiff is written
    string projection;        ///< WKT text, when known
    GeoKeys geoKeys;          ///< GeoTIFF projection keys, when known
    double pTransform[6] = {0, 1, 0, 0, 0, -1};  ///< GDAL geotransform
    double noData = -1e10;    ///< NoData value
    float cellSize = 1.0f;    ///< resolution (float, as in the reference)
    int row = 0;              ///< rows of the raster
    int column = 0;           ///< columns of the raster
    int myrank = 0;
    int processor_number = 1;
    DataType dataType = DT_Float64;
    DomDcmpType _domDcmpType = NON_DCMP;
    SpaceDims _glbDims;   ///< global dimensions
    CoordBR _MBR;         ///< this process's rectangle in global coordinates
    SpaceDims _localdims; ///< local dimensions
    CoordBR _localworkBR; ///< work rectangle in local coordinates

    int LoctoGloRow(int i) const { return _MBR.minIRow() + i; }
    int LoctoGloCol(int j) const { return _MBR.minICol() + j; }
};

}  // namespace GPRO
