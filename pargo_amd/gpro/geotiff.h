// geotiff.h -- GDAL-free GeoTIFF reading and writing for the gpro layer.
//
// Covers what the focal-operator programs exchange with the reference's
// GDAL calls (gpro/rasterLayer.h:749-778 readFileInfo, :882-903 RasterIO,
// :957-1087 createFile / rowWriteFile):
//   read:  classic or BigTIFF, little or big endian, strips or tiles, no
//          compression, LZW or PackBits, horizontal predictor, one band of
//          8/16/32-bit integers or 32/64-bit floats; the geotransform from
//          ModelPixelScale + ModelTiepoint (or ModelTransformation), the
//          GeoKey directory as raw tags, and GDAL_NODATA;
//   write: classic TIFF, one row per strip, uncompressed Float64 (what the
//          reference writes: GTiff of RasterLayer<double>, BLOCKYSIZE=1),
//          float32, int32 or uint8, with the geotransform, the GeoKeys of the input
//          and GDAL_NODATA.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace GPRO {
namespace geotiff {

struct Info {
    int rows = 0, cols = 0;
    int bits = 0;         // bits per sample
    int sample_format = 1;  // 1 uint, 2 int, 3 float
    double transform[6] = {0, 1, 0, 0, 0, -1};  // GDAL geotransform
    bool has_nodata = false;
    double nodata = 0;
    std::string nodata_text;
    // GeoTIFF projection keys copied verbatim to outputs
    std::vector<uint16_t> geokeys;
    std::vector<double> geodoubles;
    std::string geoascii;
};

// Reads band 1 of `path` converted to double (the widening GDAL RasterIO
// performs into a double buffer).  Returns false and sets `err` on failure.
bool read(const std::string &path, Info &info, std::vector<double> &data, std::string &err);

// Reads only the header (dimensions, georeferencing, NoData).
bool read_info(const std::string &path, Info &info, std::string &err);

// Sample type written by `write`.
enum class Type { Float64, Float32, Int32, UInt8 };

// Writes rows x cols values (row-major) with the georeferencing of `info`.
bool write(const std::string &path, const Info &info, const double *data, Type type, std::string &err);

}  // namespace geotiff
}  // namespace GPRO
