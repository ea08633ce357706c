// shapefile.h -- GDAL/OGR-free reading of point shapefiles for IDW.
//
// Replaces the OGR calls of IDWOperator::readSampleNums / readSamples
// (apps/spatial/idw/idwOperator.cpp:26-149): every non-deleted record of
// <name>.shp must be a point (Point, PointZ or PointM); its x and y become
// the sample position and the text of DBF field `field_index` (0-based, in
// the DBF field order, like OGRFeature::GetFieldAsDouble) parsed with
// strtod becomes the value.  The projection is the text of <name>.prj.
#pragma once

#include <string>
#include <vector>

namespace GPRO {
namespace shapefile {

struct Points {
    std::vector<double> x, y, value;
    std::string projection_wkt;  // contents of the .prj file, if any
    std::vector<std::string> field_names;
};

// `path` is the .shp file; a negative field_index reads no values (zeros).
// On a non-point record the reference prints
// "[ERROR] No point geometry" and stops reading; this returns false.
bool read_points(const std::string &path, int field_index, Points &out, std::string &err);

}  // namespace shapefile
}  // namespace GPRO
