// io.cpp -- C ABI of the GDAL-free GeoTIFF and point-shapefile codecs
// (host only; the codecs live in pargo_amd/gpro).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pargo.h"
#include "../gpro/geotiff.h"
#include "../gpro/shapefile.h"
#include "faults.h"

namespace {

void to_abi(const GPRO::geotiff::Info &in, pargo_raster_info *out)
{
    out->rows = in.rows;
    out->cols = in.cols;
    for (int i = 0; i < 6; i++) out->transform[i] = in.transform[i];
    out->has_nodata = in.has_nodata ? 1 : 0;
    out->nodata = in.nodata;
    out->bits = in.bits;
    out->sample_format = in.sample_format;
}

}  // namespace

extern "C" {

int pargo_geotiff_info(const char *path, pargo_raster_info *info)
{
    return pargo::guard(nullptr, "pargo_geotiff_info", [&]() -> int {
        if (!path || !info) return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_geotiff_info: null argument");
        GPRO::geotiff::Info gi;
        std::string err;
        if (!GPRO::geotiff::read_info(path, gi, err)) return pargo::fail(nullptr, PARGO_E_INVALID, "%s", err.c_str());
        to_abi(gi, info);
        return PARGO_OK;
    });
}

int pargo_geotiff_read(const char *path, pargo_raster_info *info, double *data)
{
    return pargo::guard(nullptr, "pargo_geotiff_read", [&]() -> int {
        if (!path || !info || !data) return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_geotiff_read: null argument");
        GPRO::geotiff::Info gi;
        std::vector<double> v;
        std::string err;
        if (!GPRO::geotiff::read(path, gi, v, err)) return pargo::fail(nullptr, PARGO_E_INVALID, "%s", err.c_str());
        to_abi(gi, info);
        std::memcpy(data, v.data(), v.size() * sizeof(double));
        return PARGO_OK;
    });
}

int pargo_geotiff_write(const char *path, const pargo_raster_info *info, const double *data, int type,
                        const char *georef_from)
{
    return pargo::guard(nullptr, "pargo_geotiff_write", [&]() -> int {
        if (!path || !info || !data || info->rows < 1 || info->cols < 1)
            return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_geotiff_write: bad arguments");
        if (type < PARGO_IO_FLOAT64 || type > PARGO_IO_UINT8)
            return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_geotiff_write: unknown sample type %d", type);
        GPRO::geotiff::Info gi;
        std::string err;
        if (georef_from && !GPRO::geotiff::read_info(georef_from, gi, err))
            return pargo::fail(nullptr, PARGO_E_INVALID, "%s", err.c_str());
        gi.rows = info->rows;
        gi.cols = info->cols;
        for (int i = 0; i < 6; i++) gi.transform[i] = info->transform[i];
        gi.has_nodata = info->has_nodata != 0;
        gi.nodata = info->nodata;
        static const GPRO::geotiff::Type types[] = {GPRO::geotiff::Type::Float64, GPRO::geotiff::Type::Float32,
                                                    GPRO::geotiff::Type::Int32, GPRO::geotiff::Type::UInt8};
        if (!GPRO::geotiff::write(path, gi, data, types[type], err))
            return pargo::fail(nullptr, PARGO_E_INVALID, "%s", err.c_str());
        return PARGO_OK;
    });
}

int pargo_points_read(const char *shp_path, int field_index, int *n, double *xs, double *ys, double *vals)
{
    return pargo::guard(nullptr, "pargo_points_read", [&]() -> int {
        if (!shp_path || !n) return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_points_read: null argument");
        GPRO::shapefile::Points p;
        std::string err;
        if (!GPRO::shapefile::read_points(shp_path, field_index, p, err))
            return pargo::fail(nullptr, PARGO_E_INVALID, "%s", err.c_str());
        const int count = (int)p.x.size();
        if (xs || ys || vals) {
            if (*n < count) return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_points_read: room for %d of %d points", *n, count);
            for (int i = 0; i < count; i++) {
                if (xs) xs[i] = p.x[i];
                if (ys) ys[i] = p.y[i];
                if (vals) vals[i] = p.value[i];
            }
        }
        *n = count;
        return PARGO_OK;
    });
}

}  // extern "C"
