// focalDevice.cpp -- see focalDevice.h.
#include "focalDevice.h"

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/pargo.h"
#include "application.h"

namespace GPRO {

namespace {

// Device buffers released on every exit path.
struct DeviceBuffers {
    pargo_ctx *ctx;
    std::vector<void *> ptrs;
    explicit DeviceBuffers(pargo_ctx *c) : ctx(c) {}
    ~DeviceBuffers()
    {
        if (!ptrs.empty()) pargo_sync(ctx);
        for (void *p : ptrs) pargo_free(ctx, p);
    }
    void *alloc(size_t bytes, const char *what)
    {
        void *p = nullptr;
        if (!deviceOk(pargo_alloc(ctx, bytes ? bytes : 1, &p), what)) return nullptr;
        ptrs.push_back(p);
        return p;
    }
};

bool isInterior(const CoordBR &w, int rows, int cols)
{
    return w.minIRow() == 1 && w.minICol() == 1 && w.maxIRow() == rows - 2 && w.maxICol() == cols - 2;
}

// The outputs of one focal call: buffers, NoData mask and where they go.
struct FocalOut {
    float *slope = nullptr;
    uint8_t *code = nullptr;
    uint32_t *mask = nullptr;
};

bool runFocal(pargo_ctx *ctx, const pargo_grid &g, const float *dem, int method, int dirType, const FocalOut &o)
{
    return deviceOk(pargo_slope_d8(ctx, &g, dem, method, PARGO_UNIT_TAN, dirType, o.slope, o.code, o.mask),
                    "pargo_slope_d8");
}

}  // namespace

bool focalWindowIsMoore(const Neighborhood<double> *nbr, const char *op)
{
    if (!nbr || nbr->empty()) {
        cerr << "[ERROR] " << op << ": the DEM layer has no neighborhood (readNeighborhood was not called)" << endl;
        return false;
    }
    const int half = (int)std::sqrt((double)nbr->size()) / 2;
    if (half != 1 || nbr->minIRow() != -1 || nbr->maxIRow() != 1 || nbr->minICol() != -1 || nbr->maxICol() != 1) {
        cerr << "[ERROR] " << op << ": only the 3x3 Moore neighborhood is supported (got " << nbr->size()
             << " cells spanning rows " << nbr->minIRow() << ".." << nbr->maxIRow() << ", columns " << nbr->minICol()
             << ".." << nbr->maxICol() << ")" << endl;
        return false;
    }
    return true;
}

bool focalOnDevice(const RasterLayer<double> &dem, const CoordBR &workBR, RasterLayer<double> *slope,
                   int slope_method, RasterLayer<double> *d8, int dir_type)
{
    const MetaData *m = dem.metaData();
    if (!m || (!dem.hasCellSpace() && !dem.deviceCopy())) {
        cerr << "[ERROR] focal operator: the DEM layer holds no data" << endl;
        return false;
    }
    const int rows = m->_localdims.nRows(), cols = m->_localdims.nCols();
    for (RasterLayer<double> *out : {slope, d8}) {
        if (out && (!out->hasCellSpace() || !out->metaData() || out->metaData()->_localdims.nRows() != rows ||
                    out->metaData()->_localdims.nCols() != cols)) {
            cerr << "[ERROR] focal operator: output layer " << out->name() << " does not match the DEM's "
                 << rows << " x " << cols << " cells (copyLayerInfo was not called)" << endl;
            return false;
        }
    }
    if (!workBR.valid(SpaceDims(rows, cols))) {
        cerr << "[ERROR] focal operator: work rectangle (" << workBR << ") outside the raster" << endl;
        return false;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    // the DEM as float32 in HBM: its current device copy, or uploaded once
    // (and kept) from the host copy
    const DeviceRaster *in = dem.deviceF32();
    if (!in) return false;
    pargo_grid g;
    g.rows = rows;
    g.cols = cols;
    g.pitch = 0;
    g.cellsize = m->cellSize;
    g.nodata = m->noData;
    const int dt = dir_type == 8 ? 8 : 256;

    if (isInterior(workBR, rows, cols)) {
        // Whole-layer outputs stay on the device: the kernels write NoData
        // on the border, which is what copyLayerInfo left there.  Slope and
        // D8 each get their own NoData mask (their NoData tests differ).
        DeviceRaster *ds = slope ? slope->deviceOutput(DeviceForm::F32, true) : nullptr;
        DeviceRaster *dc = d8 ? d8->deviceOutput(DeviceForm::Codes8, true) : nullptr;
        if ((slope && !ds) || (d8 && !dc)) return false;
        if (ds && dc) {
            if (!deviceOk(pargo_slope(ctx, &g, (const float *)in->data, slope_method, PARGO_UNIT_TAN,
                                      (float *)ds->data, ds->mask),
                          "pargo_slope"))
                return false;
            if (!deviceOk(pargo_d8(ctx, &g, (const float *)in->data, dt, (uint8_t *)dc->data, dc->mask), "pargo_d8"))
                return false;
        } else {
            FocalOut o;
            o.slope = ds ? (float *)ds->data : nullptr;
            o.code = dc ? (uint8_t *)dc->data : nullptr;
            o.mask = ds ? ds->mask : dc->mask;
            if (!runFocal(ctx, g, (const float *)in->data, slope_method, dt, o)) return false;
        }
        if (dc) dc->dirType = dt;
        // results ready when the operator returns (its run time includes
        // them, and a failure is reported here)
        if (!deviceOk(pargo_sync(ctx), "pargo_sync")) return false;
        if (slope) slope->deviceWritten();
        if (d8) d8->deviceWritten();
        return true;
    }

    // Another work rectangle: compute the whole raster, then write the
    // rectangle's cells into the host layers (the others keep their values).
    DeviceBuffers buf(ctx);
    const size_t n = (size_t)rows * cols, wpr = (size_t)(cols + 31) / 32;
    FocalOut o;
    o.mask = (uint32_t *)buf.alloc((size_t)rows * wpr * 4, "pargo_alloc(nodata mask)");
    o.slope = slope ? (float *)buf.alloc(n * sizeof(float), "pargo_alloc(slope)") : nullptr;
    o.code = d8 ? (uint8_t *)buf.alloc(n, "pargo_alloc(d8)") : nullptr;
    if (!o.mask || (slope && !o.slope) || (d8 && !o.code)) return false;
    std::vector<uint32_t> mask((size_t)rows * wpr);
    std::vector<float> hs(slope ? n : 0);
    std::vector<uint8_t> hc(d8 ? n : 0);
    // one output per call when both are wanted: their NoData masks differ
    for (RasterLayer<double> *out : {slope, d8}) {
        if (!out) continue;
        const bool is_s = out == slope;
        FocalOut q;
        q.slope = is_s ? o.slope : nullptr;
        q.code = is_s ? nullptr : o.code;
        q.mask = o.mask;
        if (!runFocal(ctx, g, (const float *)in->data, slope_method, dt, q)) return false;
        if (!deviceOk(pargo_d2h(ctx, mask.data(), o.mask, mask.size() * 4), "pargo_d2h(mask)")) return false;
        if (is_s && !deviceOk(pargo_d2h(ctx, hs.data(), o.slope, n * sizeof(float)), "pargo_d2h(slope)")) return false;
        if (!is_s && !deviceOk(pargo_d2h(ctx, hc.data(), o.code, n), "pargo_d2h(d8)")) return false;
        DeviceTransfers::d2h += mask.size() * 4 + (is_s ? n * sizeof(float) : n);
        CellSpace<double> &cs = *out->cellSpace();
        const double nd = out->metaData()->noData;
        const uint8_t flat = dt == 8 ? PARGO_D8_FLAT_8 : PARGO_D8_FLAT_256;
#pragma omp parallel for schedule(static)
        for (int r = workBR.minIRow(); r <= workBR.maxIRow(); r++) {
            for (int c = workBR.minICol(); c <= workBR.maxICol(); c++) {
                const size_t i = (size_t)r * cols + c;
                const bool isnd = (mask[r * wpr + c / 32] >> (c % 32)) & 1u;
                cs[r][c] = isnd ? nd : (is_s ? (double)hs[i] : (hc[i] == flat ? -1.0 : (double)hc[i]));
            }
        }
    }
    return true;
}

}  // namespace GPRO
