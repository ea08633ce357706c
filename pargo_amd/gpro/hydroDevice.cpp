// hydroDevice.cpp -- see hydroDevice.h.
#include "hydroDevice.h"

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/pargo.h"
#include "application.h"

namespace GPRO {

namespace {

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

bool isInteriorRect(const CoordBR &w, int rows, int cols)
{
    return w.minIRow() == 1 && w.minICol() == 1 && w.maxIRow() == rows - 2 && w.maxICol() == cols - 2;
}

// Writes the work rectangle of a device float64 result into the host layer
// (the other cells keep their values), or, for the interior rectangle, keeps
// the whole result on the device as the layer's copy.
bool deliverF64(pargo_ctx *ctx, const double *d_out, const CoordBR &workBR, int rows, int cols,
                RasterLayer<double> &layer)
{
    const size_t n = (size_t)rows * cols;
    std::vector<double> out(n);
    if (!deviceOk(pargo_d2h(ctx, out.data(), d_out, n * sizeof(double)), "pargo_d2h(result)")) return false;
    DeviceTransfers::d2h += n * sizeof(double);
    CellSpace<double> &cs = *layer.cellSpace();
#pragma omp parallel for schedule(static)
    for (int r = workBR.minIRow(); r <= workBR.maxIRow(); r++)
        for (int c = workBR.minICol(); c <= workBR.maxICol(); c++) cs[r][c] = out[(size_t)r * cols + c];
    return true;
}

}  // namespace

bool scaOnDevice(const RasterLayer<double> &d8, const CoordBR &workBR, RasterLayer<double> &sca, int mode)
{
    const MetaData *m = d8.metaData();
    if (!m || (!d8.hasCellSpace() && !d8.deviceCopy())) {
        cerr << "[ERROR] SCA operator: the D8 layer holds no data" << endl;
        return false;
    }
    const int rows = m->_localdims.nRows(), cols = m->_localdims.nCols();
    if (!sca.hasCellSpace() || !sca.metaData() || sca.metaData()->_localdims.nRows() != rows ||
        sca.metaData()->_localdims.nCols() != cols) {
        cerr << "[ERROR] SCA operator: output layer " << sca.name() << " does not match the D8 layer's " << rows
             << " x " << cols << " cells (copyLayerInfo was not called)" << endl;
        return false;
    }
    if (!workBR.valid(SpaceDims(rows, cols))) {
        cerr << "[ERROR] SCA operator: work rectangle (" << workBR << ") outside the raster" << endl;
        return false;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    DeviceBuffers buf(ctx);
    const size_t n = (size_t)rows * cols;
    pargo_grid g;
    g.rows = rows;
    g.cols = cols;
    g.pitch = 0;
    g.cellsize = m->cellSize;
    g.nodata = m->noData;
    // Flow classes: a D8 layer left on the device by the D8 operator with
    // dirType 8 already holds them (window index, 255 = flat, NoData mask);
    // otherwise the layer's float64 codes are classified on the device with
    // the reference's comparisons.
    const uint8_t *flow = nullptr;
    const uint32_t *mask = nullptr;
    const DeviceRaster *dd = d8.deviceCopy();
    if (dd && dd->form == DeviceForm::Codes8 && dd->dirType == 8) {
        flow = (const uint8_t *)dd->data;
        mask = dd->mask;
    } else {
        double *d_d8 = (double *)buf.alloc(n * sizeof(double), "pargo_alloc(d8)");
        uint8_t *d_flow = (uint8_t *)buf.alloc(n, "pargo_alloc(flow classes)");
        if (!d_d8 || !d_flow) return false;
        if (!staging::uploadF64(d8.cellSpace()->data(), n, d_d8)) return false;
        if (!deviceOk(pargo_d8_classify(ctx, &g, d_d8, d_flow), "pargo_d8_classify")) return false;
        flow = d_flow;
    }
    if (isInteriorRect(workBR, rows, cols)) {
        // the kernels write the reference's NoData on the border and on
        // NoData cells: the whole layer stays on the device
        DeviceRaster *out = sca.deviceOutput(DeviceForm::F64, false);
        if (!out) return false;
        if (!deviceOk(pargo_sca_d8(ctx, &g, flow, mask, mode, (double *)out->data), "pargo_sca_d8")) return false;
        if (!deviceOk(pargo_sync(ctx), "pargo_sync")) return false;
        sca.deviceWritten();
        return true;
    }
    double *d_sca = (double *)buf.alloc(n * sizeof(double), "pargo_alloc(sca)");
    if (!d_sca) return false;
    if (!deviceOk(pargo_sca_d8(ctx, &g, flow, mask, mode, d_sca), "pargo_sca_d8")) return false;
    return deliverF64(ctx, d_sca, workBR, rows, cols, sca);
}

bool fillOnDevice(const RasterLayer<double> &dem, const CoordBR &workBR, RasterLayer<double> &wdem)
{
    const MetaData *m = dem.metaData();
    if (!m || (!dem.hasCellSpace() && !dem.deviceCopy())) {
        cerr << "[ERROR] pit removal: the DEM layer holds no data" << endl;
        return false;
    }
    const int rows = m->_localdims.nRows(), cols = m->_localdims.nCols();
    if (!wdem.hasCellSpace() || !wdem.metaData() || wdem.metaData()->_localdims.nRows() != rows ||
        wdem.metaData()->_localdims.nCols() != cols) {
        cerr << "[ERROR] pit removal: output layer " << wdem.name() << " does not match the DEM's " << rows << " x "
             << cols << " cells (copyLayerInfo was not called)" << endl;
        return false;
    }
    if (!workBR.valid(SpaceDims(rows, cols))) {
        cerr << "[ERROR] pit removal: work rectangle (" << workBR << ") outside the raster" << endl;
        return false;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    const DeviceRaster *in = dem.deviceF32();  // the DEM's device copy, uploaded once
    if (!in) return false;
    pargo_grid g;
    g.rows = rows;
    g.cols = cols;
    g.pitch = 0;
    g.cellsize = m->cellSize;
    g.nodata = m->noData;
    int passes = 0;
    if (isInteriorRect(workBR, rows, cols)) {
        // NoData on the border, as copyLayerInfo left it: the whole layer
        // stays on the device (pargo_pit_remove returns when it is done)
        DeviceRaster *out = wdem.deviceOutput(DeviceForm::F64, false);
        if (!out) return false;
        if (!deviceOk(pargo_pit_remove(ctx, &g, (const float *)in->data, (double *)out->data, &passes),
                      "pargo_pit_remove"))
            return false;
        wdem.deviceWritten();
        return true;
    }
    DeviceBuffers buf(ctx);
    double *d_w = (double *)buf.alloc((size_t)rows * cols * sizeof(double), "pargo_alloc(filled)");
    if (!d_w) return false;
    if (!deviceOk(pargo_pit_remove(ctx, &g, (const float *)in->data, d_w, &passes), "pargo_pit_remove")) return false;
    return deliverF64(ctx, d_w, workBR, rows, cols, wdem);
}

}  // namespace GPRO
