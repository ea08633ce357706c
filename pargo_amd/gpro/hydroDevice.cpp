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

}  // namespace

bool scaOnDevice(const RasterLayer<double> &d8, const CoordBR &workBR, RasterLayer<double> &sca, int mode)
{
    const MetaData *m = d8.metaData();
    const CellSpace<double> *cs = d8.cellSpace();
    if (!m || !cs || cs->empty()) {
        cerr << "[ERROR] SCA operator: the D8 layer holds no data" << endl;
        return false;
    }
    const int rows = cs->nRows(), cols = cs->nCols();
    if (!sca.cellSpace() || sca.cellSpace()->nRows() != rows || sca.cellSpace()->nCols() != cols) {
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
    double *d_d8 = (double *)buf.alloc(n * sizeof(double), "pargo_alloc(d8)");
    uint8_t *d_flow = (uint8_t *)buf.alloc(n, "pargo_alloc(flow classes)");
    double *d_sca = (double *)buf.alloc(n * sizeof(double), "pargo_alloc(sca)");
    if (!d_d8 || !d_flow || !d_sca) return false;
    pargo_grid g;
    g.rows = rows;
    g.cols = cols;
    g.pitch = 0;
    g.cellsize = m->cellSize;
    g.nodata = m->noData;
    std::vector<double> out(n);
    if (!deviceOk(pargo_h2d(ctx, d_d8, cs->data(), n * sizeof(double)), "pargo_h2d(d8)")) return false;
    if (!deviceOk(pargo_d8_classify(ctx, &g, d_d8, d_flow), "pargo_d8_classify")) return false;
    if (!deviceOk(pargo_sca_d8(ctx, &g, d_flow, nullptr, mode, d_sca), "pargo_sca_d8")) return false;
    if (!deviceOk(pargo_d2h(ctx, out.data(), d_sca, n * sizeof(double)), "pargo_d2h(sca)")) return false;
    if (!deviceOk(pargo_sync(ctx), "pargo_sync")) return false;
    for (int r = workBR.minIRow(); r <= workBR.maxIRow(); r++)
        for (int c = workBR.minICol(); c <= workBR.maxICol(); c++) (*sca.cellSpace())[r][c] = out[(size_t)r * cols + c];
    return true;
}

bool fillOnDevice(const RasterLayer<double> &dem, const CoordBR &workBR, RasterLayer<double> &wdem)
{
    const MetaData *m = dem.metaData();
    const CellSpace<double> *cs = dem.cellSpace();
    if (!m || !cs || cs->empty()) {
        cerr << "[ERROR] pit removal: the DEM layer holds no data" << endl;
        return false;
    }
    const int rows = cs->nRows(), cols = cs->nCols();
    if (!wdem.cellSpace() || wdem.cellSpace()->nRows() != rows || wdem.cellSpace()->nCols() != cols) {
        cerr << "[ERROR] pit removal: output layer " << wdem.name() << " does not match the DEM's " << rows << " x "
             << cols << " cells (copyLayerInfo was not called)" << endl;
        return false;
    }
    if (!workBR.valid(SpaceDims(rows, cols))) {
        cerr << "[ERROR] pit removal: work rectangle (" << workBR << ") outside the raster" << endl;
        return false;
    }
    const size_t n = (size_t)rows * cols;
    std::vector<float> host(n);
    const double *src = cs->data();
    for (size_t i = 0; i < n; i++) {
        const float f = (float)src[i];
        if ((double)f != src[i] && !std::isnan(src[i])) {
            cerr << "[ERROR] pit removal: DEM value " << src[i] << " at row " << i / cols << ", column " << i % cols
                 << " is not representable in float32, the device storage type of DEMs" << endl;
            return false;
        }
        host[i] = f;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    DeviceBuffers buf(ctx);
    float *d_dem = (float *)buf.alloc(n * sizeof(float), "pargo_alloc(dem)");
    double *d_w = (double *)buf.alloc(n * sizeof(double), "pargo_alloc(filled)");
    if (!d_dem || !d_w) return false;
    pargo_grid g;
    g.rows = rows;
    g.cols = cols;
    g.pitch = 0;
    g.cellsize = m->cellSize;
    g.nodata = m->noData;
    std::vector<double> out(n);
    int passes = 0;
    if (!deviceOk(pargo_h2d(ctx, d_dem, host.data(), n * sizeof(float)), "pargo_h2d(dem)")) return false;
    if (!deviceOk(pargo_pit_remove(ctx, &g, d_dem, d_w, &passes), "pargo_pit_remove")) return false;
    if (!deviceOk(pargo_d2h(ctx, out.data(), d_w, n * sizeof(double)), "pargo_d2h(filled)")) return false;
    if (!deviceOk(pargo_sync(ctx), "pargo_sync")) return false;
    for (int r = workBR.minIRow(); r <= workBR.maxIRow(); r++)
        for (int c = workBR.minICol(); c <= workBR.maxICol(); c++) (*wdem.cellSpace())[r][c] = out[(size_t)r * cols + c];
    return true;
}

}  // namespace GPRO
