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
    const CellSpace<double> *cs = dem.cellSpace();
    if (!m || !cs || cs->empty()) {
        cerr << "[ERROR] focal operator: the DEM layer holds no data" << endl;
        return false;
    }
    const int rows = cs->nRows(), cols = cs->nCols();
    for (RasterLayer<double> *out : {slope, d8}) {
        if (out && (!out->cellSpace() || out->cellSpace()->nRows() != rows || out->cellSpace()->nCols() != cols)) {
            cerr << "[ERROR] focal operator: output layer " << out->name() << " does not match the DEM's "
                 << rows << " x " << cols << " cells (copyLayerInfo was not called)" << endl;
            return false;
        }
    }
    if (!workBR.valid(SpaceDims(rows, cols))) {
        cerr << "[ERROR] focal operator: work rectangle (" << workBR << ") outside the raster" << endl;
        return false;
    }
    // float32 staging of the DEM, refusing values float32 would round
    const size_t n = (size_t)rows * cols;
    std::vector<float> host(n);
    const double *src = cs->data();
    for (size_t i = 0; i < n; i++) {
        const float f = (float)src[i];
        if ((double)f != src[i] && !(std::isnan(src[i]))) {
            cerr << "[ERROR] focal operator: DEM value " << src[i] << " at row " << i / cols << ", column "
                 << i % cols << " is not representable in float32, the device storage type of DEMs" << endl;
            return false;
        }
        host[i] = f;
    }
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    DeviceBuffers buf(ctx);
    const size_t mask_words = (size_t)rows * ((cols + 31) / 32);
    float *d_dem = (float *)buf.alloc(n * sizeof(float), "pargo_alloc(dem)");
    uint32_t *d_mask = (uint32_t *)buf.alloc(mask_words * 4, "pargo_alloc(nodata mask)");
    float *d_slope = slope ? (float *)buf.alloc(n * sizeof(float), "pargo_alloc(slope)") : nullptr;
    uint8_t *d_code = d8 ? (uint8_t *)buf.alloc(n, "pargo_alloc(d8)") : nullptr;
    if (!d_dem || !d_mask || (slope && !d_slope) || (d8 && !d_code)) return false;

    pargo_grid g;
    g.rows = rows;
    g.cols = cols;
    g.pitch = 0;
    g.cellsize = m->cellSize;
    g.nodata = m->noData;
    if (!deviceOk(pargo_h2d(ctx, d_dem, host.data(), n * sizeof(float)), "pargo_h2d(dem)")) return false;
    if (!deviceOk(pargo_slope_d8(ctx, &g, d_dem, slope_method, PARGO_UNIT_TAN, dir_type == 8 ? 8 : 256, d_slope,
                                 d_code, d_mask),
                  "pargo_slope_d8"))
        return false;
    std::vector<uint32_t> mask(mask_words);
    std::vector<float> hs(slope ? n : 0);
    std::vector<uint8_t> hc(d8 ? n : 0);
    if (!deviceOk(pargo_d2h(ctx, mask.data(), d_mask, mask_words * 4), "pargo_d2h(mask)")) return false;
    if (slope && !deviceOk(pargo_d2h(ctx, hs.data(), d_slope, n * sizeof(float)), "pargo_d2h(slope)")) return false;
    if (d8 && !deviceOk(pargo_d2h(ctx, hc.data(), d_code, n), "pargo_d2h(d8)")) return false;
    if (!deviceOk(pargo_sync(ctx), "pargo_sync")) return false;

    const size_t wpr = (size_t)(cols + 31) / 32;
    const uint8_t flat = dir_type == 8 ? PARGO_D8_FLAT_8 : PARGO_D8_FLAT_256;
    for (int r = workBR.minIRow(); r <= workBR.maxIRow(); r++) {
        for (int c = workBR.minICol(); c <= workBR.maxICol(); c++) {
            const size_t i = (size_t)r * cols + c;
            const bool nd = (mask[r * wpr + c / 32] >> (c % 32)) & 1u;
            if (slope) (*slope->cellSpace())[r][c] = nd ? slope->metaData()->noData : (double)hs[i];
            if (d8) (*d8->cellSpace())[r][c] = nd ? d8->metaData()->noData : (hc[i] == flat ? -1.0 : (double)hc[i]);
        }
    }
    return true;
}

}  // namespace GPRO
