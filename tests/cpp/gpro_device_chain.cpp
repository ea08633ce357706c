// gpro_device_chain.cpp -- GPU checks of the device-resident gpro layer
// through the operator API (run by tests/test_gpu_gpro_chain.py):
//   * DEM -> D8Operator(8) -> SCAOperator moves no byte between host and
//     device between the two operators, and the DEM uploaded for D8 is not
//     uploaded again for SlopeOperator;
//   * the chained SCA equals, bit for bit, the SCA of the same D8 layer
//     taken through the host (the classification of its float64 codes);
//   * the device-resident slope and D8 layers, read back, equal the
//     per-rectangle host path's values bit for bit.
// Rasters: the reference's dem.tif and a 4096 x 4096 synthetic DEM.
// Usage: gpro_device_chain <dem.tif> <moore.nbr>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../pargo_amd/apps/flowdird8/d8Operator.h"
#include "../../pargo_amd/apps/scad8/scaOperator.h"
#include "../../pargo_amd/apps/slope/slopeOperator.h"
#include "../../pargo_amd/gpro/focalDevice.h"
#include "../../include/pargo.h"

static int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                          \
        }                                                                        \
    } while (0)

static bool same(const RasterLayer<double> &a, const RasterLayer<double> &b)
{
    const CellSpace<double> *x = a.cellSpace(), *y = b.cellSpace();
    if (!x || !y || x->nRows() != y->nRows() || x->nCols() != y->nCols()) return false;
    return std::memcmp(x->data(), y->data(), (size_t)x->size() * sizeof(double)) == 0;
}

// A DEM layer from pargo_synth_* (fractal terrain, 2% plateaus, 1% NoData).
static bool synthDem(RasterLayer<double> &dem, const char *nbr, int n)
{
    std::vector<float> z((size_t)n * n);
    if (pargo_synth_dem(0x5041524731ull, n, n, z.data(), 8) != PARGO_OK) return false;
    pargo_synth_plateaus(0x5041524731ull, n, n, z.data(), 0.02);
    int64_t cov = 0;
    pargo_synth_nodata_discs(0x5041524731ull, n, n, z.data(), 0, 0.01, -9999.0f, &cov);
    if (!dem.readNeighborhood(nbr)) return false;
    MetaData m;
    m.row = m.column = n;
    m.cellSize = 30.0f;
    m.noData = -9999.0;
    m.pTransform[0] = 0, m.pTransform[1] = 30, m.pTransform[3] = 30.0 * n, m.pTransform[5] = -30;
    m._glbDims = SpaceDims(n, n);
    dem.setMetaData(m);
    if (!dem.layerDcmp(ROWWISE_DCMP)) return false;
    double *d = dem.cellSpace()->data();
    for (size_t i = 0; i < z.size(); i++) d[i] = (double)z[i];
    return true;
}

static void chain(RasterLayer<double> &dem, const char *what)
{
    // DEM -> D8 (dirType 8) -> SCA, all on the device
    RasterLayer<double> d8("d8");
    d8.copyLayerInfo(dem);
    D8Operator d8op(8);
    d8op.demLayer(dem);
    d8op.d8Layer(d8);
    double t0 = Application::wtime();
    CHECK(d8op.Run());
    const double t_d8 = Application::wtime() - t0;
    CHECK(d8.deviceCopy() != nullptr);
    const unsigned long long h2d = DeviceTransfers::h2d, d2h = DeviceTransfers::d2h;
    RasterLayer<double> sca("sca");
    sca.copyLayerInfo(d8);
    SCAOperator scaop;
    scaop.d8Layer(d8);
    scaop.scaLayer(sca);
    t0 = Application::wtime();
    CHECK(scaop.Run());
    const double t_sca = Application::wtime() - t0;
    CHECK(DeviceTransfers::h2d == h2d && DeviceTransfers::d2h == d2h);  // nothing crossed PCIe
    // the slope of the same DEM reuses its device copy
    RasterLayer<double> slope("slope");
    slope.copyLayerInfo(dem);
    SlopeOperator sop;
    sop.demLayer(dem);
    sop.slopeLayer(slope);
    sop.calcAlgorithm(FD);
    CHECK(sop.Run());
    CHECK(DeviceTransfers::h2d == h2d);  // the DEM was not uploaded again

    // the same D8 layer through the host: its float64 codes classified on
    // the device; the SCA must not differ in any bit
    RasterLayer<double> d8h("d8h");
    d8h = d8;  // copies the host cells (brought up to date first)
    CHECK(d8h.deviceCopy() == nullptr);
    RasterLayer<double> scah("scah");
    scah.copyLayerInfo(d8h);
    SCAOperator scaop2;
    scaop2.d8Layer(d8h);
    scaop2.scaLayer(scah);
    CHECK(scaop2.Run());
    CHECK(same(sca, scah));

    // device-resident outputs against the per-rectangle host path (any
    // other work rectangle writes its cells into the host layers): a
    // rectangle one row short of the interior takes that path
    {
        const MetaData &m = *dem.metaData();
        const int rows = m._localdims.nRows(), cols = m._localdims.nCols();
        RasterLayer<double> s2("s2"), c2("c2");
        s2.copyLayerInfo(dem);
        c2.copyLayerInfo(dem);
        CoordBR part(CellCoord(1, 1), CellCoord(rows - 3, cols - 2));
        CHECK(focalOnDevice(dem, part, &s2, FD, &c2, 8));
        const CellSpace<double> &a = *slope.cellSpace(), &b = *s2.cellSpace();
        const CellSpace<double> &c = *d8.cellSpace(), &d = *c2.cellSpace();
        bool ok = true;
        for (int r = 1; r <= rows - 3 && ok; r++)
            ok = std::memcmp(a[r] + 1, b[r] + 1, (cols - 2) * sizeof(double)) == 0 &&
                 std::memcmp(c[r] + 1, d[r] + 1, (cols - 2) * sizeof(double)) == 0;
        CHECK(ok);
    }
    std::printf("%s (%d x %d): D8 %.2f ms, SCA %.2f ms (no transfer between them)\n", what,
                dem.metaData()->_localdims.nRows(), dem.metaData()->_localdims.nCols(), t_d8 * 1e3, t_sca * 1e3);
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: gpro_device_chain <dem.tif> <moore.nbr>\n");
        return 2;
    }
    if (!Application::START(HIP_Type, argc, argv)) return 1;
    {
        RasterLayer<double> dem("dem");
        CHECK(dem.readNeighborhood(argv[2]));
        CHECK(dem.readFile(argv[1], ROWWISE_DCMP));
        chain(dem, "dem.tif");
    }
    {
        RasterLayer<double> dem("synthetic");
        CHECK(synthDem(dem, argv[2], 4096));
        chain(dem, "synthetic DEM");
    }
    const bool ok = Application::END();
    if (failures || !ok) {
        std::fprintf(stderr, "%d gpro device chain checks failed\n", failures);
        return 1;
    }
    std::printf("gpro device chain checks passed\n");
    return 0;
}
