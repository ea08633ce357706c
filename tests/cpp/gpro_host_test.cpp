// gpro_host_test.cpp -- host-only checks of the gpro layer (no GPU):
// .nbr parsing and work rectangles, RasterLayer reading/writing and band
// reads, rowAtOtherLayer, ComputeLayer dimensions and load bands.
// Usage: gpro_host_test <dir with dem.tif mask_500.tif *.nbr> <scratch dir>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <sstream>

#include "../../pargo_amd/gpro/computeLayer.h"
#include "../../pargo_amd/gpro/focalDevice.h"
#include "../../pargo_amd/gpro/rasterLayer.h"

using namespace GPRO;

static int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                          \
        }                                                                        \
    } while (0)

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: gpro_host_test <data dir> <scratch dir>\n");
        return 2;
    }
    const std::string data = argv[1], tmp = argv[2];

    // .nbr files (neighborhood.h:626-662)
    {
        RasterLayer<double> l("l");
        CHECK(l.readNeighborhood((data + "/moore.nbr").c_str()));
        const Neighborhood<double> &n = *l.nbrhood();
        CHECK(n.size() == 9 && n.minIRow() == -1 && n.maxIRow() == 1 && n.minICol() == -1 && n.maxICol() == 1);
        CHECK(focalWindowIsMoore(l.nbrhood(), "test"));
        CoordBR w;
        CHECK(n.calcWorkBR(w, SpaceDims(124, 104)));
        CHECK(w.minIRow() == 1 && w.minICol() == 1 && w.maxIRow() == 122 && w.maxICol() == 102);
        CHECK(!n.calcWorkBR(w, SpaceDims(2, 104)));  // smaller than the window
    }
    {
        RasterLayer<double> l("l");
        CHECK(l.readNeighborhood((data + "/moore20.nbr").c_str()));  // count >= 100: 41 x 41 square
        CHECK(l.nbrhood()->size() == 41 * 41 && l.nbrhood()->minIRow() == -20 && l.nbrhood()->maxICol() == 20);
        std::fprintf(stderr, "(expected message follows)\n");
        CHECK(!focalWindowIsMoore(l.nbrhood(), "test"));
        RasterLayer<double> m("m");
        CHECK(m.readNeighborhood((data + "/local.nbr").c_str()));
        CHECK(m.nbrhood()->size() == 1 && m.nbrhood()->minIRow() == 0);
        CHECK(!m.readNeighborhood((data + "/no_such.nbr").c_str()));
    }
    {
        Neighborhood<double> n;
        std::fprintf(stderr, "(expected messages follow)\n");
        CHECK(!n.fromVect({CellCoord(0, 0), CellCoord(0, 1), CellCoord(0, 1)}));  // duplicate
        CHECK(!n.fromVect({CellCoord(0, 1)}));                                    // no centre
        std::istringstream is("3\n0 0 1.0\n0 1 2.5\n-1 0 1.0\n");
        is >> n;
        CHECK(n.size() == 3 && n.weight(1) == 2.5 && n.minIRow() == -1 && n.maxICol() == 1);
    }

    // RasterLayer read / copyLayerInfo / write (rasterLayer.h:740-1087)
    RasterLayer<double> dem("dem");
    CHECK(dem.readNeighborhood((data + "/moore.nbr").c_str()));
    CHECK(dem.readFile((data + "/dem.tif").c_str(), ROWWISE_DCMP));
    {
        const MetaData &m = *dem.metaData();
        CHECK(m.row == 124 && m.column == 104 && m.noData == -9999.0 && m.cellSize == 90.0f);
        CHECK(m._localworkBR.minIRow() == 1 && m._localworkBR.maxIRow() == 122);
        CHECK(dem.cellSpace()->count(-9999.0) == 5539);
        CHECK(std::fabs((*dem.cellSpace())[60][50] - 300.0) < 100.0);
    }
    RasterLayer<double> out("out");
    CHECK(out.copyLayerInfo(dem));
    CHECK(out.cellSpace()->count(-9999.0) == 124 * 104);
    (*out.cellSpace())[3][4] = 1.25;
    const std::string p = tmp + "/gpro_roundtrip.tif";
    CHECK(out.writeFile(p.c_str()));
    {
        RasterLayer<double> back("back");
        CHECK(back.readFile(p.c_str()));
        CHECK(back.metaData()->noData == -9999.0 && back.metaData()->pTransform[0] == dem.metaData()->pTransform[0]);
        CHECK((*back.cellSpace())[3][4] == 1.25 && back.cellSpace()->count(-9999.0) == 124 * 104 - 1);
        CHECK(back.metaData()->geoKeys.keys == dem.metaData()->geoKeys.keys);
    }

    // band read of a mask (rasterLayer.h:855-878, :927-955) and the mask
    // lookup rowAtOtherLayer (rasterLayer.h:1289-1319)
    {
        RasterLayer<int> mask("mask");
        mask.newLocalNbrhood();
        CHECK(mask.readFile((data + "/mask_500.tif").c_str(), CoordBR(CellCoord(0, 0), CellCoord(698, 862))));
        const MetaData &m = *mask.metaData();
        CHECK(m._glbDims.nRows() == 707 && m._localdims.nRows() == 699 && m._localworkBR.maxIRow() == 698);
        RasterLayer<double> grid("grid");
        MetaData g;
        g._glbDims = SpaceDims(699, 863);
        grid.setMetaData(g);
        // scale 707/699: row 698 would map to 705.99, capped at the local 698
        CHECK(grid.rowAtOtherLayer(&mask, 698) == 698);
        CHECK(grid.rowAtOtherLayer(&mask, 100) == (int)(707.0 / 699 * 100));
        CHECK(grid.colAtOtherLayer(&mask, 862) == 883);  // 862 * 885 / 863 = 883.97
        CHECK(grid.colAtOtherLayer(&mask, 0) == 0);
        RasterLayer<int> whole("whole");
        whole.newLocalNbrhood();
        CHECK(whole.readFile((data + "/mask_500.tif").c_str(), ROWWISE_DCMP));
        CHECK(whole.cellSpace()->nRows() == 707);
        for (int r = 0; r < 699; r++)
            for (int c = 0; c < 885; c += 7) CHECK((*whole.cellSpace())[r][c] == (*mask.cellSpace())[r][c]);
    }

    // ComputeLayer (computeLayer.h:208-260, 378-443)
    {
        RasterLayer<double> grid("grid");
        grid.newLocalNbrhood();
        MetaData g;
        g._glbDims = SpaceDims(699, 863);
        g.row = 699, g.column = 863, g.cellSize = 500.0f, g.noData = -9999;
        g.pTransform[0] = 100.0, g.pTransform[1] = 500.0, g.pTransform[3] = 9000.0, g.pTransform[5] = -500.0;
        grid.setMetaData(g);
        CHECK(grid.layerDcmp(ROWWISE_DCMP));
        ComputeLayer<double> cl("cl");
        cl.addRasterLayerSerial(&grid);
        CHECK(cl.init(nullptr, 10));
        CHECK(cl.metaData()->_glbDims.nRows() == 70 && cl.metaData()->_glbDims.nCols() == 87);
        CHECK(cl.metaData()->cellSize == 5000.0f && cl.metaData()->pTransform[1] == 5000.0);
        cl.cellSpace()->initVals(1.0);
        CoordBR band;
        CHECK(cl.getCompuLoad(ROWWISE_DCMP, 1, band));
        CHECK(band.minIRow() == 0 && band.maxIRow() == 698 && band.maxICol() == 862);
        vector<CoordBR> bands;
        CHECK(cl.getCompuLoadBands(2, bands));
        // load rows cut at 36 (70 uniform rows, the exact-average quirk),
        // scaled by 699/70
        CHECK(bands.size() == 2 && bands[0].minIRow() == 0 && bands[1].maxIRow() == 698);
        CHECK(bands[1].minIRow() == (int)(36 * (699 / 70.0)) && bands[0].maxIRow() == bands[1].minIRow() - 1);
        CHECK(!cl.getCompuLoadBands(71, bands));
    }
    // The D8 kernels divide by sqrt(2.0) with q0 = x*y; r = fma(-q0, S, x);
    // q = fma(r, y, q0) (pargo_amd/csrc/focal_common.h).  Check that this
    // equals IEEE x / sqrt(2.0) over every binade that a difference of two
    // float32 values can occupy (2^-149 .. 2^129), with random mantissas.
    {
        const double S = 0x1.6a09e667f3bcdp+0, Y = 0x1.6a09e667f3bccp-1;
        CHECK(S == std::sqrt(2.0));
        std::mt19937_64 rng(12345);
        long long bad = 0, n = 0;
        for (int e = -149; e <= 129; e++) {
            for (int i = 0; i < 4096; i++) {
                const uint64_t mant = rng() & ((1ull << 52) - 1);
                double x = std::ldexp(1.0 + std::ldexp((double)mant, -52), e);
                if (i & 1) x = -x;
                if (i == 0) x = std::ldexp(1.0, e);
                const volatile double xv = x;
                const double q0 = xv * Y;
                const double r = std::fma(-q0, S, xv);
                const double q = std::fma(r, Y, q0);
                const volatile double ref = xv / std::sqrt(2.0);
                bad += (q != ref);
                n++;
            }
        }
        CHECK(bad == 0);
        std::printf("division by sqrt(2): %lld of %lld values differ from IEEE\n", bad, n);
    }
    if (failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
    else std::printf("gpro host checks passed\n");
    return failures ? 1 : 0;
}
