// dcmp.cpp -- load-balanced row decomposition (host side of the C ABI).
#include <cmath>
#include <vector>

#include "../../include/pargo.h"
#include "faults.h"

namespace {
constexpr double kEps = 0.0000001;  // gpro/deComposition.h:24
}

static int load_row_bands(const double *load, int rows, int cols, double nodata, int nbands, int *band_first)
{
    if (!load || !band_first || rows < 1 || cols < 1)
        return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_load_row_bands: bad arguments");
    if (nbands < 1 || nbands > rows)
        return pargo::fail(nullptr, PARGO_E_INVALID, "invalid number of SubSpaces (%d) for %d rows", nbands, rows);
    // per-row load over the cells that are neither -9999 nor NoData
    pargo::alloc_point("bands.row_loads");
    std::vector<double> rowLoad(rows, 0.0);
    double total = 0.0;
    for (int r = 0; r < rows; r++) {
        const double *row = load + (size_t)r * cols;
        for (int c = 0; c < cols; c++)
            if ((row[c] + 9999) > kEps && (row[c] - nodata) > kEps) rowLoad[r] += row[c];
        total += rowLoad[r];
    }
    double aver = total / nbands, accu = 0.0;
    int begin = 0, nb = 0;
    band_first[nb++] = 0;
    for (int i = 0; i < nbands - 1; ++i) {
        double sub = 0.0, minDiff = 0.0;
        for (int r = begin; r < rows; r++) {
            sub += rowLoad[r];
            accu += rowLoad[r];
            const int rowsLeft = rows - 1 - r, bandsLeft = nbands - 1 - i;
            if (sub <= aver && rowsLeft > bandsLeft) {
                // still under the target: remember how far under
                if (std::fabs(minDiff) < kEps || (aver - sub) < minDiff) minDiff = aver - sub;
                continue;
            }
            if (std::fabs(minDiff) < kEps || (sub - aver) <= minDiff || rowsLeft == bandsLeft) {
                begin = r + 1;  // this row closes the band
            } else {
                accu -= rowLoad[r];  // the previous row was closer to the target
                begin = r;
            }
            aver = (total - accu) / (nbands - i - 1);
            band_first[nb++] = begin;
            break;
        }
    }
    if (nb != nbands) return pargo::fail(nullptr, PARGO_E_INVALID, "pargo_load_row_bands: placed %d of %d cuts", nb - 1,
                                         nbands - 1);  // unreachable: every pass places one cut
    band_first[nbands] = rows;
    return PARGO_OK;
}

extern "C" int pargo_load_row_bands(const double *load, int rows, int cols, double nodata, int nbands,
                                    int *band_first)
{
    return pargo::guard(nullptr, "pargo_load_row_bands",
                        [&]() -> int { return load_row_bands(load, rows, cols, nodata, nbands, band_first); });
}
