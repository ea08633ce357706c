/* oracle_san.c -- drives every function of the CPU oracle
 * (oracle/pargo_oracle.c) on small random inputs, including degenerate
 * sizes, NoData, NaN and infinities, so that an AddressSanitizer and
 * UndefinedBehaviorSanitizer build (tests/test_sanitizers.py) can check the
 * oracle's indexing.  Results are not checked here: the oracle's values are
 * pinned by the tests/test_oracle_*.py suites.
 *
 * Usage: oracle_san <cases> <seed> */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int oracle_slope(const double *dem, int rows, int cols, float cellSize, double noData, int method, double *out,
                 int nthreads);
int oracle_d8(const double *dem, int rows, int cols, double noData, int dirType, double *out, int nthreads);
int oracle_idw_dims(const double *xs, const double *ys, int n, float cellSize, int *nRows, int *nCols,
                    double *minX, double *maxY);
int oracle_idw(const double *xs, const double *ys, const double *vals, int n, float cellSize, int K, int power,
               double buffer, double blockSize, const int32_t *mask, int maskRows, int maskCols, int maskNoData,
               double *out, int64_t *row_candidates, int64_t *row_blocks, int nthreads);
int oracle_idw_load(const double *xs, const double *ys, int n, float cellSize, int K, double blockSize,
                    int granularity, const int32_t *mask, int maskRows, int maskCols, double maskNoData,
                    double noData, double *out, int *loadRows, int *loadCols);
int oracle_val_row_dcmp(const double *load, int rows, int cols, double noData, int nSubSpcs, int *first);
int oracle_intensity_load(const double *data, int nlayers, const double *nodata, int rows, int cols, int row0,
                          int col0, int work_rows, int work_cols, int g, double nodataLoad, double validLoad,
                          double *load);
int oracle_sca_d8(const double *d8, int rows, int cols, double noData, int converge, double *sca_out, int *sweeps);
int oracle_pit_remove(const double *dem, int rows, int cols, double noData, double *wdem_out, int max_sweeps,
                      int *sweeps);
void oracle_smpl1d_dcmp(int glbBegin, int glbEnd, int nSub, int iSub, int *subBegin, int *subEnd);

static uint64_t st;
static uint64_t next(void)
{
    st += 0x9e3779b97f4a7c15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static double unif(void) { return (double)(next() >> 11) * 0x1p-53; }
static int irange(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }

static double cell_value(void)
{
    const int k = irange(0, 40);
    if (k == 0) return -9999.0;
    if (k == 1) return NAN;
    if (k == 2) return INFINITY;
    if (k == 3) return 1e300;
    if (k < 10) return (double)irange(0, 3); /* ties */
    return unif() * 100.0;
}

int main(int argc, char **argv)
{
    const int cases = argc > 1 ? atoi(argv[1]) : 200;
    st = argc > 2 ? strtoull(argv[2], NULL, 10) : 1;
    for (int t = 0; t < cases; t++) {
        const int rows = irange(1, 24), cols = irange(1, 24);
        const size_t n = (size_t)rows * cols;
        double *dem = malloc(n * sizeof(double)), *out = malloc(n * sizeof(double));
        for (size_t i = 0; i < n; i++) dem[i] = cell_value();
        for (int m = 0; m < 7; m++) oracle_slope(dem, rows, cols, (float)(0.5 + unif() * 50), -9999.0, m, out, 1);
        oracle_d8(dem, rows, cols, -9999.0, 256, out, 1);
        oracle_d8(dem, rows, cols, -9999.0, 8, out, 1);
        int sweeps = 0;
        oracle_pit_remove(dem, rows, cols, -9999.0, out, 200, &sweeps);
        /* D8 layers of window codes 0..8, -1, NoData and stray values */
        for (size_t i = 0; i < n; i++) {
            const int k = irange(0, 12);
            dem[i] = k <= 8 ? (double)k : (k == 9 ? -1.0 : (k == 10 ? -9999.0 : (k == 11 ? NAN : 3.5)));
        }
        oracle_sca_d8(dem, rows, cols, -9999.0, t & 1, out, &sweeps);
        /* loads and bands */
        for (size_t i = 0; i < n; i++) dem[i] = irange(0, 9) == 0 ? -9999.0 : unif() * 10.0;
        int *first = malloc((size_t)(rows + 2) * sizeof(int));
        oracle_val_row_dcmp(dem, rows, cols, -9999.0, irange(1, rows), first);
        const double nds[2] = {-9999.0, 0.0};
        const int g = irange(1, 5);
        double *load = malloc(((size_t)rows / g + 2) * ((size_t)cols / g + 2) * sizeof(double));
        oracle_intensity_load(dem, 1, nds, rows, cols, 0, 0, rows, cols, g, 0.5, 1.0, load);
        for (int k = 0; k < 4; k++) {
            int b = 0, e = 0;
            oracle_smpl1d_dcmp(0, rows - 1, 4, k, &b, &e);
        }
        free(first);
        free(load);
        /* IDW: samples in a small extent, with coincident points */
        const int ns = irange(1, 40);
        double *xs = malloc(ns * sizeof(double)), *ys = malloc(ns * sizeof(double)), *vs = malloc(ns * sizeof(double));
        for (int i = 0; i < ns; i++) {
            xs[i] = (i && irange(0, 5) == 0) ? xs[i - 1] : unif() * 900.0;
            ys[i] = (i && irange(0, 5) == 0) ? ys[i - 1] : unif() * 700.0;
            vs[i] = unif() * 50.0;
        }
        const float cs = (float)(20.0 + unif() * 80.0);
        int gr = 0, gc = 0;
        double minx = 0, maxy = 0;
        if (oracle_idw_dims(xs, ys, ns, cs, &gr, &gc, &minx, &maxy) == 0 && (size_t)gr * gc < 100000) {
            double *grid = malloc((size_t)gr * gc * sizeof(double));
            int64_t *rc = calloc((size_t)gr, sizeof(int64_t)), *rb = calloc((size_t)gr, sizeof(int64_t));
            const int K = irange(1, 12);
            /* buffer > 0 always: with buffer 0 and n < K the reference never stops */
            const double buffer = 50.0 + unif() * 500.0, block = 100.0 + unif() * 400.0;
            int32_t mask[6] = {1, 0, 1, 1, 7, 1};
            oracle_idw(xs, ys, vs, ns, cs, K, irange(1, 3), buffer, block, (t & 1) ? mask : NULL, 2, 3, 7, grid, rc, rb,
                       1);
            int lr = 0, lc = 0;
            if (ns >= K) {  /* with fewer samples than K the reference's estimate never ends */
                const int gran = irange(1, 4);
                if (oracle_idw_load(xs, ys, ns, cs, K, block, gran, NULL, 0, 0, 0.0, -9999.0, NULL, &lr, &lc) == 0) {
                    double *lo = malloc(((size_t)lr * lc + 1) * sizeof(double));
                    oracle_idw_load(xs, ys, ns, cs, K, block, gran, (t & 1) ? mask : NULL, 2, 3, 7.0, -9999.0, lo,
                                    &lr, &lc);
                    free(lo);
                }
            }
            free(grid);
            free(rc);
            free(rb);
        }
        free(xs);
        free(ys);
        free(vs);
        free(dem);
        free(out);
    }
    printf("oracle_san: %d cases\n", cases);
    return 0;
}
