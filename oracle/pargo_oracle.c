/*
 * pargo_oracle.c -- CPU restatement of PaRGO's focal-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline ("port") for the benchmark.  Only tests/, the smoke check and the
 * benchmark's cpu_baseline leg load it.  The product library (libpargo.so)
 * never links, loads or calls anything in this directory.
 *
 * Everything is FP64 and follows the reference's operation order:
 *   - Horn slope and the other -mtd methods .... apps/morphology/slope/slopeOperator.cpp:51-139
 *   - D8 flow direction (dirType 8 / 256) ....... apps/hydrology/flowdird8/d8Operator.cpp:28-102
 *   - IDW extent, blocks, ring search, weights .. apps/spatial/idw/idwOperator.cpp:5-20,126-169,332-491
 *                                                 apps/spatial/idw/idwOperator.h:68-73,88-94
 *   - mask row/col mapping ....................... gpro/rasterLayer.h:1288-1316
 *   - per-cell virtual Operator + Work loop ...... gpro/rasterOperator.h:147-252
 *   - row bands ................................. gpro/deComposition.h:83-115 (_smpl1DDcmp)
 *   - IDW load estimate, load-balanced bands ..... apps/spatial/idw/idwTransformation.cpp:3-62,
 *                                                 gpro/computeLayer.h:208-260,
 *                                                 gpro/deComposition.h:305-393 (valRowDcmp)
 *   - intensity-ratio compute load ............... gpro/transformation.h:155-228
 *   - D8 contributing area (scad8) ............... apps/hydrology/scad8/scaOperator.cpp:5-131
 *   - depression filling (pitremove) ............. apps/hydrology/pitremove/pitRemoveOperator.cpp:3-72
 *   - bounds-checked row access .................. gpro/cellSpace.h:471-491
 *
 * Built with -O2 -ffp-contract=off and without -ffast-math so that every
 * double operation rounds exactly where the reference's C++ source rounds.
 *
 * Parity pins (tests/test_oracle_*.py): closed forms for planes and
 * constants, hand-built 3x3 known-answer windows for every D8 direction,
 * tie and the 1e-6 epsilon boundaries, the documented reference quirks (TFD
 * reuses d[6], MD divides by the last neighbour's distance), brute-force KNN
 * for IDW with buffer 0 and a brute-force "blocks inside the buffer ring"
 * KNN for buffer > 0, and hand-derived tie-order cases for the strict-<
 * replacement rule; scad8 against closed forms, brute-force path counts and
 * two independent formulations; pitremove against a priority-flood solver;
 * the intensity load against closed forms.  The reference itself cannot be compiled in this
 * container (it needs GDAL), so no function here is checked against a
 * reference binary.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* CellSpace: row-major matrix with a bounds-checked row accessor      */
/* (gpro/cellSpace.h:471-491).                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    double *m;
    int rows, cols;
} cellspace;

static double *cs_row(const cellspace *cs, int iRow)
{
    if (iRow < 0 || iRow >= cs->rows) {
        fprintf(stderr, "oracle: row index [%d] out of CellSpace boundary [%d %d]\n",
                iRow, cs->rows, cs->cols);
        return NULL;
    }
    return cs->m + (size_t)iRow * (size_t)cs->cols;
}

/* A per-cell operator, called through a function pointer like the
 * reference's virtual RasterOperator::Operator (rasterOperator.h:59). */
typedef int (*cell_op)(void *self, int iRow, int iCol);

/* _smpl1DDcmp (deComposition.h:83-115): equal split, remainder to the
 * first bands. */
void oracle_smpl1d_dcmp(int glbBegin, int glbEnd, int nSub, int iSub,
                        int *subBegin, int *subEnd)
{
    int glbRange = glbEnd - glbBegin + 1;
    int remainder = glbRange % nSub;
    int blockSize;
    if (remainder == 0) {
        blockSize = glbRange / nSub;
        *subBegin = glbBegin + iSub * blockSize;
        *subEnd = *subBegin + blockSize - 1;
    } else {
        blockSize = glbRange / nSub + 1;
        if (iSub < remainder) {
            *subBegin = glbBegin + iSub * blockSize;
            *subEnd = *subBegin + blockSize - 1;
        } else if (iSub == remainder) {
            *subBegin = glbBegin + iSub * blockSize;
            *subEnd = *subBegin + blockSize - 2;
        } else {
            *subBegin = glbBegin + remainder * blockSize + (iSub - remainder) * (blockSize - 1);
            *subEnd = *subBegin + blockSize - 2;
        }
    }
}

/* Work loop (rasterOperator.h:171-185 single thread; :220-234 OpenMP).
 * With nthreads > 1 every thread takes one row band from _smpl1DDcmp,
 * which is what P MPI ranks of ROWWISE_DCMP compute for single-pass
 * operators. Returns 0 when every Operator call succeeded. */
static int work_loop(cell_op op, void *self, int r0, int r1, int c0, int c1, int nthreads)
{
    int failed = 0;
    if (r1 < r0 || c1 < c0) return 0;
    if (nthreads <= 1) {
        for (int iRow = r0; iRow <= r1; iRow++)
            for (int iCol = c0; iCol <= c1; iCol++)
                if (!op(self, iRow, iCol)) { failed = 1; break; }
        return failed ? -1 : 0;
    }
    int nb = nthreads;
    if (nb > r1 - r0 + 1) nb = r1 - r0 + 1;
#pragma omp parallel num_threads(nb) reduction(| : failed)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int b0, b1;
        oracle_smpl1d_dcmp(r0, r1, nb, t, &b0, &b1);
        for (int iRow = b0; iRow <= b1; iRow++)
            for (int iCol = c0; iCol <= c1; iCol++)
                if (!op(self, iRow, iCol)) { failed = 1; break; }
    }
    return failed ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* Slope (apps/morphology/slope/slopeOperator.cpp:51-139)              */
/* ------------------------------------------------------------------ */
enum { M_FD = 0, M_FFD = 1, M_MD = 2, M_SD = 3, M_SFD = 4, M_TFD = 5, M_TFDW = 6 };

typedef struct {
    cellspace dem, slope;
    double cellSize; /* SlopeOperator::cellSize is double, assigned from float MetaData::cellSize */
    double noData;
    int method;
    int nbrRadius;   /* ((int)sqrt(nbr.size()))/2, slopeOperator.cpp:55 */
} slope_op;

#define SLOPE_EPS 0.0000001 /* slopeOperator.h:17 */

static int slope_cell(void *vself, int iRow, int iCol)
{
    slope_op *s = (slope_op *)vself;
    int n = s->nbrRadius;
    double d[9];
    int k = 0, tag = 0;
    for (int iRow1 = iRow - n; iRow1 <= iRow + n; iRow1++) {
        const double *row = cs_row(&s->dem, iRow1);
        if (!row) return 0;
        for (int iCol1 = iCol - n; iCol1 <= iCol + n; iCol1++) {
            d[k] = row[iCol1];
            if (fabs(d[k] - s->noData) < SLOPE_EPS) tag = 1;
            k++;
        }
    }
    double *out = cs_row(&s->slope, iRow);
    if (!out) return 0;
    if (tag == 1) {
        out[iCol] = s->noData;
        return 1;
    }
    const double cellSize = s->cellSize;
    double dx, dy;
    switch (s->method) {
    case M_FD:
        dx = (d[8] + 2 * d[5] + d[2] - d[6] - 2 * d[3] - d[0]) / (8.0 * cellSize);
        dy = (d[2] + 2 * d[1] + d[0] - d[6] - 2 * d[7] - d[8]) / (8.0 * cellSize);
        out[iCol] = sqrt(dx * dx + dy * dy);
        break;
    case M_FFD:
        dx = (d[8] - d[6] + d[2] - d[0]) / (4.0 * cellSize);
        dy = (d[0] - d[6] + d[2] - d[8]) / (4.0 * cellSize);
        out[iCol] = sqrt(dx * dx + dy * dy);
        break;
    case M_MD: {
        /* slopeOperator.cpp:91-103: the distance divisor is whatever the
         * last visited neighbour (SE) set it to. */
        double max_diff = 0.;
        double hori_dist = cellSize;
        for (int tmpr = iRow - n; tmpr <= iRow + n; tmpr++) {
            const double *row = cs_row(&s->dem, tmpr);
            if (!row) return 0;
            for (int tmpc = iCol - n; tmpc <= iCol + n; tmpc++) {
                if (tmpr == iRow && tmpc == iCol) continue;
                if (max_diff < d[4] - row[tmpc]) max_diff = d[4] - row[tmpc];
                hori_dist = cellSize * sqrt((double)((tmpr - iRow) * (tmpr - iRow) + (tmpc - iCol) * (tmpc - iCol)));
            }
        }
        out[iCol] = max_diff / hori_dist;
        break;
    }
    case M_SD:
        dx = (d[4] - d[3]) / cellSize;
        dy = (d[4] - d[7]) / cellSize;
        out[iCol] = sqrt(dx * dx + dy * dy);
        break;
    case M_SFD:
        dx = (d[5] - d[3]) / (2.0 * cellSize);
        dy = (d[1] - d[7]) / (2.0 * cellSize);
        out[iCol] = sqrt(dx * dx + dy * dy);
        break;
    case M_TFD:
        /* slopeOperator.cpp:131 subtracts d[6] twice in dy; kept. */
        dx = (d[8] + d[5] + d[2] - d[0] - d[3] - d[6]) / (6.0 * cellSize);
        dy = (d[2] + d[1] + d[0] - d[6] - d[7] - d[6]) / (6.0 * cellSize);
        out[iCol] = sqrt(dx * dx + dy * dy);
        break;
    case M_TFDW:
        dx = (d[8] + sqrt(2.0) * d[5] + d[2] - d[6] - sqrt(2.0) * d[3] - d[0]) / ((4.0 + sqrt(8.0)) * cellSize);
        dy = (d[2] + sqrt(2.0) * d[1] + d[0] - d[6] - sqrt(2.0) * d[7] - d[8]) / ((4.0 + sqrt(8.0)) * cellSize);
        out[iCol] = sqrt(dx * dx + dy * dy);
        break;
    default:
        return 0;
    }
    return 1;
}

/* Whole-layer slope.  dem and out are rows*cols doubles; the output is first
 * filled with noData like copyLayerInfo (rasterLayer.h:739-745), then the
 * work rectangle [1, rows-2] x [1, cols-2] (neighborhood.h:447-448 for a
 * radius-1 neighbourhood) is swept. */
int oracle_slope(const double *dem, int rows, int cols, float cellSize, double noData,
                 int method, double *out, int nthreads)
{
    if (rows < 3 || cols < 3) return -1;
    for (size_t i = 0; i < (size_t)rows * cols; i++) out[i] = noData;
    slope_op s;
    s.dem.m = (double *)dem; s.dem.rows = rows; s.dem.cols = cols;
    s.slope.m = out; s.slope.rows = rows; s.slope.cols = cols;
    s.cellSize = (double)cellSize;
    s.noData = noData;
    s.method = method;
    s.nbrRadius = 1;
    return work_loop(slope_cell, &s, 1, rows - 2, 1, cols - 2, nthreads);
}

/* ------------------------------------------------------------------ */
/* D8 (apps/hydrology/flowdird8/d8Operator.cpp:28-102)                  */
/* ------------------------------------------------------------------ */
#define D8_EPS 0.000001 /* d8Operator.cpp:3 redefines Eps */

typedef struct {
    cellspace dem, d8;
    double noData;
    int dirType;
    int nbrRadius;
} d8_op;

static int d8_cell(void *vself, int iRow, int iCol)
{
    d8_op *s = (d8_op *)vself;
    int n = s->nbrRadius;
    double d[9];
    int k = 0, tag = 1;
    for (int iRow1 = iRow - n; iRow1 <= iRow + n; iRow1++) {
        const double *row = cs_row(&s->dem, iRow1);
        if (!row) return 0;
        for (int iCol1 = iCol - n; iCol1 <= iCol + n; iCol1++) {
            d[k] = row[iCol1];
            if (fabs(d[k] - s->noData) < D8_EPS) tag = 0;
            k++;
        }
    }
    double *out = cs_row(&s->d8, iRow);
    if (!out) return 0;
    if (!tag) {
        out[iCol] = s->noData;
        return 1;
    }
    int valid[9] = {0};
    double max = 0.0;
    double diff[9];
    for (int i = 0; i < 9; i++) {
        diff[i] = d[4] - d[i];
        if (i == 0 || i == 2 || i == 6 || i == 8) diff[i] = diff[i] / sqrt(2.0);
        if (diff[i] > max) max = diff[i];
    }
    if (max < D8_EPS) {
        out[iCol] = -1;
        return 1;
    }
    if (s->dirType == 8) {
        /* validDir[length>>1]: the median tied window index (:78-88). */
        int validDir[9], length = 0;
        for (int i = 0; i < 9; i++)
            if (fabs(diff[i] - max) < D8_EPS) validDir[length++] = i;
        out[iCol] = validDir[length >> 1];
    } else {
        for (int i = 0; i < 9; i++)
            if (fabs(diff[i] - max) < D8_EPS) valid[i] = 1;
        out[iCol] = valid[5] * pow(2.0, 0) + valid[8] * pow(2.0, 1) + valid[7] * pow(2.0, 2) + valid[6] * pow(2.0, 3)
                  + valid[3] * pow(2.0, 4) + valid[0] * pow(2.0, 5) + valid[1] * pow(2.0, 6) + valid[2] * pow(2.0, 7);
    }
    return 1;
}

int oracle_d8(const double *dem, int rows, int cols, double noData, int dirType,
              double *out, int nthreads)
{
    if (rows < 3 || cols < 3) return -1;
    if (!(dirType == 8 || dirType == 256)) return -2; /* pD8.cpp:250-253 */
    for (size_t i = 0; i < (size_t)rows * cols; i++) out[i] = noData;
    d8_op s;
    s.dem.m = (double *)dem; s.dem.rows = rows; s.dem.cols = cols;
    s.d8.m = out; s.d8.rows = rows; s.d8.cols = cols;
    s.noData = noData;
    s.dirType = dirType;
    s.nbrRadius = 1;
    return work_loop(d8_cell, &s, 1, rows - 2, 1, cols - 2, nthreads);
}

/* ------------------------------------------------------------------ */
/* IDW (apps/spatial/idw/idwOperator.{h,cpp}), 1-rank semantics         */
/* ------------------------------------------------------------------ */
#define IDW_EPS 0.0000001 /* idwOperator.h:13 */
#define IDW_NODATA -9999  /* idwOperator.h:14 */

typedef struct { double x, y, value; } sample_point;

typedef struct {
    /* IDWOperator members (idwOperator.h:94-113) */
    float cellSize;
    int nRows, nCols;
    int nbrPoints, idwPower;
    double buffer, blockSize;
    double minX, maxX, minY, maxY; /* _glb_extent == _sub_extent at 1 rank */
    int blockRows, blockCols;
    int *blockStart;                /* CSR over blocks, samples in file order */
    sample_point *blockPts;
    /* optional mask layer (RasterLayer<int>) */
    const int32_t *mask;
    int maskRows, maskCols;
    int maskNoData;
    cellspace out;
    /* statistics (not in the reference): candidate visits per row */
    int64_t *rowCandidates;
    int64_t *rowBlocks;
    int maxRingGuard;
    int failed;
} idw_op;

/* readSamples extent rule (idwOperator.cpp:88-141).  cellSize is float:
 * "_cellSize / 2" and "totalcol * _cellSize" are float expressions. */
void oracle_idw_extent(const double *xs, const double *ys, int n, float cellSize,
                       double *minX, double *maxX, double *minY, double *maxY,
                       int *nRows, int *nCols)
{
    double mnx = xs[0], mxx = xs[0], mny = ys[0], mxy = ys[0];
    for (int i = 1; i < n; i++) {
        double x = xs[i], y = ys[i];
        if (x > mxx) mxx = x;
        if (x < mnx) mnx = x;
        if (y > mxy) mxy = y;
        if (y < mny) mny = y;
    }
    mnx = mnx - cellSize / 2;
    int totalcol = (int)((mxx - mnx) / cellSize);
    if ((mxx - mnx) != totalcol * cellSize) totalcol++;
    mxx = mnx + cellSize * totalcol;

    mny = mny - cellSize / 2;
    int totalrow = (int)((mxy - mny) / cellSize);
    if ((mxy - mny) != totalrow * cellSize) totalrow++;
    mxy = mny + cellSize * totalrow;

    *minX = mnx; *maxX = mxx; *minY = mny; *maxY = mxy;
    *nRows = totalrow; *nCols = totalcol;
}

static int idw_blk_row(const idw_op *o, double y) { return (int)((o->maxY - y) / o->blockSize); }  /* idwOperator.h:68 */
static int idw_blk_col(const idw_op *o, double x) { return (int)((x - o->minX) / o->blockSize); }  /* idwOperator.h:69 */

/* idwOperator.cpp:5-20 with _sub_extent == _glb_extent (1 rank). */
static double idw_min_dist_to_block_bound(const idw_op *o, double x, double y)
{
    double boundXY[4] = {
        idw_blk_col(o, x) * o->blockSize + o->minX,
        (idw_blk_col(o, x) + 1) * o->blockSize + o->minX,
        o->maxY - idw_blk_row(o, y) * o->blockSize,
        o->maxY - (idw_blk_row(o, y) + 1) * o->blockSize,
    };
    double distances[4] = {
        fabs(x - boundXY[0]), fabs(x - boundXY[1]), fabs(y - boundXY[2]), fabs(y - boundXY[3]),
    };
    /* std::min(a, b) is (b < a) ? b : a */
    double a = (distances[1] < distances[0]) ? distances[1] : distances[0];
    double b = (distances[3] < distances[2]) ? distances[3] : distances[2];
    return (b < a) ? b : a;
}

/* searchNbrSamples (idwOperator.cpp:332-437).  nbrSamples holds
 * [dist, value] pairs.  Returns min(found, K). */
static int idw_search(idw_op *o, int cellRow, int cellCol, double *nbrSamples,
                      int64_t *candidates, int64_t *blocksSearched)
{
    const int subMinRow = 0;
    double cellX = (double)((cellCol + 0.5) * o->cellSize + o->minX);              /* idwOperator.h:92 */
    double cellY = (double)(o->maxY - (cellRow + subMinRow + 0.5) * o->cellSize);  /* idwOperator.h:91 */
    int blockRow = idw_blk_row(o, cellY);
    int blockCol = idw_blk_col(o, cellX);
    int blockRows = o->blockRows, blockCols = o->blockCols;

    double maxDist = 0.0;
    int maxDistIdx = -1;
    int tailIdx = -1;
    int searchRad = 0;
    int isSearch = 1;
    double minDistToBound = idw_min_dist_to_block_bound(o, cellX, cellY);
    int *block2search = (int *)malloc(sizeof(int) * (size_t)(8 * (o->maxRingGuard + 2) + 1));
    while (isSearch) {
        double searchRange = (double)(searchRad - 1) * o->blockSize + minDistToBound;
        if (o->buffer > 0 && searchRange >= o->buffer) break;
        /* Past maxRingGuard every ring lies outside the block grid: the
         * reference keeps looping over empty rings until one of its two
         * stop rules fires, which changes nothing, so stop here. */
        if (searchRad > o->maxRingGuard) break;
        int blockCount = 0;
        for (int tRow = blockRow - searchRad; tRow <= blockRow + searchRad; ++tRow) {
            if (tRow < 0 || tRow >= blockRows) continue;
            if (tRow == blockRow - searchRad || tRow == blockRow + searchRad) {
                for (int tCol = blockCol - searchRad; tCol <= blockCol + searchRad; ++tCol) {
                    if (tCol < 0 || tCol >= blockCols) continue;
                    block2search[blockCount++] = tRow * blockCols + tCol;
                }
            } else {
                if (blockCol - searchRad >= 0 && blockCol - searchRad < blockCols)
                    block2search[blockCount++] = tRow * blockCols + blockCol - searchRad;
                if (blockCol + searchRad >= 0 && blockCol + searchRad < blockCols)
                    block2search[blockCount++] = tRow * blockCols + blockCol + searchRad;
            }
        }
        if (blocksSearched) *blocksSearched += blockCount;
        for (int i = 0; i < blockCount; ++i) {
            int blockIdx = block2search[i];
            for (int p = o->blockStart[blockIdx]; p < o->blockStart[blockIdx + 1]; ++p) {
                const sample_point *it = &o->blockPts[p];
                double tmpDist = sqrt(pow(it->x - cellX, 2) + pow(it->y - cellY, 2));
                if (candidates) (*candidates)++;
                if (tmpDist == 0) tmpDist = IDW_EPS;
                if (tailIdx < o->nbrPoints - 1) {
                    tailIdx++;
                    nbrSamples[tailIdx * 2] = tmpDist;
                    nbrSamples[tailIdx * 2 + 1] = it->value;
                    if (tmpDist > maxDist) {
                        maxDist = tmpDist;
                        maxDistIdx = tailIdx;
                    }
                } else {
                    tailIdx++;
                    if (tmpDist < maxDist) {
                        nbrSamples[maxDistIdx * 2] = tmpDist;
                        nbrSamples[maxDistIdx * 2 + 1] = it->value;
                        maxDist = nbrSamples[0];
                        maxDistIdx = 0;
                        for (int j = 1; j < o->nbrPoints; ++j) {
                            if (nbrSamples[j * 2] > maxDist) {
                                maxDist = nbrSamples[j * 2];
                                maxDistIdx = j;
                            }
                        }
                    }
                }
            }
        }
        if (tailIdx >= o->nbrPoints - 1) {
            if (searchRange >= maxDist) isSearch = 0;
            else ++searchRad;
        } else {
            ++searchRad;
        }
    }
    free(block2search);
    return (tailIdx + 1 < o->nbrPoints) ? tailIdx + 1 : o->nbrPoints;
}

/* rowAtOtherLayer / colAtOtherLayer (rasterLayer.h:1296-1316): the double
 * min(scale*row, maxIRow) is truncated to int on return. */
static int idw_mask_index(int otherDim, int thisDim, int idx)
{
    double scale = (double)otherDim / thisDim;
    double v = scale * idx;
    double lim = (double)(otherDim - 1);
    double m = (lim < v) ? lim : v;
    return (int)m;
}

/* IDWOperator::Operator (idwOperator.cpp:439-491). */
static int idw_cell(void *vself, int iRow, int iCol)
{
    idw_op *o = (idw_op *)vself;
    double *outRow = cs_row(&o->out, iRow);
    if (!outRow) return 0;
    if (o->mask) {
        int mr = idw_mask_index(o->maskRows, o->nRows, iRow);
        int mc = idw_mask_index(o->maskCols, o->nCols, iCol);
        int mask = o->mask[(size_t)mr * o->maskCols + mc];
        if ((mask == 0) | (mask == o->maskNoData)) {
            outRow[iCol] = IDW_NODATA;
            return 1;
        }
    }
    double *pNbrSamples = (double *)malloc(sizeof(double) * 2 * (size_t)o->nbrPoints);
    int64_t cand = 0, blk = 0;
    int sampleNum = idw_search(o, iRow, iCol, pNbrSamples, &cand, &blk);
    if (o->rowCandidates) o->rowCandidates[iRow] += cand;
    if (o->rowBlocks) o->rowBlocks[iRow] += blk;
    if (sampleNum < 0) {
        free(pNbrSamples);
        o->failed = 1;
        return 0;
    }
    double weightSum = 0.0;
    double *pWeight = (double *)malloc(sizeof(double) * (size_t)o->nbrPoints);
    outRow[iCol] = 0;
    for (int i = 0; i < sampleNum; ++i) {
        pWeight[i] = 1 / pow(pNbrSamples[i * 2], (double)o->idwPower);
        weightSum += pWeight[i];
    }
    for (int i = 0; i < sampleNum; ++i) {
        double value = pNbrSamples[i * 2 + 1] * pWeight[i] / weightSum;
        outRow[iCol] += value;
    }
    free(pNbrSamples);
    free(pWeight);
    return 1;
}

/* Grid dimensions that oracle_idw will produce for these samples. */
int oracle_idw_dims(const double *xs, const double *ys, int n, float cellSize,
                    int *nRows, int *nCols, double *minX, double *maxY)
{
    double a, b, c, d;
    if (n < 1) return -1;
    oracle_idw_extent(xs, ys, n, cellSize, &a, &b, &c, &d, nRows, nCols);
    *minX = a;
    *maxY = d;
    return 0;
}

/* Full IDW.  xs/ys/vals in file order.  out has nRows*nCols doubles (use
 * oracle_idw_dims).  row_candidates (nullable, nRows int64) receives the
 * number of (cell, sample) distance evaluations per output row. */
int oracle_idw(const double *xs, const double *ys, const double *vals, int n,
               float cellSize, int K, int power, double buffer, double blockSize,
               const int32_t *mask, int maskRows, int maskCols, int maskNoData,
               double *out, int64_t *row_candidates, int64_t *row_blocks, int nthreads)
{
    if (n < 1 || K < 1 || blockSize <= 0) return -1;
    if (buffer <= 0 && n < K) return -3; /* the reference loops forever here */
    idw_op o;
    memset(&o, 0, sizeof(o));
    o.cellSize = cellSize;
    o.nbrPoints = K;
    o.idwPower = power;
    o.buffer = buffer;
    o.blockSize = blockSize;
    oracle_idw_extent(xs, ys, n, cellSize, &o.minX, &o.maxX, &o.minY, &o.maxY, &o.nRows, &o.nCols);
    /* creatSampleBlocks (idwOperator.cpp:152-169) */
    o.blockRows = (int)ceil((o.maxY - o.minY) / o.blockSize);
    o.blockCols = (int)ceil((o.maxX - o.minX) / o.blockSize);
    int nb = o.blockRows * o.blockCols;
    int *count = (int *)calloc((size_t)nb + 1, sizeof(int));
    int *bidx = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) {
        int iCol = idw_blk_col(&o, xs[i]);
        int iRow = idw_blk_row(&o, ys[i]);
        int index = iRow * o.blockCols + iCol;
        if (index < 0 || index >= nb) {
            free(count); free(bidx);
            return -4;
        }
        bidx[i] = index;
        count[index + 1]++;
    }
    for (int b = 0; b < nb; b++) count[b + 1] += count[b];
    o.blockStart = count;
    o.blockPts = (sample_point *)malloc(sizeof(sample_point) * (size_t)n);
    int *fill = (int *)malloc(sizeof(int) * (size_t)nb);
    for (int b = 0; b < nb; b++) fill[b] = count[b];
    for (int i = 0; i < n; i++) {
        sample_point p = {xs[i], ys[i], vals[i]};
        o.blockPts[fill[bidx[i]]++] = p;
    }
    free(fill);
    free(bidx);
    o.mask = mask;
    o.maskRows = maskRows;
    o.maskCols = maskCols;
    o.maskNoData = maskNoData;
    o.out.m = out; o.out.rows = o.nRows; o.out.cols = o.nCols;
    o.rowCandidates = row_candidates;
    o.rowBlocks = row_blocks;
    if (row_candidates) memset(row_candidates, 0, sizeof(int64_t) * (size_t)o.nRows);
    if (row_blocks) memset(row_blocks, 0, sizeof(int64_t) * (size_t)o.nRows);
    o.maxRingGuard = (o.blockRows > o.blockCols ? o.blockRows : o.blockCols);
    /* idwLayer with a local neighbourhood: workBR is the whole grid. */
    int rc = work_loop(idw_cell, &o, 0, o.nRows - 1, 0, o.nCols - 1, nthreads);
    free(o.blockPts);
    free(count);
    if (o.failed) return -5;
    return rc;
}

/* ------------------------------------------------------------------ */
/* IDW compute-load estimate and load-balanced row bands               */
/* ------------------------------------------------------------------ */

/* IdwTransformation::Operator (apps/spatial/idw/idwTransformation.cpp:3-62)
 * over every cell of the ComputeLayer built by ComputeLayer::_init
 * (gpro/computeLayer.h:208-260) from the IDW layer, driven like
 * Transformation::run (gpro/transformation.h:204-228): row-major, serial.
 * Returns -3 when a cell's walk can never reach K samples (the reference's
 * while loop never ends there), -4 when a sample falls outside the block
 * grid.  With out == NULL only *loadRows / *loadCols are set. */
int oracle_idw_load(const double *xs, const double *ys, int n, float cellSize, int K, double blockSize,
                    int granularity, const int32_t *mask, int maskRows, int maskCols, double maskNoData,
                    double noData, double *out, int *loadRows, int *loadCols)
{
    if (n < 1 || K < 1 || blockSize <= 0 || granularity < 1) return -1;
    double minX, maxX, minY, maxY;
    int nRows, nCols;
    oracle_idw_extent(xs, ys, n, cellSize, &minX, &maxX, &minY, &maxY, &nRows, &nCols);
    /* compute layer: ceil of the IDW work rectangle over the granularity */
    int lr = nRows / granularity + ((nRows % granularity) ? 1 : 0);
    int lc = nCols / granularity + ((nCols % granularity) ? 1 : 0);
    *loadRows = lr;
    *loadCols = lc;
    if (!out) return 0;
    float comptCellSize = cellSize * granularity; /* float * int (computeLayer.h:221) */
    int blockRows = (int)ceil((maxY - minY) / blockSize);
    int blockCols = (int)ceil((maxX - minX) / blockSize);
    int nb = blockRows * blockCols;
    int *cnt = (int *)calloc((size_t)nb, sizeof(int));
    for (int i = 0; i < n; i++) {
        int iCol = (int)((xs[i] - minX) / blockSize), iRow = (int)((maxY - ys[i]) / blockSize);
        int index = iRow * blockCols + iCol;
        if (index < 0 || index >= nb) { free(cnt); return -4; }
        cnt[index]++;
    }
    int guard = blockRows > blockCols ? blockRows : blockCols;
    int rc = 0;
    for (int cRow = 0; cRow < lr && rc == 0; cRow++) {
        for (int cCol = 0; cCol < lc; cCol++) {
            double *o = &out[(size_t)cRow * lc + cCol];
            if (mask) {
                /* _pComptLayer->rowAtOtherLayer(maskLayer, cRow) */
                int maskRow = idw_mask_index(maskRows, lr, cRow);
                int maskCol = idw_mask_index(maskCols, lc, cCol);
                int m = mask[(size_t)maskRow * maskCols + maskCol];
                if ((m == 0) | (m == maskNoData)) { *o = noData; continue; }
            }
            /* getBlockRow/ColIndexByCellIndex(i, cellSize) (idwOperator.h:70-71) */
            int a = (int)((cRow + 0.5) * (double)comptCellSize / blockSize);
            int b = (int)((cCol + 0.5) * (double)comptCellSize / blockSize);
            int blockRow = a < blockRows - 1 ? a : blockRows - 1;
            int blockCol = b < blockCols - 1 ? b : blockCols - 1;
            int isPointEnough = 0, searchRadius = 0, searchBlockNums = 0;
            int nbrNums = cnt[blockRow * blockCols + blockCol];
            while (nbrNums < K) {
                if (searchRadius > guard) { rc = -3; break; }
                for (int tRow = blockRow - searchRadius; tRow <= blockRow + searchRadius; ++tRow) {
                    if (isPointEnough) break;
                    if (tRow < 0 || tRow >= blockRows) continue;
                    if (tRow == blockRow - searchRadius || tRow == blockRow + searchRadius) {
                        for (int tCol = blockCol - searchRadius; tCol <= blockCol + searchRadius; ++tCol) {
                            if (isPointEnough) break;
                            if (tCol < 0 || tCol >= blockCols) continue;
                            nbrNums += cnt[tRow * blockCols + tCol];
                            searchBlockNums++;
                            if (nbrNums >= K) isPointEnough = 1;
                        }
                    }
                }
                if (isPointEnough) break;
                ++searchRadius;
            }
            if (rc) break;
            double load_1 = searchBlockNums;
            double load_2 = nbrNums * 2;
            *o = load_1 + load_2;
        }
    }
    free(cnt);
    return rc;
}

/* DeComposition::valRowDcmp (gpro/deComposition.h:305-393).  first[i] is
 * the first row of band i (nSub + 1 entries, first[nSub] = rows).  The
 * reference only ever uses the bands' first rows (computeLayer.h:396-428). */
int oracle_val_row_dcmp(const double *load, int rows, int cols, double noData, int nSubSpcs, int *first)
{
    const double Eps = 0.0000001; /* deComposition.h:24 */
    if (nSubSpcs < 1 || nSubSpcs > rows) return -1;
    double *rowComptLoad = (double *)malloc(sizeof(double) * (size_t)rows);
    double totalComptLoad = 0.0;
    for (int cRow = 0; cRow < rows; cRow++) {
        rowComptLoad[cRow] = 0.0;
        for (int cCol = 0; cCol < cols; cCol++) {
            double v = load[(size_t)cRow * cols + cCol];
            if ((v + 9999) > Eps && (v - noData) > Eps) rowComptLoad[cRow] += v;
        }
        totalComptLoad += rowComptLoad[cRow];
    }
    double averComptLoad = totalComptLoad / (nSubSpcs);
    double accuComptLoad = 0;
    int subBegin = 0, nb = 0;
    first[nb++] = 0;
    for (int i = 0; i < nSubSpcs - 1; ++i) {
        double subComptLoad = 0.0;
        double minComptDiff = 0.0;
        for (int cRow = subBegin; cRow < rows; cRow++) {
            int rowGap = cRow - 0;
            subComptLoad += rowComptLoad[rowGap];
            accuComptLoad += rowComptLoad[rowGap];
            if (subComptLoad <= averComptLoad && rows - 1 - cRow > nSubSpcs - 1 - i) {
                if (fabs(minComptDiff) < Eps || (averComptLoad - subComptLoad) < minComptDiff) {
                    minComptDiff = averComptLoad - subComptLoad;
                }
            } else {
                if (fabs(minComptDiff) < Eps || (subComptLoad - averComptLoad) <= minComptDiff ||
                    rows - 1 - cRow == nSubSpcs - 1 - i) {
                    averComptLoad = (totalComptLoad - accuComptLoad) / (nSubSpcs - i - 1);
                    subBegin = cRow + 1;
                } else {
                    accuComptLoad -= rowComptLoad[rowGap];
                    subComptLoad -= rowComptLoad[rowGap];
                    averComptLoad = (totalComptLoad - accuComptLoad) / (nSubSpcs - i - 1);
                    subBegin = cRow;
                }
                first[nb++] = subBegin;
                break;
            }
        }
    }
    free(rowComptLoad);
    first[nSubSpcs] = rows;
    return nb == nSubSpcs ? 0 : -2;
}

/* ------------------------------------------------------------------ */
/* Intensity-ratio compute load: Transformation::Operator and run()     */
/* (gpro/transformation.h:155-228) over a ComputeLayer of granularity g */
/* (gpro/computeLayer.h:208-260).                                       */
/* ------------------------------------------------------------------ */
#define CL_EPS 0.0000001 /* computeLayer.h:24 */

/* data: nlayers raster layers of rows x cols doubles, back to back, each
 * with its own NoData; the work rectangle starts at (row0, col0) and spans
 * work_rows x work_cols; load: ceil(work_rows/g) x ceil(work_cols/g). */
int oracle_intensity_load(const double *data, int nlayers, const double *nodata, int rows, int cols, int row0,
                          int col0, int work_rows, int work_cols, int g, double nodataLoad, double validLoad,
                          double *load)
{
    if (nodataLoad < 0 || validLoad < 0) return -1; /* transformation.h:168-171 */
    const int lr = work_rows / g + (work_rows % g ? 1 : 0);
    const int lc = work_cols / g + (work_cols % g ? 1 : 0);
    const double grain = (double)g; /* _computGrain is a double */
    for (int cRow = 0; cRow < lr; cRow++) {
        for (int cCol = 0; cCol < lc; cCol++) {
            double *out = load + (size_t)cRow * lc + cCol;
            *out = 0.0;
            for (int L = 0; L < nlayers; L++) {
                const double *d = data + (size_t)L * rows * cols;
                for (int dRow = (int)(cRow * grain + row0); dRow < (cRow + 1) * grain + row0; ++dRow) {
                    for (int dCol = (int)(cCol * grain + col0); dCol < (cCol + 1) * grain + col0; ++dCol) {
                        if (dRow > rows - 1 || dRow >= rows || dCol > cols - 1 || dCol >= cols) continue;
                        if (fabs(d[(size_t)dRow * cols + dCol] - nodata[L]) > CL_EPS) *out += validLoad;
                        else *out += nodataLoad;
                    }
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* D8 contributing area (apps/hydrology/scad8/scaOperator.cpp:5-131)    */
/* run by RasterOperator::Work's MPI_Type loop at one rank              */
/* (gpro/rasterOperator.h:147-216).                                     */
/* ------------------------------------------------------------------ */
#define SCA_EPS 0.000001 /* scaOperator.cpp:3 */

typedef struct {
    cellspace d8, sca, degree; /* degree is a RasterLayer<double> (scaOperator.h:37) */
    double noData;
    int num, maxRow, maxCol;
    int termination;
    int converge; /* 0: the reference; 1: phase-2 sweeps until nothing changes */
    int64_t changes;
} sca_op;

static int sca_cell(void *vself, int iRow, int iCol)
{
    sca_op *o = (sca_op *)vself;
    double *d8c = cs_row(&o->d8, iRow), *scac = cs_row(&o->sca, iRow), *degc = cs_row(&o->degree, iRow);
    if (!d8c || !scac || !degc) return 0;
    const int last = (iRow == o->maxRow && iCol == o->maxCol);

    if (o->num == 0) { /* scaOperator.cpp:46-63 */
        if (fabs(d8c[iCol] - o->noData) < SCA_EPS) {
            degc[iCol] = -2;
            scac[iCol] = o->noData;
        } else {
            degc[iCol] = 0;
            scac[iCol] = 1;
        }
        if (last) {
            o->num = 1;
            o->termination = 0;
        }
        return 1;
    }
    if (o->num == 1) { /* scaOperator.cpp:65-84: count the neighbours that drain here */
        int dir = 8;
        for (int tRow = iRow - 1; tRow <= iRow + 1; tRow++) {
            double *d8t = cs_row(&o->d8, tRow);
            if (!d8t) return 0;
            for (int tCol = iCol - 1; tCol <= iCol + 1; tCol++) {
                if (d8t[tCol] == dir && (dir + d8c[iCol] != 8) && dir != 4 &&
                    fabs(d8c[iCol] - o->noData) > SCA_EPS)
                    degc[iCol]++;
                dir--;
            }
        }
        if (last) {
            o->num = 2;
            o->termination = 0;
        }
        return 1;
    }
    /* scaOperator.cpp:86-128: pull finished upstream neighbours.  The last
     * work cell always runs the loop in the reference; the converging
     * variant applies the degree test to it like to every other cell. */
    if (degc[iCol] <= 0 && !(last && !o->converge)) return 1;
    int dir = 8;
    for (int tRow = iRow - 1; tRow <= iRow + 1; tRow++) {
        double *d8t = cs_row(&o->d8, tRow), *scat = cs_row(&o->sca, tRow), *degt = cs_row(&o->degree, tRow);
        if (!d8t || !scat || !degt) return 0;
        for (int tCol = iCol - 1; tCol <= iCol + 1; tCol++) {
            if ((degt[tCol] == 0 || degt[tCol] == -1) && d8t[tCol] == dir) {
                if ((dir + d8c[iCol] != 8) && dir != 4) {
                    scac[iCol] += scat[tCol];
                    degc[iCol]--;
                    if (degc[iCol] == 0) degc[iCol] = -1;
                    o->changes++;
                }
                if (degt[tCol] == -1) {
                    degt[tCol] = -2;
                    o->changes++;
                }
            }
            dir--;
        }
    }
    if (last && !o->converge) {
        /* end of the sweep (scaOperator.cpp:110-127); Termination stays 1,
         * so Work leaves its loop after this single accumulation sweep */
        for (int i = 1; i <= o->maxRow; ++i) {
            double *degi = cs_row(&o->degree, i);
            if (!degi) return 0;
            for (int j = 1; j <= o->maxCol; ++j) {
                if (degi[j] == 0) degi[j] = -2;
                else if (degi[j] == -1) degi[j] = 0;
            }
        }
        o->termination = 1;
    }
    return 1;
}

/* Contributing area from a D8 layer of window-index codes (dirType 8,
 * flowdird8).  The SCA layer and the degree layer start as copies of the D8
 * layer's metadata filled with its NoData value (rasterLayer.h:739-745), so
 * border cells come out as NoData.  converge = 0 reproduces the reference:
 * one initial sweep, one degree sweep and a single accumulation sweep
 * (every phase-2 sweep ends with Termination = 1, scaOperator.cpp:126).
 * converge = 1 repeats the accumulation sweep, with the reference's
 * end-of-sweep degree update, until a sweep changes nothing, and treats
 * the last work cell like the others.  *sweeps (may be NULL) receives the
 * number of Work sweeps.  Returns 0. */
int oracle_sca_d8(const double *d8, int rows, int cols, double noData, int converge,
                  double *sca_out, int *sweeps)
{
    size_t n = (size_t)rows * (size_t)cols;
    double *deg = (double *)malloc(n * sizeof(double) + 1);
    if (!deg) return -1;
    for (size_t i = 0; i < n; i++) {
        sca_out[i] = noData;
        deg[i] = noData;
    }
    sca_op o;
    memset(&o, 0, sizeof(o));
    o.d8.m = (double *)d8; o.d8.rows = rows; o.d8.cols = cols;
    o.sca.m = sca_out; o.sca.rows = rows; o.sca.cols = cols;
    o.degree.m = deg; o.degree.rows = rows; o.degree.cols = cols;
    o.noData = noData;
    o.maxRow = rows - 2;
    o.maxCol = cols - 2;
    o.converge = converge;
    int nsweep = 0, rc = 0;
    if (!converge) {
        do { /* rasterOperator.h:170-216 */
            o.termination = 1;
            if (work_loop(sca_cell, &o, 1, rows - 2, 1, cols - 2, 1) != 0) { rc = -1; break; }
            nsweep++;
        } while (!o.termination);
    } else {
        for (int ph = 0; ph < 2 && rc == 0; ph++) {
            if (work_loop(sca_cell, &o, 1, rows - 2, 1, cols - 2, 1) != 0) rc = -1;
            nsweep++;
        }
        if (rows < 3 || cols < 3) o.num = 2;
        while (rc == 0) {
            o.changes = 0;
            if (work_loop(sca_cell, &o, 1, rows - 2, 1, cols - 2, 1) != 0) { rc = -1; break; }
            nsweep++;
            for (int i = 1; i <= rows - 2; ++i)
                for (int j = 1; j <= cols - 2; ++j) {
                    double *dg = deg + (size_t)i * cols + j;
                    if (*dg == 0) { *dg = -2; o.changes++; }
                    else if (*dg == -1) { *dg = 0; o.changes++; }
                }
            if (o.changes == 0) break;
        }
    }
    free(deg);
    if (sweeps) *sweeps = nsweep;
    return rc;
}

/* ------------------------------------------------------------------ */
/* Depression filling (apps/hydrology/pitremove/pitRemoveOperator.cpp:  */
/* 3-72), Work's MPI_Type loop at one rank.                             */
/* ------------------------------------------------------------------ */
#define PIT_EPS 0.000001 /* pitRemoveOperator.h:85 */

typedef struct {
    cellspace dem, wdem;
    double noData;
    int xSize, ySize, nbrRadius, num;
    int termination;
} pit_op;

static int pit_cell(void *vself, int iRow, int iCol)
{
    pit_op *o = (pit_op *)vself;
    double *demc = cs_row(&o->dem, iRow), *wc = cs_row(&o->wdem, iRow);
    if (!demc || !wc) return 0;
    const double gap = 0.0005; /* pitRemoveOperator.cpp:36 */
    if (o->num == 0) {
        if (fabs(demc[iCol] - o->noData) > PIT_EPS)
            wc[iCol] = 10000.0;
        else
            wc[iCol] = o->noData;
        if (iRow == o->xSize - 2 && iCol == o->ySize - 2) {
            o->num = 1;
            o->termination = 0;
        }
        return 1;
    }
    if (fabs(demc[iCol] - o->noData) <= PIT_EPS || wc[iCol] <= demc[iCol]) return 1;
    for (int i = iRow - o->nbrRadius; i <= iRow + o->nbrRadius; i++) {
        double *wi = cs_row(&o->wdem, i);
        if (!wi) return 0;
        for (int j = iCol - o->nbrRadius; j <= iCol + o->nbrRadius; j++) {
            if ((demc[iCol] >= (wi[j] + gap)) || fabs(wi[j] - o->noData) < PIT_EPS) {
                wc[iCol] = demc[iCol];
                o->termination = 0;
            } else if (wc[iCol] > (wi[j] + gap)) {
                wc[iCol] = wi[j] + gap;
                o->termination = 0;
            }
        }
    }
    return 1;
}

/* Planchon-Darboux style filling with a 0.0005 gap, swept in row-major
 * order until a sweep assigns nothing.  wdem starts as a copy of the DEM
 * layer's metadata filled with NoData (rasterLayer.h:739-745), so border
 * cells stay NoData and act as outlets.  max_sweeps <= 0 means no limit;
 * returns 1 if the limit stopped the loop, 0 when it converged. */
int oracle_pit_remove(const double *dem, int rows, int cols, double noData, double *wdem_out,
                      int max_sweeps, int *sweeps)
{
    size_t n = (size_t)rows * (size_t)cols;
    for (size_t i = 0; i < n; i++) wdem_out[i] = noData;
    pit_op o;
    memset(&o, 0, sizeof(o));
    o.dem.m = (double *)dem; o.dem.rows = rows; o.dem.cols = cols;
    o.wdem.m = wdem_out; o.wdem.rows = rows; o.wdem.cols = cols;
    o.noData = noData;
    o.xSize = rows;
    o.ySize = cols;
    o.nbrRadius = 1;
    int nsweep = 0, rc = 0;
    do {
        o.termination = 1;
        if (work_loop(pit_cell, &o, 1, rows - 2, 1, cols - 2, 1) != 0) { rc = -1; break; }
        nsweep++;
        if (max_sweeps > 0 && nsweep >= max_sweeps && !o.termination) { rc = 1; break; }
    } while (!o.termination);
    if (sweeps) *sweeps = nsweep;
    return rc;
}

/* ------------------------------------------------------------------ */
/* Helpers used by the tests.                                          */
/* ------------------------------------------------------------------ */

/* x / sqrt(2.0) with IEEE division, as the reference evaluates it. */
void oracle_div_sqrt2(const double *x, double *q, int64_t n)
{
    for (int64_t i = 0; i < n; i++) q[i] = x[i] / sqrt(2.0);
}

/* Widen float32 to double exactly (what GDAL RasterIO into a double buffer
 * does for a Float32 band, rasterLayer.h:892-894). */
void oracle_widen(const float *src, double *dst, int64_t n)
{
    for (int64_t i = 0; i < n; i++) dst[i] = (double)src[i];
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
