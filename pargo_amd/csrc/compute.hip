// compute.hip -- intensity-ratio compute load (gpro's Transformation) on
// the device.
//
// Transformation::Operator (gpro/transformation.h:155-197) gives each cell
// of the coarse compute grid the sum, over the data cells of its granule,
// of validLoad (fabs(v - noData) > 1e-7, computeLayer.h:24) or nodataLoad.
// Granules start at the data work rectangle's first row and column and are
// clipped only by the raster's extent, so the last granule of a row or
// column also covers the border beyond the work rectangle.  One thread per
// coarse cell adds in the reference's row-major order, so sums are the
// reference's doubles.
#include <cmath>

#include "internal.h"

namespace pargo {

namespace {

struct IntensityParams {
    const double *__restrict__ data;
    double *__restrict__ out;
    long long pitch;
    int rows, cols, row0, col0, g, load_rows, load_cols, accumulate;
    double nodata, nodata_load, valid_load;
};

__global__ __launch_bounds__(256) void intensity_load_kernel(IntensityParams p)
{
    const int j = blockIdx.x * 256 + threadIdx.x, i = blockIdx.y;
    if (j >= p.load_cols) return;
    const long long o = (long long)i * p.load_cols + j;
    double acc = p.accumulate ? p.out[o] : 0.0;
    const int r0 = i * p.g + p.row0, c0 = j * p.g + p.col0;
    for (int r = r0; r < r0 + p.g && r < p.rows; r++) {
        const double *row = p.data + (long long)r * p.pitch;
        for (int c = c0; c < c0 + p.g && c < p.cols; c++)
            acc += fabs(row[c] - p.nodata) > 0.0000001 ? p.valid_load : p.nodata_load;
    }
    p.out[o] = acc;
}

}  // namespace

}  // namespace pargo

using pargo::fail;

extern "C" int pargo_intensity_load(pargo_ctx *ctx, const pargo_grid *g, const double *data, int work_row0,
                                    int work_col0, int work_rows, int work_cols, int granularity, double nodata_load,
                                    double valid_load, int accumulate, double *load_out, int *load_rows,
                                    int *load_cols)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!g) return fail(ctx, PARGO_E_INVALID, "pargo_intensity_load: null grid");
    if (granularity < 1) return fail(ctx, PARGO_E_INVALID, "granularity must be >= 1, got %d", granularity);
    if (work_rows < 1 || work_cols < 1 || work_row0 < 0 || work_col0 < 0 || work_row0 + work_rows > g->rows ||
        work_col0 + work_cols > g->cols)
        return fail(ctx, PARGO_E_INVALID, "work rectangle (%d, %d) + (%d x %d) outside the %d x %d raster", work_row0,
                    work_col0, work_rows, work_cols, g->rows, g->cols);
    // "The load specified cannot be negative." (transformation.h:168-171)
    if (nodata_load < 0 || valid_load < 0 || std::isnan(nodata_load) || std::isnan(valid_load))
        return fail(ctx, PARGO_E_INVALID, "The load specified cannot be negative.");
    const int lr = work_rows / granularity + (work_rows % granularity ? 1 : 0);
    const int lc = work_cols / granularity + (work_cols % granularity ? 1 : 0);
    if (load_rows) *load_rows = lr;
    if (load_cols) *load_cols = lc;
    if (!load_out) return PARGO_OK;
    if (!data) return fail(ctx, PARGO_E_INVALID, "pargo_intensity_load: null data");
    const long long pitch = g->pitch ? g->pitch : g->cols;
    if (pitch < g->cols) return fail(ctx, PARGO_E_INVALID, "pitch %lld < cols %d", pitch, g->cols);
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    pargo::IntensityParams p{};
    p.data = data;
    p.out = load_out;
    p.pitch = pitch;
    p.rows = g->rows;
    p.cols = g->cols;
    p.row0 = work_row0;
    p.col0 = work_col0;
    p.g = granularity;
    p.load_rows = lr;
    p.load_cols = lc;
    p.accumulate = accumulate ? 1 : 0;
    p.nodata = g->nodata;
    p.nodata_load = nodata_load;
    p.valid_load = valid_load;
    hipLaunchKernelGGL(pargo::intensity_load_kernel, dim3((lc + 255) / 256, lr), dim3(256), 0, ctx->stream, p);
    return pargo::check_launch(ctx, "pargo_intensity_load");
}
