/*
 * pargo.h -- C ABI of the MI355X focal-operator library (libpargo.so).
 *
 * This is the device seam that PaRGO's RasterOperator<T>::Work leaves empty
 * for CUDA_Type (gpro/rasterOperator.h:347-348).  Instead of calling a
 * virtual Operator() per cell (rasterOperator.h:174-185), a host operator
 * hands its layer descriptors to one entry point below, which runs the whole
 * work rectangle on the GPU.
 *
 * Conventions
 *   - Every entry point returns an int status: PARGO_OK (0) or a negative
 *     PARGO_E_* code.  Nothing throws and nothing calls exit().  The text of
 *     the last failure on a context is returned by pargo_last_error(), the
 *     counterpart of the reference's cerr messages.
 *   - Raster and output pointers are DEVICE pointers unless a parameter says
 *     "host".  Device memory belongs to the caller; the library never frees
 *     caller memory.  pargo_alloc/pargo_free are conveniences.
 *   - Calls on a context are ordered on the context's stream and return
 *     before the GPU work finishes; pargo_sync() waits.  One context per host
 *     thread.  The library keeps no global mutable state apart from the
 *     per-device scratch of the no-allocation entry point, if present.
 *   - Rasters are row-major float32 with a pitch in elements (0 means
 *     pitch == cols).  Output rasters of a call use the same pitch as its
 *     input.  NoData bitmasks hold ceil(cols/32) uint32 words per row, bit
 *     (c % 32) of word (c / 32) set when cell c is NoData (including the
 *     border rows and columns, which the reference never computes).
 */
#ifndef PARGO_H
#define PARGO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PARGO_ABI_VERSION 1

/* status codes */
#define PARGO_OK 0
#define PARGO_E_INVALID -1     /* an argument is out of range or NULL          */
#define PARGO_E_DEVICE -2      /* a HIP runtime call failed                    */
#define PARGO_E_NOMEM -3       /* device or host allocation failed            */
#define PARGO_E_UNSUPPORTED -4 /* valid request this build does not implement */
#define PARGO_E_NONTERM -5     /* the reference's loop would never terminate  */

/* Slope methods, in the order of enum SlopeAlgor
 * (apps/morphology/slope/slopeOperator.h:20-28). */
#define PARGO_SLOPE_FD 0   /* Horn: 3rd-order FD weighted by 1/d^2 */
#define PARGO_SLOPE_FFD 1  /* frame finite difference */
#define PARGO_SLOPE_MD 2   /* maximum downslope */
#define PARGO_SLOPE_SD 3   /* simple difference */
#define PARGO_SLOPE_SFD 4  /* second-order finite difference */
#define PARGO_SLOPE_TFD 5  /* third-order finite difference */
#define PARGO_SLOPE_TFDW 6 /* third-order FD weighted by 1/d */

/* Slope output units.  The reference writes the gradient magnitude
 * (tan of the slope angle, slopeOperator.cpp:86). */
#define PARGO_UNIT_TAN 0
#define PARGO_UNIT_DEGREES 1

/* D8 device codes (uint8 per cell).
 *   dirtype 256: ESRI bit-sum of all tied steepest directions
 *                (E=1 SE=2 S=4 SW=8 W=16 NW=32 N=64 NE=128), 0 = flat/pit
 *                (the reference's -1), d8Operator.cpp:73-76,89-97.
 *   dirtype 8:   median tied window index 0..8 (0=NW 1=N 2=NE 3=W 5=E
 *                6=SW 7=S 8=SE), 255 = flat/pit (the reference's -1 read
 *                as int8), d8Operator.cpp:78-88.
 * NoData cells (any NoData value in the 3x3 window, and the border) are
 * reported through the NoData bitmask or the slope output, never through
 * the code byte, which carries no NoData value. */
#define PARGO_D8_FLAT_256 0
#define PARGO_D8_FLAT_8 255

typedef struct pargo_ctx pargo_ctx;

/* Raster descriptor: the fields of gpro's MetaData that the focal kernels
 * use (gpro/metaData.h:55-58).  cellsize is rounded to float inside the
 * library, because MetaData::cellSize is a float (metaData.h:56). */
typedef struct {
    int32_t rows;
    int32_t cols;
    int64_t pitch;   /* elements between row starts; 0 = cols */
    double cellsize;
    double nodata;
} pargo_grid;

/* ---- context ------------------------------------------------------- */

/* Opens a context on HIP device `device` with its own stream.  Replaces
 * Application::START's program-type switch (gpro/application.cpp:5-18). */
int pargo_create(int device, pargo_ctx **out);
int pargo_destroy(pargo_ctx *ctx);
/* Message of the last failure on ctx; with ctx == NULL, of the last failure
 * on this thread that had no context (e.g. pargo_create). */
const char *pargo_last_error(const pargo_ctx *ctx);
int pargo_abi_version(void);
/* Runs subsequent work on `hip_stream` (a hipStream_t; NULL is HIP's
 * default stream) instead of the context's own stream.
 * pargo_set_stream(ctx, pargo_own_stream(ctx)) switches back.
 * One stream at a time: a context's device state (fixup queue, scratch,
 * batch queue) is shared by everything queued through it, so a switch makes
 * the new stream wait (hipStreamWaitEvent) for the work already queued on
 * the old one.  Work for two streams that must overlap needs two contexts. */
int pargo_set_stream(pargo_ctx *ctx, void *hip_stream);
void *pargo_own_stream(const pargo_ctx *ctx);
int pargo_sync(pargo_ctx *ctx);

int pargo_alloc(pargo_ctx *ctx, size_t bytes, void **dptr);
int pargo_free(pargo_ctx *ctx, void *dptr);
/* Page-locked host memory (hipHostMalloc) for staging copies at full PCIe
 * rate; the gpro layer stages rasters through it.  The reference moves
 * rasters host-side only (GDAL RasterIO into CellSpace). */
int pargo_alloc_host(pargo_ctx *ctx, size_t bytes, void **hptr);
int pargo_free_host(pargo_ctx *ctx, void *hptr);
int pargo_h2d(pargo_ctx *ctx, void *dst, const void *host_src, size_t bytes);
int pargo_d2h(pargo_ctx *ctx, void *host_dst, const void *src, size_t bytes);

/* ---- focal operators (3x3 Moore neighbourhood) --------------------- */

/* Slope of every interior cell; border cells and cells with NoData in their
 * window get grid->nodata.  Replaces SlopeOperator::Operator
 * (apps/morphology/slope/slopeOperator.cpp:51-139) driven by
 * RasterOperator::Work (gpro/rasterOperator.h:171-185).
 * nodata_mask_out may be NULL. */
int pargo_slope(pargo_ctx *ctx, const pargo_grid *grid, const float *dem,
                int method, int unit, float *slope_out, uint32_t *nodata_mask_out);

/* D8 flow direction codes (see PARGO_D8_*).  Replaces D8Operator::Operator
 * (apps/hydrology/flowdird8/d8Operator.cpp:28-102); dirtype as in
 * D8Operator(int) (d8Operator.h:16-21), 8 or 256.  Decisions are made on
 * the same FP64 values as the reference, so codes are bit-exact. */
int pargo_d8(pargo_ctx *ctx, const pargo_grid *grid, const float *dem, int dirtype,
             uint8_t *code_out, uint32_t *nodata_mask_out);

/* Slope and D8 in one pass over the DEM (the reference runs them as two
 * programs, pSlope.cpp and pD8.cpp).  Either output may be NULL. */
int pargo_slope_d8(pargo_ctx *ctx, const pargo_grid *grid, const float *dem,
                   int method, int unit, int dirtype, float *slope_out,
                   uint8_t *code_out, uint32_t *nodata_mask_out);

/* pargo_slope_d8 without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, outputs and errors as pargo_slope_d8; the fixup queue it needs
 * (two counters and 16384 strip entries, 262,400 bytes) lives in a static
 * device array of the library, one per device, resolved by pargo_create.
 * That fixed size covers every raster: strips beyond the 16384 entries are
 * handled by the fixup pass's full scan, as in pargo_slope_d8.  Every call
 * on one device shares the array, so calls that overlap on a device (two
 * streams, two contexts, two threads) race; it is cleared by a kernel at the
 * start of each call, never by a memset, so the call can be captured. */
int pargo_slope_d8_noalloc(pargo_ctx *ctx, const pargo_grid *grid, const float *dem,
                           int method, int unit, int dirtype, float *slope_out,
                           uint8_t *code_out, uint32_t *nodata_mask_out);

/* ---- many rasters ---------------------------------------------------- */

/* One raster of a batch: its descriptor and device buffers.  All tiles of a
 * batch must agree on which outputs are NULL. */
typedef struct {
    pargo_grid grid;
    const float *dem;
    float *slope_out;
    uint8_t *code_out;
} pargo_tile;

#define PARGO_BATCH_PER_TILE 0   /* one pargo_slope_d8 per tile, stream-ordered */
#define PARGO_BATCH_PERSISTENT 1 /* one persistent launch over a device work queue */
#define PARGO_BATCH_GRAPH 2      /* the persistent launch captured once in a HIP graph and replayed */

/* Slope and/or D8 of every tile (`tiles` is a host array).  The reference
 * spreads one raster over MPI ranks by row bands (DeComposition::rowDcmp,
 * gpro/deComposition.h:148-182) and balances bands by estimated load
 * (ComputeLayer, gpro/computeLayer.h:378-443); here (tile, column band, row
 * strip) items of many rasters are handed to waves from one queue, largest
 * tiles first.  The three modes give identical bits.  They equal per-tile
 * pargo_slope_d8 calls bit for bit, except that a batch holding a tile whose
 * cell size is outside [2^-20, 2^40] takes the generic FP64 Horn path for
 * every tile: then the other tiles' slopes are within 1e-5 of pargo_slope_d8
 * (which takes the FP32-root path for them) and their D8 codes identical. */
int pargo_batch_slope_d8(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                         int mode);

/* Work orders from the reference's load models (batch queue, IDW tiles). */
#define PARGO_ORDER_ROWS 0     /* as given / row-major */
#define PARGO_ORDER_ESTIMATE 1 /* heaviest first by the estimated load */
#define PARGO_ORDER_MEASURED 2 /* heaviest first by a measured pass */

/* Queue order of later batches on ctx, heaviest tile first by a load
 * model (the reference's ComputeLayer load modes, gpro/computeLayer.h:
 * 378-443, which valRowDcmp turns into bands, gpro/deComposition.h:
 * 305-393): PARGO_ORDER_ESTIMATE (the default) by cell count, the
 * intensity-ratio load with equal NoData and valid weights, which is the
 * fused kernel's per-cell cost; PARGO_ORDER_MEASURED by each tile's kernel
 * time in one per-tile pass over `tiles` (the preliminary experiment);
 * PARGO_ORDER_ROWS in the given order.  Outputs do not depend on the order.
 * Synchronous. */
int pargo_batch_order(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                      int source);

/* Times the persistent launch for each candidate shape (columns per lane 2
 * or 4, rows per strip 16, 32, 64 or 128) on this batch, keeps the fastest for
 * later batches on ctx and reports it (outputs may be NULL). */
int pargo_batch_autotune(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                         int *cols_per_lane, int *rows_per_strip);

/* ---- D8 contributing area (scad8) ------------------------------------ */

/* Flow classes, one uint8 per cell (what pargo_sca_d8 reads):
 *   0..8   a valid cell draining to window index k (0=NW 1=N 2=NE 3=W 5=E
 *          6=SW 7=S 8=SE), the dirType-8 codes of pargo_d8; 4 drains nowhere
 *   PARGO_FLOW_EXCL | k   a valid cell that drains nowhere but, like a cell
 *          draining to k, takes no inflow from window cell k (a D8 value v
 *          with dir + v == 8 in double, dir = 8 - k, that is not exactly k)
 *   PARGO_FLOW_NODATA     fabs(v - nodata) < 1e-6
 *   PARGO_FLOW_INERT      neither < nor > 1e-6 from nodata (NaN, or exactly
 *          1e-6 away): area 1,
 *          takes no inflow (scaOperator.cpp:48 and :71 disagree on it)
 *   any other value       a valid cell that drains nowhere (the reference's
 *          -1 for flats and pits; pargo_d8's 255). */
#define PARGO_FLOW_EXCL 0x10
#define PARGO_FLOW_INERT 253
#define PARGO_FLOW_NODATA 254
#define PARGO_FLOW_NONE 255

#define PARGO_SCA_REFERENCE 0 /* the reference: a single accumulation sweep */
#define PARGO_SCA_CONVERGE 1  /* accumulation sweeps repeated until nothing changes */

/* Flow classes of a reference D8 layer (device double, dirType-8 codes,
 * NoData grid->nodata) with the reference's own comparisons.  NaN NoData,
 * or NoData within 2e-6 of an integer in -1..8, is refused: the reference
 * would let such NoData cells drain and count border cells as finished. */
int pargo_d8_classify(pargo_ctx *ctx, const pargo_grid *grid, const double *d8, uint8_t *flow_out);

/* Specific catchment area in cells of every interior cell: SCAOperator
 * (apps/hydrology/scad8/scaOperator.cpp:37-131) driven by
 * RasterOperator::Work (gpro/rasterOperator.h:170-216).  flow holds flow
 * classes (pargo_d8 dirtype-8 output qualifies); nodata_mask (may be NULL)
 * marks further NoData cells, e.g. pargo_d8's mask.  sca_out (device double,
 * grid->pitch) gets NoData on the border and on NoData cells, as the
 * reference's SCA layer.  PARGO_SCA_REFERENCE reproduces the reference's
 * result bit for bit, including that it stops after one accumulation sweep
 * in row-major order and that its last work cell (rows-2, cols-2) adds
 * inflows to NoData; PARGO_SCA_CONVERGE gives the area that repeated sweeps
 * converge to (full D8 flow accumulation).  grid->cellsize is unused.
 * Working memory: the context's scratch grows to 2 bytes per cell (rows and
 * columns rounded up to 8) plus 16,640 bytes, kept until pargo_destroy. */
int pargo_sca_d8(pargo_ctx *ctx, const pargo_grid *grid, const uint8_t *flow, const uint32_t *nodata_mask,
                 int mode, double *sca_out);

/* ---- depression filling (pitremove) ------------------------------------ */

/* Filled DEM: PitRemoveOperator (apps/hydrology/pitremove/
 * pitRemoveOperator.cpp:26-72) driven by RasterOperator::Work
 * (gpro/rasterOperator.h:170-216) until a sweep changes nothing.  dem is a
 * device float32 raster (grid->pitch), filled_out a device double raster of
 * the same pitch: valid interior cells get the reference's value bit for
 * bit, the border and NoData cells (fabs(dem - nodata) <= 1e-6) get NoData.
 * The device applies the same updates in another order, which reaches the
 * same result when no filled value can come within 1e-6 of NoData; a DEM
 * whose valid elevations reach down to NoData (NoData inside
 * [min(dem, 10000), 10000]) is refused with PARGO_E_UNSUPPORTED, as is a
 * DEM holding NaN (the reference never settles next to NoData there).
 * Synchronous.  *passes (host, may be NULL) receives the number of device
 * passes.  Working memory: the context's scratch grows to 524 bytes per
 * 64 x 32 tile (worklists, stamps and a copy of each tile's edge columns),
 * kept until pargo_destroy. */
int pargo_pit_remove(pargo_ctx *ctx, const pargo_grid *grid, const float *dem, double *filled_out, int *passes);

/* ---- IDW ------------------------------------------------------------ */

/* Output grid of IDW for these samples (host arrays, file order), by the
 * extent rule of IDWOperator::readSamples (idwOperator.cpp:126-141). */
int pargo_idw_grid(const double *xs, const double *ys, int n, float cellsize,
                   int *rows, int *cols, double *min_x, double *max_y);
/* The same rule with the whole extent: extent = {minX, maxX, minY, maxY}
 * after the reference's half-cell shift and rounding up to whole cells. */
int pargo_idw_extent(const double *xs, const double *ys, int n, float cellsize,
                     double extent[4], int *rows, int *cols);

typedef struct pargo_idw_plan pargo_idw_plan;

/* Buckets the samples (host arrays, file order) into blockSize blocks like
 * IDWOperator::creatSampleBlocks (idwOperator.cpp:152-169) and uploads them.
 * The plan belongs to the context that made it. */
int pargo_idw_prepare(pargo_ctx *ctx, const double *xs, const double *ys, const double *vals,
                      int n, float cellsize, double block_size, pargo_idw_plan **out);
int pargo_idw_plan_free(pargo_ctx *ctx, pargo_idw_plan *plan);
/* Replaces new sample values (host array, file order) without re-bucketing. */
int pargo_idw_set_values(pargo_ctx *ctx, pargo_idw_plan *plan, const double *vals);
/* The same from a device array (file order, n doubles): a gather into the
 * plan's bucket order, queued on the context's stream with no host
 * synchronisation, so a loop that updates the values and reruns IDW never
 * waits for the host.  Replaces the host-side value reload between two
 * Run() calls of IDWOperator (idwOperator.cpp:63-149 reads values once). */
int pargo_idw_set_values_device(pargo_ctx *ctx, pargo_idw_plan *plan, const double *d_vals);

/* IDW of every output cell: IDWOperator::Operator + searchNbrSamples
 * (idwOperator.cpp:332-491) with 1-rank semantics.  out is a device float
 * grid of rows x cols from pargo_idw_grid.  mask (device int32,
 * mask_rows x mask_cols, may be NULL) is looked up like
 * rowAtOtherLayer/colAtOtherLayer (gpro/rasterLayer.h:1296-1316); masked
 * cells get -9999 (idwOperator.h:14).  Returns PARGO_E_NONTERM when
 * buffer <= 0 and n < K, where the reference loops forever.
 * candidates_out (device uint64, may be NULL) receives the number of
 * (cell, sample) distance evaluations. */
int pargo_idw_run(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                  const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata,
                  float *out, unsigned long long *candidates_out);

/* Launch order of IDW's 16 x 16 cell tiles for every later pargo_idw_run,
 * pargo_idw_work and pargo_idw_adjoint on this plan, from the reference's
 * load models (ComputeLayer::getCompuLoad, gpro/computeLayer.h:378-443):
 * PARGO_ORDER_ROWS drops the order (row-major tiles); PARGO_ORDER_ESTIMATE
 * launches the heaviest tiles first by the estimate of IdwTransformation
 * (apps/spatial/idw/idwTransformation.cpp:3-62) over coarse cells of 16 x 16
 * cells; PARGO_ORDER_MEASURED by the samples each tile's searches examine
 * in one pass (the preliminary experiment, idwOperator.cpp:439-491).  On
 * one GPU the load balance the reference gets from valRowDcmp's row bands
 * (gpro/deComposition.h:305-393) becomes the order workgroups start in;
 * values do not depend on it.  Synchronous. */
int pargo_idw_order_tiles(pargo_ctx *ctx, pargo_idw_plan *plan, int k, int power, double buffer,
                          const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, int source);

/* Measured work of IDW per cell, the device form of the reference's
 * preliminary experiment: with -dcmp space -writeLoad the reference stores
 * each cell's CPU time in the output layer (apps/spatial/idw/
 * idwOperator.cpp:439-491, pIDW.cpp:274-281); here work_out (device double,
 * rows x cols) receives the number of samples the cell's search examined
 * (0 for masked cells), a deterministic measure of the same cost.  out
 * (may be NULL) receives the IDW values as pargo_idw_run would. */
int pargo_idw_work(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                   const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata,
                   float *out, double *work_out);

/* ---- load-balanced decomposition ------------------------------------ */

/* Intensity-ratio compute load of one data layer: Transformation::Operator
 * (gpro/transformation.h:155-197) over the coarse grid of ComputeLayer::init
 * (gpro/computeLayer.h:208-260).  The coarse grid has
 * ceil(work_rows / g) x ceil(work_cols / g) cells; cell (i, j) sums, over
 * the data cells of rows work_row0 + i*g .. work_row0 + (i+1)*g - 1 and the
 * matching columns that lie inside the raster, valid_load for a value with
 * fabs(v - grid->nodata) > 1e-7 and nodata_load otherwise, in row-major
 * order.  data: device double raster (grid->pitch); load_out: device double
 * array of load_rows x load_cols.  accumulate != 0 adds to load_out (the
 * reference sums its data layers in order).  Loads must be >= 0.  With
 * load_out NULL only *load_rows / *load_cols are set. */
int pargo_intensity_load(pargo_ctx *ctx, const pargo_grid *grid, const double *data, int work_row0, int work_col0,
                         int work_rows, int work_cols, int granularity, double nodata_load, double valid_load,
                         int accumulate, double *load_out, int *load_rows, int *load_cols);

/* Estimated compute load of IDW per cell of a coarse grid of
 * ceil(rows/granularity) x ceil(cols/granularity) cells, cell size
 * (float)(cellsize * granularity): IdwTransformation::Operator
 * (apps/spatial/idw/idwTransformation.cpp:3-62) over ComputeLayer
 * (gpro/computeLayer.h:208-260).  load = blocks visited + 2 * samples
 * counted by a ring walk that visits only the full top and bottom rows of
 * each ring and counts the centre block twice, as the reference does.
 * mask as in pargo_idw_run but compared with mask_nodata as a double
 * (idwTransformation.cpp:13-19); masked cells get `nodata`.  With load_out
 * NULL only *load_rows / *load_cols are set.  load_out is a device double
 * array; the call waits for the result and returns PARGO_E_NONTERM when a
 * cell's walk can never reach k samples (the reference loops forever). */
int pargo_idw_load(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int granularity,
                   const int32_t *mask, int mask_rows, int mask_cols, double mask_nodata,
                   double nodata, double *load_out, int *load_rows, int *load_cols);

/* Splits the rows of a load grid (host array, rows x cols) into nbands
 * bands of near-equal summed load: DeComposition::valRowDcmp
 * (gpro/deComposition.h:305-393), which skips cells equal to -9999 or to
 * `nodata` (one-sided 1e-7 tests) and places each cut greedily.
 * band_first (host, nbands + 1 entries) receives the first row of every
 * band and rows as the end.  nbands must be in [1, rows]. */
int pargo_load_row_bands(const double *load, int rows, int cols, double nodata, int nbands,
                         int *band_first);

/* ---- adjoints (no counterpart in the reference) ---------------------- */

/* dL/dz of the Horn slope (PARGO_SLOPE_FD, PARGO_UNIT_TAN) given dL/dslope
 * per cell.  Cells whose slope is NoData or a border cell take no gradient;
 * a cell with zero gradient magnitude contributes 0.  Only the DEM is read:
 * the stencil intermediates are recomputed. */
int pargo_slope_adjoint(pargo_ctx *ctx, const pargo_grid *grid, const float *dem,
                        const float *grad_slope, float *grad_dem);

/* dL/d(sample value) of IDW given dL/d(out) per cell.  grad_vals is a
 * device double array of n entries in file order, overwritten. */
int pargo_idw_adjoint(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                      const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata,
                      const float *grad_out, double *grad_vals);

/* ---- files (host only; no GPU needed) --------------------------------- */

/* Raster header of a GeoTIFF: what RasterLayer::readFileInfo takes from
 * GDAL (gpro/rasterLayer.h:749-778). */
typedef struct {
    int32_t rows;
    int32_t cols;
    double transform[6]; /* GDAL geotransform */
    int32_t has_nodata;  /* GDAL_NODATA tag present */
    double nodata;
    int32_t bits;          /* bits per sample of the file */
    int32_t sample_format; /* 1 unsigned, 2 signed, 3 float */
} pargo_raster_info;

/* Sample types of pargo_geotiff_write. */
#define PARGO_IO_FLOAT64 0
#define PARGO_IO_FLOAT32 1
#define PARGO_IO_INT32 2
#define PARGO_IO_UINT8 3

/* Header only.  Errors are reported through pargo_last_error(NULL). */
int pargo_geotiff_info(const char *path, pargo_raster_info *info);
/* Band 1 widened to double (host array of rows * cols), as GDAL RasterIO
 * into a double buffer does (rasterLayer.h:882-903).  Reads classic and
 * BigTIFF, strips or tiles, uncompressed, LZW or PackBits, with or without
 * a horizontal predictor. */
int pargo_geotiff_read(const char *path, pargo_raster_info *info, double *data);
/* Writes one band, one row per strip, uncompressed, with the geotransform,
 * GDAL_NODATA and, when georef_from names a GeoTIFF, its GeoKeys
 * (rasterLayer.h:957-1087 writes Float64 for RasterLayer<double>). */
int pargo_geotiff_write(const char *path, const pargo_raster_info *info, const double *data, int type,
                        const char *georef_from);
/* Point shapefile <name>.shp/.dbf: x, y and the DBF field field_index
 * (0-based, strtod of the field text) of every non-deleted record, the
 * replacement of IDWOperator::readSamples' OGR calls
 * (apps/spatial/idw/idwOperator.cpp:63-149).  *n is the room in the output
 * arrays on entry and the point count on return; with all three arrays
 * NULL only the count is returned.  A non-point record is an error. */
int pargo_points_read(const char *shp_path, int field_index, int *n, double *xs, double *ys, double *vals);

/* ---- synthetic inputs (deterministic; SplitMix64 seeded) ------------ */

/* Fractal value-noise DEM on the host: 8 octaves, persistence 0.5, base
 * wavelength max(rows,cols)/4 cells, normalised to [0, 2000] m, float32. */
int pargo_synth_dem(uint64_t seed, int rows, int cols, float *host_out, int nthreads);
/* Rectangular plateaus at constant elevation until `fraction` of the cells
 * are covered (sides 3..50 cells). */
int pargo_synth_plateaus(uint64_t seed, int rows, int cols, float *host_dem, double fraction);
/* NoData discs (radius 5..40 cells): n_discs > 0 places that many discs,
 * otherwise discs are added until `fraction` of the cells are covered.
 * Returns the number of NoData cells through covered (may be NULL). */
int pargo_synth_nodata_discs(uint64_t seed, int rows, int cols, float *host_dem, int n_discs,
                             double fraction, float nodata, int64_t *covered);
/* IDW samples: n/2 uniform over the extent and n - n/2 from a 10-component
 * Gaussian mixture (sigma 1-5 km), values from the same noise field plus
 * N(0, 1 m), all float32-representable; the first two samples are the
 * anchors (x0+cs/2, y0+cs/2) and (x0+W-cs/2, y0+H-cs/2). */
int pargo_synth_samples(uint64_t seed, int n, double x0, double y0, double width, double height,
                        double cellsize, double *xs, double *ys, double *vals);

#ifdef __cplusplus
}
#endif

#endif /* PARGO_H */
