// fill.hip -- depression filling (pitremove) on the device.
//
// The reference (apps/hydrology/pitremove/pitRemoveOperator.cpp:26-72, run
// by RasterOperator::Work, gpro/rasterOperator.h:170-216) starts from
// W = 10000 on valid cells and NoData elsewhere (the border too, through
// copyLayerInfo) and sweeps the interior in row-major order until a sweep
// assigns nothing.  A valid cell with W > dem looks at its 3x3 window:
//   if dem >= W(n) + 0.0005 or W(n) is NoData for some n:  W = dem
//   otherwise                                              W = min(W, min_n W(n) + 0.0005)
// (the window loop never breaks, and once W = dem no later neighbour
// changes it again).
//
// That update is monotone in the neighbours' values as long as no valid
// cell's W can come within 1e-6 of NoData (W stays in [min(dem, 10000),
// 10000]; pargo_pit_remove checks this and refuses other inputs).  The
// reference's loop then ends at the greatest fixed point below the start,
// and so does any order of the same updates: every update reads values that
// are >= that fixed point, and a state in which no update changes anything
// is a fixed point.  The result is therefore bit-identical to the
// reference's while the order is free.
//
// Device order: 64 x 32 tiles.  A workgroup loads its tile and a one-cell
// halo of W into LDS (the halo columns from per-tile copies of the
// neighbours' edge columns) and sweeps the update over the tile (each wave
// down and up its 8 rows) until nothing in it changes; then it writes the
// changed cells back (and the edge copies of those in its first or last
// column) and schedules for the next pass the tile across each changed
// edge and the diagonal tile at each changed corner.  Passes repeat until
// one changes nothing.  Tiles read each other's edges while they change;
// that only feeds the monotone iteration earlier values, never wrong ones.
//
// Only scheduled tiles run: a pass's workgroups walk a device worklist with
// a fixed stride, and a tile enters the next pass's list when it is first
// stamped for it.  Pass 0's list holds the tiles with a NoData or border cell within
// their one-cell halo: with every valid W at 10000, nothing else can change
// (a cell moves only on a NoData neighbour or a neighbour below 10000).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace pargo {

namespace {

constexpr int kTW = 64, kTH = 32;  // tile: 64 columns x 32 rows, 8 cells per thread
constexpr int kRW = kTH / 4;       // rows per wave (4 waves of kTW lanes)
static_assert(kTW == 64, "fill_init_kernel takes a wave's 64 columns for one tile column");
constexpr double kGap = 0.0005;    // pitRemoveOperator.cpp:36
constexpr double kStart = 10000.0; // pitRemoveOperator.cpp:41

struct FillParams {
    const float *__restrict__ dem;
    double *__restrict__ w;
    int *__restrict__ stamp;         // per tile: the pass that must process it (-1: none yet)
    unsigned *__restrict__ changed;  // per pass of the current batch: tiles that changed
    unsigned *__restrict__ count;    // [2]: entries of the worklists of even and odd passes
    int *__restrict__ list;          // [2][tiles]: tiles of the pass (even, odd)
    double *__restrict__ edge;       // [tiles][2][kTH]: W of each tile's first and last column
    long long pitch;
    int rows, cols, tiles_x, tiles_y;
    float nd_lo, nd_hi;  // float32 values with fabs((double)v - nodata) <= 1e-6
    double nodata;
};

__device__ __forceinline__ bool is_nd(const FillParams &p, float v) { return v >= p.nd_lo && v <= p.nd_hi; }

// Lane i gets lane i-1's (from_left) or lane i+1's (from_right) v; the lane
// without a source (0, resp. 63) gets `edge`.  Two 32-bit DPP wave shifts.
__device__ __forceinline__ double dpp_from_left(double v, double edge)
{
    const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)b, 0x138, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(e >> 32), (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double dpp_from_right(double v, double edge)
{
    const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)e, (int)(unsigned)b, 0x130, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(e >> 32), (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Each tile's first and last column of W, also kept contiguously per tile:
// a tile's halo columns are its neighbours' edge columns, read as two short
// runs instead of one extra cache line per halo row on each side.
__device__ __forceinline__ long long edge_index(const FillParams &p, int ty, int tx, int right, int row_in_tile)
{
    return (((long long)ty * p.tiles_x + tx) * 2 + right) * kTH + row_in_tile;
}

// Puts tile t on the worklist of pass `pass` unless it is there already
// (its stamp says so).
__device__ __forceinline__ void schedule(const FillParams &p, int t, int pass)
{
    if (atomicExch(&p.stamp[t], pass) != pass) {
        const unsigned k = atomicAdd(&p.count[pass & 1], 1u);
        p.list[(long long)(pass & 1) * p.tiles_x * p.tiles_y + k] = t;
    }
}

// W before the first sweep (pitRemoveOperator.cpp:39-45 plus copyLayerInfo),
// and pass 0's worklist: every tile holding a NoData or border cell in its
// one-cell halo.
// kInitRows rows of 256 columns per workgroup: one workgroup per row
// segment spent most of the pass starting workgroups (a million of them at
// 16384^2).
constexpr int kInitRows = 16;

__device__ __forceinline__ void fill_init_cell(const FillParams &p, int r, int c)
{
    const long long i = (long long)r * p.pitch + c;
    const bool interior = r >= 1 && r <= p.rows - 2 && c >= 1 && c <= p.cols - 2;
    const bool valid = interior && !is_nd(p, p.dem[i]);
    const double w0 = valid ? kStart : p.nodata;
    // streamed past the caches: the passes read W back much later
    __builtin_nontemporal_store(w0, &p.w[i]);
    if (c % kTW == 0 || c % kTW == kTW - 1)
        p.edge[edge_index(p, r / kTH, c / kTW, c % kTW != 0, r % kTH)] = w0;
    // The tiles whose cells or halo hold an excluded cell of this wave.  A
    // wave is 64 columns of one row starting at a multiple of kTW, so they
    // are its own tile column w, w - 1 when its first cell is excluded and
    // w + 1 when its last one is, in the tile rows of r - 1 .. r + 1.  Lane 0
    // (the first cell, active whenever any lane is) schedules them, once per
    // wave rather than once per cell: same-address atomics serialise.
    const unsigned long long nd = __builtin_amdgcn_ballot_w64(!valid);
    if (nd == 0 || (threadIdx.x & 63) != 0) return;
    const int w = c / kTW;
    const int tx0 = (nd & 1ull) && w > 0 ? w - 1 : w;
    const int tx1 = (nd >> 63) && (w + 1) * kTW <= p.cols - 1 ? w + 1 : w;
    const int ty0 = max(r - 1, 0) / kTH, ty1 = min(r + 1, p.rows - 1) / kTH;
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1; tx++) {
            const int t = ty * p.tiles_x + tx;
            // a plain read skips the atomic once the stamp is set: stamps only
            // go from -1 to 0 here, so a stale -1 costs an atomic, nothing more
            if (p.stamp[t] != 0) schedule(p, t, 0);
        }
}

__global__ __launch_bounds__(256) void fill_init_kernel(FillParams p)
{
    const int c = blockIdx.x * 256 + threadIdx.x, rb = blockIdx.y * kInitRows;
    if (c >= p.cols) return;  // whole waves: 256-column segments hold 4 of them
    const int re = min(rb + kInitRows, p.rows);
    for (int r = rb; r < re; r++) fill_init_cell(p, r, c);
}

// Stamps to -1 and both worklists empty (before the first pass), or the
// per-pass change counters of a batch to zero.
__global__ void fill_stamp_init_kernel(int *stamp, int n, unsigned *changed, int npass, unsigned *count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) stamp[i] = -1;
    if (i < npass) changed[i] = 0;
    if (n > 0 && i < 2) count[i] = 0;
}

// Empties the worklist that pass `pass` + 1 fills (the one pass - 1 read).
__global__ void fill_list_reset_kernel(unsigned *count, int pass) { count[(pass + 1) & 1] = 0; }

// One pass over one tile (see the top of the file).
__device__ __forceinline__ void fill_tile(const FillParams &p, int tile, int pass, int batch_first,
                                          double (&sw)[kTH + 2][kTW + 2], int &edge_flags, int &run, int tid)
{
    // one read of the stamp for the whole workgroup: a neighbour may rewrite
    // it meanwhile (to pass + 1, which defers this tile by one pass)
    if (tid == 0) {
        edge_flags = 0;
        run = p.stamp[tile] == pass;
    }
    __syncthreads();
    if (!run) return;  // moved on to the next pass meanwhile
    const int ty = tile / p.tiles_x, tx = tile - ty * p.tiles_x;
    const int r0 = ty * kTH, c0 = tx * kTW;
    {
        // Rows r0-1 .. r0+kTH of the tile's columns from W (aligned runs of
        // kTW values), the halo columns from the neighbours' edge columns.
        // Every load of the thread is issued before the first store to LDS,
        // so that their latencies overlap.
        constexpr int N = (kTH + 2) * kTW, PER = (N + 255) / 256;
        double v[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + k * 256;
            const int r = r0 - 1 + i / kTW, c = c0 + i % kTW;
            v[k] = (i < N && r >= 0 && r < p.rows && c < p.cols) ? p.w[(long long)r * p.pitch + c] : p.nodata;
        }
        // threads 0 .. 2(kTH+2)-1: left halo column, then right
        double h = p.nodata;
        const int side = tid / (kTH + 2), hr = tid % (kTH + 2);
        if (side < 2) {
            const int r = r0 - 1 + hr, c = side == 0 ? c0 - 1 : c0 + kTW;
            if (r >= 0 && r < p.rows && c >= 0 && c < p.cols)
                h = p.edge[edge_index(p, r / kTH, c / kTW, side == 0, r % kTH)];
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + k * 256;
            if (i < N) sw[i / kTW][i % kTW + 1] = v[k];
        }
        if (side < 2) sw[hr][side == 0 ? 0 : kTW + 1] = h;
    }
    // own cells: wave w owns tile rows 8w .. 8w+7, lane = column
    const int lc = tid % kTW + 1;
    const int rbase = (tid / kTW) * kRW;  // first tile row of this wave
    float dem_f[kRW];  // the cells' DEM values (widened to double where used)
    unsigned live = 0;  // bit k: tile row rbase + k holds a valid interior cell
#pragma unroll
    for (int k = 0; k < kRW; k++) {
        const int r = r0 + rbase + k, c = c0 + lc - 1;
        dem_f[k] = 0.0f;
        if (r >= 1 && r <= p.rows - 2 && c >= 1 && c <= p.cols - 2) {
            const float f = p.dem[(long long)r * p.pitch + c];
            if (!is_nd(p, f)) {
                dem_f[k] = f;
                live |= 1u << k;
            }
        }
    }
    __syncthreads();
    // A NoData neighbour sends a cell to its DEM value whatever the other
    // neighbours hold.  NoData cells keep W = NoData and no valid W comes
    // within 1e-6 of it, so which cells have one is fixed during the visit.
    unsigned nd_nbr = 0;  // bit k: the cell in tile row rbase + k has a NoData neighbour
#pragma unroll
    for (int k = 0; k < kRW; k++) {
        const int lr = rbase + k + 1;
        bool any = false;
#pragma unroll
        for (int n = 0; n < 9; n++)
            if (n != 4) any |= fabs(sw[lr + n / 3 - 1][lc + n % 3 - 1] - p.nodata) < 0.000001;
        nd_nbr |= (unsigned)any << k;
    }
    // The wave's 8 cells per lane live in registers during the sweeps; left
    // and right neighbours come from the adjacent lanes by DPP, the halo
    // columns (lane 0: the left one, lane 63: the right one) from LDS, and
    // the rows above and below the wave's band from LDS, where the waves
    // publish their first and last rows after every half-sweep.
    const int lane = tid & 63;
    double wv[kRW];
#pragma unroll
    for (int k = 0; k < kRW; k++) wv[k] = sw[rbase + k + 1][lc];
    // halo column of tile row j-1 (j = sw row): the left one for lane 0, the
    // right one for lane 63 (static during the visit; a broadcast LDS read)
    const int ecol = lane == 0 ? 0 : kTW + 1;
    auto eg = [&](int j) -> double { return sw[rbase + j][ecol]; };
    double up_b = sw[rbase][lc], dn_b = sw[rbase + kRW + 1][lc];
    unsigned touched = 0;  // bit k: the cell changed during this pass
    // One update of the cell in tile row rbase + k; true if it changed.
    // Without a NoData neighbour the reference's window loop gives
    // W = dem if dem >= W(n) + gap for some n, else min(W, min_n W(n) + gap);
    // the rounded addition is monotone, so min_n (W(n) + gap) is
    // (min_n W(n)) + gap with one rounding, and "for some n" is a test
    // against that minimum.  Every lane runs every step (the DPP moves read
    // the neighbouring lanes), and the cells that must not change keep
    // their value by selection.
    auto update = [&](int k) -> bool {
        const double up = k == 0 ? up_b : wv[k > 0 ? k - 1 : 0];
        const double dn = k == kRW - 1 ? dn_b : wv[k < kRW - 1 ? k + 1 : 0];
        const double w = wv[k];
        double wmin = fmin(fmin(dpp_from_left(up, eg(k)), up), dpp_from_right(up, eg(k)));
        wmin = fmin(wmin, fmin(dpp_from_left(w, eg(k + 1)), dpp_from_right(w, eg(k + 1))));
        wmin = fmin(wmin, fmin(fmin(dpp_from_left(dn, eg(k + 2)), dn), dpp_from_right(dn, eg(k + 2))));
        const double dk = (double)dem_f[k];
        const double sm = wmin + kGap;
        const double nw = ((nd_nbr >> k) & 1u) ? dk : (dk >= sm ? dk : fmin(w, sm));
        const bool ch = ((live >> k) & 1u) && w > dk && nw != w;
        wv[k] = ch ? nw : w;
        touched |= (unsigned)ch << k;
        return ch;
    };
    // Each iteration sweeps the wave's 8 rows down and then up, one row per
    // step (a lowered value travels 8 rows and up to 16 columns), with the
    // band edges exchanged between the half-sweeps.  A wave may read a
    // neighbour's row before or after that neighbour's latest change; either
    // feeds the monotone iteration a value at or above the fixed point, and
    // the loop ends only after an iteration in which no cell changed.
    for (;;) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < kRW; k++) any |= update(k);
        sw[rbase + 1][lc] = wv[0];
        sw[rbase + kRW][lc] = wv[kRW - 1];
        __syncthreads();
        up_b = sw[rbase][lc];
        dn_b = sw[rbase + kRW + 1][lc];
#pragma unroll
        for (int k = kRW - 1; k >= 0; k--) any |= update(k);
        sw[rbase + 1][lc] = wv[0];
        sw[rbase + kRW][lc] = wv[kRW - 1];
        const bool more = __syncthreads_or(any);
        up_b = sw[rbase][lc];
        dn_b = sw[rbase + kRW + 1][lc];
        if (!more) break;
    }
    unsigned flags = 0;
#pragma unroll
    for (int k = 0; k < kRW; k++) {
        if (!((touched >> k) & 1u)) continue;
        const int tr = rbase + k, tc = lc - 1;
        const int r = r0 + tr, c = c0 + tc;
        const double nw = wv[k];
        p.w[(long long)r * p.pitch + c] = nw;
        if (tc == 0 || tc == kTW - 1) p.edge[edge_index(p, ty, tx, tc != 0, tr)] = nw;
        const unsigned top = tr == 0, bottom = tr == kTH - 1, left = tc == 0, right = tc == kTW - 1;
        // edges (bits 0-3) schedule the tile across them; a corner (bits
        // 4-7) is the only cell in a diagonal neighbour's halo
        flags |= top | (bottom << 1) | (left << 2) | (right << 3) | ((top & left) << 4) | ((top & right) << 5) |
                 ((bottom & left) << 6) | ((bottom & right) << 7) | (1u << 8);
    }
    if (flags) atomicOr(&edge_flags, (int)flags);
    __syncthreads();
    if (tid == 0 && edge_flags) {
        const int f = edge_flags;
        atomicAdd(&p.changed[pass - batch_first], 1u);
        const int next = pass + 1;
        for (int dy = -1; dy <= 1; dy++) {
            for (int dx = -1; dx <= 1; dx++) {
                if (dx == 0 && dy == 0) continue;
                const int corner = 4 + (dy == 1 ? 2 : 0) + (dx == 1 ? 1 : 0);
                const bool need = dx == 0   ? (f >> (dy == -1 ? 0 : 1)) & 1
                                  : dy == 0 ? (f >> (dx == -1 ? 2 : 3)) & 1
                                            : (f >> corner) & 1;
                const int nty = ty + dy, ntx = tx + dx;
                if (need && nty >= 0 && nty < p.tiles_y && ntx >= 0 && ntx < p.tiles_x)
                    schedule(p, nty * p.tiles_x + ntx, next);
            }
        }
    }
}


__global__ __launch_bounds__(256, 4) void fill_pass_kernel(FillParams p, int pass, int batch_first)
{
    __shared__ double sw[kTH + 2][kTW + 2];
    __shared__ int edge_flags, run;
    const int tid = threadIdx.x;
    const unsigned n = p.count[pass & 1];
    const int *list = p.list + (long long)(pass & 1) * p.tiles_x * p.tiles_y;
    // A fixed stride over the list keeps the tiles in flight apart (tiles
    // next to each other in the list are neighbours); taking them one at a
    // time from a cursor ran neighbours together, which then revisited each
    // other more (27 against 24 ms).  The grid is what fits on the GPU at
    // once, so no workgroup starts late with a full share.
    for (unsigned e = blockIdx.x; e < n; e += gridDim.x) {
        fill_tile(p, list[e], pass, batch_first, sw, edge_flags, run, tid);
        __syncthreads();  // the next tile reuses the LDS tile
    }
}
// Smallest valid elevation (as ordered bits of a float) and a NaN count,
// for the monotonicity check: grid-stride over the interior, one atomic per
// workgroup.
__global__ __launch_bounds__(256) void fill_range_kernel(FillParams p, unsigned *minbits, unsigned *nans)
{
    __shared__ unsigned smin[4], sbad[4];
    unsigned u = 0xffffffffu, bad = 0;
    // interior rows by workgroup, their interior columns by thread (a flat
    // index would need a 64-bit division per cell)
    for (int r = 1 + (int)blockIdx.x; r <= p.rows - 2; r += (int)gridDim.x) {
        const float *row = p.dem + (long long)r * p.pitch;
        for (int c = 1 + (int)threadIdx.x; c <= p.cols - 2; c += 256) {
            const float f = row[c];
            if (f != f) {
                bad++;
            } else if (!is_nd(p, f)) {
                unsigned k = __float_as_uint(f);  // order-preserving map of floats to unsigned
                k = (k & 0x80000000u) ? ~k : (k | 0x80000000u);
                u = k < u ? k : u;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(u, off);
        u = o < u ? o : u;
        bad += __shfl_xor(bad, off);
    }
    if ((threadIdx.x & 63) == 0) {
        smin[threadIdx.x >> 6] = u;
        sbad[threadIdx.x >> 6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = smin[0], b = sbad[0];
        for (int k = 1; k < 4; k++) {
            m = smin[k] < m ? smin[k] : m;
            b += sbad[k];
        }
        atomicMin(minbits, m);
        if (b) atomicAdd(nans, b);
    }
}

}  // namespace

}  // namespace pargo

using pargo::fail;

extern "C" int pargo_pit_remove(pargo_ctx *ctx, const pargo_grid *g, const float *dem, double *filled_out,
                                int *passes_out)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!g || !dem || !filled_out) return fail(ctx, PARGO_E_INVALID, "pargo_pit_remove: null argument");
    if (g->rows < 1 || g->cols < 1) return fail(ctx, PARGO_E_INVALID, "empty grid (%d x %d)", g->rows, g->cols);
    const long long pitch = g->pitch ? g->pitch : g->cols;
    if (pitch < g->cols) return fail(ctx, PARGO_E_INVALID, "pitch %lld < cols %d", pitch, g->cols);
    if (std::isnan(g->nodata)) return fail(ctx, PARGO_E_INVALID, "NaN NoData is not supported");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    pargo::FillParams p{};
    p.dem = dem;
    p.w = filled_out;
    p.pitch = pitch;
    p.rows = g->rows;
    p.cols = g->cols;
    p.tiles_x = (g->cols + pargo::kTW - 1) / pargo::kTW;
    p.tiles_y = (g->rows + pargo::kTH - 1) / pargo::kTH;
    p.nodata = g->nodata;
    // the reference keeps a cell when fabs(dem - noData) > 1e-6
    // (pitRemoveOperator.cpp:40,53): NoData is the float interval with
    // fabs <= 1e-6
    pargo::nodata_interval(g->nodata, 0.000001, &p.nd_lo, &p.nd_hi, true);
    if (p.nd_lo > p.nd_hi) {
        p.nd_lo = INFINITY;
        p.nd_hi = -INFINITY;
    }
    const int ntiles = p.tiles_x * p.tiles_y;
    constexpr int kBatch = 8;      // passes between checks
    constexpr int kMaxPasses = 1 << 20;
    // scratch: [0] min bits, [1] NaN count, per-pass counters, the two
    // worklist counts, then stamps, the two worklists and the tiles' edge
    // columns
    const size_t counters = 2 + (size_t)kBatch + 2;
    const size_t edge_off = (sizeof(unsigned) * counters + sizeof(int) * 3 * (size_t)ntiles + 255) / 256 * 256;
    rc = pargo::ensure_scratch(ctx, edge_off + sizeof(double) * 2 * pargo::kTH * (size_t)ntiles);
    if (rc) return rc;
    unsigned *scr = static_cast<unsigned *>(ctx->scratch);
    unsigned *minbits = scr, *nans = scr + 1, *changed = scr + 2;
    p.changed = changed;
    p.count = scr + 2 + kBatch;
    p.stamp = reinterpret_cast<int *>(scr + counters);
    p.list = p.stamp + ntiles;
    p.edge = reinterpret_cast<double *>(static_cast<char *>(ctx->scratch) + edge_off);

    // monotonicity check: every valid W lies in [min(dem, 10000), 10000]
    unsigned host2[2] = {0xffffffffu, 0u};
    PARGO_HIP(ctx, hipMemcpyAsync(scr, host2, sizeof(host2), hipMemcpyHostToDevice, ctx->stream));
    dim3 initgrid((g->cols + 255) / 256, (g->rows + pargo::kInitRows - 1) / pargo::kInitRows);
    const int cus = ctx->num_cu > 0 ? ctx->num_cu : 256;
    hipLaunchKernelGGL(pargo::fill_range_kernel, dim3(cus * 8), dim3(256), 0, ctx->stream, p, minbits, nans);
    PARGO_HIP(ctx, hipMemcpyAsync(host2, scr, sizeof(host2), hipMemcpyDeviceToHost, ctx->stream));
    PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (host2[1]) return fail(ctx, PARGO_E_INVALID, "the DEM holds %u NaN cells: the reference never settles there", host2[1]);
    {
        unsigned u = host2[0];
        u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
        float mn;
        std::memcpy(&mn, &u, sizeof(mn));
        // no valid cell leaves the smallest key's image at NaN
        const bool any_valid = !std::isnan(mn);
        const double lo = any_valid ? std::min((double)mn, 10000.0) : 0.0;
        const bool ok = !any_valid || g->nodata < lo - 0.000001 || g->nodata > 10000.0 + 0.000001;
        if (!ok)
            return fail(ctx, PARGO_E_UNSUPPORTED,
                        "NoData %.17g lies inside the range of filled values [%.17g, 10000]: the reference's "
                        "result would depend on its sweep order",
                        g->nodata, lo);
    }
    // stamps at -1 and empty worklists, then W and pass 0's worklist
    hipLaunchKernelGGL(pargo::fill_stamp_init_kernel, dim3((std::max(ntiles, kBatch) + 255) / 256), dim3(256), 0,
                       ctx->stream, p.stamp, ntiles, changed, kBatch, p.count);
    hipLaunchKernelGGL(pargo::fill_init_kernel, initgrid, dim3(256), 0, ctx->stream, p);
    // a fixed grid walks each pass's worklist: as many workgroups as are
    // resident at once
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pargo::fill_pass_kernel, 256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    const int pass_blocks = std::min(ntiles, cus * per_cu);
    int pass = 0;
    for (;;) {
        // change counters per batch
        if (pass > 0)
            hipLaunchKernelGGL(pargo::fill_stamp_init_kernel, dim3(1), dim3(256), 0, ctx->stream, p.stamp, 0, changed,
                               kBatch, p.count);
        for (int k = 0; k < kBatch; k++) {
            hipLaunchKernelGGL(pargo::fill_list_reset_kernel, dim3(1), dim3(1), 0, ctx->stream, p.count, pass + k);
            hipLaunchKernelGGL(pargo::fill_pass_kernel, dim3(pass_blocks), dim3(256), 0, ctx->stream, p, pass + k, pass);
        }
        unsigned h[kBatch];
        PARGO_HIP(ctx, hipMemcpyAsync(h, changed, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
        rc = pargo::check_launch(ctx, "pargo_pit_remove");
        if (rc) return rc;
        int settled = -1;
        for (int k = 0; k < kBatch; k++)
            if (h[k] == 0) {
                settled = k;
                break;
            }
        if (settled >= 0) {
            pass += settled + 1;
            break;
        }
        pass += kBatch;
        if (pass >= kMaxPasses)
            return fail(ctx, PARGO_E_NONTERM, "depression filling did not settle within %d passes", kMaxPasses);
    }
    if (passes_out) *passes_out = pass;
    return PARGO_OK;
}
