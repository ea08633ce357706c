// sca.hip -- D8 contributing area (scad8) on the device.
//
// The reference (apps/hydrology/scad8/scaOperator.cpp:37-131, run by
// RasterOperator::Work, gpro/rasterOperator.h:170-216) sweeps the grid in
// row-major order three times: NoData/initial values, then the number of
// neighbours draining into each cell ("degree"), then ONE accumulation sweep
// in which a cell pulls the area of every inflowing neighbour whose degree
// is 0 (a cell nothing drains into) or -1 (finished earlier in the same
// sweep).  Every accumulation sweep ends with Termination = 1
// (scaOperator.cpp:126), so Work stops after that single sweep.
//
// What one sweep computes, stated without the sweep order:
//   In(c)   = neighbours u whose code points at c, unless c's own code
//             points back at u (scaOperator.cpp:71,94);
//   source  = interior valid cell with In empty (degree 0 after sweep 2);
//   avail(u,c) = u is a source, or u precedes c in row-major order
//             (u is the NW, N, NE or W neighbour) and u is done;
//   done(c) = every u in In(c) is avail(u,c);
//   sca(c)  = 1 + sum of sca(u) over avail(u,c).
// Border cells are never visited: their degree keeps the layer's NoData
// value, so a border cell draining inward is never available.
// The last work cell runs the pull loop even when its degree is <= 0
// (scaOperator.cpp:86): a NoData cell there adds its available inflows to
// NoData.
//
// The "converge" variant repeats the accumulation sweep until nothing
// changes, which lifts the row-major restriction: every inflow is pending
// and sca(c) = 1 + sum over done inflows.
//
// Device formulation: the dependencies that matter (non-source inflows the
// cell waits for) form a forest, each cell feeding at most its one target.
// One 64-bit word per cell, kept in the output buffer, holds the cell's
// static facts and an arrival accumulator:
//   bits  0-39  sum of the areas of done arrivals
//   bits 40-43  arrivals that were not done
//   bits 44-47  arrivals
//   bits 48-51  npend: inflows the cell waits for
//   bits 52-55  nsrc: inflows from sources (each adds 1)
//   bit  56     an inflow can never be available (border, or behind in
//               the one-sweep order)
//   bit  57     feeds: the cell is one of its target's pending inflows
//   bits 58-61  the cell's code (window index of its target)
//   bit  62     the cell has an area (interior, not NoData)
//   bit  63     the last-cell rule applies (value in the context scratch)
// Every cell with npend == 0 that feeds starts a walk: it adds its result
// to its target with one atomicAdd; the arrival that completes the target
// (the returned word holds everything else) carries on from there.  Each
// cell is finished exactly once, so the work is linear in the cells; the
// walk is as long as the longest chain of completions.
#include <cmath>

#include "internal.h"

namespace pargo {

namespace {

constexpr unsigned kNone = 15;
constexpr int kTileC = 64, kTileR = 32;  // cells per workgroup: 64 x 32
constexpr uint64_t kAccMask = (1ull << 40) - 1;

__device__ __forceinline__ unsigned flow_of(unsigned b) { return (b <= 8u && b != 4u) ? b : kNone; }
__device__ __forceinline__ unsigned excl_of(unsigned b)
{
    if (b <= 8u) return b;
    if ((b & 0xF0u) == PARGO_FLOW_EXCL && (b & 15u) <= 8u) return b & 15u;
    return kNone;
}
__device__ __forceinline__ bool is_valid(unsigned b) { return b != PARGO_FLOW_NODATA && b != PARGO_FLOW_INERT; }

struct ScaParams {
    const uint8_t *__restrict__ flow;
    const uint32_t *__restrict__ mask;  // may be null
    unsigned long long *__restrict__ w; // the output buffer, as words
    long long pitch;
    int rows, cols, mask_words;
    int converge;
};

__device__ __forceinline__ unsigned load_class(const ScaParams &p, int r, int c)
{
    if (r < 0 || r >= p.rows || c < 0 || c >= p.cols) return PARGO_FLOW_NODATA;
    if (p.mask && ((p.mask[(long long)r * p.mask_words + (c >> 5)] >> (c & 31)) & 1u)) return PARGO_FLOW_NODATA;
    return p.flow[(long long)r * p.pitch + c];
}

// Static words of a 64 x 32 tile.  Classes of the tile plus a 2-cell halo
// go to LDS, then the inflow mask of the tile plus a 1-cell halo (bit j set
// when the neighbour at window position j < 4 ? j : j + 1 drains into the
// cell and is not excluded; a cell needs to know which of its inflows are
// sources, and whether its target counts it).
__device__ __forceinline__ int win_of_bit(int j) { return j < 4 ? j : j + 1; }
__device__ __forceinline__ int bit_of_win(int k) { return k < 4 ? k : k - 1; }

__global__ __launch_bounds__(256) void sca_init_kernel(ScaParams p, unsigned long long *cursor)
{
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *cursor = 0;  // for the walk kernel
    constexpr int CW = kTileC + 4, CH = kTileR + 4;  // classes, halo 2
    constexpr int NW = kTileC + 2, NH = kTileR + 2;  // inflow masks, halo 1
    __shared__ uint8_t cls[CH][CW];
    __shared__ uint8_t inm[NH][NW];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.y * kTileR, c0 = blockIdx.x * kTileC;
    for (int i = tid; i < CH * CW; i += 256) {
        const int lr = i / CW, lc = i % CW;
        cls[lr][lc] = (uint8_t)load_class(p, r0 - 2 + lr, c0 - 2 + lc);
    }
    __syncthreads();
    for (int i = tid; i < NH * NW; i += 256) {
        const int lr = i / NW + 1, lc = i % NW + 1;  // position in cls
        const unsigned b = cls[lr][lc];
        unsigned m = 0;
        if (is_valid(b)) {
            const unsigned ex = excl_of(b);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = win_of_bit(j);
                const unsigned u = cls[lr + k / 3 - 1][lc + k % 3 - 1];
                m |= (flow_of(u) == (unsigned)(8 - k) && ex != (unsigned)k) ? 1u << j : 0u;
            }
        }
        inm[lr - 1][lc - 1] = (uint8_t)m;
    }
    __syncthreads();
    for (int i = tid; i < kTileR * kTileC; i += 256) {
        const int tr = i / kTileC, tc = i % kTileC;
        const int r = r0 + tr, c = c0 + tc;
        if (r >= p.rows || c >= p.cols) continue;
        const bool interior = r >= 1 && r <= p.rows - 2 && c >= 1 && c <= p.cols - 2;
        const unsigned b = cls[tr + 2][tc + 2];
        unsigned long long word = 0;
        if (interior && b == PARGO_FLOW_INERT) {
            word = 1ull << 62;
        } else if (interior && b != PARGO_FLOW_NODATA) {
            const unsigned mine = inm[tr + 1][tc + 1];
            unsigned npend = 0, nsrc = 0, poison = 0;
            for (unsigned m = mine; m != 0; m &= m - 1) {
                const int k = win_of_bit(__builtin_ctz(m));
                const int ur = r + k / 3 - 1, uc = c + k % 3 - 1;
                const bool u_interior = ur >= 1 && ur <= p.rows - 2 && uc >= 1 && uc <= p.cols - 2;
                if (!u_interior) poison = 1;
                else if (inm[tr + 1 + k / 3 - 1][tc + 1 + k % 3 - 1] == 0) nsrc++;
                else if (p.converge || k < 4) npend++;
                else poison = 1;
            }
            const unsigned f = flow_of(b);
            unsigned feeds = 0;
            if (f != kNone && mine != 0 && (p.converge || f > 4)) {
                // the target counts this cell when its inflow mask has it
                const int trr = r + (int)f / 3 - 1, tcc = c + (int)f % 3 - 1;
                const bool t_interior = trr >= 1 && trr <= p.rows - 2 && tcc >= 1 && tcc <= p.cols - 2;
                const unsigned tm = inm[tr + 1 + (int)f / 3 - 1][tc + 1 + (int)f % 3 - 1];
                feeds = (t_interior && ((tm >> bit_of_win(8 - (int)f)) & 1u)) ? 1u : 0u;
            }
            word = ((unsigned long long)npend << 48) | ((unsigned long long)nsrc << 52) |
                   ((unsigned long long)poison << 56) | ((unsigned long long)feeds << 57) |
                   ((unsigned long long)(f == kNone ? 0u : f) << 58) | (1ull << 62);
        }
        p.w[(long long)r * p.pitch + c] = word;
    }
}

__device__ __forceinline__ unsigned npend_of(unsigned long long w) { return (unsigned)(w >> 48) & 15u; }
__device__ __forceinline__ unsigned nsrc_of(unsigned long long w) { return (unsigned)(w >> 52) & 15u; }
__device__ __forceinline__ bool poisoned_of(unsigned long long w) { return (w >> 56) & 1ull; }
__device__ __forceinline__ bool feeds_of(unsigned long long w) { return (w >> 57) & 1ull; }
__device__ __forceinline__ int code_of(unsigned long long w) { return (int)(w >> 58) & 15; }

// Walks, persistent: every lane runs one hop per iteration; a lane without
// a walk takes the next cell of its wave's chunk (chunks of kChunk cells come
// from a global cursor) and starts a walk there if the cell waits for
// nothing and feeds its target.  Lanes stay busy until the cursor passes the
// last cell and every walk of the wave has ended.  Checking one cell per
// idle lane and iteration keeps the number of hops in flight, and so the
// queueing delay on the longest chain, moderate: starting every walk at
// once was slower.
//
// A hop first reads the target's word.  A target with one pending inflow
// has this walk as its only writer, so the walk completes it with a plain
// store; only confluences (two or more pending inflows, a few percent of
// the hops on D8 grids) take the atomicAdd.
constexpr long long kChunk = 4096;

__device__ __forceinline__ long long first_lane(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffll));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__global__ __launch_bounds__(256) void sca_walk_kernel(ScaParams p, unsigned long long *cursor)
{
    const long long ncells = (long long)p.rows * p.cols;
    long long chunk_next = 0, chunk_end = 0;  // the same in every lane of a wave
    bool more = true;                         // the same in every lane of a wave
    long long cur = -1;                       // cell this lane's walk stands on
    int code = 0;
    unsigned long long val = 0;
    bool done = false;
    for (;;) {
        // the target's word is requested first, so that its latency
        // overlaps the refill below
        long long t = -1;
        unsigned long long tw = 0;
        if (cur >= 0) {
            t = cur + (long long)(code / 3 - 1) * p.pitch + (code % 3 - 1);
            tw = p.w[t];
        }
        const unsigned long long need = __ballot(cur < 0);
        long long fresh = -1;
        unsigned long long fw = 0;
        if (need != 0 && more) {
            if (chunk_next >= chunk_end) {
                long long b = 0;
                if (__lane_id() == 0) b = (long long)atomicAdd(cursor, (unsigned long long)kChunk);
                b = first_lane(b);
                if (b >= ncells) {
                    more = false;
                } else {
                    chunk_next = b;
                    chunk_end = b + kChunk < ncells ? b + kChunk : ncells;
                }
            }
            if (more) {
                const unsigned rank =
                    __builtin_amdgcn_mbcnt_hi((unsigned)(need >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)need, 0u));
                const long long npop = __popcll(need);
                const long long take = npop < chunk_end - chunk_next ? npop : chunk_end - chunk_next;
                if (cur < 0 && (long long)rank < take) {
                    const long long idx = chunk_next + rank;
                    const int r = (int)(idx / p.cols), c = (int)(idx - (long long)r * p.cols);
                    fresh = (long long)r * p.pitch + c;
                    fw = p.w[fresh];
                }
                chunk_next += take;
            }
        }
        if (cur >= 0) {
            const unsigned long long add = (done ? val : 0ull) | ((unsigned long long)(!done) << 40) | (1ull << 44);
            unsigned long long old;
            bool complete;
            if (npend_of(tw) == 1) {
                old = tw;  // nobody else writes this word
                p.w[t] = tw + add;
                complete = true;
            } else {
                old = atomicAdd(&p.w[t], add);
                complete = ((unsigned)(old >> 44) & 15u) + 1u == npend_of(old);
            }
            if (!complete) {
                cur = -1;  // an inflow is still on its way
            } else {
                const unsigned long long acc = (old & kAccMask) + (done ? val : 0ull);
                const unsigned not_done = ((unsigned)(old >> 40) & 15u) + (done ? 0u : 1u);
                done = !poisoned_of(old) && not_done == 0;
                val = 1ull + nsrc_of(old) + acc;
                if (feeds_of(old)) {
                    cur = t;
                    code = code_of(old);
                } else {
                    cur = -1;
                }
            }
        } else if (fresh >= 0 && ((fw >> 62) & 1ull) && npend_of(fw) == 0 && feeds_of(fw)) {
            cur = fresh;
            code = code_of(fw);
            val = 1ull + nsrc_of(fw);
            done = !poisoned_of(fw);
        }
        if (!more && __ballot(cur >= 0) == 0) break;
    }
}

// The reference's last work cell (rows-2, cols-2) always runs the pull
// loop (scaOperator.cpp:86-108).  When it is NoData (or a cell that takes no
// inflow) it adds, in window order, the area of every interior inflowing
// neighbour that is done; all of those precede it in row-major order.
__global__ void sca_last_cell_kernel(ScaParams p, double nodata, double *slot)
{
    const int r = p.rows - 2, c = p.cols - 2;
    const unsigned b = load_class(p, r, c);
    if (is_valid(b)) return;
    double v = b == PARGO_FLOW_NODATA ? nodata : 1.0;
    for (int k = 0; k < 9; k++) {
        if (k == 4) continue;
        const int ur = r + k / 3 - 1, uc = c + k % 3 - 1;
        if (ur < 1 || ur > p.rows - 2 || uc < 1 || uc > p.cols - 2) continue;
        if (flow_of(load_class(p, ur, uc)) != (unsigned)(8 - k)) continue;
        const unsigned long long w = p.w[(long long)ur * p.pitch + uc];
        const bool done = !poisoned_of(w) && ((w >> 40) & 15ull) == 0 && ((w >> 44) & 15ull) == npend_of(w);
        if (done) v += (double)(1ull + nsrc_of(w) + (w & kAccMask));
    }
    slot[0] = v;
    p.w[(long long)r * p.pitch + c] |= 1ull << 63;
}

// Words to areas, in place.
__global__ __launch_bounds__(256) void sca_final_kernel(ScaParams p, double nodata, const double *slot)
{
    const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
    if (c >= p.cols) return;
    const long long i = (long long)r * p.pitch + c;
    const unsigned long long w = p.w[i];
    double v;
    if (w >> 63) v = slot[0];
    else if ((w >> 62) & 1ull) v = (double)(1ull + nsrc_of(w) + (w & kAccMask));
    else v = nodata;
    reinterpret_cast<double *>(p.w)[i] = v;
}

// Reference D8 layer (double) to flow classes: NoData when
// fabs(v - nodata) < 1e-6, valid when > 1e-6 (scaOperator.cpp:48,71), and
// neither (NaN) marks a cell that takes no inflow.  A valid cell flows to
// window index k when v == k; it rejects inflow from window cell k when
// dir + v == 8 in double for dir = 8 - k (scaOperator.cpp:71,94).
__global__ __launch_bounds__(256) void d8_classify_kernel(const double *__restrict__ d8, uint8_t *__restrict__ out,
                                                          long long pitch, int rows, int cols, double nodata)
{
    const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
    if (c >= cols) return;
    const long long i = (long long)r * pitch + c;
    const double v = d8[i];
    const double dv = fabs(v - nodata);
    unsigned b;
    if (dv < 0.000001) {
        b = PARGO_FLOW_NODATA;
    } else if (dv > 0.000001) {
        unsigned ex = kNone;
        for (int dir = 0; dir <= 8; dir++)
            if (dir != 4 && (double)dir + v == 8.0) ex = 8u - (unsigned)dir;
        b = PARGO_FLOW_NONE;
        for (int k = 0; k <= 8; k++)
            if (k != 4 && v == (double)k) b = (unsigned)k;
        if (b == PARGO_FLOW_NONE && ex != kNone) b = PARGO_FLOW_EXCL | ex;
    } else {
        b = PARGO_FLOW_INERT;
    }
    out[i] = (uint8_t)b;
}

// NoData values the flow classes cannot carry: the reference would treat
// NoData cells holding a code as draining, and border degrees equal to 0 or
// -1 as available.
bool sca_nodata_ok(double nodata)
{
    if (std::isnan(nodata)) return false;
    for (int k = -1; k <= 8; k++)
        if (std::fabs(nodata - (double)k) <= 0.000002) return false;
    return true;
}

}  // namespace

}  // namespace pargo

using pargo::fail;

extern "C" int pargo_d8_classify(pargo_ctx *ctx, const pargo_grid *g, const double *d8, uint8_t *flow_out)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!g || !d8 || !flow_out) return fail(ctx, PARGO_E_INVALID, "pargo_d8_classify: null argument");
    if (g->rows < 1 || g->cols < 1) return fail(ctx, PARGO_E_INVALID, "empty grid (%d x %d)", g->rows, g->cols);
    const long long pitch = g->pitch ? g->pitch : g->cols;
    if (pitch < g->cols) return fail(ctx, PARGO_E_INVALID, "pitch %lld < cols %d", pitch, g->cols);
    if (!pargo::sca_nodata_ok(g->nodata))
        return fail(ctx, PARGO_E_INVALID,
                    "D8 NoData value %.17g is NaN or within 2e-6 of a code in -1..8, which the flow classes "
                    "cannot represent",
                    g->nodata);
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    dim3 grid((g->cols + 255) / 256, g->rows);
    hipLaunchKernelGGL(pargo::d8_classify_kernel, grid, dim3(256), 0, ctx->stream, d8, flow_out, pitch, g->rows,
                       g->cols, g->nodata);
    return pargo::check_launch(ctx, "pargo_d8_classify");
}

extern "C" int pargo_sca_d8(pargo_ctx *ctx, const pargo_grid *g, const uint8_t *flow, const uint32_t *nodata_mask,
                            int mode, double *sca_out)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!g || !flow || !sca_out) return fail(ctx, PARGO_E_INVALID, "pargo_sca_d8: null argument");
    if (g->rows < 1 || g->cols < 1) return fail(ctx, PARGO_E_INVALID, "empty grid (%d x %d)", g->rows, g->cols);
    if (mode != PARGO_SCA_REFERENCE && mode != PARGO_SCA_CONVERGE)
        return fail(ctx, PARGO_E_INVALID, "unknown SCA mode %d", mode);
    const long long pitch = g->pitch ? g->pitch : g->cols;
    if (pitch < g->cols) return fail(ctx, PARGO_E_INVALID, "pitch %lld < cols %d", pitch, g->cols);
    if ((long long)g->rows * g->cols >= (1ll << 40))
        return fail(ctx, PARGO_E_INVALID, "grid of %lld cells exceeds the 2^40 area counter",
                    (long long)g->rows * g->cols);
    if (!pargo::sca_nodata_ok(g->nodata))
        return fail(ctx, PARGO_E_INVALID,
                    "D8 NoData value %.17g is NaN or within 2e-6 of a code in -1..8, which the flow classes "
                    "cannot represent",
                    g->nodata);
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    const bool last_rule = mode == PARGO_SCA_REFERENCE && g->rows >= 3 && g->cols >= 3;
    // The final kernel's read of the slot is guarded per cell, but the
    // compiler issues it as a scalar load that runs whatever the lanes
    // decide, so the slot must be a valid address in every mode.
    rc = pargo::ensure_scratch(ctx, 2 * sizeof(double));
    if (rc) return rc;
    double *slot = static_cast<double *>(ctx->scratch);
    unsigned long long *cursor = reinterpret_cast<unsigned long long *>(slot + 1);
    pargo::ScaParams p{};
    p.flow = flow;
    p.mask = nodata_mask;
    p.w = reinterpret_cast<unsigned long long *>(sca_out);
    p.pitch = pitch;
    p.rows = g->rows;
    p.cols = g->cols;
    p.mask_words = (g->cols + 31) / 32;
    p.converge = mode == PARGO_SCA_CONVERGE;
    dim3 tiles((g->cols + pargo::kTileC - 1) / pargo::kTileC, (g->rows + pargo::kTileR - 1) / pargo::kTileR);
    hipLaunchKernelGGL(pargo::sca_init_kernel, tiles, dim3(256), 0, ctx->stream, p, cursor);
    // persistent walk: 8 workgroups of 4 waves per CU
    const int cus = ctx->num_cu > 0 ? ctx->num_cu : 256;
    hipLaunchKernelGGL(pargo::sca_walk_kernel, dim3(cus * 8), dim3(256), 0, ctx->stream, p, cursor);
    dim3 rowsgrid((g->cols + 255) / 256, g->rows);
    if (last_rule) hipLaunchKernelGGL(pargo::sca_last_cell_kernel, dim3(1), dim3(1), 0, ctx->stream, p, g->nodata, slot);
    hipLaunchKernelGGL(pargo::sca_final_kernel, rowsgrid, dim3(256), 0, ctx->stream, p, g->nodata, slot);
    return pargo::check_launch(ctx, "pargo_sca_d8");
}
