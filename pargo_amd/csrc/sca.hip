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
// A 16-bit static word per cell holds what a walk needs:
//   bits  0-3   npend: inflows the cell waits for
//   bits  4-7   nsrc: inflows from sources (each adds 1)
//   bit   8     an inflow can never be available (border, or behind in
//               the one-sweep order)
//   bit   9     feeds: the cell is one of its target's pending inflows
//   bits 10-13  the cell's code (window index of its target)
//   bit  14     the cell has an area (interior, not NoData)
// The static words sit in the context scratch in 8 x 8 blocks of 128 bytes,
// one cache line each, so that the hops of a walk, which move to a
// neighbour, mostly stay in a line that is already cached.
// The output holds each cell's area as a double from the start when the
// cell waits for nothing (1 + nsrc; NoData; 1 for a cell without a class).
// A cell that waits (npend >= 1) holds an arrival accumulator instead until
// it completes:
//   bits  0-39  sum of the areas of done arrivals
//   bits 40-43  arrivals that were not done
//   bits 44-47  arrivals
//   bit  63     the cell has not completed (a finished area is >= 1)
// Every cell with npend == 0 that feeds starts a walk.  A hop reads the
// target's static word: a target that waits for this walk alone completes
// at once; a confluence (npend >= 2) takes an atomicAdd on its accumulator
// and completes on the last arrival (the returned word holds the others).
// The completing walk writes the target's area and carries on from there.
// Each cell completes once, so the work is linear in the cells; the walk is
// as long as the longest chain of completions.  Waiting cells and
// completions are counted; cells that never complete (a cycle of codes,
// which the converging mode admits for arbitrary codes) leave the counts
// apart, and only then a last pass turns their accumulators into areas.
#include <cmath>

#include "internal.h"
#include "sca_swar.h"

namespace pargo {

namespace {

constexpr unsigned kNone = 15;
constexpr int kTileC = 64, kTileR = 32;  // cells per workgroup: 64 x 32
constexpr uint64_t kAccMask = (1ull << 40) - 1;
constexpr uint64_t kOpen = 1ull << 63;   // accumulator tag: not completed
// The waiting cells are counted by one atomic per init workgroup, spread
// over 64 counters 256 bytes apart: one counter took 131,072 device-scope
// atomics per pass at 16384^2, serialised at one address (about 0.3 ms).
constexpr int kWaitSlots = 64, kWaitStride = 32;  // stride in 8-byte words

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
    unsigned long long *__restrict__ w; // the output buffer: areas (doubles) and accumulators
    uint16_t *__restrict__ st;          // static words, 8 x 8 blocks
    unsigned long long *__restrict__ counters;  // [0] walk cursor, [1] waiting cells, [2] completions
    unsigned *__restrict__ last_done;   // done flags of the 3 x 3 cells ending at (rows-1, cols-1)
    unsigned long long *__restrict__ wait_slots;  // waiting cells, kWaitSlots partial counts kWaitStride apart
    long long pitch;
    int rows, cols, mask_words;
    int bx, by;  // 8 x 8 blocks per block row, block rows
    int converge;
    double nodata;
};

__device__ __forceinline__ unsigned npend_of(unsigned s) { return s & 15u; }
__device__ __forceinline__ unsigned nsrc_of(unsigned s) { return (s >> 4) & 15u; }
__device__ __forceinline__ bool poisoned_of(unsigned s) { return (s >> 8) & 1u; }
__device__ __forceinline__ bool feeds_of(unsigned s) { return (s >> 9) & 1u; }
__device__ __forceinline__ int code_of(unsigned s) { return (int)(s >> 10) & 15; }
__device__ __forceinline__ bool area_of(unsigned s) { return (s >> 14) & 1u; }

__device__ __forceinline__ long long st_index(const ScaParams &p, int r, int c)
{
    return ((long long)(r >> 3) * p.bx + (c >> 3)) * 64 + ((r & 7) << 3) + (c & 7);
}

__device__ __forceinline__ unsigned long long bits_of(double v) { return (unsigned long long)__double_as_longlong(v); }

__device__ __forceinline__ unsigned load_class(const ScaParams &p, int r, int c)
{
    if (r < 0 || r >= p.rows || c < 0 || c >= p.cols) return PARGO_FLOW_NODATA;
    if (p.mask && ((p.mask[(long long)r * p.mask_words + (c >> 5)] >> (c & 31)) & 1u)) return PARGO_FLOW_NODATA;
    return p.flow[(long long)r * p.pitch + c];
}

__device__ __forceinline__ unsigned wave_sum(unsigned v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Counters and last-cell flags to zero, before the init kernel counts.
__global__ void sca_reset_kernel(ScaParams p)
{
    const int i = threadIdx.x;
    if (i < 3) p.counters[i] = 0;
    if (i < 9) p.last_done[i] = 0;
    if (i < kWaitSlots) p.wait_slots[i * kWaitStride] = 0;
}

__device__ __forceinline__ void add_waiting(const ScaParams &p, unsigned n)
{
    if (n) atomicAdd(&p.wait_slots[(blockIdx.x & (kWaitSlots - 1)) * kWaitStride], (unsigned long long)n);
}

// Static words and starting values of a 64 x 32 tile.  Classes of the tile
// plus a 2-cell halo go to LDS, then the inflow mask of the tile plus a
// 1-cell halo (bit j set when the neighbour at window position j < 4 ? j :
// j + 1 drains into the cell and is not excluded; a cell needs to know which
// of its inflows are sources, and whether its target counts it).  Tiles are
// dealt to the 8 XCDs in contiguous runs (workgroup g runs on XCD g % 8), so
// neighbouring tiles share their halo lines in one L2.
__device__ __forceinline__ int win_of_bit(int j) { return j < 4 ? j : j + 1; }
__device__ __forceinline__ int bit_of_win(int k) { return k < 4 ? k : k - 1; }

typedef unsigned long long u64x2_a8 __attribute__((ext_vector_type(2), aligned(8)));

// A tile at least 2 cells from every border whose class rows can be read
// as words, four cells per word (sca_swar.h): the halo columns c0 - 8 ..
// c0 + 71, each word from the aligned pair around it, so bytes c0 - 11 ..
// c0 + 75 of each row are read.
__device__ __forceinline__ bool sca_swar_tile(const ScaParams &p, int r0, int c0)
{
    return r0 >= 2 && r0 + kSwarTileR + 2 <= p.rows && c0 >= 11 && c0 + 76 <= p.cols;
}

// The init pass of such a tile, four cells per lane: the same static words
// and starting values as the per-cell code of sca_init_kernel, with half its
// VALU instructions and a sixth of its LDS accesses.
__device__ __forceinline__ void sca_init_words(const ScaParams &p, SwarTile &T, unsigned *waiting, int r0, int c0)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        constexpr int N = kSwarClsRows * kSwarClsWords, PER = (N + 255) / 256;
        uint32_t lo[PER], hi[PER], mw[PER];
        unsigned sh[PER];
        // every load of the thread is issued before the first is used
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + 256 * k;
            lo[k] = hi[k] = mw[k] = sh[k] = 0;
            if (i < N) {
                const int rr = i / kSwarClsWords, w = i - rr * kSwarClsWords;
                const int r = r0 - 2 + rr, col = c0 - 8 + 4 * w;
                const uintptr_t a = (uintptr_t)(p.flow + (long long)r * p.pitch + col);
                const uint32_t *al = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
                sh[k] = (unsigned)(a & 3);
                lo[k] = al[0];
                hi[k] = al[1];
                if (p.mask) mw[k] = p.mask[(long long)r * p.mask_words + (col >> 5)] >> (col & 31);
            }
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + 256 * k;
            if (i < N) {
                const int rr = i / kSwarClsWords, w = i - rr * kSwarClsWords;
                const uint32_t v = swar_select(swar_bits_to_mask(mw[k]), swar_splat(PARGO_FLOW_NODATA),
                                               __builtin_amdgcn_alignbyte(hi[k], lo[k], sh[k]));
                T.cls[rr][w] = v;
                T.tgt[rr][w] = swar_flow_of(v);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < kSwarInmRows * kSwarInmWords; i += 256) {
        const int lr = i / kSwarInmWords, q = i - lr * kSwarInmWords;
        T.inm[lr][q] = swar_inflow(T, lr, q);
    }
    __syncthreads();
    unsigned nwait = 0;
    // 16 lanes per 8 x 8 block of static words (128 contiguous bytes)
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int v = tid + 256 * k, blk = v >> 4, within = v & 15;
        const int tr = (blk >> 3) * 8 + (within >> 1), g = (blk & 7) * 2 + (within & 1);
        const SwarCells o = swar_cells(T, tr, g, p.converge != 0);
        const int r = r0 + tr, c = c0 + 4 * g;
        *reinterpret_cast<uint2 *>(p.st + st_index(p, r, c)) = make_uint2(o.s01, o.s23);
        unsigned long long val[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const unsigned npend = (o.npend >> (8 * j)) & 0xFFu, nsrc = (o.nsrc >> (8 * j)) & 0xFFu;
            val[j] = ((o.nodata >> (8 * j + 7)) & 1u)  ? bits_of(p.nodata)
                     : ((o.inert >> (8 * j + 7)) & 1u) ? bits_of(1.0)
                     : npend                           ? kOpen
                                                       : bits_of((double)(1u + nsrc));
        }
        // streamed past the caches: the walk reads these words later
        u64x2_a8 *dst = reinterpret_cast<u64x2_a8 *>(p.w + (long long)r * p.pitch + c);
        const u64x2_a8 v01 = {val[0], val[1]}, v23 = {val[2], val[3]};
        __builtin_nontemporal_store(v01, dst);
        __builtin_nontemporal_store(v23, dst + 1);
        nwait += (unsigned)__builtin_popcount(o.waiting);
    }
    nwait = wave_sum(nwait);
    if (lane == 0) waiting[wave] = nwait;
    __syncthreads();
    if (tid == 0) add_waiting(p, waiting[0] + waiting[1] + waiting[2] + waiting[3]);
}

__global__ __launch_bounds__(256) void sca_init_kernel(ScaParams p, int tiles_x, int ntiles)
{
    constexpr int CW = kTileC + 4, CH = kTileR + 4;  // classes, halo 2
    constexpr int NW = kTileC + 2, NH = kTileR + 2;  // inflow masks, halo 1
    __shared__ uint8_t cls[CH][CW];
    __shared__ uint8_t tgt[CH][CW];  // flow_of(class): window index of the target, kNone if none
    __shared__ uint8_t inm[NH][NW];
    __shared__ uint16_t stw[kTileR][kTileC];
    __shared__ unsigned waiting[4];
    const int tid = threadIdx.x;
    const int g = blockIdx.x;
    const int per = (ntiles + 7) / 8;  // tiles per XCD
    const int tile = (g & 7) * per + (g >> 3);
    if (tile >= ntiles) return;  // the whole workgroup
    const int r0 = (tile / tiles_x) * kTileR, c0 = (tile % tiles_x) * kTileC;
    static_assert(kSwarTileR == kTileR && kSwarTileC == kTileC, "one tile shape");
    if (sca_swar_tile(p, r0, c0)) {
        __shared__ SwarTile T;
        sca_init_words(p, T, waiting, r0, c0);
        return;
    }
    {
        // every load of the thread is issued before the first store to LDS,
        // so that their latencies overlap
        constexpr int PER = (CH * CW + 255) / 256;
        unsigned v[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + k * 256;
            v[k] = PARGO_FLOW_NODATA;
            if (i < CH * CW) {
                const int r = r0 - 2 + i / CW, c = c0 - 2 + i % CW;
                if (r >= 0 && r < p.rows && c >= 0 && c < p.cols) {
                    const unsigned b = p.flow[(long long)r * p.pitch + c];
                    const unsigned nd = p.mask ? (p.mask[(long long)r * p.mask_words + (c >> 5)] >> (c & 31)) & 1u : 0u;
                    v[k] = nd ? (unsigned)PARGO_FLOW_NODATA : b;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + k * 256;
            if (i < CH * CW) {
                cls[i / CW][i % CW] = (uint8_t)v[k];
                tgt[i / CW][i % CW] = (uint8_t)flow_of(v[k]);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < NH * NW; i += 256) {
        const int lr = i / NW + 1, lc = i % NW + 1;  // position in cls
        const unsigned b = cls[lr][lc];
        unsigned m = 0;
        if (is_valid(b)) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = win_of_bit(j);
                m |= (tgt[lr + k / 3 - 1][lc + k % 3 - 1] == (unsigned)(8 - k)) ? 1u << j : 0u;
            }
            // the excluded window position (at most one) takes no inflow
            const unsigned ex = excl_of(b);
            if (ex != kNone && ex != 4u) m &= ~(1u << bit_of_win((int)ex));
        }
        inm[lr - 1][lc - 1] = (uint8_t)m;
    }
    __syncthreads();
    unsigned nwait = 0;
    for (int i = tid; i < kTileR * kTileC; i += 256) {
        const int tr = i / kTileC, tc = i % kTileC;
        const int r = r0 + tr, c = c0 + tc;
        unsigned s = 0;
        if (r < p.rows && c < p.cols) {
            const bool interior = r >= 1 && r <= p.rows - 2 && c >= 1 && c <= p.cols - 2;
            const unsigned b = cls[tr + 2][tc + 2];
            unsigned long long value = bits_of(p.nodata);
            if (interior && b == PARGO_FLOW_INERT) {
                s = 1u << 14;
                value = bits_of(1.0);
            } else if (interior && b != PARGO_FLOW_NODATA) {
                const unsigned mine = inm[tr + 1][tc + 1];
                // per inflow bit j: the neighbour is interior (intn), is a
                // source (srcn: nothing drains into it), may be waited for
                // (it precedes the cell in the one-sweep order, or any in
                // the converging mode).  Eight unconditional reads and mask
                // arithmetic, so that no wave loops over its busiest cell.
                unsigned srcn = 0, intn = 0xFFu;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int k = win_of_bit(j);
                    srcn |= (inm[tr + 1 + k / 3 - 1][tc + 1 + k % 3 - 1] == 0 ? 1u : 0u) << j;
                }
                // only a cell next to the border has neighbours outside the interior
                if (r <= 1 || r >= p.rows - 2 || c <= 1 || c >= p.cols - 2) {
                    intn = 0;
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int k = win_of_bit(j);
                        const int ur = r + k / 3 - 1, uc = c + k % 3 - 1;
                        intn |= (ur >= 1 && ur <= p.rows - 2 && uc >= 1 && uc <= p.cols - 2 ? 1u : 0u) << j;
                    }
                }
                const unsigned waitable = p.converge ? 0xFFu : 0x0Fu;  // window bits 0-3: NW, N, NE, W
                const unsigned in_int = mine & intn;
                const unsigned nsrc = __builtin_popcount(in_int & srcn);
                const unsigned npend = __builtin_popcount(in_int & ~srcn & waitable);
                const unsigned poison = ((mine & ~intn) | (in_int & ~srcn & ~waitable)) != 0 ? 1u : 0u;
                const unsigned f = flow_of(b);
                unsigned feeds = 0;
                if (f != kNone && mine != 0 && (p.converge || f > 4)) {
                    // the target counts this cell when its inflow mask has it
                    const int trr = r + (int)f / 3 - 1, tcc = c + (int)f % 3 - 1;
                    const bool t_interior = trr >= 1 && trr <= p.rows - 2 && tcc >= 1 && tcc <= p.cols - 2;
                    const unsigned tm = inm[tr + 1 + (int)f / 3 - 1][tc + 1 + (int)f % 3 - 1];
                    feeds = (t_interior && ((tm >> bit_of_win(8 - (int)f)) & 1u)) ? 1u : 0u;
                }
                s = npend | (nsrc << 4) | (poison << 8) | (feeds << 9) | ((f == kNone ? 0u : f) << 10) | (1u << 14);
                value = npend ? kOpen : bits_of((double)(1u + nsrc));
                nwait += npend ? 1u : 0u;
            }
            p.w[(long long)r * p.pitch + c] = value;
        }
        stw[tr][tc] = (uint16_t)s;
    }
    nwait = wave_sum(nwait);
    if ((tid & 63) == 0) waiting[tid >> 6] = nwait;
    __syncthreads();
    if (tid == 0) {
        const unsigned n = waiting[0] + waiting[1] + waiting[2] + waiting[3];
        add_waiting(p, n);
    }
    // the tile's 4 x 8 blocks, one wave-wide 128-byte line per block
    for (int i = tid; i < kTileR * kTileC; i += 256) {
        const int blk = i >> 6, within = i & 63;
        const int tr = (blk >> 3) * 8 + (within >> 3), tc = (blk & 7) * 8 + (within & 7);
        const int r = r0 + tr, c = c0 + tc;
        if ((r >> 3) < p.by && (c >> 3) < p.bx) p.st[st_index(p, r, c)] = stw[tr][tc];
    }
}

// Walks, persistent: every lane runs one hop per iteration; a lane without
// a walk takes the next cell of its wave's chunk (chunks of kChunk cells in
// block order come from a global cursor) and starts a walk there if the cell
// waits for nothing and feeds its target.  Lanes stay busy until the cursor
// passes the last cell and every walk of the wave has ended.  Checking one
// cell per idle lane and iteration keeps the number of hops in flight, and
// so the queueing delay on the longest chain, moderate: starting every walk
// at once was slower.
//
// The static words do not change during the walk, so their loads may hit
// any cache.  Accumulators are only touched by atomics until their cell
// completes; after that only the completing lane writes the area there.
constexpr long long kChunk = 4096;

__device__ __forceinline__ long long first_lane(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffll));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__global__ __launch_bounds__(256) void sca_walk_kernel(ScaParams p)
{
    const long long nwords = (long long)p.by * p.bx * 64;
    long long chunk_next = 0, chunk_end = 0;  // the same in every lane of a wave
    bool more = true;                         // the same in every lane of a wave
    int cr = -1, cc = 0;                      // cell this lane's walk stands on (cr < 0: none)
    int code = 0;
    unsigned long long val = 0;
    bool done = false;
    unsigned completions = 0;
    for (;;) {
        // the target's static word is requested first, so that its latency
        // overlaps the refill below
        int tr = 0, tc = 0;
        unsigned ts = 0;
        if (cr >= 0) {
            tr = cr + code / 3 - 1;
            tc = cc + code % 3 - 1;
            ts = p.st[st_index(p, tr, tc)];
        }
        const unsigned long long need = __ballot(cr < 0);
        int fr = -1, fc = 0;
        unsigned fs = 0;
        if (need != 0 && more) {
            if (chunk_next >= chunk_end) {
                long long b = 0;
                if (__lane_id() == 0) b = (long long)atomicAdd(&p.counters[0], (unsigned long long)kChunk);
                b = first_lane(b);
                if (b >= nwords) {
                    more = false;
                } else {
                    chunk_next = b;
                    chunk_end = b + kChunk < nwords ? b + kChunk : nwords;
                }
            }
            if (more) {
                const unsigned rank =
                    __builtin_amdgcn_mbcnt_hi((unsigned)(need >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)need, 0u));
                const long long npop = __popcll(need);
                const long long take = npop < chunk_end - chunk_next ? npop : chunk_end - chunk_next;
                if (cr < 0 && (long long)rank < take) {
                    const long long idx = chunk_next + rank;
                    fs = p.st[idx];
                    const long long blk = idx >> 6;
                    const int within = (int)(idx & 63);
                    const int br = (int)(blk / p.bx);
                    fr = br * 8 + (within >> 3);
                    fc = (int)(blk - (long long)br * p.bx) * 8 + (within & 7);
                }
                chunk_next += take;
            }
        }
        if (cr >= 0) {
            const long long t = (long long)tr * p.pitch + tc;
            bool complete = true;
            unsigned long long old = 0;  // other arrivals: none when the walk is the only one awaited
            if (npend_of(ts) != 1) {
                const unsigned long long add =
                    (done ? val : 0ull) | ((unsigned long long)(!done) << 40) | (1ull << 44);
                old = atomicAdd(&p.w[t], add);
                complete = ((unsigned)(old >> 44) & 15u) + 1u == npend_of(ts);
            }
            if (!complete) {
                cr = -1;  // an inflow is still on its way
            } else {
                const unsigned long long acc = (old & kAccMask) + (done ? val : 0ull);
                const unsigned not_done = ((unsigned)(old >> 40) & 15u) + (done ? 0u : 1u);
                done = !poisoned_of(ts) && not_done == 0;
                val = 1ull + nsrc_of(ts) + acc;
                p.w[t] = bits_of((double)val);
                completions++;
                if (tr >= p.rows - 3 && tc >= p.cols - 3)
                    p.last_done[(tr - (p.rows - 3)) * 3 + (tc - (p.cols - 3))] = done ? 1u : 0u;
                if (feeds_of(ts)) {
                    cr = tr;
                    cc = tc;
                    code = code_of(ts);
                } else {
                    cr = -1;
                }
            }
        } else if (fr >= 0 && area_of(fs) && npend_of(fs) == 0 && feeds_of(fs)) {
            cr = fr;
            cc = fc;
            code = code_of(fs);
            val = 1ull + nsrc_of(fs);
            done = !poisoned_of(fs);
        }
        if (!more && __ballot(cr >= 0) == 0) break;
    }
    completions = wave_sum(completions);
    if (__lane_id() == 0 && completions) atomicAdd(&p.counters[2], (unsigned long long)completions);
}

// Cells that never completed (their counts disagree only then): the area of
// the done arrivals they received, as the converging semantics defines it.
__global__ __launch_bounds__(256) void sca_open_kernel(ScaParams p)
{
    unsigned long long waiting = 0;
    for (int i = 0; i < kWaitSlots; i++) waiting += p.wait_slots[i * kWaitStride];
    if (waiting == p.counters[2]) return;
    const long long n = (long long)p.rows * p.cols;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int r = (int)(i / p.cols), c = (int)(i - (long long)r * p.cols);
        const unsigned s = p.st[st_index(p, r, c)];
        if (!area_of(s) || npend_of(s) == 0) continue;
        const long long o = (long long)r * p.pitch + c;
        const unsigned long long w = p.w[o];
        if (w & kOpen) p.w[o] = bits_of((double)(1ull + nsrc_of(s) + (w & kAccMask)));
    }
}

// The reference's last work cell (rows-2, cols-2) always runs the pull
// loop (scaOperator.cpp:86-108).  When it is NoData (or a cell that takes no
// inflow) it adds, in window order, the area of every interior inflowing
// neighbour that is done; all of those precede it in row-major order.
__global__ void sca_last_cell_kernel(ScaParams p)
{
    const int r = p.rows - 2, c = p.cols - 2;
    const unsigned b = load_class(p, r, c);
    if (is_valid(b)) return;
    double v = b == PARGO_FLOW_NODATA ? p.nodata : 1.0;
    for (int k = 0; k < 9; k++) {
        if (k == 4) continue;
        const int ur = r + k / 3 - 1, uc = c + k % 3 - 1;
        if (ur < 1 || ur > p.rows - 2 || uc < 1 || uc > p.cols - 2) continue;
        if (flow_of(load_class(p, ur, uc)) != (unsigned)(8 - k)) continue;
        const unsigned s = p.st[st_index(p, ur, uc)];
        const bool done = npend_of(s) == 0 ? !poisoned_of(s)
                                           : p.last_done[(ur - (p.rows - 3)) * 3 + (uc - (p.cols - 3))] != 0;
        if (done) v += __longlong_as_double((long long)p.w[(long long)ur * p.pitch + uc]);
    }
    p.w[(long long)r * p.pitch + c] = bits_of(v);
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
    const int tiles_x = (g->cols + pargo::kTileC - 1) / pargo::kTileC;
    const long long ntiles = (long long)tiles_x * ((g->rows + pargo::kTileR - 1) / pargo::kTileR);
    if (ntiles > (1ll << 24))
        return fail(ctx, PARGO_E_INVALID, "grid of %d x %d cells needs %lld tiles, more than 2^24", g->rows, g->cols,
                    ntiles);
    // scratch: counters and last-cell flags, the waiting-cell partial
    // counts, then the static words in 8 x 8 blocks (2 bytes per cell, rows
    // and columns rounded up to 8)
    const int bx = (g->cols + 7) / 8, by = (g->rows + 7) / 8;
    const size_t head = 256 + (size_t)pargo::kWaitSlots * pargo::kWaitStride * sizeof(unsigned long long);
    rc = pargo::ensure_scratch(ctx, head + (size_t)bx * by * 64 * sizeof(uint16_t));
    if (rc) return rc;
    char *scr = static_cast<char *>(ctx->scratch);
    pargo::ScaParams p{};
    p.flow = flow;
    p.mask = nodata_mask;
    p.w = reinterpret_cast<unsigned long long *>(sca_out);
    p.st = reinterpret_cast<uint16_t *>(scr + head);
    p.counters = reinterpret_cast<unsigned long long *>(scr);
    p.last_done = reinterpret_cast<unsigned *>(scr + 64);
    p.wait_slots = reinterpret_cast<unsigned long long *>(scr + 256);
    p.pitch = pitch;
    p.rows = g->rows;
    p.cols = g->cols;
    p.mask_words = (g->cols + 31) / 32;
    p.bx = bx;
    p.by = by;
    p.converge = mode == PARGO_SCA_CONVERGE;
    p.nodata = g->nodata;
    hipLaunchKernelGGL(pargo::sca_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, p);
    const int per_xcd = (int)((ntiles + 7) / 8);
    hipLaunchKernelGGL(pargo::sca_init_kernel, dim3(8 * per_xcd), dim3(256), 0, ctx->stream, p, tiles_x, (int)ntiles);
    // persistent walk: 4 workgroups of 4 waves per CU.  Fewer walks in flight
    // queue less on the longest chains; measured at 16384^2 (reference /
    // converging mode): 1 per CU 8.0 / 11.6 ms, 2 5.4 / 8.3, 4 4.4 / 7.6,
    // 8 4.6 / 8.3, 16 4.6 / 8.3
    const int cus = ctx->num_cu > 0 ? ctx->num_cu : 256;
    hipLaunchKernelGGL(pargo::sca_walk_kernel, dim3(cus * 4), dim3(256), 0, ctx->stream, p);
    hipLaunchKernelGGL(pargo::sca_open_kernel, dim3(cus * 4), dim3(256), 0, ctx->stream, p);
    if (last_rule) hipLaunchKernelGGL(pargo::sca_last_cell_kernel, dim3(1), dim3(1), 0, ctx->stream, p);
    return pargo::check_launch(ctx, "pargo_sca_d8");
}
