// idw.hip -- inverse-distance-weighted interpolation with the reference's
// block ring search (1-rank semantics).
//
// Reference behaviour reproduced here:
//   extent          apps/spatial/idw/idwOperator.cpp:126-141 (float cellSize)
//   sample blocks   idwOperator.cpp:152-169, idwOperator.h:68-69 (file order)
//   cell centres    idwOperator.h:91-92
//   ring search     idwOperator.cpp:332-437 (buffer stop before a ring,
//                   strict-< replacement of the first maximum slot, stop
//                   once K are found and (r-1)*B + m >= maxDist)
//   block bound m   idwOperator.cpp:5-20 with the global extent (one rank)
//   weights         idwOperator.cpp:470-480, zero distance -> 1e-7
//   mask            idwOperator.cpp:446-461, gpro/rasterLayer.h:1296-1316
//
// Device layout: samples bucketed by block into CSR order (file order inside
// a block), each bucket padded to a multiple of 4 slots with NaN samples
// (which never pass a distance test), x/y interleaved as double2; values
// and the file index of each bucket slot in separate arrays, the real
// sample count of each block beside the padded starts.
//
// Work split: one lane per output cell, a wave covers an 8x8 cell tile.
// The lanes of a wave almost always share one sample block, so the wave
// walks the ring sequence of one centre block at a time (once in the common
// case, twice or four times for a tile on a block edge or corner), every
// lane with that centre taking part: candidate addresses are wave-uniform,
// so candidates arrive through scalar loads, eight per wait (a bucket's
// last four alone), and each lane only does its distance test.  A lane
// whose own search has ended (ring stop, buffer stop) stays in the walk
// with a threshold no distance passes.  Each lane's K nearest (distance,
// bucket slot) pairs live in LDS.
//
// Exactness: squared distances are formed as RN(RN(dx*dx) + RN(dy*dy)) from
// dx = RN(xs - cx), the operands of the reference's sqrt(pow(dx,2) +
// pow(dy,2)).  A candidate is pre-filtered by an FMA form of its squared
// distance against maxDist^2 * (1 + 2^-40), a necessary condition for
// tmpDist < maxDist; survivors get the reference's squared distance, the
// correctly rounded FP64 sqrt and the reference's exact comparisons, so the
// K-set and its slot order equal the reference's.  A block whose samples'
// bounding box is too far for every lane is skipped without changing either.
#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "internal.h"

struct pargo_idw_plan {
    pargo_ctx *owner = nullptr;
    int n = 0;
    float cs = 0;
    double block = 0;
    double minX = 0, maxX = 0, minY = 0, maxY = 0;
    int rows = 0, cols = 0;
    int blockRows = 0, blockCols = 0;
    void *d_mem = nullptr;
    double2 *d_xy = nullptr;   // bucket order
    double *d_v = nullptr;     // bucket order
    int *d_start = nullptr;    // blockRows*blockCols + 1, padded slot offsets
    int *d_count = nullptr;    // blockRows*blockCols, samples per block
    double4 *d_bbox = nullptr; // blockRows*blockCols, (min x, max x, min y, max y) of the samples
    int *d_perm = nullptr;     // bucket slot -> file index (-1: padding)
    int slots = 0;             // padded slot count
    std::vector<int> h_perm;
    // launch order of the 16 x 16 cell tiles, heaviest estimated or
    // measured load first (pargo_idw_order_tiles); null: row-major order
    int *d_order = nullptr;
    int order_source = 0;
};

// Frees a plan with its device memory (after the work queued on it).
struct PlanDeleter {
    void operator()(pargo_idw_plan *p) const
    {
        if (!p) return;
        if (p->owner) hipStreamSynchronize(p->owner->stream);
        if (p->d_mem) hipFree(p->d_mem);
        if (p->d_order) hipFree(p->d_order);
        delete p;
    }
};

namespace pargo {

// readSamples extent rule.  "_cellSize / 2" and "totalcol * _cellSize" are
// float expressions in the reference and are kept float here.
static void idw_extent(const double *xs, const double *ys, int n, float cs, double &minX, double &maxX,
                       double &minY, double &maxY, int &rows, int &cols)
{
    double mnx = xs[0], mxx = xs[0], mny = ys[0], mxy = ys[0];
    for (int i = 1; i < n; i++) {
        if (xs[i] > mxx) mxx = xs[i];
        if (xs[i] < mnx) mnx = xs[i];
        if (ys[i] > mxy) mxy = ys[i];
        if (ys[i] < mny) mny = ys[i];
    }
    mnx = mnx - (double)(cs / 2);
    int tc = (int)((mxx - mnx) / (double)cs);
    if ((mxx - mnx) != (double)((float)tc * cs)) tc++;
    mxx = mnx + (double)(cs * (float)tc);
    mny = mny - (double)(cs / 2);
    int tr = (int)((mxy - mny) / (double)cs);
    if ((mxy - mny) != (double)((float)tr * cs)) tr++;
    mxy = mny + (double)(cs * (float)tr);
    minX = mnx; maxX = mxx; minY = mny; maxY = mxy;
    rows = tr; cols = tc;
}

struct IdwParams {
    const double2 *__restrict__ xy;
    const double *__restrict__ v;
    const int *__restrict__ start;
    const int *__restrict__ count;
    const double4 *__restrict__ bbox;
    const int32_t *__restrict__ mask;
    float *__restrict__ out;             // may be null (work only)
    double *__restrict__ work_out;       // per cell: samples examined, may be null
    unsigned long long *__restrict__ cand_out;
    const float *__restrict__ grad_out;  // adjoint: dL/d(out), rows x cols
    double *__restrict__ grad_acc;       // adjoint: per bucket slot
    const int *__restrict__ order;       // tile launch order (null: row-major)
    double *__restrict__ tile_work;      // per tile: samples examined (may be null)
    double minX, maxY, block, buffer, cs;
    int rows, cols, blockRows, blockCols, maxRing;
    int k, power;
    int mask_rows, mask_cols, mask_nodata;
    int slots;  // bucket slots (padded sample count)
};

constexpr double kIdwEps = 0.0000001;  // idwOperator.h:13
constexpr double kPrefilterSlack = 1.0 + 0x1p-40;

// A load at a wave-uniform address of data no kernel of this launch writes
// (the sample buckets): through the constant address space the compiler
// can issue it as a scalar load, which the block walk waits on far less
// than on a vector load per block.
template <typename T>
__device__ __forceinline__ T load_uniform(const T *ptr)
{
    typedef const __attribute__((address_space(4))) T constant_t;
    return *(constant_t *)(unsigned long long)ptr;
}

// The reference's distance from its squared form (idwOperator.cpp:398-402):
// the square root, or 1e-7 for a coincident sample.
//
// The square root for s in [2^-767, 2^1000] is the hardware reciprocal
// square root estimate refined by Goldschmidt's iteration and two Newton
// corrections with fused multiply-adds (Markstein's method): the operations
// the compiler issues for an IEEE double sqrt, without the scaling it adds
// for tiny inputs and the selects for zero and infinity, which this range
// excludes (the same result bits).  Other s take sqrt().
__device__ __forceinline__ double idw_sqrt(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    // a wave-uniform test first, and the argument made opaque inside it
    // (an empty asm), so that the compiler cannot compute the full sqrt
    // for every lane alongside and select
    const bool in_range = x >= 0x1p-767 && x <= 0x1p+1000;
    if (__builtin_amdgcn_ballot_w64(!in_range) != 0) {
        double xo = x;
        asm volatile("" : "+v"(xo));
        if (!in_range) g = sqrt(xo);
    }
    return g;
}
__device__ __forceinline__ double idw_dist(double s) { return s == 0.0 ? kIdwEps : idw_sqrt(s); }

// rowAtOtherLayer: (int)min(scale * idx, other - 1) with double scale.
__device__ __forceinline__ int mask_index(int other, int self, int idx)
{
    double scale = (double)other / (double)self;
    double v = scale * (double)idx;
    double lim = (double)(other - 1);
    return (int)((lim < v) ? lim : v);
}

// 1/x within an ulp for the weights: the hardware reciprocal estimate and
// two Newton steps, which need no range scaling for x in [2^-1000, 2^1000];
// other x (and 0, Inf, NaN) take the IEEE division.
__device__ __forceinline__ double idw_recip(double x)
{
    if (!(x >= 0x1p-1000 && x <= 0x1p1000)) return 1.0 / x;
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double powi(double d, int p)
{
    // pow(d, p) for integer p; the result only feeds weights (1e-5 bar).
    if (p == 2) return d * d;
    if (p == 1) return d;
    double r = 1.0, b = d;
    unsigned e = p < 0 ? (unsigned)(-p) : (unsigned)p;
    while (e) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return p < 0 ? 1.0 / r : r;
}

// Eight candidates (128 bytes at a wave-uniform address) as two scalar
// loads in flight together and one wait for both.
typedef unsigned int u32x16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ void load8_uniform(const double2 *src, double2 c[8])
{
    // the address is wave-uniform; say so, so that it sits in SGPRs
    const unsigned long long a64 = (unsigned long long)src;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a64);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a64 >> 32));
    const unsigned long long base = ((unsigned long long)hi << 32) | lo;
    u32x16 a, b;
    asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx16 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
                 : "=&s"(a), "=&s"(b)  // early clobber: the first load must not land on the address
                 : "s"(base)
                 : "memory");
#pragma unroll
    for (int i = 0; i < 4; i++) {
        c[i].x = __hiloint2double((int)a[4 * i + 1], (int)a[4 * i]);
        c[i].y = __hiloint2double((int)a[4 * i + 3], (int)a[4 * i + 2]);
        c[i + 4].x = __hiloint2double((int)b[4 * i + 1], (int)b[4 * i]);
        c[i + 4].y = __hiloint2double((int)b[4 * i + 3], (int)b[4 * i + 2]);
    }
}

// The K-nearest search of the lanes whose cell lies in sample block
// (blockRow, blockCol) (wave-uniform) and that have `mine` set.  Each lane's
// list of (distance, bucket slot) pairs lives in LDS at kd[i * NT] /
// ks[i * NT], i < K (the caller offsets the pointers by the lane's thread
// index): the list is written on an insertion and read on a replacement's
// rescan and at the end, while the test of a candidate needs only the
// running threshold in registers.  Fills the list like nbrSamples and
// returns min(found, K).  Counts the samples of the visited blocks in
// *ncand.
template <int NT, typename SlotT>
__device__ __forceinline__ int idw_search_group(const IdwParams &p, int blockRow, int blockCol, double cellX,
                                                double cellY, bool mine, double *kd, SlotT *ks, unsigned &ncand)
{
    const int K = p.k;
    // getMinDistanceToBlockBound (idwOperator.cpp:5-20), one rank
    const double bx0 = blockCol * p.block + p.minX, bx1 = (blockCol + 1) * p.block + p.minX;
    const double by0 = p.maxY - blockRow * p.block, by1 = p.maxY - (blockRow + 1) * p.block;
    const double d0 = fabs(cellX - bx0), d1 = fabs(cellX - bx1), d2 = fabs(cellY - by0), d3 = fabs(cellY - by1);
    const double a = (d1 < d0) ? d1 : d0, b = (d3 < d2) ? d3 : d2;
    const double minDistToBound = (b < a) ? b : a;

    double maxDist = 0.0;
    int maxIdx = -1;
    int tail = -1;
    bool act = mine;
    // prefilter on squared distance: everything passes while the list fills,
    // nothing once this lane's search has ended
    double kthr = act ? INFINITY : -INFINITY;
    // The prefilter's squared distance takes one FMA: it is within 2^-50
    // (relative) of the reference's RN(RN(dx^2) + RN(dy^2)), far inside the
    // 2^-40 slack of kthr, so it still passes every candidate the exact
    // test can take; that test recomputes the reference's value.
    auto sqdist = [&](double2 xy) {
        const double dx = xy.x - cellX, dy = xy.y - cellY;
        return __builtin_fma(dx, dx, dy * dy);
    };
    auto visit = [&](double2 xy, double pre, int q) {
        if (pre <= kthr) {
            const double dx = xy.x - cellX, dy = xy.y - cellY;
            const double s = dx * dx + dy * dy;
            const double tmpDist = idw_dist(s);
            if (tail < K - 1) {
                tail++;
                kd[tail * NT] = tmpDist;
                ks[tail * NT] = q;
                if (tmpDist > maxDist) {
                    maxDist = tmpDist;
                    maxIdx = tail;
                }
                if (tail == K - 1) kthr = maxDist * maxDist * kPrefilterSlack;
            } else if (tmpDist < maxDist) {
                // (the reference also counts tailIdx up here; past K-1 only
                // "tail >= K-1" is ever used)
                kd[maxIdx * NT] = tmpDist;
                ks[maxIdx * NT] = q;
                maxDist = kd[0];
                maxIdx = 0;
                for (int i = 1; i < K; i++) {
                    const double di = kd[i * NT];
                    if (di > maxDist) {
                        maxDist = di;
                        maxIdx = i;
                    }
                }
                kthr = maxDist * maxDist * kPrefilterSlack;
            }
        }
    };
    for (int rad = 0;; rad++) {
        const double searchRange = (double)(rad - 1) * p.block + minDistToBound;
        if (act && p.buffer > 0 && searchRange >= p.buffer) {
            act = false;
            kthr = -INFINITY;
        }
        if (rad > p.maxRing) break;  // every further ring is outside the block grid
        if (__builtin_amdgcn_ballot_w64(act) == 0) break;
        for (int tRow = blockRow - rad; tRow <= blockRow + rad; ++tRow) {
            if (tRow < 0 || tRow >= p.blockRows) continue;
            const bool full = (tRow == blockRow - rad) || (tRow == blockRow + rad);
            // full top/bottom rows, only the two end columns in between
            const int step = full ? 1 : 2 * rad;
            for (int tCol = blockCol - rad; tCol <= blockCol + rad; tCol += step) {
                if (tCol < 0 || tCol >= p.blockCols) continue;
                const int bi = tRow * p.blockCols + tCol;
                const int s0 = p.start[bi], s1 = p.start[bi + 1];
                if (act) ncand += (unsigned)p.count[bi];
                if (s0 == s1) continue;
                // Skip the block when no lane can take any of its samples:
                // RN is monotone, so RN(RN(dxb^2) + RN(dyb^2)) from the
                // distances to the samples' bounding box is a lower bound of
                // every sample's squared distance as computed below, and a
                // sample above kthr is never taken (kthr is +inf while a
                // lane's list fills and -inf once its search has ended).
                {
                    const double4 bb = p.bbox[bi];
                    const double dxb = fmax(fmax(bb.x - cellX, cellX - bb.y), 0.0);
                    const double dyb = fmax(fmax(bb.z - cellY, cellY - bb.w), 0.0);
                    const double lb2 = dxb * dxb + dyb * dyb;
                    if (__builtin_amdgcn_ballot_w64(lb2 <= kthr) == 0) continue;
                }
                // Eight candidates per scalar-load wait (scalar loads return
                // out of order, so every wait is for all of them), then the
                // bucket's last four if its padded length is not a multiple
                // of eight.  Usually none passes once the list is full;
                // otherwise they go in order (each may lower kthr).
                int q = s0;
                for (; q + 8 <= s1; q += 8) {
                    double2 c[8];
                    load8_uniform(p.xy + q, c);
                    double e[8];
#pragma unroll
                    for (int i = 0; i < 8; i++) e[i] = sqdist(c[i]);
                    bool any = false;
#pragma unroll
                    for (int i = 0; i < 8; i++) any |= e[i] <= kthr;
                    if (__builtin_amdgcn_ballot_w64(any) != 0) {
#pragma unroll
                        for (int i = 0; i < 8; i++) visit(c[i], e[i], q + i);
                    }
                }
                if (q < s1) {
                    double2 c[4];
                    double e[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        c[i] = p.xy[q + i];
                        e[i] = sqdist(c[i]);
                    }
                    bool any = false;
#pragma unroll
                    for (int i = 0; i < 4; i++) any |= e[i] <= kthr;
                    if (__builtin_amdgcn_ballot_w64(any) != 0) {
#pragma unroll
                        for (int i = 0; i < 4; i++) visit(c[i], e[i], q + i);
                    }
                }
            }
        }
        if (act && tail >= K - 1 && searchRange >= maxDist) {
            act = false;
            kthr = -INFINITY;
        }
    }
    return (tail + 1 < K) ? tail + 1 : K;
}

// v of lane l (l wave-uniform) in every lane: two v_readlane_b32 into
// scalar registers, where a shuffle would take the LDS crossbar twice
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The same search with each ring's candidates visited nearest-first, for
// the lanes with `mine` set.  The reference's result is the K smallest
// distances among the samples of the visited rings, whatever order they are
// visited in, except when a sample outside the list ties the list's largest
// distance exactly: then its visiting order picks the member.  So the
// visited rings are the reference's (same stop tests, which depend only on
// the K-th distance and the sample count), each ring's candidates are
// staged across the wave (candidate i of the ring, blocks in ring order and
// slots in file order, in lane i % 64, register i / 64) and visited in bins
// of increasing distance from the group's cells, and the smallest distance
// seen outside the list is kept: `tie` reports a lane whose list depends on
// the order, which the caller then searches again in the reference's order.
// A bin's candidates are visited only while some lane's threshold reaches
// the bin's inner edge, so far candidates are never loaded into a test.
constexpr int kStage = 4;     // staged candidates per lane: 256 per wave
constexpr int kBins = 48;     // bins of block / 16 each, the last one open
constexpr int kBinBits = 6;   // a staged entry packs (slot << 6) | bin

template <int MAXK, int NT, typename SlotT>
__device__ __forceinline__ int idw_search_binned(const IdwParams &p, int blockRow, int blockCol, double cellX,
                                                 double cellY, bool mine, double *kd, SlotT *ks, unsigned &ncand,
                                                 bool &tie)
{
    const int K = p.k;
    const int lane = threadIdx.x & 63;
    const double bx0 = blockCol * p.block + p.minX, bx1 = (blockCol + 1) * p.block + p.minX;
    const double by0 = p.maxY - blockRow * p.block, by1 = p.maxY - (blockRow + 1) * p.block;
    const double d0 = fabs(cellX - bx0), d1 = fabs(cellX - bx1), d2 = fabs(cellY - by0), d3 = fabs(cellY - by1);
    const double a = (d1 < d0) ? d1 : d0, b = (d3 < d2) ? d3 : d2;
    const double minDistToBound = (b < a) ? b : a;

    double maxDist = 0.0;
    int maxIdx = -1;
    int tail = -1;
    bool act = mine;
    double kthr = act ? INFINITY : -INFINITY;
    double runner = INFINITY;  // smallest distance visited outside the list
    // bounding box of the group's cells (every lane's cell lies inside it)
    double gx0 = mine ? cellX : INFINITY, gx1 = mine ? cellX : -INFINITY;
    double gy0 = mine ? cellY : INFINITY, gy1 = mine ? cellY : -INFINITY;
    for (int o = 32; o >= 1; o >>= 1) {
        gx0 = fmin(gx0, __shfl_xor(gx0, o, 64));
        gx1 = fmax(gx1, __shfl_xor(gx1, o, 64));
        gy0 = fmin(gy0, __shfl_xor(gy0, o, 64));
        gy1 = fmax(gy1, __shfl_xor(gy1, o, 64));
    }
    const float inv_w = fminf((float)(16.0 / p.block), 3.4e38f);  // finite: no 0 * inf in the bins
    const double w = p.block / 16.0;

    auto visit = [&](double2 xy, double pre, int q) {
        if (pre <= kthr) {
            const double dx = xy.x - cellX, dy = xy.y - cellY;
            const double s = dx * dx + dy * dy;
            const double tmpDist = idw_dist(s);
            if (tail < K - 1) {
                tail++;
                kd[tail * NT] = tmpDist;
                ks[tail * NT] = q;
                if (tmpDist > maxDist) {
                    maxDist = tmpDist;
                    maxIdx = tail;
                }
                if (tail == K - 1) kthr = maxDist * maxDist * kPrefilterSlack;
            } else if (tmpDist < maxDist) {
                runner = fmin(runner, maxDist);  // the evicted distance
                kd[maxIdx * NT] = tmpDist;
                ks[maxIdx * NT] = q;
                // the new largest distance and its first slot, as the
                // reference's strict-> rescan finds them; every entry is
                // read before the first compare (one LDS wait, not K).
                // Slots K .. MAXK - 1 hold 0, below every distance.
                double di[MAXK];
#pragma unroll
                for (int i = 0; i < MAXK; i++) di[i] = kd[i * NT];
                double mx = di[0];
#pragma unroll
                for (int i = 1; i < MAXK; i++) mx = fmax(mx, di[i]);
                int mi = 0;
#pragma unroll
                for (int i = MAXK - 1; i >= 0; i--)
                    if (di[i] == mx) mi = i;
                maxDist = mx;
                maxIdx = mi;
                kthr = maxDist * maxDist * kPrefilterSlack;
            } else {
                runner = fmin(runner, tmpDist);
            }
        }
    };

    // the staging registers and their fill level (wave-uniform)
    double sx[kStage], sy[kStage];
    int sqb[kStage];  // the bucket slot while staging, then (slot << kBinBits) | bin
    int staged = 0;
    // visits the staged candidates bin by bin, nearest bins first; stops at
    // the first bin whose inner edge no lane's threshold reaches.  A
    // candidate's bin b satisfies sqrt(lb2) >= b w (1 - 3e-7) (three float
    // roundings and the hardware square root's 1 ulp: at most 2.25 ulp),
    // lb2 <= its squared distance as the test computes it (1 + 2^-51), so
    // with the margins below no skipped candidate passes.
    // positions and bins of the staged slots: every lane's loads are
    // requested before the first is used, one wait per flush rather than one
    // per block
    auto prepare = [&]() {
        double2 xy[kStage];
#pragma unroll
        for (int k = 0; k < kStage; k++)
            if (64 * k < staged) xy[k] = 64 * k + lane < staged ? p.xy[sqb[k]] : make_double2(NAN, NAN);
#pragma unroll
        for (int k = 0; k < kStage; k++) {
            if (64 * k >= staged) continue;
            const double ex = fmax(fmax(gx0 - xy[k].x, xy[k].x - gx1), 0.0);
            const double ey = fmax(fmax(gy0 - xy[k].y, xy[k].y - gy1), 0.0);
            const float r = __builtin_amdgcn_sqrtf((float)(ex * ex + ey * ey)) * inv_w;
            sx[k] = xy[k].x;
            sy[k] = xy[k].y;
            sqb[k] = (sqb[k] << kBinBits) | (r < (float)(kBins - 1) ? (int)r : kBins - 1);  // NaN: last bin
        }
    };
    auto flush = [&]() {
        prepare();
        for (int bn = 0; bn < kBins; bn++) {
            const double edge = (double)bn * w * (1.0 - 1e-5);
            if (__builtin_amdgcn_ballot_w64(edge * edge * (1.0 - 0x1p-40) <= kthr) == 0) break;
#pragma unroll
            for (int k = 0; k < kStage; k++) {
                if (64 * k >= staged) break;
                unsigned long long m =
                    __builtin_amdgcn_ballot_w64(64 * k + lane < staged && (sqb[k] & ((1 << kBinBits) - 1)) == bn);
                while (m) {
                    const int l = (int)__builtin_ctzll(m);
                    m &= m - 1;
                    const double2 xy = make_double2(lane_value(sx[k], l), lane_value(sy[k], l));
                    const int q = __builtin_amdgcn_readlane(sqb[k], l) >> kBinBits;
                    const double dx = xy.x - cellX, dy = xy.y - cellY;
                    visit(xy, __builtin_fma(dx, dx, dy * dy), q);
                }
            }
        }
        staged = 0;
    };

    for (int rad = 0;; rad++) {
        const double searchRange = (double)(rad - 1) * p.block + minDistToBound;
        if (act && p.buffer > 0 && searchRange >= p.buffer) {
            act = false;
            kthr = -INFINITY;
        }
        if (rad > p.maxRing) break;
        if (__builtin_amdgcn_ballot_w64(act) == 0) break;
        for (int tRow = blockRow - rad; tRow <= blockRow + rad; ++tRow) {
            if (tRow < 0 || tRow >= p.blockRows) continue;
            const bool full = (tRow == blockRow - rad) || (tRow == blockRow + rad);
            const int step = full ? 1 : 2 * rad;
            for (int tCol = blockCol - rad; tCol <= blockCol + rad; tCol += step) {
                if (tCol < 0 || tCol >= p.blockCols) continue;
                const int bi = tRow * p.blockCols + tCol;
                // the block's start, end, count and bounding box are all
                // requested before any is used (one wait per block)
                const int s0 = load_uniform(p.start + bi), s1 = load_uniform(p.start + bi + 1);
                const int cnt = load_uniform(p.count + bi);
                const double *bbp = reinterpret_cast<const double *>(p.bbox + bi);
                const double4 bb = make_double4(load_uniform(bbp), load_uniform(bbp + 1), load_uniform(bbp + 2),
                                                load_uniform(bbp + 3));
                ncand += act ? (unsigned)cnt : 0u;
                const double dxb = fmax(fmax(bb.x - cellX, cellX - bb.y), 0.0);
                const double dyb = fmax(fmax(bb.z - cellY, cellY - bb.w), 0.0);
                const double lb2 = dxb * dxb + dyb * dyb;
                const unsigned long long reach = __builtin_amdgcn_ballot_w64(lb2 <= kthr);
                if (s0 == s1 || reach == 0) continue;
                // stage the block's slots (padding slots hold NaN positions,
                // which no test passes)
                for (int q0 = s0; q0 < s1;) {
                    const int take = min(s1 - q0, 64 * kStage - staged);
#pragma unroll
                    for (int k = 0; k < kStage; k++) {
                        // registers the block's slots do not reach (wave-uniform)
                        if (64 * k + 64 <= staged || 64 * k >= staged + take) continue;
                        const int i = 64 * k + lane;
                        // only the slot is recorded here; positions and bins
                        // follow for all staged slots at once (prepare)
                        if (i >= staged && i < staged + take) sqb[k] = q0 + (i - staged);
                    }
                    staged += take;
                    q0 += take;
                    if (staged == 64 * kStage) flush();
                }
            }
        }
        // Ring 0 never ends a search (its range is negative), so its
        // candidates may wait for ring 1's flush when no lane would leave
        // before ring 1 (the buffer rule at range minDistToBound) and a ring
        // 1 exists: the visited set is the same, one position-load wait
        // fewer.
        const bool defer = rad == 0 && p.maxRing >= 1 &&
                           __builtin_amdgcn_ballot_w64(act && p.buffer > 0 && minDistToBound >= p.buffer) == 0;
        if (staged && !defer) flush();
        if (act && tail >= K - 1 && searchRange >= maxDist) {
            act = false;
            kthr = -INFINITY;
        }
    }
    if (staged) flush();  // never reached with candidates left: ring 1 flushes
    tie = mine && tail >= K - 1 && runner == maxDist;
    return (tail + 1 < K) ? tail + 1 : K;
}

// ADJOINT: instead of writing the grid, add g(cell) * w_i / W to the
// gradient of each selected sample (bucket order).  The cells of a workgroup
// select few distinct samples, so their contributions are first summed in
// an LDS hash table (open addressing on the bucket slot) and flushed with
// one global FP64 atomic per distinct sample; a crowded table falls back to
// global atomics.
// 64 entries (768 bytes) keep the table and the K-lists of a one-wave
// workgroup within 10 KiB of LDS (K = 12), four waves per SIMD; an 8 x 8
// quarter tile selects a few dozen distinct samples at most.
constexpr int kAdjTableBits = 6;
constexpr int kAdjTable = 1 << kAdjTableBits;
constexpr int kAdjProbe = 16;

// One wave per workgroup (an 8 x 8 quarter of a 16 x 16 tile, four
// consecutive workgroups per tile), so that a wave's registers and LDS are
// freed when it ends rather than when the slowest wave of its tile does.
// kIdwLanes: lanes per workgroup, the stride of the K-lists in LDS.
constexpr int kIdwLanes = 64;

// SlotT: the type of the K-lists' bucket slots in LDS (16 bits when every
// slot index fits).
template <int MAXK, bool ADJOINT, typename SlotT>
__device__ __forceinline__ void idw_body(const IdwParams &p)
{
    constexpr int NT = kIdwLanes;
    __shared__ int tkey[ADJOINT ? kAdjTable : 1];
    __shared__ double tval[ADJOINT ? kAdjTable : 1];
    __shared__ double kdist[MAXK * NT];  // [slot i][thread]
    __shared__ SlotT kslot[MAXK * NT];
    if constexpr (ADJOINT) {
        for (int i = threadIdx.x; i < kAdjTable; i += blockDim.x) {
            tkey[i] = -1;
            tval[i] = 0.0;
        }
        __syncthreads();
    }
    auto adj_add = [&](int q, double v) {
        unsigned h = ((unsigned)q * 2654435761u) >> (32 - kAdjTableBits);
        for (int t = 0; t < kAdjProbe; t++, h = (h + 1) & (kAdjTable - 1)) {
            const int k = atomicCAS(&tkey[h], -1, q);
            if (k == -1 || k == q) {
                unsafeAtomicAdd(&tval[h], v);
                return;
            }
        }
        atomicAdd(&p.grad_acc[q], v);
    };
    // list slots beyond K: 0, so that the unrolled rescan of a full list
    // never picks one (every distance is at least 1e-7)
    for (int i = p.k; i < MAXK; i++) kdist[i * NT + threadIdx.x] = 0.0;
    const int lane = threadIdx.x;
    const int wave = (int)(blockIdx.x & 3);  // the quarter of the tile
    // 16 x 16 tiles, launched in row-major tile order or in the order of a
    // load model (heaviest first)
    const unsigned tslot = blockIdx.x >> 2;
    const int tiles_x = (p.cols + 15) >> 4;
    const int tile = p.order ? p.order[tslot] : (int)tslot;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int row = ty * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int col = tx * 16 + (wave & 1) * 8 + (lane & 7);
    const bool inside = row < p.rows && col < p.cols;
    unsigned ncand = 0;
    float result = -9999.0f;
    bool pending = inside;
    int num = 0;  // this lane's list length once its group is searched
    if (inside && p.mask) {
        const int mr = mask_index(p.mask_rows, p.rows, row);
        const int mc = mask_index(p.mask_cols, p.cols, col);
        const int m = p.mask[(long long)mr * p.mask_cols + mc];
        pending = !((m == 0) | (m == p.mask_nodata));
    }
    const double cellX = (double)((col + 0.5) * (double)p.cs + p.minX);
    const double cellY = (double)(p.maxY - (row + 0.5) * (double)p.cs);
    const int blockRow = (int)((p.maxY - cellY) / p.block);
    const int blockCol = (int)((cellX - p.minX) / p.block);
    // one search per distinct centre block among the pending lanes; each
    // lane's K-list stays in LDS until the weights below
    const bool searched = pending;
    for (;;) {
        const unsigned long long m = __builtin_amdgcn_ballot_w64(pending);
        if (m == 0) break;
        const int leader = (int)__builtin_ctzll(m);
        const int gRow = __builtin_amdgcn_readlane(blockRow, leader);
        const int gCol = __builtin_amdgcn_readlane(blockCol, leader);
        const bool mine = pending && blockRow == gRow && blockCol == gCol;
        double *kd = kdist + threadIdx.x;
        SlotT *ks = kslot + threadIdx.x;
        bool tie = false;
        const int n1 = idw_search_binned<MAXK, NT, SlotT>(p, gRow, gCol, cellX, cellY, mine, kd, ks, ncand, tie);
        if (mine) num = n1;
        // lanes whose list depends on the visiting order (an exact distance
        // tie at the K-th place) search again in the reference's order
        if (__builtin_amdgcn_ballot_w64(tie) != 0) {
            unsigned again = 0;
            const int num2 = idw_search_group<NT, SlotT>(p, gRow, gCol, cellX, cellY, tie, kd, ks, again);
            if (tie) num = num2;
        }
        pending = pending && !mine;
    }
    // weights and output once for the whole wave (a wave whose cells span
    // several centre blocks would otherwise run them once per group)
    {
        double *kd = kdist + threadIdx.x;
        const SlotT *ks = kslot + threadIdx.x;
        if (searched) {
            // weights 1 / d^p summed in slot order (idwOperator.cpp:472-480).
            // The reference forms each weight twice and divides every term
            // by the sum; here each weight is formed once (kept in the
            // list's distance slot, 1/d^p within an ulp) and the weighted
            // sum divided once, which moves the result by a few FP64 ulps,
            // far inside the float output's rounding.
            // unrolled to the list capacity and masked by num, so that the
            // LDS reads and divisions of all members overlap
            double wsum = 0.0;
#pragma unroll
            for (int i = 0; i < MAXK; i++) {
                if (i < num) {
                    const double wi = idw_recip(powi(kd[i * NT], p.power));
                    kd[i * NT] = wi;
                    wsum += wi;
                }
            }
            if (ADJOINT) {
                const double gc = (double)p.grad_out[(long long)row * p.cols + col];
                if (gc != 0.0 && num > 0) {
                    const double scale = gc / wsum;
                    // Neighbouring cells share most members, mostly in the
                    // same list slots, and LDS atomics of one instruction
                    // on one address are serialized: each lane starts at
                    // slot lane mod num, so one instruction's lanes spread
                    // over the slots.
                    int i = lane % num;
                    for (int t = 0; t < num; t++) {
                        adj_add(ks[i * NT], kd[i * NT] * scale);
                        i = i + 1 == num ? 0 : i + 1;
                    }
                }
            } else {
                // every member's value is requested before the first is used
                // (a loop to num would wait for each dependent load in turn)
                double vals[MAXK];
#pragma unroll
                for (int i = 0; i < MAXK; i++) vals[i] = i < num ? p.v[ks[i * NT]] : 0.0;
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i < MAXK; i++)
                    if (i < num) acc += vals[i] * kd[i * NT];
                result = num > 0 ? (float)(acc / wsum) : 0.0f;  // no sample in range: the empty sum, 0
            }
        }
    }
    if (!ADJOINT && inside && p.out) p.out[(long long)row * p.cols + col] = result;
    if (!ADJOINT && inside && p.work_out) p.work_out[(long long)row * p.cols + col] = (double)ncand;
    if constexpr (ADJOINT) {
        __syncthreads();
        for (int i = threadIdx.x; i < kAdjTable; i += blockDim.x) {
            const int q = tkey[i];
            if (q >= 0) atomicAdd(&p.grad_acc[q], tval[i]);
        }
    }
    if (p.cand_out || p.tile_work) {
        unsigned long long c = ncand;
        for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0 && c && p.cand_out) atomicAdd(p.cand_out, c);
        if (lane == 0 && c && p.tile_work) atomicAdd(&p.tile_work[tile], (double)c);
    }
}

template <int MAXK, bool ADJOINT>
__global__ __launch_bounds__(kIdwLanes) void idw_kernel(IdwParams p)
{
    idw_body<MAXK, ADJOINT, int>(p);
}

// The forward with 16-bit list slots (every bucket slot below 65536):
// 7.5 KiB of LDS per wave at K = 12, so five waves per SIMD fit, with 96
// VGPRs and a few spilled (1.31 -> 1.25 ms on 20,000 samples onto 4096^2;
// the adjoint, whose hash table needs more LDS, gained nothing).
template <int MAXK>
__global__ __launch_bounds__(kIdwLanes) __attribute__((amdgpu_waves_per_eu(5, 5))) void idw_kernel_s16(IdwParams p)
{
    idw_body<MAXK, false, unsigned short>(p);
}

template <int MAXK, bool ADJOINT>
static void launch_idw_k(const IdwParams &p, hipStream_t s)
{
    dim3 block(kIdwLanes);
    dim3 grid((unsigned)(((p.cols + 15) / 16) * ((p.rows + 15) / 16)) * 4u);
    if constexpr (!ADJOINT && MAXK <= 12) {
        if (p.slots <= 65536) {
            hipLaunchKernelGGL((idw_kernel_s16<MAXK>), grid, block, 0, s, p);
            return;
        }
    }
    hipLaunchKernelGGL((idw_kernel<MAXK, ADJOINT>), grid, block, 0, s, p);
}

template <bool ADJOINT>
static int launch_idw(const IdwParams &p, hipStream_t s)
{
    if (p.k <= 4) launch_idw_k<4, ADJOINT>(p, s);
    else if (p.k <= 8) launch_idw_k<8, ADJOINT>(p, s);
    else if (p.k <= 12) launch_idw_k<12, ADJOINT>(p, s);
    else if (p.k <= 16) launch_idw_k<16, ADJOINT>(p, s);
    else if (p.k <= 32) launch_idw_k<32, ADJOINT>(p, s);
    else return PARGO_E_UNSUPPORTED;
    return PARGO_OK;
}

// grad_vals[file index] = grad_acc[bucket slot] (padding slots skipped)
__global__ void idw_unbucket_kernel(const double *__restrict__ acc, const int *__restrict__ perm, int slots,
                                    double *__restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < slots && perm[q] >= 0) out[perm[q]] = acc[q];
}

// v[slot] = vals[perm[slot]] (padding slots 0): new sample values in
// bucket order, from a device array in file order
__global__ void idw_bucket_values_kernel(const double *__restrict__ vals, const int *__restrict__ perm, int slots,
                                         double *__restrict__ v)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < slots) v[q] = perm[q] >= 0 ? vals[perm[q]] : 0.0;
}

// Compute-load estimate of one coarse cell (IdwTransformation::Operator,
// idwTransformation.cpp:3-62): the sample count of the cell's block, then
// rings r = 0, 1, ... of which only the full top and bottom block rows are
// visited (the centre block again at r = 0), until K samples are counted.
// load = blocks visited + 2 * samples counted.
struct IdwLoadParams {
    const int *__restrict__ count;
    const int32_t *__restrict__ mask;
    double *__restrict__ out;
    unsigned *__restrict__ nonterm;
    double block, ccs, mask_nodata, nodata;
    int rows, cols, blockRows, blockCols, k, maxRing;
    int mask_rows, mask_cols;
};

__global__ __launch_bounds__(256) void idw_load_kernel(IdwLoadParams p)
{
    const int cCol = blockIdx.x * 16 + (threadIdx.x & 15);
    const int cRow = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (cRow >= p.rows || cCol >= p.cols) return;
    double *o = p.out + (long long)cRow * p.cols + cCol;
    if (p.mask) {
        const int mr = mask_index(p.mask_rows, p.rows, cRow);
        const int mc = mask_index(p.mask_cols, p.cols, cCol);
        const int m = p.mask[(long long)mr * p.mask_cols + mc];
        if (m == 0 || (double)m == p.mask_nodata) {
            *o = p.nodata;
            return;
        }
    }
    // getBlockRowIndexByCellIndex(iRow, cellSize) (idwOperator.h:70-71)
    const int bR = min((int)((cRow + 0.5) * p.ccs / p.block), p.blockRows - 1);
    const int bC = min((int)((cCol + 0.5) * p.ccs / p.block), p.blockCols - 1);
    auto count = [&](int r, int c) { return p.count[r * p.blockCols + c]; };
    long long nbr = count(bR, bC);
    long long blocks = 0;
    bool enough = false;
    for (int rad = 0; nbr < p.k; rad++) {
        if (rad > p.maxRing) {  // no block is left to add: the reference never returns
            atomicAdd(p.nonterm, 1u);
            *o = p.nodata;
            return;
        }
        for (int tRow = bR - rad; tRow <= bR + rad && !enough; ++tRow) {
            if (tRow < 0 || tRow >= p.blockRows) continue;
            if (tRow != bR - rad && tRow != bR + rad) continue;
            for (int tCol = bC - rad; tCol <= bC + rad && !enough; ++tCol) {
                if (tCol < 0 || tCol >= p.blockCols) continue;
                nbr += count(tRow, tCol);
                blocks++;
                if (nbr >= p.k) enough = true;
            }
        }
        if (enough) break;
    }
    *o = (double)blocks + (double)(nbr * 2);
}

int idw_fill_params(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                    const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, IdwParams &p)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!plan || plan->owner != ctx) return fail(ctx, PARGO_E_INVALID, "IDW plan is null or from another context");
    if (k < 1) return fail(ctx, PARGO_E_INVALID, "searchPointNum must be >= 1, got %d", k);
    if (k > 32) return fail(ctx, PARGO_E_UNSUPPORTED, "searchPointNum %d > 32 is not supported", k);
    if (!(buffer > 0) && plan->n < k)
        return fail(ctx, PARGO_E_NONTERM,
                    "searchRange <= 0 with %d samples < searchPointNum %d: the reference's search never ends",
                    plan->n, k);
    if (mask && (mask_rows < 1 || mask_cols < 1)) return fail(ctx, PARGO_E_INVALID, "bad mask dimensions");
    if ((long long)((plan->rows + 15) / 16) * ((plan->cols + 15) / 16) > 0x7fffffffLL)
        return fail(ctx, PARGO_E_INVALID, "grid of %d x %d cells has too many 16 x 16 tiles", plan->rows, plan->cols);
    p = IdwParams{};
    p.xy = plan->d_xy;
    p.slots = plan->slots;
    p.v = plan->d_v;
    p.start = plan->d_start;
    p.count = plan->d_count;
    p.bbox = plan->d_bbox;
    p.mask = mask;
    p.minX = plan->minX;
    p.maxY = plan->maxY;
    p.block = plan->block;
    p.buffer = buffer;
    p.cs = (double)plan->cs;
    p.rows = plan->rows;
    p.cols = plan->cols;
    p.blockRows = plan->blockRows;
    p.blockCols = plan->blockCols;
    p.maxRing = std::max(plan->blockRows, plan->blockCols);
    p.k = k;
    p.power = power;
    p.mask_rows = mask_rows;
    p.mask_cols = mask_cols;
    p.mask_nodata = mask_nodata;
    p.order = plan->d_order;
    return PARGO_OK;
}

}  // namespace pargo

using pargo::fail;

extern "C" {

int pargo_idw_grid(const double *xs, const double *ys, int n, float cellsize, int *rows, int *cols, double *min_x,
                   double *max_y)
{
    if (!xs || !ys || n < 1 || !(cellsize > 0) || !rows || !cols) return PARGO_E_INVALID;
    double a, b, c, d;
    pargo::idw_extent(xs, ys, n, cellsize, a, b, c, d, *rows, *cols);
    if (min_x) *min_x = a;
    if (max_y) *max_y = d;
    return PARGO_OK;
}

int pargo_idw_extent(const double *xs, const double *ys, int n, float cellsize, double extent[4], int *rows,
                     int *cols)
{
    if (!xs || !ys || n < 1 || !(cellsize > 0) || !extent || !rows || !cols) return PARGO_E_INVALID;
    pargo::idw_extent(xs, ys, n, cellsize, extent[0], extent[1], extent[2], extent[3], *rows, *cols);
    return PARGO_OK;
}

static int idw_prepare(pargo_ctx *ctx, const double *xs, const double *ys, const double *vals, int n, float cellsize,
                       double block_size, pargo_idw_plan **out)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!out || !xs || !ys || !vals || n < 1) return fail(ctx, PARGO_E_INVALID, "pargo_idw_prepare: bad arguments");
    if (!(cellsize > 0) || !std::isfinite(cellsize)) return fail(ctx, PARGO_E_INVALID, "resolution must be > 0");
    if (!(block_size > 0) || !std::isfinite(block_size)) return fail(ctx, PARGO_E_INVALID, "blockSize must be > 0");
    for (int i = 0; i < n; i++)
        if (!std::isfinite(xs[i]) || !std::isfinite(ys[i]))
            return fail(ctx, PARGO_E_INVALID, "sample %d has a non-finite coordinate", i);
    *out = nullptr;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    pargo::alloc_point("IDW plan (host)");
    // owned until it is handed out: every early return frees it
    std::unique_ptr<pargo_idw_plan, PlanDeleter> pl(new pargo_idw_plan());
    pl->owner = ctx;
    pl->n = n;
    pl->cs = cellsize;
    pl->block = block_size;
    pargo::idw_extent(xs, ys, n, cellsize, pl->minX, pl->maxX, pl->minY, pl->maxY, pl->rows, pl->cols);
    if (pl->rows < 1 || pl->cols < 1) {
        return fail(ctx, PARGO_E_INVALID, "empty output grid");
    }
    // creatSampleBlocks: ceil of the extent over the block size
    const double br = std::ceil((pl->maxY - pl->minY) / block_size), bc = std::ceil((pl->maxX - pl->minX) / block_size);
    if (br * bc > 1 << 26) {
        return fail(ctx, PARGO_E_INVALID, "too many sample blocks (%.0f x %.0f)", br, bc);
    }
    pl->blockRows = (int)br;
    pl->blockCols = (int)bc;
    const int nb = pl->blockRows * pl->blockCols;
    std::vector<int> start(nb + 1, 0), bidx(n);
    for (int i = 0; i < n; i++) {
        const int iCol = (int)((xs[i] - pl->minX) / block_size);
        const int iRow = (int)((pl->maxY - ys[i]) / block_size);
        const long long index = (long long)iRow * pl->blockCols + iCol;
        if (iCol < 0 || iRow < 0 || iCol >= pl->blockCols || index >= nb) {
                return fail(ctx, PARGO_E_INVALID, "Index out of range. An error in creatSampleBlocks() (sample %d)", i);
        }
        bidx[i] = (int)index;
        start[index + 1]++;
    }
    // counts, then slot offsets with every bucket padded to a multiple of 4
    std::vector<int> count(nb);
    for (int b = 0; b < nb; b++) count[b] = start[b + 1];
    start[0] = 0;
    for (int b = 0; b < nb; b++) start[b + 1] = start[b] + ((count[b] + 3) & ~3);
    const int slots = start[nb];
    if (slots >= (1 << 25))  // the search packs a slot and a 6-bit bin in 31 bits
        return fail(ctx, PARGO_E_UNSUPPORTED, "%d samples are more than the 2^25 bucket slots supported", n);
    pl->slots = slots;
    std::vector<int> fill(start.begin(), start.end() - 1);
    // 4 trailing padding slots: a lane may read one candidate past its bucket
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::vector<double2> xy((size_t)slots + 4, make_double2(nan, nan));
    std::vector<double> v((size_t)slots, 0.0);
    pl->h_perm.assign((size_t)slots, -1);
    const double inf = std::numeric_limits<double>::infinity();
    std::vector<double4> bbox((size_t)nb, make_double4(inf, -inf, inf, -inf));
    for (int i = 0; i < n; i++) {
        const int q = fill[bidx[i]]++;
        xy[q] = make_double2(xs[i], ys[i]);
        v[q] = vals[i];
        pl->h_perm[q] = i;
        double4 &bb = bbox[bidx[i]];
        bb.x = std::min(bb.x, xs[i]);
        bb.y = std::max(bb.y, xs[i]);
        bb.z = std::min(bb.z, ys[i]);
        bb.w = std::max(bb.w, ys[i]);
    }
    const size_t bytes_xy = sizeof(double2) * ((size_t)slots + 4), bytes_v = sizeof(double) * slots;
    const size_t bytes_start = sizeof(int) * (nb + 1), bytes_count = sizeof(int) * nb, bytes_perm = sizeof(int) * slots;
    const size_t bytes_bbox = sizeof(double4) * nb;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = al(bytes_xy) + al(bytes_v) + al(bytes_start) + al(bytes_count) + al(bytes_perm) + al(bytes_bbox);
    hipError_t e = pargo::dev_malloc(&pl->d_mem, total, "IDW samples");
    if (e != hipSuccess)
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc(%zu) for IDW samples: %s%s", total, hipGetErrorString(e),
                    pargo::injected_note());
    char *base = (char *)pl->d_mem;
    pl->d_xy = (double2 *)base;
    base += al(bytes_xy);
    pl->d_v = (double *)base;
    base += al(bytes_v);
    pl->d_start = (int *)base;
    base += al(bytes_start);
    pl->d_count = (int *)base;
    base += al(bytes_count);
    pl->d_perm = (int *)base;
    base += al(bytes_perm);
    pl->d_bbox = (double4 *)base;
    e = hipMemcpyAsync(pl->d_xy, xy.data(), bytes_xy, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pl->d_v, v.data(), bytes_v, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pl->d_start, start.data(), bytes_start, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pl->d_count, count.data(), bytes_count, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pl->d_perm, pl->h_perm.data(), bytes_perm, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pl->d_bbox, bbox.data(), bytes_bbox, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, PARGO_E_DEVICE, "uploading IDW samples: %s", hipGetErrorString(e));
    *out = pl.release();
    return PARGO_OK;
}

int pargo_idw_prepare(pargo_ctx *ctx, const double *xs, const double *ys, const double *vals, int n,
                      float cellsize, double block_size, pargo_idw_plan **out)
{
    return pargo::guard(ctx, "pargo_idw_prepare",
                        [&]() -> int { return idw_prepare(ctx, xs, ys, vals, n, cellsize, block_size, out); });
}


int pargo_idw_plan_free(pargo_ctx *ctx, pargo_idw_plan *plan)
{
    if (!ctx || !plan) return PARGO_E_INVALID;
    if (plan->owner != ctx) return fail(ctx, PARGO_E_INVALID, "IDW plan belongs to another context");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    PlanDeleter()(plan);
    return PARGO_OK;
}

int pargo_idw_set_values(pargo_ctx *ctx, pargo_idw_plan *plan, const double *vals)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!plan || !vals || plan->owner != ctx) return fail(ctx, PARGO_E_INVALID, "pargo_idw_set_values: bad arguments");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    return pargo::guard(ctx, "pargo_idw_set_values", [&]() -> int {
        pargo::alloc_point("IDW values (host)");
        std::vector<double> v((size_t)plan->slots, 0.0);
        for (int q = 0; q < plan->slots; q++)
            if (plan->h_perm[q] >= 0) v[q] = vals[plan->h_perm[q]];
        PARGO_HIP(ctx, hipMemcpyAsync(plan->d_v, v.data(), sizeof(double) * plan->slots, hipMemcpyHostToDevice,
                                      ctx->stream));
        PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return PARGO_OK;
    });
}

int pargo_idw_set_values_device(pargo_ctx *ctx, pargo_idw_plan *plan, const double *d_vals)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!plan || !d_vals || plan->owner != ctx)
        return fail(ctx, PARGO_E_INVALID, "pargo_idw_set_values_device: bad arguments");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(pargo::idw_bucket_values_kernel, dim3((plan->slots + 255) / 256), dim3(256), 0, ctx->stream,
                       d_vals, (const int *)plan->d_perm, plan->slots, plan->d_v);
    return pargo::check_launch(ctx, "pargo_idw_set_values_device");
}

int pargo_idw_run(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                  const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, float *out,
                  unsigned long long *candidates_out)
{
    pargo::IdwParams p;
    int rc = pargo::idw_fill_params(ctx, plan, k, power, buffer, mask, mask_rows, mask_cols, mask_nodata, p);
    if (rc) return rc;
    if (!out) return fail(ctx, PARGO_E_INVALID, "pargo_idw_run: null output");
    rc = pargo::bind(ctx);
    if (rc) return rc;
    p.out = out;
    p.cand_out = candidates_out;
    if (candidates_out) PARGO_HIP(ctx, hipMemsetAsync(candidates_out, 0, sizeof(unsigned long long), ctx->stream));
    rc = pargo::launch_idw<false>(p, ctx->stream);
    if (rc) return fail(ctx, rc, "unsupported searchPointNum %d", k);
    return pargo::check_launch(ctx, "pargo_idw_run");
}

int pargo_idw_work(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                   const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, float *out,
                   double *work_out)
{
    pargo::IdwParams p;
    int rc = pargo::idw_fill_params(ctx, plan, k, power, buffer, mask, mask_rows, mask_cols, mask_nodata, p);
    if (rc) return rc;
    if (!work_out) return fail(ctx, PARGO_E_INVALID, "pargo_idw_work: null work output");
    rc = pargo::bind(ctx);
    if (rc) return rc;
    p.out = out;
    p.work_out = work_out;
    rc = pargo::launch_idw<false>(p, ctx->stream);
    if (rc) return fail(ctx, rc, "unsupported searchPointNum %d", k);
    return pargo::check_launch(ctx, "pargo_idw_work");
}

int pargo_idw_load(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int granularity, const int32_t *mask,
                   int mask_rows, int mask_cols, double mask_nodata, double nodata, double *load_out,
                   int *load_rows, int *load_cols)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!plan || plan->owner != ctx) return fail(ctx, PARGO_E_INVALID, "IDW plan is null or from another context");
    if (k < 1) return fail(ctx, PARGO_E_INVALID, "searchPointNum must be >= 1, got %d", k);
    if (granularity < 1) return fail(ctx, PARGO_E_INVALID, "granularity must be >= 1, got %d", granularity);
    if (mask && (mask_rows < 1 || mask_cols < 1)) return fail(ctx, PARGO_E_INVALID, "bad mask dimensions");
    // ComputeLayer::_init (gpro/computeLayer.h:208-260): ceil of the work
    // rectangle over the granularity, cell size cellSize * granularity in float
    const int rows = plan->rows / granularity + (plan->rows % granularity ? 1 : 0);
    const int cols = plan->cols / granularity + (plan->cols % granularity ? 1 : 0);
    if (load_rows) *load_rows = rows;
    if (load_cols) *load_cols = cols;
    if (!load_out) return PARGO_OK;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    rc = pargo::ensure_scratch(ctx, sizeof(unsigned));
    if (rc) return rc;
    pargo::IdwLoadParams p{};
    p.count = plan->d_count;
    p.mask = mask;
    p.out = load_out;
    p.nonterm = (unsigned *)ctx->scratch;
    p.block = plan->block;
    p.ccs = (double)(plan->cs * (float)granularity);
    p.mask_nodata = mask_nodata;
    p.nodata = nodata;
    p.rows = rows;
    p.cols = cols;
    p.blockRows = plan->blockRows;
    p.blockCols = plan->blockCols;
    p.k = k;
    p.maxRing = std::max(plan->blockRows, plan->blockCols);
    p.mask_rows = mask_rows;
    p.mask_cols = mask_cols;
    PARGO_HIP(ctx, hipMemsetAsync(p.nonterm, 0, sizeof(unsigned), ctx->stream));
    hipLaunchKernelGGL(pargo::idw_load_kernel, dim3((cols + 15) / 16, (rows + 15) / 16), dim3(256), 0, ctx->stream, p);
    rc = pargo::check_launch(ctx, "pargo_idw_load");
    if (rc) return rc;
    unsigned bad = 0;
    PARGO_HIP(ctx, hipMemcpyAsync(&bad, p.nonterm, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (bad)
        return fail(ctx, PARGO_E_NONTERM,
                    "%u compute cells can never count searchPointNum %d samples: the reference's estimate never ends",
                    bad, k);
    return PARGO_OK;
}

static int idw_order_tiles(pargo_ctx *ctx, pargo_idw_plan *plan, int k, int power, double buffer,
                           const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, int source)
{
    pargo::IdwParams p;
    int rc = pargo::idw_fill_params(ctx, plan, k, power, buffer, mask, mask_rows, mask_cols, mask_nodata, p);
    if (rc) return rc;
    if (source != PARGO_ORDER_ROWS && source != PARGO_ORDER_ESTIMATE && source != PARGO_ORDER_MEASURED)
        return fail(ctx, PARGO_E_INVALID, "unknown tile-order source %d", source);
    rc = pargo::bind(ctx);
    if (rc) return rc;
    PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));  // runs queued on the old order have finished
    if (plan->d_order) {
        PARGO_HIP(ctx, hipFree(plan->d_order));
        plan->d_order = nullptr;
    }
    plan->order_source = PARGO_ORDER_ROWS;
    if (source == PARGO_ORDER_ROWS) return PARGO_OK;
    const int tiles_x = (plan->cols + 15) / 16, tiles_y = (plan->rows + 15) / 16;
    const size_t tiles = (size_t)tiles_x * tiles_y;
    double *d_load = nullptr;
    if (pargo::dev_malloc((void **)&d_load, tiles * sizeof(double), "tile loads") != hipSuccess)
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc(%zu) for the tile loads%s", tiles * sizeof(double),
                    pargo::injected_note());
    // the load buffer is freed on every path, exceptions included
    std::unique_ptr<double, decltype(&hipFree)> load_mem(d_load, &hipFree);
    pargo::alloc_point("tile order (host)");
    std::vector<double> load(tiles);
    if (source == PARGO_ORDER_ESTIMATE) {
        // the reference's estimate (IdwTransformation over a ComputeLayer of
        // granularity 16): one coarse cell per tile
        int lr = 0, lc = 0;
        rc = pargo_idw_load(ctx, plan, k, 16, mask, mask_rows, mask_cols, (double)mask_nodata, -1.0, d_load, &lr, &lc);
        if (rc == PARGO_OK && (lr != tiles_y || lc != tiles_x))
            rc = fail(ctx, PARGO_E_INVALID, "load grid %d x %d does not match %d x %d tiles", lr, lc, tiles_y, tiles_x);
    } else {
        // the preliminary experiment: samples each tile's searches examined in one pass
        hipError_t e = hipMemsetAsync(d_load, 0, tiles * sizeof(double), ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, PARGO_E_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
        p.order = nullptr;
        p.tile_work = d_load;
        if (rc == PARGO_OK && pargo::launch_idw<false>(p, ctx->stream) != PARGO_OK)
            rc = fail(ctx, PARGO_E_UNSUPPORTED, "unsupported searchPointNum %d", k);
        if (rc == PARGO_OK) rc = pargo::check_launch(ctx, "pargo_idw_order_tiles");
    }
    if (rc == PARGO_OK) {
        hipError_t e = hipMemcpyAsync(load.data(), d_load, tiles * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, PARGO_E_DEVICE, "reading the tile loads: %s", hipGetErrorString(e));
    }
    load_mem.reset();
    if (rc) return rc;
    // heaviest first; equal loads keep row-major order
    std::vector<int> order(tiles);
    for (size_t t = 0; t < tiles; t++) order[t] = (int)t;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return load[a] > load[b]; });
    if (pargo::dev_malloc((void **)&plan->d_order, tiles * sizeof(int), "tile order") != hipSuccess) {
        plan->d_order = nullptr;
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc(%zu) for the tile order%s", tiles * sizeof(int),
                    pargo::injected_note());
    }
    hipError_t e = hipMemcpyAsync(plan->d_order, order.data(), tiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipFree(plan->d_order);
        plan->d_order = nullptr;
        return fail(ctx, PARGO_E_DEVICE, "uploading the tile order: %s", hipGetErrorString(e));
    }
    plan->order_source = source;
    return PARGO_OK;
}

int pargo_idw_order_tiles(pargo_ctx *ctx, pargo_idw_plan *plan, int k, int power, double buffer,
                          const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, int source)
{
    return pargo::guard(ctx, "pargo_idw_order_tiles", [&]() -> int {
        return idw_order_tiles(ctx, plan, k, power, buffer, mask, mask_rows, mask_cols, mask_nodata, source);
    });
}

int pargo_idw_adjoint(pargo_ctx *ctx, const pargo_idw_plan *plan, int k, int power, double buffer,
                      const int32_t *mask, int mask_rows, int mask_cols, int mask_nodata, const float *grad_out,
                      double *grad_vals)
{
    pargo::IdwParams p;
    int rc = pargo::idw_fill_params(ctx, plan, k, power, buffer, mask, mask_rows, mask_cols, mask_nodata, p);
    if (rc) return rc;
    if (!grad_out || !grad_vals) return fail(ctx, PARGO_E_INVALID, "pargo_idw_adjoint: null gradient pointer");
    rc = pargo::bind(ctx);
    if (rc) return rc;
    rc = pargo::ensure_scratch(ctx, sizeof(double) * (size_t)plan->slots);
    if (rc) return rc;
    p.grad_out = grad_out;
    p.grad_acc = (double *)ctx->scratch;
    PARGO_HIP(ctx, hipMemsetAsync(p.grad_acc, 0, sizeof(double) * (size_t)plan->slots, ctx->stream));
    rc = pargo::launch_idw<true>(p, ctx->stream);
    if (rc) return fail(ctx, rc, "unsupported searchPointNum %d", k);
    rc = pargo::check_launch(ctx, "pargo_idw_adjoint");
    if (rc) return rc;
    hipLaunchKernelGGL(pargo::idw_unbucket_kernel, dim3((plan->slots + 255) / 256), dim3(256), 0, ctx->stream,
                       (const double *)p.grad_acc, (const int *)plan->d_perm, plan->slots, grad_vals);
    return pargo::check_launch(ctx, "pargo_idw_adjoint (unbucket)");
}

}  // extern "C"
