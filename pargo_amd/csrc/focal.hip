// focal.hip -- Horn slope, the other -mtd slope methods and D8 flow
// direction on a 3x3 Moore neighbourhood, fused into one pass over the DEM.
//
// Reference behaviour reproduced here:
//   slope  apps/morphology/slope/slopeOperator.cpp:51-139
//   D8     apps/hydrology/flowdird8/d8Operator.cpp:28-102
//   sweep  gpro/rasterOperator.h:171-185 over the work rectangle
//          [1, rows-2] x [1, cols-2] (gpro/neighborhood.h:447-448); the
//          border keeps the NoData that copyLayerInfo wrote
//          (gpro/rasterLayer.h:739-745).
//
// Layout and work split (one GPU, no ranks): the DEM is row-major float32 in
// HBM.  A lane owns CPT adjacent columns and walks a strip of rows from top
// to bottom, keeping the rows above, at and below the output row in
// registers, so every DEM value is read from HBM once per strip (plus one
// halo row above and below the strip).  Neighbour columns across lanes come
// from cross-lane shuffles; only lanes 0 and 63 read their outer halo
// column from memory.  Rows of a wave are coalesced 64*CPT*4-byte segments.
//
// D8 codes are those of the doubles the reference compares: drops
// RN(z4 - zi) for cardinal neighbours and RN(RN(z4 - zi) / sqrt(2.0)) for
// diagonal ones, max over them and 0, then the 1e-6 tests.  An FP32 filter
// on the same drops settles every cell whose code it can prove (almost
// all); a lane holding any other cell recomputes its cells on the doubles.
// A drop shared by two neighbouring cells is computed once (RN(a-b) =
// -RN(b-a), and the division routine is odd), and the division by
// sqrt(2.0) is the correctly rounded Markstein sequence of focal_common.h.
// Windows that hold Inf or |z| >= 2^40 take a per-cell path with IEEE
// division instead.
//
// The Horn numerators are formed in FP64 (exact for float32 input of
// moderate range), the square root in FP32: relative error below 5e-7.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "focal_common.h"


#include "internal.h"

namespace pargo {

namespace {

struct Lut8Init {
    uint8_t v[256];
};

// ESRI bit -> window index (d8Operator.cpp:96-97 weights; window order
// 0=NW 1=N 2=NE 3=W 4=C 5=E 6=SW 7=S 8=SE).
constexpr Lut8Init make_lut8()
{
    Lut8Init t{};
    const int win_of_bit[8] = {5, 8, 7, 6, 3, 0, 1, 2};
    for (int code = 0; code < 256; code++) {
        int list[9];
        int n = 0;
        for (int w = 0; w < 9; w++)
            for (int b = 0; b < 8; b++)
                if (((code >> b) & 1) && win_of_bit[b] == w) list[n++] = w;
        t.v[code] = n ? (uint8_t)list[n >> 1] : (uint8_t)255;
    }
    return t;
}

}  // namespace

__constant__ Lut8Init c_lut8 = make_lut8();

struct FocalParams {
    const float *__restrict__ dem;
    float *__restrict__ slope;
    uint8_t *__restrict__ code;
    uint32_t *__restrict__ mask;
    long long pitch;
    int rows, cols;
    int rows_per_strip;
    int mask_words;       // uint32 words per mask row
    float nd_lo, nd_hi;   // slope NoData interval (eps 1e-7)
    float nd8_lo, nd8_hi; // D8 NoData interval (eps 1e-6)
    float ndu_lo, ndu_hi; // smallest interval holding both (+inf,-inf if empty)
    float slope_nodata;   // value written to NoData slope cells
    double inv64cs2;      // 1 / (64 cs^2) for the Horn fast path
    double fast_scale;    // other methods' fast path: 1 / (k cs)^2 (MD: 1 / (cs sqrt 2)); 0: per cell in FP64
    double cs;            // float-rounded cell size
    int method;
    int unit;
    int dirtype;          // 0: no D8, 8 or 256
    int mask_from_d8;     // NoData mask reports the D8 NoData set
    // Strips (one wave's columns over a row strip) with a window that holds
    // Inf or |z| >= 2^40 are queued here; focal_fixup_kernel recomputes those
    // windows with the reference-order FP64 code.
    unsigned *fix_count;  // [0] queued strips, [1] finished fixup blocks
    uint4 *fix_list;      // (tile index, first row, end row, first column)
    unsigned fix_cap;
    unsigned tile_id;     // index of this raster in a batch (0 otherwise)
    int vec;              // 4-byte aligned rows: vector loads and stores
    int waves_x;          // batch: wave-wide column bands of this tile
};

// Row vectors at 4-byte alignment and code words at byte alignment: gfx950
// runs global dwordx4/dwordx2 accesses at any dword address and dword
// stores at any byte address as single instructions, so a row takes the
// vector path whatever its pitch (rasters whose width is not a multiple of
// 4, and batch tiles, included).
typedef float f32x4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float f32x2 __attribute__((ext_vector_type(2), aligned(4)));
typedef unsigned u32_b1 __attribute__((aligned(1)));
typedef unsigned short u16_b1 __attribute__((aligned(1)));

template <int CPT>
struct RawRow {
    float v[CPT];
    float hl;  // outer halo column: c0-1 on lane 0, c0+CPT on lane 63
};

// Loads DEM row r (clamped into the raster) for the lane's CPT columns and
// its outer halo column.  Every load is unconditional: rows above the top or
// below the bottom and columns past the right edge read the nearest row or
// column instead, which only ever feeds border cells (NoData anyway) or
// columns that are not stored.  Keeping the loads out of branches lets the
// compiler wait only for the oldest of the rows in flight.
// VECLOAD: the wave's 64*CPT columns are inside the raster (one dwordx4 /
// dwordx2 per lane at any dword address).  The halo load is
// issued by every lane: lane 0 fetches column c0-1, lane 63 column c0+CPT,
// the others re-read a column of their own segment (an L1 hit).
template <int CPT, bool VECLOAD>
__device__ __forceinline__ void load_raw(const FocalParams &p, int r, int c0, int lane, RawRow<CPT> &o)
{
    const int rr = r < 0 ? 0 : (r >= p.rows ? p.rows - 1 : r);
    const float *rowp = p.dem + (long long)rr * p.pitch;
    if constexpr (VECLOAD && CPT == 4) {
        const f32x4 t = *reinterpret_cast<const f32x4 *>(rowp + (unsigned)c0);
        o.v[0] = t.x; o.v[1] = t.y; o.v[2] = t.z; o.v[3] = t.w;
    } else if constexpr (VECLOAD && CPT == 2) {
        const f32x2 t = *reinterpret_cast<const f32x2 *>(rowp + (unsigned)c0);
        o.v[0] = t.x; o.v[1] = t.y;
    } else {
#pragma unroll
        for (int k = 0; k < CPT; k++) o.v[k] = rowp[(unsigned)min(c0 + k, p.cols - 1)];
    }
    int hc = (lane == 0) ? c0 - 1 : (lane == 63 ? c0 + CPT : c0);
    hc = hc < 0 ? 0 : (hc > p.cols - 1 ? p.cols - 1 : hc);
    o.hl = rowp[(unsigned)hc];
}

// Cross-lane moves by DPP wavefront shifts (lane i reads lane i-1 / i+1).
// The lane without a source (0, resp. 63) is not written by the move and
// returns `edge`, the old value of the destination.
__device__ __forceinline__ float from_lower_lane(float v, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_upper_lane(float v, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// OR of v over each aligned group of N lanes (N = 4, 8 or 16), every lane
// of the group getting the result: two quad permutations (lane ^ 1,
// lane ^ 2), then the half-row mirror (lane i <-> 7 - i within 8 lanes)
// and the row mirror (i <-> 15 - i within 16), each pairing a lane with one
// of the other quad / half-row.
template <int N>
__device__ __forceinline__ unsigned or_lanes(unsigned v)
{
    static_assert(N == 4 || N == 8 || N == 16, "group of 4, 8 or 16 lanes");
    v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    if (N >= 8) v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    if (N >= 16) v |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xf, 0xf, false); // row_mirror
    return v;
}

// FP64 max/min without the quieting "v_max x,x" LLVM inserts for values it
// cannot prove canonical (loop-carried ones): every operand here is the
// result of an FP64 add/fma, so it is never a signalling NaN, and the
// instructions' IEEE maxNum/minNum semantics (a quiet NaN operand yields the
// other operand) are exactly what the reference's "if (diff > max)" does.
__device__ __forceinline__ double vmax(double a, double b)
{
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double vmin(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// max(a, -b, 0)
__device__ __forceinline__ double vmax_neg_zero(double a, double b)
{
    double r;
    asm("v_max_f64 %0, %1, -%2\n\tv_max_f64 %0, %0, 0" : "=&v"(r) : "v"(a), "v"(b));
    return r;
}

// FP32 three-way max/min as single instructions, again without the quieting
// LLVM adds in front of fmaxf/fminf.  IEEE mode: a quiet NaN operand is
// skipped and a signalling one makes the result a quiet NaN; every caller
// either feeds results of FP32 arithmetic (quiet) or treats a NaN result as
// "take the careful path".
__device__ __forceinline__ float max3f(float a, float b, float c)
{
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float min3f(float a, float b, float c)
{
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// max(-a, -b, -c)
__device__ __forceinline__ float max3f_neg(float a, float b, float c)
{
    float r;
    asm("v_max3_f32 %0, -%1, -%2, -%3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// max(a, -b, 0)
__device__ __forceinline__ float max3f_pos_neg_zero(float a, float b)
{
    float r;
    asm("v_max3_f32 %0, %1, -%2, 0" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// max(a, -b)
__device__ __forceinline__ float maxf_neg(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, -%2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// c = 2c + (lane bit of `mask`): one v_addc with the lane mask as carry-in.
__device__ __forceinline__ unsigned push_bit(unsigned c, bool cond)
{
    const unsigned long long m = __builtin_amdgcn_ballot_w64(cond);
    unsigned r;
    asm("v_addc_co_u32_e64 %0, vcc, %1, %1, %2" : "=v"(r) : "v"(c), "s"(m) : "vcc");
    return r;
}

// One DEM row of the rolling window (columns c0-1 .. c0+CPT) and, once the
// row below is known, the drops from this row to it.  A kernel keeps three
// slots and rotates their roles statically, so nothing is copied per row.
template <int CPT>
struct Slot {
    double z[CPT + 2];   // the row widened to FP64 (non-Horn slope methods only)
    float zf[CPT + 2];   // the row as loaded (float32)
    float V[CPT];        // RN32(z[j+1] - below[j+1])
    float SE[CPT + 1];   // RN32(RN32(z[i] - below[i+1]) * RN32(1/sqrt 2)), i = 0..CPT
    float SW[CPT + 2];   // the same towards below[i-1], i = 1..CPT+1
    // per-position bits of this lane's CPT+2 values, filled only for rows
    // holding NoData or a special value: slope NoData (bits 0-7), D8 NoData
    // (8-15), |z| >= 2^40 or Inf (16-23)
    unsigned flags;
};

// Converts a loaded row.  The common case costs a min and a max over the
// CPT+2 values (three-way instructions) and max(mx, -mn) = max |z|; only
// rows holding NoData or values with |z| >= 2^40 (incl. Inf) get
// per-position bits.  A NaN that the min/max skip needs no flag: the FP32
// filter and the FP64 tests skip NaN drops exactly as the reference's > and
// < comparisons do, and NaN is never NoData.  A signalling NaN turns mn/mx
// into NaN, which only sends the row to the per-position test.
template <int CPT, bool KEEP_Z>
__device__ __forceinline__ void finish_row(const FocalParams &p, const RawRow<CPT> &raw, int lane, Slot<CPT> &o)
{
    float f[CPT + 2];
    // lanes 0 and 63 have no source lane and keep the halo value they loaded
    f[0] = from_lower_lane(raw.v[CPT - 1], raw.hl);
    f[CPT + 1] = from_upper_lane(raw.v[0], raw.hl);
#pragma unroll
    for (int k = 0; k < CPT; k++) f[k + 1] = raw.v[k];
    float mn = f[0], mx = f[0];
#pragma unroll
    for (int i = 1; i < CPT + 2; i += 2) {
        const float b = f[i], c = (i + 1 < CPT + 2) ? f[i + 1] : f[i];
        mn = min3f(mn, b, c);
        mx = max3f(mx, b, c);
    }
    const float am = maxf_neg(mx, mn);
    const bool exc = !(mx < p.ndu_lo || mn > p.ndu_hi) || !(am < 0x1p40f);
    o.flags = 0;
    if (exc) {
        unsigned fl = 0;
#pragma unroll
        for (int i = 0; i < CPT + 2; i++) {
            fl |= (unsigned)in_interval(f[i], p.nd_lo, p.nd_hi) << i;
            fl |= (unsigned)in_interval(f[i], p.nd8_lo, p.nd8_hi) << (i + 8);
            fl |= (unsigned)(!(__builtin_fabsf(f[i]) < 0x1p40f)) << (i + 16);
        }
        o.flags = fl;
    }
    if (KEEP_Z) {
#pragma unroll
        for (int i = 0; i < CPT + 2; i++) o.z[i] = (double)f[i];
    }
#pragma unroll
    for (int i = 0; i < CPT + 2; i++) o.zf[i] = f[i];
}

// Per-cell slope for any method, in the reference's FP64 operation order
// (slopeOperator.cpp:83-136).  d[] is the window in row-major order.
__device__ double slope_generic(const double d[9], int method, double cellSize)
{
    double dx, dy;
    switch (method) {
    case PARGO_SLOPE_FFD:
        dx = (d[8] - d[6] + d[2] - d[0]) / (4.0 * cellSize);
        dy = (d[0] - d[6] + d[2] - d[8]) / (4.0 * cellSize);
        return sqrt(dx * dx + dy * dy);
    case PARGO_SLOPE_MD: {
        // max over the 8 raw drops (initial 0), divided by the distance of
        // the last neighbour visited, cellSize*sqrt(2) (slopeOperator.cpp:91-103).
        double m = 0.;
        const int order[8] = {0, 1, 2, 3, 5, 6, 7, 8};
        for (int t = 0; t < 8; t++) {
            double v = d[4] - d[order[t]];
            if (m < v) m = v;
        }
        return m / (cellSize * kSqrt2);
    }
    case PARGO_SLOPE_SD:
        dx = (d[4] - d[3]) / cellSize;
        dy = (d[4] - d[7]) / cellSize;
        return sqrt(dx * dx + dy * dy);
    case PARGO_SLOPE_SFD:
        dx = (d[5] - d[3]) / (2.0 * cellSize);
        dy = (d[1] - d[7]) / (2.0 * cellSize);
        return sqrt(dx * dx + dy * dy);
    case PARGO_SLOPE_TFD:
        dx = (d[8] + d[5] + d[2] - d[0] - d[3] - d[6]) / (6.0 * cellSize);
        dy = (d[2] + d[1] + d[0] - d[6] - d[7] - d[6]) / (6.0 * cellSize);  // d[6] twice, as the reference
        return sqrt(dx * dx + dy * dy);
    case PARGO_SLOPE_TFDW: {
        const double s8 = 0x1.6a09e667f3bcdp+1;  // sqrt(8.0)
        dx = (d[8] + kSqrt2 * d[5] + d[2] - d[6] - kSqrt2 * d[3] - d[0]) / ((4.0 + s8) * cellSize);
        dy = (d[2] + kSqrt2 * d[1] + d[0] - d[6] - kSqrt2 * d[7] - d[8]) / ((4.0 + s8) * cellSize);
        return sqrt(dx * dx + dy * dy);
    }
    default:  // PARGO_SLOPE_FD
        dx = (d[8] + 2 * d[5] + d[2] - d[6] - 2 * d[3] - d[0]) / (8.0 * cellSize);
        dy = (d[2] + 2 * d[1] + d[0] - d[6] - 2 * d[7] - d[8]) / (8.0 * cellSize);
        return sqrt(dx * dx + dy * dy);
    }
}

// D8 of one window with IEEE division, exactly d8Operator.cpp:61-97.
// Returns the ESRI code (0 = flat).
__device__ unsigned d8_exact(const double d[9])
{
    double diff[9];
    double mx = 0.0;
    for (int i = 0; i < 9; i++) {
        diff[i] = d[4] - d[i];
        if (i == 0 || i == 2 || i == 6 || i == 8) diff[i] = diff[i] / kSqrt2;
        if (diff[i] > mx) mx = diff[i];
    }
    if (mx < kD8Eps) return 0;
    const int bit_of_win[9] = {5, 6, 7, 4, -1, 0, 3, 2, 1};
    unsigned code = 0;
    for (int i = 0; i < 9; i++)
        if (i != 4 && fabs(diff[i] - mx) < kD8Eps) code |= 1u << bit_of_win[i];
    return code;
}

// D8 of one window on the reference's doubles (d8Operator.cpp:61-97) with
// the correctly rounded division by sqrt(2.0) of div_sqrt2; window order
// 0=NW 1=N 2=NE 3=W 4=C 5=E 6=SW 7=S 8=SE.  Returns the ESRI code (0 = flat).
__device__ __forceinline__ unsigned d8_window_fp64(double d0, double d1, double d2, double d3, double d4, double d5,
                                                   double d6, double d7, double d8)
{
    const double nw = div_sqrt2(d4 - d0), n = d4 - d1, ne = div_sqrt2(d4 - d2), w = d4 - d3;
    const double e = d4 - d5, sw = div_sqrt2(d4 - d6), s_ = d4 - d7, se = div_sqrt2(d4 - d8);
    const double M = vmax(vmax(vmax(vmax(nw, n), vmax(ne, w)), vmax(vmax(e, sw), vmax(s_, se))), 0.0);
    unsigned c = 0;
    c = push_bit(c, fabs(ne - M) < kD8Eps);  // NE  128
    c = push_bit(c, fabs(n - M) < kD8Eps);   // N    64
    c = push_bit(c, fabs(nw - M) < kD8Eps);  // NW   32
    c = push_bit(c, fabs(w - M) < kD8Eps);   // W    16
    c = push_bit(c, fabs(sw - M) < kD8Eps);  // SW    8
    c = push_bit(c, fabs(s_ - M) < kD8Eps);  // S     4
    c = push_bit(c, fabs(se - M) < kD8Eps);  // SE    2
    c = push_bit(c, fabs(e - M) < kD8Eps);   // E     1
    return (M < kD8Eps) ? 0u : c;
}

// The same methods with the two divisions and the FP64 root replaced: the
// numerators of dx and dy are the reference's FP64 expressions, in its
// order, over the float-widened window; the sum of their squares is scaled
// once by 1 / (k cs)^2 (k = 8, 4, 1, 2, 6 or 4 + sqrt 8 as the method's
// divisor has it) and its root taken in FP32.  MD's maximum drop is scaled
// by 1 / (cs sqrt 2).  Against the reference's double result this moves the
// float output by at most about one float ulp.  With |z| < 2^40 (larger
// windows go per cell) and 2^-20 <= cs <= 2^40 the scaled sum stays below
// 2^123, inside the float range.
__device__ __forceinline__ float slope_fast(const double d[9], int method, double scale)
{
    double nx, ny;
    switch (method) {
    case PARGO_SLOPE_FFD:
        nx = d[8] - d[6] + d[2] - d[0];
        ny = d[0] - d[6] + d[2] - d[8];
        break;
    case PARGO_SLOPE_MD: {
        double m = 0.;
        const int order[8] = {0, 1, 2, 3, 5, 6, 7, 8};
        for (int t = 0; t < 8; t++) {
            const double v = d[4] - d[order[t]];
            if (m < v) m = v;
        }
        return (float)(m * scale);
    }
    case PARGO_SLOPE_SD:
        nx = d[4] - d[3];
        ny = d[4] - d[7];
        break;
    case PARGO_SLOPE_SFD:
        nx = d[5] - d[3];
        ny = d[1] - d[7];
        break;
    case PARGO_SLOPE_TFD:
        nx = d[8] + d[5] + d[2] - d[0] - d[3] - d[6];
        ny = d[2] + d[1] + d[0] - d[6] - d[7] - d[6];  // d[6] twice, as the reference
        break;
    case PARGO_SLOPE_TFDW:
        nx = d[8] + kSqrt2 * d[5] + d[2] - d[6] - kSqrt2 * d[3] - d[0];
        ny = d[2] + kSqrt2 * d[1] + d[0] - d[6] - kSqrt2 * d[7] - d[8];
        break;
    default:  // PARGO_SLOPE_FD
        nx = d[8] + 2 * d[5] + d[2] - d[6] - 2 * d[3] - d[0];
        ny = d[2] + 2 * d[1] + d[0] - d[6] - 2 * d[7] - d[8];
        break;
    }
    return __builtin_amdgcn_sqrtf((float)((nx * nx + ny * ny) * scale));
}

__device__ __forceinline__ float slope_unit(float s, int unit)
{
    return unit == PARGO_UNIT_DEGREES ? atanf(s) * 57.295779513082323f : s;
}

// Output row r with up = row r-1, mid = row r, dn = the slot that receives
// row r+1 from `raw`; `raw` is then refilled with row r+4, or with row
// rlast (the strip's last input row) once r+4 passes it, so that a strip
// fetches no row of the next one (a repeated row is a cache hit).
// `special` is set when a window of this lane holds a value the fast path
// does not handle.  MASK false compiles the NoData mask output out (p.mask
// is then ignored); MASK true writes it when p.mask is set.
template <int CPT, bool VECLOAD, bool DO_SLOPE, bool DO_D8, int SM, bool MASK>
__device__ __forceinline__ void focal_step(const FocalParams &p, int r, int rlast, int c0, int lane, unsigned colb,
                                           Slot<CPT> &up, Slot<CPT> &mid, Slot<CPT> &dn, RawRow<CPT> &raw,
                                           bool &special)
{
    finish_row<CPT, DO_SLOPE && SM == 0>(p, raw, lane, dn);
    load_raw<CPT, VECLOAD>(p, min(r + 4, rlast), c0, lane, raw);

    float sl[CPT];
    unsigned cd[CPT];
    // ---- slope ------------------------------------------------------------
    if (DO_SLOPE) {
        if constexpr (SM == 1) {
            // Horn numerators: nx from column sums up + 2 mid + dn, ny from the
            // vertical differences up - dn (both FP64; the float inputs make
            // them exact or within an ulp of the reference's sums)
            double col[CPT + 2], vd[CPT + 2];
#pragma unroll
            for (int i = 0; i < CPT + 2; i++) {
                const double u = up.zf[i], m = mid.zf[i], d = dn.zf[i];
                col[i] = __builtin_fma(2.0, m, d) + u;
                vd[i] = u - d;
            }
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const double nx = col[j + 2] - col[j];
                const double ny = __builtin_fma(2.0, vd[j + 1], vd[j] + vd[j + 2]);
                const double q = __builtin_fma(nx, nx, ny * ny) * p.inv64cs2;
                sl[j] = __builtin_amdgcn_sqrtf((float)q);
            }
        } else if constexpr (SM == 2) {
            // the other methods' fast path on the float-widened window
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const int i = j + 1;
                const double d[9] = {up.zf[i - 1], up.zf[i], up.zf[i + 1], mid.zf[i - 1], mid.zf[i],
                                     mid.zf[i + 1], dn.zf[i - 1], dn.zf[i], dn.zf[i + 1]};
                sl[j] = slope_fast(d, p.method, p.fast_scale);
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const int i = j + 1;
                const double d[9] = {up.z[i - 1], up.z[i], up.z[i + 1], mid.z[i - 1], mid.z[i],
                                     mid.z[i + 1], dn.z[i - 1], dn.z[i], dn.z[i + 1]};
                sl[j] = (float)slope_generic(d, p.method, p.cs);
            }
        }
    }
    // ---- D8 ---------------------------------------------------------------
    // FP32 drops decide every cell whose FP64 answer they prove.  With x64
    // the reference's drop and x32 ours, |x64 - x32| <= 2^-22 |x32|
    // (+2^-149 near underflow) for cardinal and diagonal drops alike, and
    // x32, x64 have the same sign.  Let M = max(0, x32) over the 8 drops,
    // G = the float just below 0.9999e-6, K = RN32(1.0001e-6) and
    // L = max(RN32(M (1 - 2^-20) - K), G); a drop is a candidate when
    // x32 > L (a NaN drop never is: the reference's > and < skip it too).
    //  * M <= G proves the reference's max below 1e-6 (flat, code 0), and
    //    then no drop exceeds L: the candidate set is empty.
    //  * M >= 2.00025e-6 gives L = RN32(M (1 - 2^-20) - K) >= G and M > K.
    //    A negative non-candidate has x64 <= 0, so M64 - x64 >= M64 > 1e-6;
    //    a non-negative one has x32 <= L, and with the rounding of L
    //    M64 - x64 >= M (1.75 2^-22) + K > 1e-6.  So if exactly one drop is
    //    a candidate it is the reference's unique max, and its code bit is
    //    the reference's code.
    //  * M in between (within 0.50025e-6 of 1.5e-6), or two or more
    //    candidates: the lane recomputes its cells in FP64.
    // Every FP64 code bit is a candidate, so the candidate set is the code
    // itself whenever it has one bit.
    bool amb = false;  // this lane recomputes its cells with the FP64 windows
    unsigned popmax = 0;
    float band = 0.0f;
    if (DO_D8) {
        float H[CPT + 1];
#pragma unroll
        for (int i = 0; i <= CPT; i++) H[i] = mid.zf[i] - mid.zf[i + 1];
#pragma unroll
        for (int j = 0; j < CPT; j++) mid.V[j] = mid.zf[j + 1] - dn.zf[j + 1];
#pragma unroll
        for (int i = 0; i <= CPT; i++) mid.SE[i] = (mid.zf[i] - dn.zf[i + 1]) * kInvSqrt2f;
#pragma unroll
        for (int i = 1; i <= CPT + 1; i++) mid.SW[i] = (mid.zf[i] - dn.zf[i - 1]) * kInvSqrt2f;
        band = __builtin_huge_valf();
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int i = j + 1;
            // drops  E=H[i]  W=-H[i-1]  S=mid.V[j]  N=-up.V[j]
            //        SE=mid.SE[i]  NW=-up.SE[i-1]  SW=mid.SW[i]  NE=-up.SW[i+1]
            const float M = max3f(max3f(H[i], mid.V[j], mid.SE[i]), max3f_neg(H[i - 1], up.V[j], up.SE[i - 1]),
                                  max3f_pos_neg_zero(mid.SW[i], up.SW[i + 1]));
            const float L = fmaxf(__builtin_fmaf(M, kD8Shrink, -kD8SureF), kD8FloorF);
            // candidates, in the order that leaves the ESRI code
            unsigned c = 0;
            c = push_bit(c, -up.SW[i + 1] > L);  // NE  128
            c = push_bit(c, -up.V[j] > L);       // N    64
            c = push_bit(c, -up.SE[i - 1] > L);  // NW   32
            c = push_bit(c, -H[i - 1] > L);      // W    16
            c = push_bit(c, mid.SW[i] > L);      // SW    8
            c = push_bit(c, mid.V[j] > L);       // S     4
            c = push_bit(c, mid.SE[i] > L);      // SE    2
            c = push_bit(c, H[i] > L);           // E     1
            cd[j] = c;
            popmax = max(popmax, (unsigned)__builtin_popcount(c));
            band = fminf(band, __builtin_fabsf(M - kD8BandMid));
        }
    }

    // ---- NoData, border and special windows (rare) -------------------------
    // (windows with a special value keep their fast-path values until the
    // fixup pass recomputes them)
    const bool rowb = (r == 0) || (r == p.rows - 1);
    const unsigned flw = up.flags | mid.flags | dn.flags;
    unsigned ndflags = 0;  // bit j: slope NoData; bit j+8: D8 NoData
    if (__builtin_expect((flw != 0) | (colb != 0) | rowb, 0)) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const bool border = rowb || ((colb >> j) & 1u);
            const bool nds = border || ((flw >> j) & 7u), n8 = border || ((flw >> (j + 8)) & 7u);
            ndflags |= (unsigned)nds << j;
            ndflags |= (unsigned)n8 << (j + 8);
            // a special value in a window that is not NoData for every output
            special |= ((flw >> (j + 16)) & 7u) && !((!DO_SLOPE || nds) && (!DO_D8 || n8));
        }
    }
    if (DO_D8) {
        // NoData windows take NoData whatever their candidates
        if (__builtin_expect((ndflags >> 8) != 0, 0)) {
            popmax = 0;
#pragma unroll
            for (int j = 0; j < CPT; j++)
                if (!((ndflags >> (j + 8)) & 1u)) popmax = max(popmax, (unsigned)__builtin_popcount(cd[j]));
        }
        amb = popmax > 1 || band < kD8BandHalf;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(amb) != 0, 0)) {
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const int i = j + 1;
                const unsigned c64 = d8_window_fp64(up.zf[i - 1], up.zf[i], up.zf[i + 1], mid.zf[i - 1], mid.zf[i],
                                                    mid.zf[i + 1], dn.zf[i - 1], dn.zf[i], dn.zf[i + 1]);
                if (amb) cd[j] = c64;
            }
        }
    }
    if (DO_SLOPE && p.unit == PARGO_UNIT_DEGREES) {
#pragma unroll
        for (int j = 0; j < CPT; j++) sl[j] = slope_unit(sl[j], p.unit);
    }
    unsigned ndbits = 0;  // NoData mask bits of this lane's cells
    if (__builtin_expect(ndflags != 0, 0)) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const bool nds = (ndflags >> j) & 1u, n8 = (ndflags >> (j + 8)) & 1u;
            if (DO_SLOPE && nds) sl[j] = p.slope_nodata;
            if (DO_D8 && n8) cd[j] = 0;
            const bool m = (DO_D8 && p.mask_from_d8) ? n8 : nds;
            ndbits |= (unsigned)(m && (c0 + j) < p.cols) << j;
        }
    }
    if (DO_D8 && p.dirtype == 8) {
#pragma unroll
        for (int j = 0; j < CPT; j++) cd[j] = cd[j] ? (unsigned)c_lut8.v[cd[j]] : 255u;
    }

    // ---- stores --------------------------------------------------------------
    // Full-wave rows go out as non-temporal stores: the outputs are not read
    // again by this pass, and streaming them past the caches measured 1-1.5 %
    // faster at every strip height on 16384^2 (values unchanged).
    const long long rbase = (long long)r * p.pitch;
    const unsigned uc0 = (unsigned)c0;  // column indexes are never negative
    if (DO_SLOPE) {
        float *srow = p.slope + rbase;
        if constexpr (VECLOAD) {
            if constexpr (CPT == 4) {
                const f32x4 t = {sl[0], sl[1], sl[2], sl[3]};
                __builtin_nontemporal_store(t, reinterpret_cast<f32x4 *>(srow + uc0));
            } else if constexpr (CPT == 2) {
                const f32x2 t = {sl[0], sl[1]};
                __builtin_nontemporal_store(t, reinterpret_cast<f32x2 *>(srow + uc0));
            } else {
#pragma unroll
                for (int j = 0; j < CPT; j++) srow[uc0 + j] = sl[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; j++)
                if (c0 + j < p.cols) srow[uc0 + j] = sl[j];
        }
    }
    if (DO_D8) {
        uint8_t *crow = p.code + rbase;
        if constexpr (VECLOAD) {
            if constexpr (CPT == 4) {
                const unsigned packed = __builtin_amdgcn_perm(
                    __builtin_amdgcn_perm(cd[3], cd[2], 0x0c0c0400u), __builtin_amdgcn_perm(cd[1], cd[0], 0x0c0c0400u),
                    0x05040100u);  // byte 0 of each code, three byte permutes
                __builtin_nontemporal_store(packed, reinterpret_cast<u32_b1 *>(crow + uc0));
            } else if constexpr (CPT == 2) {
                __builtin_nontemporal_store((unsigned short)(cd[0] | (cd[1] << 8)),
                                            reinterpret_cast<u16_b1 *>(crow + uc0));
            } else {
#pragma unroll
                for (int j = 0; j < CPT; j++) crow[uc0 + j] = (uint8_t)cd[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; j++)
                if (c0 + j < p.cols) crow[uc0 + j] = (uint8_t)cd[j];
        }
    }
    if (MASK && p.mask) {
        // gather the CPT-bit groups of 32/CPT lanes into one word (nothing
        // to gather when no lane of the wave has a NoData cell in this row)
        // The lane's shift is re-derived from the lane id where it is used:
        // a per-lane value kept across the row loop can end up spilled, and
        // its reload waits for every load in flight, the row prefetches
        // included.
        constexpr int LPW = 32 / CPT;
        unsigned w = 0;
        if (__builtin_amdgcn_ballot_w64(ndbits != 0) != 0) {
            unsigned ln;
            asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(ln));
            w = or_lanes<LPW>(ndbits << (CPT * (ln % LPW)));
        }
        uint32_t *mrow = p.mask + (long long)r * p.mask_words;
        if ((lane % LPW) == 0 && c0 < p.cols) mrow[uc0 >> 5] = w;
    }
}

// Pairs from a row to the row below (the prologue of a strip).
template <int CPT, bool DO_SLOPE, bool DO_D8, int SM>
__device__ __forceinline__ void pair_rows(Slot<CPT> &up, const Slot<CPT> &below)
{
    if (DO_D8) {
#pragma unroll
        for (int j = 0; j < CPT; j++) up.V[j] = up.zf[j + 1] - below.zf[j + 1];
#pragma unroll
        for (int i = 0; i <= CPT; i++) up.SE[i] = (up.zf[i] - below.zf[i + 1]) * kInvSqrt2f;
#pragma unroll
        for (int i = 1; i <= CPT + 1; i++) up.SW[i] = (up.zf[i] - below.zf[i - 1]) * kInvSqrt2f;
    }
}

// One wave computes output rows [rbeg, rend) of the 64*CPT columns that
// start at wave_c0 (wave-uniform).
template <int CPT, bool VECLOAD, bool DO_SLOPE, bool DO_D8, int SM, bool MASK>
__device__ __forceinline__ void focal_strip_impl(const FocalParams &p, int wave_c0, int lane, int rbeg, int rend)
{
    const int c0 = wave_c0 + lane * CPT;
    unsigned colb = 0;  // bit j: column c0+j is on (or past) the border
#pragma unroll
    for (int j = 0; j < CPT; j++) colb |= (unsigned)(c0 + j == 0 || c0 + j >= p.cols - 1) << j;

    Slot<CPT> A, B, C;
    RawRow<CPT> R0, R1, R2;
    {
        RawRow<CPT> ra, rb;
        load_raw<CPT, VECLOAD>(p, rbeg - 1, c0, lane, ra);
        load_raw<CPT, VECLOAD>(p, rbeg, c0, lane, rb);
        load_raw<CPT, VECLOAD>(p, min(rbeg + 1, rend), c0, lane, R0);
        load_raw<CPT, VECLOAD>(p, min(rbeg + 2, rend), c0, lane, R1);
        load_raw<CPT, VECLOAD>(p, min(rbeg + 3, rend), c0, lane, R2);
        finish_row<CPT, DO_SLOPE && SM == 0>(p, ra, lane, A);
        finish_row<CPT, DO_SLOPE && SM == 0>(p, rb, lane, B);
    }
    pair_rows<CPT, DO_SLOPE, DO_D8, SM>(A, B);
    bool special = false;
    // Three slots rotate roles statically: (up, mid, dn) = (A,B,C), (B,C,A), (C,A,B).
    for (int r = rbeg; r < rend; r += 3) {
        focal_step<CPT, VECLOAD, DO_SLOPE, DO_D8, SM, MASK>(p, r, rend, c0, lane, colb, A, B, C, R0, special);
        if (r + 1 >= rend) break;
        focal_step<CPT, VECLOAD, DO_SLOPE, DO_D8, SM, MASK>(p, r + 1, rend, c0, lane, colb, B, C, A, R1, special);
        if (r + 2 >= rend) break;
        focal_step<CPT, VECLOAD, DO_SLOPE, DO_D8, SM, MASK>(p, r + 2, rend, c0, lane, colb, C, A, B, R2, special);
    }
    // queue the strip for the exact pass (rare: Inf or |z| >= 2^40 nearby)
    if (__builtin_amdgcn_ballot_w64(special) != 0 && lane == 0) {
        const unsigned k = atomicAdd(p.fix_count, 1u);
        if (k < p.fix_cap) p.fix_list[k] = make_uint4(p.tile_id, (unsigned)rbeg, (unsigned)rend, (unsigned)wave_c0);
    }
}

// VEC: the raster's pitch and base allow vector access; a wave uses it when
// all its columns are inside the raster.
template <int CPT, bool VEC, bool DO_SLOPE, bool DO_D8, int SM, bool MASK = true>
__device__ __forceinline__ void focal_strip(const FocalParams &p, int wave_c0, int lane, int rbeg, int rend)
{
    if constexpr (VEC) {
        if (wave_c0 + 64 * CPT <= p.cols) {
            focal_strip_impl<CPT, true, DO_SLOPE, DO_D8, SM, MASK>(p, wave_c0, lane, rbeg, rend);
            return;
        }
    }
    focal_strip_impl<CPT, false, DO_SLOPE, DO_D8, SM, MASK>(p, wave_c0, lane, rbeg, rend);
}

// One raster: block = 4 waves side by side, one row strip per block.
template <int CPT, bool VEC, bool DO_SLOPE, bool DO_D8, int SM, bool MASK = true>
__device__ __forceinline__ void focal_raster(const FocalParams &p)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // The kernels that may write a NoData mask take the XCD-contiguous strip
    // order (5 % less read traffic; 1.3 % faster for the fused pass with the
    // mask at 16384^2, 9 % for slope alone at 8192^2).  The fused kernel
    // without the mask keeps the launch order, which spreads the strips in
    // flight over the whole raster: renumbered, it read as little but ran
    // 5 % slower.
    unsigned cb = blockIdx.x, strip = blockIdx.y;
    if (MASK) xcd_strip_order(cb, strip);
    const int wave_c0 = (int)(cb * 4 + wave) * 64 * CPT;
    const int rbeg = (int)strip * p.rows_per_strip;
    const int rend = min(rbeg + p.rows_per_strip, p.rows);
    focal_strip<CPT, VEC, DO_SLOPE, DO_D8, SM, MASK>(p, wave_c0, lane, rbeg, rend);
}

template <int CPT, bool VEC, bool DO_SLOPE, bool DO_D8, int SM>
__global__ __launch_bounds__(256) void focal_kernel(FocalParams p)
{
    focal_raster<CPT, VEC, DO_SLOPE, DO_D8, SM>(p);
}

// The fused Horn slope + D8 pass without a NoData mask, at four waves per
// SIMD (at most 128 VGPRs, a few spilled to scratch): it ran faster than at
// three waves without spills, since the pass issues VALU work most of the
// time and a fourth wave fills the dependency stalls.  That balance is
// fragile: any extra state in the loop (the mask's, for one) brings spill
// reloads into it, and each reload waits for every load in flight.  The
// variant that writes the mask therefore runs at three waves, like the
// other kernels.
template <int CPT, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void focal_kernel_fused(FocalParams p)
{
    focal_raster<CPT, VEC, true, true, 1, false>(p);
}

template <int CPT, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void focal_kernel_fused_mask(FocalParams p)
{
    focal_raster<CPT, VEC, true, true, 1, true>(p);
}

// Many rasters in one persistent launch: every wave takes (tile, column
// band, row strip) items from a device queue until it is empty.  Items are
// numbered tile by tile (first[t] = first item of tile t; tiles are in
// decreasing size so that the tail of the queue holds the small items).
// Batch tiles carry no NoData mask, so the mask code is compiled out.
template <int CPT, bool DO_SLOPE, bool DO_D8, int SM>
__device__ __forceinline__ void focal_batch_body(const FocalParams *__restrict__ tiles,
                                                 const unsigned *__restrict__ first, int ntiles,
                                                 unsigned *__restrict__ queue, unsigned total)
{
    const int lane = threadIdx.x & 63;
    // the tile of the previous item: a wave's next item usually falls in the
    // same (large) tile, and then needs no search and no parameter reload
    unsigned f0 = 1, f1 = 0;
    FocalParams p;
    for (;;) {
        unsigned item = 0;
        if (lane == 0) item = atomicAdd(queue, 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= total) break;
        if (!(item >= f0 && item < f1)) {
            int lo = 0, hi = ntiles - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (first[mid] <= item) lo = mid;
                else hi = mid - 1;
            }
            f0 = first[lo];
            f1 = lo + 1 < ntiles ? first[lo + 1] : total;
            p = tiles[lo];
        }
        const unsigned local = item - f0;
        const int wx = (int)(local % (unsigned)p.waves_x), strip = (int)(local / (unsigned)p.waves_x);
        const int rbeg = strip * p.rows_per_strip;
        const int rend = min(rbeg + p.rows_per_strip, p.rows);
        if (p.vec) focal_strip<CPT, true, DO_SLOPE, DO_D8, SM, false>(p, wx * 64 * CPT, lane, rbeg, rend);
        else focal_strip<CPT, false, DO_SLOPE, DO_D8, SM, false>(p, wx * 64 * CPT, lane, rbeg, rend);
    }
}

template <int CPT, bool DO_SLOPE, bool DO_D8, int SM>
__global__ __launch_bounds__(256) void focal_batch_kernel(const FocalParams *__restrict__ tiles,
                                                          const unsigned *__restrict__ first, int ntiles,
                                                          unsigned *__restrict__ queue, unsigned total)
{
    focal_batch_body<CPT, DO_SLOPE, DO_D8, SM>(tiles, first, ntiles, queue, total);
}

// The fused Horn slope + D8 batch at 4 columns per lane runs at four waves
// per SIMD, like the single-raster fused kernel.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void focal_batch_kernel_fused(
    const FocalParams *__restrict__ tiles, const unsigned *__restrict__ first, int ntiles,
    unsigned *__restrict__ queue, unsigned total)
{
    focal_batch_body<4, true, true, true>(tiles, first, ntiles, queue, total);
}

// Zeroes the batch queue head (a kernel rather than a memset, so that the
// sequence can be captured in a HIP graph and replayed).
__global__ void focal_queue_reset(unsigned *queue) { queue[0] = 0; }

// Recomputes, with the reference-order FP64 code, every window holding Inf
// or |z| >= 2^40 in the queued strips (width: columns per strip).  If the
// queue overflowed it scans every raster instead.  The last block to finish
// resets the queue for the next call on the stream.  `tiles` is NULL for a
// single raster (then `one` describes it).
template <bool DO_SLOPE, bool DO_D8>
__global__ __launch_bounds__(256) void focal_fixup_kernel(FocalParams one, const FocalParams *__restrict__ tiles,
                                                          int ntiles, int width)
{
    const unsigned *cnt = one.fix_count;
    const unsigned n = __atomic_load_n(&cnt[0], __ATOMIC_RELAXED);
    auto fix = [&](const FocalParams &p, int r, int c) {
        if (r < 1 || r > p.rows - 2 || c < 1 || c > p.cols - 2) return;
        double d[9];
        bool nd = false, nd8 = false, sp = false;
        for (int k = 0; k < 9; k++) {
            const float f = p.dem[(long long)(r + k / 3 - 1) * p.pitch + (c + k % 3 - 1)];
            nd |= in_interval(f, p.nd_lo, p.nd_hi);
            nd8 |= in_interval(f, p.nd8_lo, p.nd8_hi);
            sp |= !(__builtin_fabsf(f) < 0x1p40f);
            d[k] = (double)f;
        }
        if (!sp) return;
        const long long o = (long long)r * p.pitch + c;
        if (DO_D8 && !nd8) {
            unsigned v = d8_exact(d);
            if (p.dirtype == 8) v = v ? (unsigned)c_lut8.v[v] : 255u;
            p.code[o] = (uint8_t)v;
        }
        if (DO_SLOPE && !nd) p.slope[o] = slope_unit((float)slope_generic(d, p.method, p.cs), p.unit);
    };
    if (n <= one.fix_cap) {
        for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
            const uint4 e = one.fix_list[k];
            const FocalParams &p = tiles ? tiles[e.x] : one;
            const int r0 = (int)e.y, nr = (int)e.z - (int)e.y, c0 = (int)e.w;
            for (int t = threadIdx.x; t < nr * width; t += blockDim.x) fix(p, r0 + t / width, c0 + t % width);
        }
    } else {
        const unsigned stride = gridDim.x * blockDim.x;
        const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
        for (int t = 0; t < (tiles ? ntiles : 1); t++) {
            const FocalParams &p = tiles ? tiles[t] : one;
            const long long cells = (long long)(p.rows - 2) * (p.cols - 2);
            for (long long k = tid; k < cells; k += stride) fix(p, 1 + (int)(k / (p.cols - 2)), 1 + (int)(k % (p.cols - 2)));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned *c = one.fix_count;
        const unsigned done = atomicAdd(&c[1], 1u);
        if (done == gridDim.x - 1) {
            c[0] = 0;
            c[1] = 0;
            __threadfence();
        }
    }
}

// The fixup pass runs after every focal launch and its queue is nearly
// always empty: 64 workgroups cost a few microseconds less per step than
// 256 (their grid-stride loops cover any queue).
constexpr int kFixupBlocks = 64;
constexpr unsigned kFixupCap = 1u << 14;  // queued strips

struct LaunchCfg {
    int cpt;
    int rows_per_strip;
};

template <bool DO_SLOPE, bool DO_D8, int SM>
static void launch_t(const FocalParams &p, const LaunchCfg &cfg, bool vec, hipStream_t s)
{
    const int threads = 256;
    dim3 block(threads);
    const int cols_per_block = threads * cfg.cpt;
    dim3 grid((p.cols + cols_per_block - 1) / cols_per_block, (p.rows + cfg.rows_per_strip - 1) / cfg.rows_per_strip);
    if (cfg.cpt == 4) {
        if constexpr (DO_SLOPE && DO_D8 && SM == 1) {
            if (p.mask) {
                if (vec) hipLaunchKernelGGL((focal_kernel_fused_mask<4, true>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((focal_kernel_fused_mask<4, false>), grid, block, 0, s, p);
            } else {
                if (vec) hipLaunchKernelGGL((focal_kernel_fused<4, true>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((focal_kernel_fused<4, false>), grid, block, 0, s, p);
            }
        } else {
            if (vec) hipLaunchKernelGGL((focal_kernel<4, true, DO_SLOPE, DO_D8, SM>), grid, block, 0, s, p);
            else hipLaunchKernelGGL((focal_kernel<4, false, DO_SLOPE, DO_D8, SM>), grid, block, 0, s, p);
        }
    } else {
        if (vec) hipLaunchKernelGGL((focal_kernel<2, true, DO_SLOPE, DO_D8, SM>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((focal_kernel<2, false, DO_SLOPE, DO_D8, SM>), grid, block, 0, s, p);
    }
}

static void launch_fixup(const FocalParams &p, const FocalParams *tiles, int ntiles, bool do_slope, bool do_d8,
                         int cpt, hipStream_t s)
{
    dim3 block(256), grid(kFixupBlocks);
    const int width = 64 * cpt;
    if (do_slope && do_d8) hipLaunchKernelGGL((focal_fixup_kernel<true, true>), grid, block, 0, s, p, tiles, ntiles, width);
    else if (do_slope) hipLaunchKernelGGL((focal_fixup_kernel<true, false>), grid, block, 0, s, p, tiles, ntiles, width);
    else hipLaunchKernelGGL((focal_fixup_kernel<false, true>), grid, block, 0, s, p, tiles, ntiles, width);
}

// Slope mode of the kernels: 1 the Horn fast path (FD), 2 the other
// methods' fast path, 0 every cell in FP64 with the reference's divisions.
// The fast paths form the scaled sum of squares in FP64 and take the root in
// FP32; with |z| < 2^40 (larger values go per cell) and 2^-20 <= cs <= 2^40
// it stays below 2^127 (fast_scale is 0 outside that range).
static int slope_mode(const FocalParams &p)
{
    if (!(p.fast_scale > 0.0)) return 0;
    return p.method == PARGO_SLOPE_FD ? 1 : 2;
}

static void launch_focal_h(const FocalParams &p, const LaunchCfg &cfg, bool vec, bool do_slope, bool do_d8,
                           int sm, hipStream_t s)
{
    if (do_slope && do_d8) {
        if (sm == 1) launch_t<true, true, 1>(p, cfg, vec, s);
        else if (sm == 2) launch_t<true, true, 2>(p, cfg, vec, s);
        else launch_t<true, true, 0>(p, cfg, vec, s);
    } else if (do_slope) {
        if (sm == 1) launch_t<true, false, 1>(p, cfg, vec, s);
        else if (sm == 2) launch_t<true, false, 2>(p, cfg, vec, s);
        else launch_t<true, false, 0>(p, cfg, vec, s);
    } else {
        launch_t<false, true, 1>(p, cfg, vec, s);
    }
}

static void launch_focal(const FocalParams &p, const LaunchCfg &cfg, bool vec, bool do_slope, bool do_d8,
                         hipStream_t s)
{
    launch_focal_h(p, cfg, vec, do_slope, do_d8, slope_mode(p), s);
}

static LaunchCfg default_cfg(int rows)
{
    LaunchCfg c;
    // 4 columns per lane and 16-row strips measured best for the fused
    // kernel on 16384^2 (MI355X, 4 waves per SIMD): 0.495 ms against 0.500
    // at 32 rows, 0.504 at 12, 0.52-0.58 at 48-128 rows (fewer, longer work
    // items leave a ragged tail); the extra halo rows cost less than that.
    c.cpt = 4;
    c.rows_per_strip = 16;
    if (const char *e = getenv("PARGO_FOCAL_CPT")) c.cpt = atoi(e) == 2 ? 2 : 4;
    if (const char *e = getenv("PARGO_FOCAL_ROWS")) {
        int v = atoi(e);
        if (v >= 1 && v <= 4096) c.rows_per_strip = v;
    }
    (void)rows;
    return c;
}

// The fixup queue: two counters and kFixupCap strip entries, zeroed once at
// allocation; every fixup launch leaves the counters at zero again.
static int ensure_fixup(pargo_ctx *ctx)
{
    if (ctx->fix_mem) return PARGO_OK;
    const size_t bytes = 64 * sizeof(unsigned) + (size_t)kFixupCap * sizeof(uint4);
    hipError_t e = dev_malloc(&ctx->fix_mem, bytes, "fixup queue");
    if (e != hipSuccess)
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc(%zu) for the fixup queue: %s%s", bytes, hipGetErrorString(e),
                    injected_note());
    e = hipMemsetAsync(ctx->fix_mem, 0, 64 * sizeof(unsigned), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, PARGO_E_DEVICE, "clearing the fixup queue: %s", hipGetErrorString(e));
    return PARGO_OK;
}

static void attach_fixup_at(void *mem, FocalParams &p)
{
    p.fix_count = (unsigned *)mem;
    p.fix_list = (uint4 *)((unsigned *)mem + 64);
    p.fix_cap = kFixupCap;
}

static void attach_fixup(pargo_ctx *ctx, FocalParams &p) { attach_fixup_at(ctx->fix_mem, p); }

// The fixup queue of pargo_slope_d8_noalloc, in a file-scope device array
// instead of memory the context allocates: the same two counters (in a
// 256-byte header) and kFixupCap strip entries, 262,400 bytes per device,
// held until the process ends.  Every call on a device shares it.  Host code
// never uses the array's name as an address: resolve_noalloc_scratch keeps
// the device address hipGetSymbolAddress gives for the context's device.
constexpr size_t kNoallocScratchBytes = 64 * sizeof(unsigned) + (size_t)kFixupCap * sizeof(uint4);
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

int resolve_noalloc_scratch(pargo_ctx *ctx)
{
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(ctx, PARGO_E_DEVICE, "hipGetSymbolAddress of the static fixup queue on device %d: %s",
                    ctx->device, hipGetErrorString(e));
    ctx->noalloc_scratch = ptr;
    return PARGO_OK;
}

// Zeroes the two queue counters (a kernel rather than a memset, so that a
// caller may capture the call in a HIP graph).
__global__ void fixup_counters_reset(unsigned *c)
{
    c[0] = 0;
    c[1] = 0;
}

// Validates the request and fills the kernel parameters.
static int prepare(pargo_ctx *ctx, const pargo_grid *g, const float *dem, int method, int unit, int dirtype,
                   float *slope_out, uint8_t *code_out, uint32_t *mask_out, FocalParams &p, bool &vec)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!g || !dem) return fail(ctx, PARGO_E_INVALID, "null grid or DEM pointer");
    if (g->rows < 3 || g->cols < 3)
        return fail(ctx, PARGO_E_INVALID,
                    "the dimensions (%d %d) are smaller than the 3x3 neighbourhood", g->rows, g->cols);
    const long long pitch = g->pitch ? g->pitch : g->cols;
    if (pitch < g->cols) return fail(ctx, PARGO_E_INVALID, "pitch %lld < cols %d", pitch, g->cols);
    if ((long long)g->rows * pitch > (1LL << 40)) return fail(ctx, PARGO_E_INVALID, "raster too large");
    if (method < PARGO_SLOPE_FD || method > PARGO_SLOPE_TFDW)
        return fail(ctx, PARGO_E_INVALID, "unknown slope method %d", method);
    if (unit != PARGO_UNIT_TAN && unit != PARGO_UNIT_DEGREES) return fail(ctx, PARGO_E_INVALID, "unknown unit %d", unit);
    if (code_out && !(dirtype == 8 || dirtype == 256))
        return fail(ctx, PARGO_E_INVALID, "please input right parameter: dirType must be 8 or 256, got %d", dirtype);
    if (!(g->cellsize > 0) || !std::isfinite(g->cellsize)) return fail(ctx, PARGO_E_INVALID, "cell size must be > 0");
    const float csf = (float)g->cellsize;
    if (!(csf > 0) || !std::isfinite(csf)) return fail(ctx, PARGO_E_INVALID, "cell size not representable as float");
    p = FocalParams{};
    p.dem = dem;
    p.slope = slope_out;
    p.code = code_out;
    p.mask = mask_out;
    p.pitch = pitch;
    p.rows = g->rows;
    p.cols = g->cols;
    p.mask_words = (g->cols + 31) / 32;
    nodata_interval(g->nodata, 0.0000001, &p.nd_lo, &p.nd_hi);
    nodata_interval(g->nodata, 0.000001, &p.nd8_lo, &p.nd8_hi);
    {
        const bool e1 = p.nd_lo > p.nd_hi, e2 = p.nd8_lo > p.nd8_hi;
        if (e1 && e2) {
            p.ndu_lo = INFINITY;
            p.ndu_hi = -INFINITY;
        } else {
            p.ndu_lo = e1 ? p.nd8_lo : (e2 ? p.nd_lo : std::min(p.nd_lo, p.nd8_lo));
            p.ndu_hi = e1 ? p.nd8_hi : (e2 ? p.nd_hi : std::max(p.nd_hi, p.nd8_hi));
        }
    }
    p.slope_nodata = (float)g->nodata;
    p.cs = (double)csf;
    p.inv64cs2 = 1.0 / (64.0 * (double)csf * (double)csf);
    p.fast_scale = 0.0;
    if (p.cs >= 0x1p-20 && p.cs <= 0x1p40) {
        if (method == PARGO_SLOPE_MD) {
            p.fast_scale = 1.0 / (p.cs * kSqrt2);
        } else {
            const double k = method == PARGO_SLOPE_FFD    ? 4.0
                             : method == PARGO_SLOPE_SD   ? 1.0
                             : method == PARGO_SLOPE_SFD  ? 2.0
                             : method == PARGO_SLOPE_TFD  ? 6.0
                             : method == PARGO_SLOPE_TFDW ? 4.0 + 0x1.6a09e667f3bcdp+1
                                                          : 8.0;
            const double kcs = k * p.cs;
            p.fast_scale = 1.0 / (kcs * kcs);
        }
    }
    p.method = method;
    p.unit = unit;
    p.dirtype = code_out ? dirtype : 0;
    p.mask_from_d8 = code_out ? 1 : 0;
    auto aligned = [](const void *ptr, size_t a) { return ((uintptr_t)ptr % a) == 0; };
    vec = aligned(dem, 4) && (!slope_out || aligned(slope_out, 4));
    p.vec = vec ? 1 : 0;
    return PARGO_OK;
}

static int run_focal(pargo_ctx *ctx, const pargo_grid *g, const float *dem, int method, int unit, int dirtype,
                     float *slope_out, uint8_t *code_out, uint32_t *mask_out, const char *name, bool noalloc = false)
{
    FocalParams p;
    bool vec = false;
    int rc = prepare(ctx, g, dem, method, unit, dirtype, slope_out, code_out, mask_out, p, vec);
    if (rc) return rc;
    if (!slope_out && !code_out && !mask_out) return PARGO_OK;
    rc = bind(ctx);
    if (rc) return rc;
    LaunchCfg cfg = default_cfg(g->rows);
    p.rows_per_strip = cfg.rows_per_strip;
    const bool do_slope = slope_out != nullptr;
    const bool do_d8 = code_out != nullptr;
    if (!do_slope && !do_d8) return fail(ctx, PARGO_E_INVALID, "%s: a NoData mask needs a slope or code output", name);
    if (noalloc) {
        // the static queue: no allocation; its counters may hold what an
        // interrupted earlier call left, so they are cleared first
        if (!ctx->noalloc_scratch) return fail(ctx, PARGO_E_DEVICE, "%s: the static fixup queue is unresolved", name);
        attach_fixup_at(ctx->noalloc_scratch, p);
        hipLaunchKernelGGL(fixup_counters_reset, dim3(1), dim3(1), 0, ctx->stream, p.fix_count);
    } else {
        rc = ensure_fixup(ctx);
        if (rc) return rc;
        attach_fixup(ctx, p);
    }
    launch_focal(p, cfg, vec, do_slope, do_d8, ctx->stream);
    rc = check_launch(ctx, name);
    if (rc) return rc;
    launch_fixup(p, nullptr, 1, do_slope, do_d8, cfg.cpt, ctx->stream);
    return check_launch(ctx, name);
}

// ---- batches ------------------------------------------------------------

struct BatchPlan {
    std::vector<unsigned char> key;  // tiles + parameters the plan was built for
    void *d_mem = nullptr;
    FocalParams *d_tiles = nullptr;
    unsigned *d_first = nullptr;
    unsigned *d_queue = nullptr;
    std::vector<FocalParams> h_tiles;  // original order, for per-tile launches
    unsigned total = 0;
    int ntiles = 0;
    int cpt = 4;
    int blocks = 0;
    bool do_slope = false, do_d8 = false;
    int sm = 0;  // slope mode shared by every tile (slope_mode)
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
};

void free_batch_plan(pargo_ctx *ctx)
{
    BatchPlan *b = ctx->batch;
    if (!b) return;
    hipStreamSynchronize(ctx->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->d_mem) hipFree(b->d_mem);
    delete b;
    ctx->batch = nullptr;
}

template <int CPT, bool DO_SLOPE, bool DO_D8, int SM>
static int batch_blocks(pargo_ctx *ctx)
{
    int per_cu = 0;
    hipError_t e;
    if constexpr (CPT == 4 && DO_SLOPE && DO_D8 && SM == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, focal_batch_kernel_fused, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, focal_batch_kernel<CPT, DO_SLOPE, DO_D8, SM>, 256, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 2;
    const int cus = ctx->num_cu > 0 ? ctx->num_cu : 256;
    return per_cu * cus;
}

template <int CPT, bool DO_SLOPE, bool DO_D8, int SM>
static void launch_batch_t(const BatchPlan &b, hipStream_t s)
{
    if constexpr (CPT == 4 && DO_SLOPE && DO_D8 && SM == 1)
        hipLaunchKernelGGL(focal_batch_kernel_fused, dim3(b.blocks), dim3(256), 0, s, (const FocalParams *)b.d_tiles,
                           (const unsigned *)b.d_first, b.ntiles, b.d_queue, b.total);
    else
        hipLaunchKernelGGL((focal_batch_kernel<CPT, DO_SLOPE, DO_D8, SM>), dim3(b.blocks), dim3(256), 0, s,
                           (const FocalParams *)b.d_tiles, (const unsigned *)b.d_first, b.ntiles, b.d_queue, b.total);
}

template <int CPT>
static void launch_batch_c(const BatchPlan &b, hipStream_t s)
{
    if (b.do_slope && b.do_d8) {
        if (b.sm == 1) launch_batch_t<CPT, true, true, 1>(b, s);
        else if (b.sm == 2) launch_batch_t<CPT, true, true, 2>(b, s);
        else launch_batch_t<CPT, true, true, 0>(b, s);
    } else if (b.do_slope) {
        if (b.sm == 1) launch_batch_t<CPT, true, false, 1>(b, s);
        else if (b.sm == 2) launch_batch_t<CPT, true, false, 2>(b, s);
        else launch_batch_t<CPT, true, false, 0>(b, s);
    } else {
        launch_batch_t<CPT, false, true, 1>(b, s);
    }
}

template <int CPT>
static int blocks_c(pargo_ctx *ctx, const BatchPlan &b)
{
    if (b.do_slope && b.do_d8)
        return b.sm == 1   ? batch_blocks<CPT, true, true, 1>(ctx)
               : b.sm == 2 ? batch_blocks<CPT, true, true, 2>(ctx)
                           : batch_blocks<CPT, true, true, 0>(ctx);
    if (b.do_slope)
        return b.sm == 1   ? batch_blocks<CPT, true, false, 1>(ctx)
               : b.sm == 2 ? batch_blocks<CPT, true, false, 2>(ctx)
                           : batch_blocks<CPT, true, false, 0>(ctx);
    return batch_blocks<CPT, false, true, 1>(ctx);
}

// The persistent sequence: queue reset, batch kernel, fixup.
static void enqueue_batch(const BatchPlan &b, hipStream_t s)
{
    hipLaunchKernelGGL(focal_queue_reset, dim3(1), dim3(1), 0, s, b.d_queue);
    if (b.cpt == 2) launch_batch_c<2>(b, s);
    else launch_batch_c<4>(b, s);
    launch_fixup(b.h_tiles[0], b.d_tiles, b.ntiles, b.do_slope, b.do_d8, b.cpt, s);
}

static int build_batch(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype, int cpt,
                       int rows_per_strip, BatchPlan **out)
{
    if (!tiles || ntiles < 1) return fail(ctx, PARGO_E_INVALID, "empty batch");
    if (ntiles > (1 << 24)) return fail(ctx, PARGO_E_INVALID, "too many tiles");
    std::vector<unsigned char> key(sizeof(pargo_tile) * (size_t)ntiles + 7 * sizeof(int));
    std::memcpy(key.data(), tiles, sizeof(pargo_tile) * (size_t)ntiles);
    const int extra[7] = {ntiles, method, unit, dirtype, cpt, rows_per_strip, ctx->batch_order};
    std::memcpy(key.data() + sizeof(pargo_tile) * (size_t)ntiles, extra, sizeof(extra));
    if (ctx->batch && ctx->batch->key == key) {
        *out = ctx->batch;
        return PARGO_OK;
    }
    free_batch_plan(ctx);
    int rc = ensure_fixup(ctx);
    if (rc) return rc;
    alloc_point("batch plan (host)");
    BatchPlan *b = new (std::nothrow) BatchPlan();
    if (!b) return fail(ctx, PARGO_E_NOMEM, "out of host memory");
    b->key = std::move(key);
    b->cpt = cpt;
    b->ntiles = ntiles;
    b->h_tiles.resize(ntiles);
    const bool do_slope = tiles[0].slope_out != nullptr, do_d8 = tiles[0].code_out != nullptr;
    if (!do_slope && !do_d8) {
        delete b;
        return fail(ctx, PARGO_E_INVALID, "batch tiles have no outputs");
    }
    for (int t = 0; t < ntiles; t++) {
        if ((tiles[t].slope_out != nullptr) != do_slope || (tiles[t].code_out != nullptr) != do_d8) {
            delete b;
            return fail(ctx, PARGO_E_INVALID, "tile %d: all tiles must request the same outputs", t);
        }
        bool vec = false;
        rc = prepare(ctx, &tiles[t].grid, tiles[t].dem, method, unit, dirtype, tiles[t].slope_out, tiles[t].code_out,
                     nullptr, b->h_tiles[t], vec);
        if (rc) {
            delete b;
            return rc;
        }
        attach_fixup(ctx, b->h_tiles[t]);
        b->h_tiles[t].rows_per_strip = rows_per_strip;
        b->h_tiles[t].waves_x = (tiles[t].grid.cols + 64 * cpt - 1) / (64 * cpt);
    }
    b->do_slope = do_slope;
    b->do_d8 = do_d8;
    // One slope path for the whole batch, in every mode: a fast path only
    // when every tile's cell size allows it (slope_mode), so that the
    // per-tile, persistent and graph modes give the same bits.
    b->sm = slope_mode(b->h_tiles[0]);
    for (int t = 1; t < ntiles; t++)
        if (slope_mode(b->h_tiles[t]) != b->sm) b->sm = 0;
    // device order, heaviest first by the load model of ctx->batch_order:
    // cells (the intensity-ratio load with equal NoData and valid weights,
    // which is this kernel's cost: a NoData cell costs what a valid one
    // does), per-tile kernel times measured in one pass, or the given order
    std::vector<int> order(ntiles);
    for (int t = 0; t < ntiles; t++) order[t] = t;
    if (ctx->batch_order == PARGO_ORDER_MEASURED && (int)ctx->batch_load.size() == ntiles) {
        const std::vector<double> &ld = ctx->batch_load;
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return ld[a] > ld[c]; });
    } else if (ctx->batch_order != PARGO_ORDER_ROWS) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
            return (long long)tiles[a].grid.rows * tiles[a].grid.cols >
                   (long long)tiles[c].grid.rows * tiles[c].grid.cols;
        });
    }
    std::vector<FocalParams> dev(ntiles);
    std::vector<unsigned> first(ntiles);
    unsigned long long total = 0;
    for (int i = 0; i < ntiles; i++) {
        dev[i] = b->h_tiles[order[i]];
        dev[i].tile_id = (unsigned)i;
        first[i] = (unsigned)total;
        const int strips = (dev[i].rows + rows_per_strip - 1) / rows_per_strip;
        total += (unsigned long long)strips * (unsigned long long)dev[i].waves_x;
    }
    if (total >= (1ull << 32)) {
        delete b;
        return fail(ctx, PARGO_E_INVALID, "batch has too many work items");
    }
    b->total = (unsigned)total;
    const size_t bt = sizeof(FocalParams) * ntiles, bf = sizeof(unsigned) * ntiles;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    hipError_t e = dev_malloc(&b->d_mem, al(bt) + al(bf) + 256, "batch plan");
    if (e != hipSuccess) {
        delete b;
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc for the batch plan: %s%s", hipGetErrorString(e), injected_note());
    }
    b->d_tiles = (FocalParams *)b->d_mem;
    b->d_first = (unsigned *)((char *)b->d_mem + al(bt));
    b->d_queue = (unsigned *)((char *)b->d_mem + al(bt) + al(bf));
    e = hipMemcpyAsync(b->d_tiles, dev.data(), bt, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_first, first.data(), bf, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipFree(b->d_mem);
        delete b;
        return fail(ctx, PARGO_E_DEVICE, "uploading the batch plan: %s", hipGetErrorString(e));
    }
    b->blocks = cpt == 2 ? blocks_c<2>(ctx, *b) : blocks_c<4>(ctx, *b);
    ctx->batch = b;
    *out = b;
    return PARGO_OK;
}

static int run_batch(pargo_ctx *ctx, BatchPlan *b, int mode, int rows_per_strip)
{
    const hipStream_t s = ctx->stream;
    if (mode == PARGO_BATCH_PER_TILE) {
        LaunchCfg cfg{b->cpt, rows_per_strip};
        for (int t = 0; t < b->ntiles; t++) {
            const FocalParams &p = b->h_tiles[t];
            launch_focal_h(p, cfg, p.vec != 0, b->do_slope, b->do_d8, b->sm, s);
            launch_fixup(p, nullptr, 1, b->do_slope, b->do_d8, b->cpt, s);
        }
        return check_launch(ctx, "pargo_batch_slope_d8 (per tile)");
    }
    if (mode == PARGO_BATCH_PERSISTENT) {
        enqueue_batch(*b, s);
        return check_launch(ctx, "pargo_batch_slope_d8 (persistent)");
    }
    if (!b->exec) {
        // capture on a private stream so that a caller's legacy default
        // stream can be used for everything else
        hipStream_t cs;
        PARGO_HIP(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            enqueue_batch(*b, cs);
            e = hipStreamEndCapture(cs, &b->graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        hipStreamDestroy(cs);
        if (e != hipSuccess) return fail(ctx, PARGO_E_DEVICE, "capturing the batch graph: %s", hipGetErrorString(e));
    }
    PARGO_HIP(ctx, hipGraphLaunch(b->exec, s));
    return PARGO_OK;
}

}  // namespace pargo

extern "C" {

int pargo_slope(pargo_ctx *ctx, const pargo_grid *grid, const float *dem, int method, int unit, float *slope_out,
                uint32_t *nodata_mask_out)
{
    if (ctx && !slope_out) return pargo::fail(ctx, PARGO_E_INVALID, "pargo_slope: null slope output");
    return pargo::run_focal(ctx, grid, dem, method, unit, 256, slope_out, nullptr, nodata_mask_out, "pargo_slope");
}

int pargo_d8(pargo_ctx *ctx, const pargo_grid *grid, const float *dem, int dirtype, uint8_t *code_out,
             uint32_t *nodata_mask_out)
{
    if (ctx && !code_out) return pargo::fail(ctx, PARGO_E_INVALID, "pargo_d8: null code output");
    return pargo::run_focal(ctx, grid, dem, PARGO_SLOPE_FD, PARGO_UNIT_TAN, dirtype, nullptr, code_out,
                            nodata_mask_out, "pargo_d8");
}

int pargo_slope_d8(pargo_ctx *ctx, const pargo_grid *grid, const float *dem, int method, int unit, int dirtype,
                   float *slope_out, uint8_t *code_out, uint32_t *nodata_mask_out)
{
    return pargo::run_focal(ctx, grid, dem, method, unit, dirtype, slope_out, code_out, nodata_mask_out,
                            "pargo_slope_d8");
}

int pargo_slope_d8_noalloc(pargo_ctx *ctx, const pargo_grid *grid, const float *dem, int method, int unit,
                           int dirtype, float *slope_out, uint8_t *code_out, uint32_t *nodata_mask_out)
{
    return pargo::run_focal(ctx, grid, dem, method, unit, dirtype, slope_out, code_out, nodata_mask_out,
                            "pargo_slope_d8_noalloc", true);
}

int pargo_batch_slope_d8(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                         int mode)
{
    if (!ctx) return PARGO_E_INVALID;
    if (mode < PARGO_BATCH_PER_TILE || mode > PARGO_BATCH_GRAPH)
        return pargo::fail(ctx, PARGO_E_INVALID, "unknown batch mode %d", mode);
    return pargo::guard(ctx, "pargo_batch_slope_d8", [&]() -> int {
        int rc = pargo::bind(ctx);
        if (rc) return rc;
        pargo::BatchPlan *b = nullptr;
        rc = pargo::build_batch(ctx, tiles, ntiles, method, unit, dirtype, ctx->batch_cpt, ctx->batch_rows, &b);
        if (rc) return rc;
        return pargo::run_batch(ctx, b, mode, ctx->batch_rows);
    });
}

static int batch_autotune(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                          int *cols_per_lane, int *rows_per_strip)
{
    if (!ctx) return PARGO_E_INVALID;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    const int cpts[2] = {2, 4}, rps[4] = {16, 32, 64, 128};
    float best = INFINITY;
    int bc = ctx->batch_cpt, br = ctx->batch_rows;
    hipEvent_t e0, e1;
    PARGO_HIP(ctx, hipEventCreate(&e0));
    PARGO_HIP(ctx, hipEventCreate(&e1));
    for (int c : cpts)
        for (int r : rps) {
            pargo::BatchPlan *b = nullptr;
            rc = pargo::build_batch(ctx, tiles, ntiles, method, unit, dirtype, c, r, &b);
            if (rc) break;
            for (int k = 0; k < 2 && !rc; k++) rc = pargo::run_batch(ctx, b, PARGO_BATCH_PERSISTENT, r);  // warm-up
            if (rc) break;
            // the best of three timings of two runs: shapes 10 % apart were
            // picked in either order by a single timing
            float ms = INFINITY;
            for (int t = 0; t < 3; t++) {
                hipEventRecord(e0, ctx->stream);
                for (int k = 0; k < 2; k++) pargo::run_batch(ctx, b, PARGO_BATCH_PERSISTENT, r);
                hipEventRecord(e1, ctx->stream);
                hipEventSynchronize(e1);
                float m = 0;
                hipEventElapsedTime(&m, e0, e1);
                ms = fminf(ms, m);
            }
            if (ms < best) {
                best = ms;
                bc = c;
                br = r;
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    ctx->batch_cpt = bc;
    ctx->batch_rows = br;
    ctx->batch_tuned = 1;
    if (cols_per_lane) *cols_per_lane = bc;
    if (rows_per_strip) *rows_per_strip = br;
    return PARGO_OK;
}

int pargo_batch_autotune(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                         int *cols_per_lane, int *rows_per_strip)
{
    return pargo::guard(ctx, "pargo_batch_autotune", [&]() -> int {
        return batch_autotune(ctx, tiles, ntiles, method, unit, dirtype, cols_per_lane, rows_per_strip);
    });
}

static int batch_order(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                       int source)
{
    if (!ctx) return PARGO_E_INVALID;
    if (source != PARGO_ORDER_ROWS && source != PARGO_ORDER_ESTIMATE && source != PARGO_ORDER_MEASURED)
        return pargo::fail(ctx, PARGO_E_INVALID, "unknown batch order source %d", source);
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    ctx->batch_load.clear();
    if (source == PARGO_ORDER_MEASURED) {
        // the preliminary experiment: one per-tile pass with an event pair
        // around every tile's launches
        pargo::BatchPlan *b = nullptr;
        rc = pargo::build_batch(ctx, tiles, ntiles, method, unit, dirtype, ctx->batch_cpt, ctx->batch_rows, &b);
        if (rc) return rc;
        pargo::alloc_point("batch tile loads (host)");
        std::vector<hipEvent_t> ev(2 * (size_t)ntiles, nullptr);
        hipError_t e = hipSuccess;
        for (size_t i = 0; i < ev.size() && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
        const pargo::LaunchCfg cfg{b->cpt, ctx->batch_rows};
        for (int t = 0; t < ntiles && e == hipSuccess; t++) {
            const pargo::FocalParams &p = b->h_tiles[t];
            e = hipEventRecord(ev[2 * t], ctx->stream);
            pargo::launch_focal_h(p, cfg, p.vec != 0, b->do_slope, b->do_d8, b->sm, ctx->stream);
            pargo::launch_fixup(p, nullptr, 1, b->do_slope, b->do_d8, b->cpt, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[2 * t + 1], ctx->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        std::vector<double> load((size_t)ntiles, 0.0);
        for (int t = 0; t < ntiles && e == hipSuccess; t++) {
            float ms = 0;
            e = hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]);
            load[t] = ms;
        }
        for (hipEvent_t x : ev)
            if (x) hipEventDestroy(x);
        if (e != hipSuccess) return pargo::fail(ctx, PARGO_E_DEVICE, "timing the batch tiles: %s", hipGetErrorString(e));
        ctx->batch_load = std::move(load);
    }
    ctx->batch_order = source;
    pargo::free_batch_plan(ctx);  // the next batch call builds the queue in the new order
    return PARGO_OK;
}

int pargo_batch_order(pargo_ctx *ctx, const pargo_tile *tiles, int ntiles, int method, int unit, int dirtype,
                      int source)
{
    return pargo::guard(ctx, "pargo_batch_order", [&]() -> int {
        return batch_order(ctx, tiles, ntiles, method, unit, dirtype, source);
    });
}

}  // extern "C"
