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
// D8 decisions are made on exactly the doubles the reference compares:
// drops RN(z4 - zi) for cardinal neighbours and RN(RN(z4 - zi) / sqrt(2.0))
// for diagonal ones, max over them and 0, then the 1e-6 tests.  A drop
// shared by two neighbouring cells is computed once (RN(a-b) = -RN(b-a),
// and the division routine is odd), and the division by sqrt(2.0) is the
// correctly rounded Markstein sequence of focal_common.h.  Windows that hold
// NaN, Inf or |z| >= 2^40 take a per-cell path with IEEE division instead.
//
// The Horn numerators are formed in FP64 (exact for float32 input of
// moderate range), the square root in FP32: relative error below 5e-7.
#include <cmath>

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
    float slope_nodata;   // value written to NoData slope cells
    double inv64cs2;      // 1 / (64 cs^2) for the Horn fast path
    double cs;            // float-rounded cell size
    int method;
    int unit;
    int dirtype;          // 0: no D8, 8 or 256
    int mask_from_d8;     // NoData mask reports the D8 NoData set
};

template <int CPT>
struct RawRow {
    float v[CPT];
    float hl, hr;  // outer halo, valid on lanes 0 and 63
};

template <int CPT, bool VEC>
__device__ __forceinline__ void load_raw(const FocalParams &p, int r, int c0, int lane, RawRow<CPT> &o)
{
#pragma unroll
    for (int k = 0; k < CPT; k++) o.v[k] = 0.f;
    o.hl = 0.f;
    o.hr = 0.f;
    if (r < 0 || r >= p.rows) return;  // wave-uniform
    const float *rowp = p.dem + (long long)r * p.pitch;
    if (VEC && c0 + CPT <= p.cols) {
        if constexpr (CPT == 4) {
            float4 t = *reinterpret_cast<const float4 *>(rowp + c0);
            o.v[0] = t.x; o.v[1] = t.y; o.v[2] = t.z; o.v[3] = t.w;
        } else if constexpr (CPT == 2) {
            float2 t = *reinterpret_cast<const float2 *>(rowp + c0);
            o.v[0] = t.x; o.v[1] = t.y;
        } else {
#pragma unroll
            for (int k = 0; k < CPT; k++) o.v[k] = rowp[c0 + k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < CPT; k++)
            if (c0 + k < p.cols) o.v[k] = rowp[c0 + k];
    }
    if (lane == 0 && c0 >= 1 && c0 - 1 < p.cols) o.hl = rowp[c0 - 1];
    if (lane == 63 && c0 + CPT < p.cols) o.hr = rowp[c0 + CPT];
}

// Values of columns c0-1 .. c0+CPT as doubles, with NoData and "special"
// flags per position (bit i for position i).
template <int CPT>
struct Row {
    double z[CPT + 2];
    unsigned nd;   // slope NoData interval
    unsigned nd8;  // D8 NoData interval
    unsigned sp;   // NaN, Inf or |z| >= 2^40
};

template <int CPT>
__device__ __forceinline__ void finish_row(const FocalParams &p, const RawRow<CPT> &raw, int lane, Row<CPT> &o)
{
    float f[CPT + 2];
    float left = __shfl_up(raw.v[CPT - 1], 1, 64);
    float right = __shfl_down(raw.v[0], 1, 64);
    f[0] = (lane == 0) ? raw.hl : left;
    f[CPT + 1] = (lane == 63) ? raw.hr : right;
#pragma unroll
    for (int k = 0; k < CPT; k++) f[k + 1] = raw.v[k];
    unsigned nd = 0, nd8 = 0, sp = 0;
#pragma unroll
    for (int i = 0; i < CPT + 2; i++) {
        o.z[i] = (double)f[i];
        nd |= (unsigned)in_interval(f[i], p.nd_lo, p.nd_hi) << i;
        nd8 |= (unsigned)in_interval(f[i], p.nd8_lo, p.nd8_hi) << i;
        // NaN, Inf and |z| >= 2^40 take the per-cell FP64 path: the fast
        // path's FP32 magnitude stage would overflow for such values.
        sp |= (unsigned)(!(__builtin_fabsf(f[i]) < 0x1p40f)) << i;
    }
    o.nd = nd;
    o.nd8 = nd8;
    o.sp = sp;
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

__device__ __forceinline__ float slope_unit(float s, int unit)
{
    return unit == PARGO_UNIT_DEGREES ? atanf(s) * 57.295779513082323f : s;
}

template <int CPT, bool VEC, bool DO_SLOPE, bool DO_D8, bool HORN>
__global__ __launch_bounds__(256) void focal_kernel(FocalParams p)
{
    const int lane = threadIdx.x & 63;
    const int c0 = (blockIdx.x * blockDim.x + threadIdx.x) * CPT;
    const int rbeg = blockIdx.y * p.rows_per_strip;
    const int rend = min(rbeg + p.rows_per_strip, p.rows);

    RawRow<CPT> ra, rb, rc;
    load_raw<CPT, VEC>(p, rbeg - 1, c0, lane, ra);
    load_raw<CPT, VEC>(p, rbeg, c0, lane, rb);
    load_raw<CPT, VEC>(p, rbeg + 1, c0, lane, rc);  // first "down" row
    RawRow<CPT> pf;
    load_raw<CPT, VEC>(p, rbeg + 2, c0, lane, pf);  // prefetch

    Row<CPT> up, mid;
    finish_row<CPT>(p, ra, lane, up);
    finish_row<CPT>(p, rb, lane, mid);

    // Drops between the rows above and at the output row (see header):
    //   Vp[j]  = RN(up[j] - mid[j])              j = 0..CPT-1  (position j+1)
    //   SEp[i] = G(up[i] - mid[i+1])             i = 0..CPT
    //   SWp[i] = G(up[i] - mid[i-1])             i = 1..CPT+1
    double Vp[CPT], SEp[CPT + 1], SWp[CPT + 2];
    if (DO_D8) {
#pragma unroll
        for (int j = 0; j < CPT; j++) Vp[j] = up.z[j + 1] - mid.z[j + 1];
#pragma unroll
        for (int i = 0; i <= CPT; i++) SEp[i] = div_sqrt2(up.z[i] - mid.z[i + 1]);
#pragma unroll
        for (int i = 1; i <= CPT + 1; i++) SWp[i] = div_sqrt2(up.z[i] - mid.z[i - 1]);
        SWp[0] = 0.0;
    }
    // Horn row sums (z[i-1] + z[i+1]) + 2 z[i] of the rows above / at.
    double rsU[CPT], rsM[CPT];
    if (DO_SLOPE && HORN) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            rsU[j] = __builtin_fma(2.0, up.z[j + 1], up.z[j] + up.z[j + 2]);
            rsM[j] = __builtin_fma(2.0, mid.z[j + 1], mid.z[j] + mid.z[j + 2]);
        }
    }

    for (int r = rbeg; r < rend; r++) {
        Row<CPT> dn;
        finish_row<CPT>(p, rc, lane, dn);
        rc = pf;
        load_raw<CPT, VEC>(p, r + 3, c0, lane, pf);

        const bool row_border = (r == 0) || (r == p.rows - 1);
        const unsigned ndw = up.nd | mid.nd | dn.nd;
        const unsigned nd8w = up.nd8 | mid.nd8 | dn.nd8;
        const unsigned spw = up.sp | mid.sp | dn.sp;

        float sl[CPT];
        unsigned cd[CPT];
        unsigned ndbits = 0;  // bit j: cell j is NoData (for the mask)

        // ---- D8 --------------------------------------------------------
        double V[CPT], SE[CPT + 1], SW[CPT + 2];
        if (DO_D8) {
            double H[CPT + 1];
#pragma unroll
            for (int i = 0; i <= CPT; i++) H[i] = mid.z[i] - mid.z[i + 1];
#pragma unroll
            for (int j = 0; j < CPT; j++) V[j] = mid.z[j + 1] - dn.z[j + 1];
#pragma unroll
            for (int i = 0; i <= CPT; i++) SE[i] = div_sqrt2(mid.z[i] - dn.z[i + 1]);
#pragma unroll
            for (int i = 1; i <= CPT + 1; i++) SW[i] = div_sqrt2(mid.z[i] - dn.z[i - 1]);
            SW[0] = 0.0;
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const int i = j + 1;
                const double E = H[i], W = -H[i - 1], S = V[j], N = -Vp[j];
                const double dSE = SE[i], dNW = -SEp[i - 1], dSW = SW[i], dNE = -SWp[i + 1];
                double M = fmax(fmax(fmax(E, W), fmax(S, N)), fmax(fmax(dSE, dNW), fmax(dSW, dNE)));
                M = fmax(M, 0.0);
                unsigned c = 0;
                c = c + c + (unsigned)(fabs(dNE - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(N - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(dNW - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(W - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(dSW - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(S - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(dSE - M) < kD8Eps);
                c = c + c + (unsigned)(fabs(E - M) < kD8Eps);
                cd[j] = (M < kD8Eps) ? 0u : c;
            }
        }

        // ---- slope -------------------------------------------------------
        if (DO_SLOPE) {
            if (HORN) {
                double rsD[CPT];
#pragma unroll
                for (int j = 0; j < CPT; j++) rsD[j] = __builtin_fma(2.0, dn.z[j + 1], dn.z[j] + dn.z[j + 2]);
                double col[CPT + 2];
#pragma unroll
                for (int i = 0; i < CPT + 2; i++) col[i] = __builtin_fma(2.0, mid.z[i], dn.z[i]) + up.z[i];
#pragma unroll
                for (int j = 0; j < CPT; j++) {
                    const double nx = col[j + 2] - col[j];
                    const double ny = rsU[j] - rsD[j];
                    const double q = __builtin_fma(nx, nx, ny * ny) * p.inv64cs2;
                    sl[j] = __builtin_amdgcn_sqrtf((float)q);
                }
#pragma unroll
                for (int j = 0; j < CPT; j++) {
                    rsU[j] = rsM[j];
                    rsM[j] = rsD[j];
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

        // ---- windows with special values: per-cell IEEE path ---------------
        if (__builtin_expect(__any(spw != 0), 0)) {
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                if ((spw >> j) & 7u) {
                    const int i = j + 1;
                    const double d[9] = {up.z[i - 1], up.z[i], up.z[i + 1], mid.z[i - 1], mid.z[i],
                                         mid.z[i + 1], dn.z[i - 1], dn.z[i], dn.z[i + 1]};
                    if (DO_D8) cd[j] = d8_exact(d);
                    if (DO_SLOPE) sl[j] = (float)slope_generic(d, p.method, p.cs);
                }
            }
        }

        // ---- NoData, border, units, stores --------------------------------
        const bool row_ok = r < p.rows;
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = c0 + j;
            const bool border = row_border || c == 0 || c >= p.cols - 1;
            const bool nds = border || ((ndw >> j) & 7u);
            const bool nd8 = border || ((nd8w >> j) & 7u);
            if (DO_SLOPE) sl[j] = nds ? p.slope_nodata : slope_unit(sl[j], p.unit);
            if (DO_D8) {
                unsigned v = nd8 ? 0u : cd[j];
                if (p.dirtype == 8) v = (nd8 || v == 0) ? 255u : (unsigned)c_lut8.v[v];
                cd[j] = v;
            }
            const bool m = (DO_D8 && p.mask_from_d8) ? nd8 : nds;
            ndbits |= (unsigned)(m && c < p.cols) << j;
        }
        if (row_ok) {
            const long long base = (long long)r * p.pitch + c0;
            if (DO_SLOPE) {
                if (VEC && c0 + CPT <= p.cols) {
                    if constexpr (CPT == 4) {
                        *reinterpret_cast<float4 *>(p.slope + base) = make_float4(sl[0], sl[1], sl[2], sl[3]);
                    } else if constexpr (CPT == 2) {
                        *reinterpret_cast<float2 *>(p.slope + base) = make_float2(sl[0], sl[1]);
                    } else {
#pragma unroll
                        for (int j = 0; j < CPT; j++) p.slope[base + j] = sl[j];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < CPT; j++)
                        if (c0 + j < p.cols) p.slope[base + j] = sl[j];
                }
            }
            if (DO_D8) {
                if (VEC && c0 + CPT <= p.cols) {
                    if constexpr (CPT == 4) {
                        const unsigned w = cd[0] | (cd[1] << 8) | (cd[2] << 16) | (cd[3] << 24);
                        *reinterpret_cast<unsigned *>(p.code + base) = w;
                    } else if constexpr (CPT == 2) {
                        *reinterpret_cast<unsigned short *>(p.code + base) = (unsigned short)(cd[0] | (cd[1] << 8));
                    } else {
#pragma unroll
                        for (int j = 0; j < CPT; j++) p.code[base + j] = (uint8_t)cd[j];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < CPT; j++)
                        if (c0 + j < p.cols) p.code[base + j] = (uint8_t)cd[j];
                }
            }
        }
        if (p.mask) {
            // Gather the CPT-bit groups of 32/CPT lanes into one word.
            constexpr int LPW = 32 / CPT;
            unsigned w = ndbits << (CPT * (lane % LPW));
#pragma unroll
            for (int o = 1; o < LPW; o <<= 1) w |= __shfl_xor(w, o, 64);
            if (row_ok && (lane % LPW) == 0 && c0 < p.cols)
                p.mask[(long long)r * p.mask_words + (c0 >> 5)] = w;
        }

        // ---- rotate --------------------------------------------------------
        if (DO_D8) {
#pragma unroll
            for (int j = 0; j < CPT; j++) Vp[j] = V[j];
#pragma unroll
            for (int i = 0; i <= CPT; i++) SEp[i] = SE[i];
#pragma unroll
            for (int i = 0; i <= CPT + 1; i++) SWp[i] = SW[i];
        }
        up = mid;
        mid = dn;
    }
}

struct LaunchCfg {
    int cpt;
    int rows_per_strip;
};

template <bool DO_SLOPE, bool DO_D8, bool HORN>
static void launch_t(const FocalParams &p, const LaunchCfg &cfg, bool vec, hipStream_t s)
{
    const int threads = 256;
    dim3 block(threads);
    const int cols_per_block = threads * cfg.cpt;
    dim3 grid((p.cols + cols_per_block - 1) / cols_per_block, (p.rows + cfg.rows_per_strip - 1) / cfg.rows_per_strip);
    if (cfg.cpt == 4) {
        if (vec) hipLaunchKernelGGL((focal_kernel<4, true, DO_SLOPE, DO_D8, HORN>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((focal_kernel<4, false, DO_SLOPE, DO_D8, HORN>), grid, block, 0, s, p);
    } else {
        if (vec) hipLaunchKernelGGL((focal_kernel<2, true, DO_SLOPE, DO_D8, HORN>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((focal_kernel<2, false, DO_SLOPE, DO_D8, HORN>), grid, block, 0, s, p);
    }
}

static void launch_focal(const FocalParams &p, const LaunchCfg &cfg, bool vec, bool do_slope, bool do_d8,
                         hipStream_t s)
{
    // The Horn fast path forms (nx^2 + ny^2) / (64 cs^2) in FP64 and takes
    // the root in FP32; with |z| < 2^40 (larger values go per cell) and
    // 2^-20 <= cs <= 2^40 that quotient stays below 2^127.
    const bool horn = p.method == PARGO_SLOPE_FD && p.cs >= 0x1p-20 && p.cs <= 0x1p40;
    if (do_slope && do_d8) {
        if (horn) launch_t<true, true, true>(p, cfg, vec, s);
        else launch_t<true, true, false>(p, cfg, vec, s);
    } else if (do_slope) {
        if (horn) launch_t<true, false, true>(p, cfg, vec, s);
        else launch_t<true, false, false>(p, cfg, vec, s);
    } else {
        launch_t<false, true, true>(p, cfg, vec, s);
    }
}

static LaunchCfg default_cfg(int rows)
{
    LaunchCfg c;
    c.cpt = 4;
    c.rows_per_strip = 64;
    if (const char *e = getenv("PARGO_FOCAL_CPT")) c.cpt = atoi(e) == 2 ? 2 : 4;
    if (const char *e = getenv("PARGO_FOCAL_ROWS")) {
        int v = atoi(e);
        if (v >= 1 && v <= 4096) c.rows_per_strip = v;
    }
    (void)rows;
    return c;
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
    p.slope_nodata = (float)g->nodata;
    p.cs = (double)csf;
    p.inv64cs2 = 1.0 / (64.0 * (double)csf * (double)csf);
    p.method = method;
    p.unit = unit;
    p.dirtype = code_out ? dirtype : 0;
    p.mask_from_d8 = code_out ? 1 : 0;
    auto aligned = [](const void *ptr, size_t a) { return ((uintptr_t)ptr % a) == 0; };
    vec = (pitch % 4 == 0) && aligned(dem, 16) && (!slope_out || aligned(slope_out, 16)) &&
          (!code_out || aligned(code_out, 4));
    return PARGO_OK;
}

static int run_focal(pargo_ctx *ctx, const pargo_grid *g, const float *dem, int method, int unit, int dirtype,
                     float *slope_out, uint8_t *code_out, uint32_t *mask_out, const char *name)
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
    launch_focal(p, cfg, vec, do_slope, do_d8, ctx->stream);
    return check_launch(ctx, name);
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

}  // extern "C"
