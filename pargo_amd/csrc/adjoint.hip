// adjoint.hip -- reverse-mode derivative of the Horn slope with respect to
// the elevations (no counterpart in the reference, which has no gradients).
//
// Forward (slopeOperator.cpp:84-86):  s_i = |n_i| / (8 cs),
//   nx_i = sum_d dc*(2-|dr|) z(i+d),  ny_i = sum_d -dr*(2-|dc|) z(i+d).
// With p_i = g_i n_i / (8 cs |n_i|) (zero where slope is NoData, on the
// border or where |n_i| = 0), the adjoint is the correlation with the
// negated (antisymmetric) stencils:
//   dL/dz(r,c) = [Px(c-1) - Px(c+1)] + [Qy(r+1) - Qy(r-1)],
//   Px(k) = px(r-1,k) + 2 px(r,k) + px(r+1,k),
//   Qy(k) = py(k,c-1) + 2 py(k,c) + py(k,c+1).
// Only the DEM and dL/dslope are read; n_i is recomputed (12 bytes per
// cell of HBM traffic: z in, g in, dL/dz out).
//
// Work split as in focal.hip: a lane owns 4 adjacent columns and walks a
// strip of rows; it computes p for its columns plus one halo column on each
// side (so it needs z two columns out), which keeps every wave independent.
#include <cmath>

#include "focal_common.h"
#include "internal.h"

namespace pargo {

namespace {

constexpr int ACPT = 4;       // columns per lane
constexpr int AZW = ACPT + 4; // z columns c0-2 .. c0+5
constexpr int APW = ACPT + 2; // p columns c0-1 .. c0+4

struct AdjParams {
    const float *__restrict__ dem;
    const float *__restrict__ g;
    float *__restrict__ gdem;
    long long pitch;
    int rows, cols, rows_per_strip;
    float nd_lo, nd_hi;
    double inv8cs;
};

__device__ __forceinline__ float lower_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float upper_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ int clampc(int c, int cols) { return c < 0 ? 0 : (c > cols - 1 ? cols - 1 : c); }

struct ZRow {
    double z[AZW];
    unsigned nd;  // bit t: column c0-2+t holds NoData
};

// A row's raw floats: the lane's 4 columns and two halo values (lane 0:
// the two columns before the wave, lane 63: the two after; other lanes
// duplicate their own columns).  Loading is split from the conversion so
// that the next row's loads are in flight while this row is computed.
// Row vectors need only 4-byte alignment: gfx950 runs a global dwordx4
// access at any dword address as one instruction.
typedef float f32x4 __attribute__((ext_vector_type(4), aligned(4)));

struct RawRow {
    float v[ACPT];
    float h0, h1;
};

template <bool VEC, bool TWO_HALO>
__device__ __forceinline__ void load_raw(const float *__restrict__ base, const AdjParams &p, int r, int c0,
                                         int halo_first, bool full, RawRow &o)
{
#pragma unroll
    for (int k = 0; k < ACPT; k++) o.v[k] = 0.f;
    o.h0 = o.h1 = 0.f;
    if (r >= 0 && r < p.rows) {
        const float *rowp = base + (long long)r * p.pitch;
        if (VEC && full) {
            const f32x4 t = *reinterpret_cast<const f32x4 *>(rowp + c0);
            o.v[0] = t.x; o.v[1] = t.y; o.v[2] = t.z; o.v[3] = t.w;
        } else {
#pragma unroll
            for (int k = 0; k < ACPT; k++)
                if (c0 + k < p.cols) o.v[k] = rowp[c0 + k];
        }
        const int a = halo_first;
        o.h0 = (a >= 0 && a < p.cols) ? rowp[a] : 0.f;
        if (TWO_HALO) o.h1 = (a + 1 >= 0 && a + 1 < p.cols) ? rowp[a + 1] : 0.f;
    }
}

// The DEM row as doubles for columns c0-2 .. c0+5 (zeros outside) and its
// NoData bits.  halo loaded from c0-2 (lane 0) or c0+4 (lane 63).
__device__ __forceinline__ void finish_z(const AdjParams &p, const RawRow &w, int lane, ZRow &o)
{
    float f[AZW];
    const float l0 = lower_lane(w.v[2]), l1 = lower_lane(w.v[3]);
    const float r0 = upper_lane(w.v[0]), r1 = upper_lane(w.v[1]);
    f[0] = lane == 0 ? w.h0 : l0;
    f[1] = lane == 0 ? w.h1 : l1;
#pragma unroll
    for (int k = 0; k < ACPT; k++) f[k + 2] = w.v[k];
    f[6] = lane == 63 ? w.h0 : r0;
    f[7] = lane == 63 ? w.h1 : r1;
    unsigned nd = 0;
#pragma unroll
    for (int t = 0; t < AZW; t++) {
        o.z[t] = (double)f[t];
        nd |= (unsigned)in_interval(f[t], p.nd_lo, p.nd_hi) << t;
    }
    o.nd = nd;
}

// dL/dslope of the row, columns c0-1 .. c0+4 (zeros outside the raster);
// halo loaded from c0-1 (lane 0) or c0+4 (lane 63), in h0.
__device__ __forceinline__ void finish_g(const RawRow &w, int lane, float gg[APW])
{
    const float l = lower_lane(w.v[3]), rr = upper_lane(w.v[0]);
    gg[0] = lane == 0 ? w.h0 : l;
#pragma unroll
    for (int k = 0; k < ACPT; k++) gg[k + 1] = w.v[k];
    gg[APW - 1] = lane == 63 ? w.h0 : rr;
}

// p at one column: t is the column's index among c0-1 .. c0+4 (z indices
// t .. t+2 hold its window's columns).
__device__ __forceinline__ void p_at(const AdjParams &p, bool row_ok, unsigned ndw, int c0, int t, const ZRow &a,
                                     const ZRow &b, const ZRow &c, float g, float &px, float &py)
{
    const int col = c0 - 1 + t;
    const bool ok = row_ok && col >= 1 && col <= p.cols - 2 && !((ndw >> t) & 7u);
    const double nx = (__builtin_fma(2.0, b.z[t + 2], a.z[t + 2]) + c.z[t + 2]) -
                      (__builtin_fma(2.0, b.z[t], a.z[t]) + c.z[t]);
    const double ny = (__builtin_fma(2.0, a.z[t + 1], a.z[t]) + a.z[t + 2]) -
                      (__builtin_fma(2.0, c.z[t + 1], c.z[t]) + c.z[t + 2]);
    const double q = __builtin_fma(nx, nx, ny * ny);
    const double f = (double)g * p.inv8cs * __builtin_amdgcn_rsq(q);
    px = (ok && q > 0) ? (float)(f * nx) : 0.f;
    py = (ok && q > 0) ? (float)(f * ny) : 0.f;
}

// Lane i's value of v in lane i+1 (i-1); the lane without a source (0,
// resp. 63) gets `edge`.
__device__ __forceinline__ float from_left(float v, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float v, float edge)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// p of row k (columns c0-1 .. c0+4) from z rows k-1, k, k+1 and g row k.
// Each lane evaluates its own 4 columns; the halo columns are its
// neighbours' (DPP moves), except at the wave's two ends, where lane 0 needs
// column c0-1 and lane 63 column c0+4: those take one more evaluation,
// shared (each lane evaluates one of the two), so 5 per lane instead of 6.
__device__ __forceinline__ void p_row(const AdjParams &p, int k, int c0, int lane, const ZRow &a, const ZRow &b,
                                      const ZRow &c, const float gg[APW], float px[APW], float py[APW])
{
    const bool row_ok = k >= 1 && k <= p.rows - 2;
    const unsigned ndw = a.nd | b.nd | c.nd;
#pragma unroll
    for (int t = 1; t <= ACPT; t++) p_at(p, row_ok, ndw, c0, t, a, b, c, gg[t], px[t], py[t]);
    // the wave-end column: t = 5 in lane 63, t = 0 in every other lane
    // (only lane 0's is used), with z indices selected per lane
    float ex, ey;
    {
        const bool right = lane == 63;
        ZRow ae, be, ce;
#pragma unroll
        for (int u = 0; u < 3; u++) {
            ae.z[u] = right ? a.z[u + 5] : a.z[u];
            be.z[u] = right ? b.z[u + 5] : b.z[u];
            ce.z[u] = right ? c.z[u + 5] : c.z[u];
        }
        const unsigned nde = right ? (ndw >> 5) : ndw;
        p_at(p, row_ok, nde, right ? c0 + 5 : c0, 0, ae, be, ce, right ? gg[APW - 1] : gg[0], ex, ey);
    }
    px[0] = from_left(px[ACPT], ex);
    py[0] = from_left(py[ACPT], ey);
    px[APW - 1] = from_right(px[1], ex);
    py[APW - 1] = from_right(py[1], ey);
}

template <bool VEC>
__global__ __launch_bounds__(256) void slope_adjoint_kernel(AdjParams p)
{
    const int lane = threadIdx.x & 63;
    unsigned cb, strip;
    xcd_strip_order(cb, strip);
    const int c0 = (int)(cb * blockDim.x + threadIdx.x) * ACPT;
    const int rbeg = (int)strip * p.rows_per_strip;
    const int rend = min(rbeg + p.rows_per_strip, p.rows);
    const int wave_c0 = __builtin_amdgcn_readfirstlane(c0);
    const bool full = wave_c0 + 64 * ACPT <= p.cols;

    // halo columns of the z rows (two) and of the g rows (one)
    const int zh = lane == 0 ? c0 - 2 : (lane == 63 ? c0 + ACPT : c0);
    const int gh = lane == 0 ? c0 - 1 : (lane == 63 ? c0 + ACPT : c0);
    // z rows r-2.. and p rows r-1, r for the first output row r = rbeg
    ZRow zA, zB, zC;  // rows r, r+1 kept; third slot takes r+2
    RawRow nz, ng;    // z row r+2 and g row r+1, loaded one iteration ahead
    float pxm[APW], pym[APW], px0[APW], py0[APW], px1[APW], py1[APW], gg[APW];
    {
        ZRow zm2, zm1;
        RawRow w;
        load_raw<VEC, true>(p.dem, p, rbeg - 2, c0, zh, full, w);
        finish_z(p, w, lane, zm2);
        load_raw<VEC, true>(p.dem, p, rbeg - 1, c0, zh, full, w);
        finish_z(p, w, lane, zm1);
        load_raw<VEC, true>(p.dem, p, rbeg, c0, zh, full, w);
        finish_z(p, w, lane, zA);
        load_raw<VEC, true>(p.dem, p, rbeg + 1, c0, zh, full, w);
        finish_z(p, w, lane, zB);
        load_raw<VEC, false>(p.g, p, rbeg - 1, c0, gh, full, w);
        finish_g(w, lane, gg);
        p_row(p, rbeg - 1, c0, lane, zm2, zm1, zA, gg, pxm, pym);
        load_raw<VEC, false>(p.g, p, rbeg, c0, gh, full, w);
        finish_g(w, lane, gg);
        p_row(p, rbeg, c0, lane, zm1, zA, zB, gg, px0, py0);
        load_raw<VEC, true>(p.dem, p, rbeg + 2, c0, zh, full, nz);
        load_raw<VEC, false>(p.g, p, rbeg + 1, c0, gh, full, ng);
    }
    for (int r = rbeg; r < rend; r++) {
        finish_z(p, nz, lane, zC);
        finish_g(ng, lane, gg);
        // the next iteration's rows (past the strip's last need: skipped)
        if (r + 1 < rend) {
            load_raw<VEC, true>(p.dem, p, r + 3, c0, zh, full, nz);
            load_raw<VEC, false>(p.g, p, r + 2, c0, gh, full, ng);
        }
        p_row(p, r + 1, c0, lane, zA, zB, zC, gg, px1, py1);
        float out[ACPT];
        float Px[APW];
#pragma unroll
        for (int t = 0; t < APW; t++) Px[t] = __builtin_fmaf(2.f, px0[t], pxm[t] + px1[t]);
#pragma unroll
        for (int j = 0; j < ACPT; j++) {
            const float gx = Px[j] - Px[j + 2];
            const float qdn = __builtin_fmaf(2.f, py1[j + 1], py1[j] + py1[j + 2]);
            const float qup = __builtin_fmaf(2.f, pym[j + 1], pym[j] + pym[j + 2]);
            out[j] = gx + (qdn - qup);
        }
        float *dst = p.gdem + (long long)r * p.pitch + c0;
        if (VEC && full) {
            // streamed past the caches: the gradient is not read again here
            const f32x4 t = {out[0], out[1], out[2], out[3]};
            __builtin_nontemporal_store(t, reinterpret_cast<f32x4 *>(dst));
        } else {
#pragma unroll
            for (int j = 0; j < ACPT; j++)
                if (c0 + j < p.cols) dst[j] = out[j];
        }
        // rotate
        zA = zB;
        zB = zC;
#pragma unroll
        for (int t = 0; t < APW; t++) {
            pxm[t] = px0[t];
            pym[t] = py0[t];
            px0[t] = px1[t];
            py0[t] = py1[t];
        }
    }
}

}  // namespace

}  // namespace pargo

using pargo::fail;

extern "C" int pargo_slope_adjoint(pargo_ctx *ctx, const pargo_grid *g, const float *dem, const float *grad_slope,
                                   float *grad_dem)
{
    if (!ctx) return PARGO_E_INVALID;
    if (!g || !dem || !grad_slope || !grad_dem) return fail(ctx, PARGO_E_INVALID, "pargo_slope_adjoint: null argument");
    if (g->rows < 3 || g->cols < 3)
        return fail(ctx, PARGO_E_INVALID, "the dimensions (%d %d) are smaller than the 3x3 neighbourhood", g->rows,
                    g->cols);
    const long long pitch = g->pitch ? g->pitch : g->cols;
    if (pitch < g->cols) return fail(ctx, PARGO_E_INVALID, "pitch %lld < cols %d", pitch, g->cols);
    const float csf = (float)g->cellsize;
    if (!(csf > 0) || !std::isfinite(csf)) return fail(ctx, PARGO_E_INVALID, "cell size must be > 0");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    pargo::AdjParams p{};
    p.dem = dem;
    p.g = grad_slope;
    p.gdem = grad_dem;
    p.pitch = pitch;
    p.rows = g->rows;
    p.cols = g->cols;
    p.rows_per_strip = 32;
    pargo::nodata_interval(g->nodata, 0.0000001, &p.nd_lo, &p.nd_hi);
    if (p.nd_lo > p.nd_hi) {
        p.nd_lo = INFINITY;
        p.nd_hi = -INFINITY;
    }
    p.inv8cs = 1.0 / (8.0 * (double)csf);
    auto al = [](const void *q) { return ((uintptr_t)q % 4) == 0; };
    const bool vec = al(dem) && al(grad_slope) && al(grad_dem);
    dim3 block(256);
    dim3 grid((g->cols + 256 * pargo::ACPT - 1) / (256 * pargo::ACPT), (g->rows + p.rows_per_strip - 1) / p.rows_per_strip);
    if (vec) hipLaunchKernelGGL(pargo::slope_adjoint_kernel<true>, grid, block, 0, ctx->stream, p);
    else hipLaunchKernelGGL(pargo::slope_adjoint_kernel<false>, grid, block, 0, ctx->stream, p);
    return pargo::check_launch(ctx, "pargo_slope_adjoint");
}
