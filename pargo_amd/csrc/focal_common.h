// focal_common.h -- device helpers shared by the focal (3x3) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pargo {

// sqrt(2.0) as the reference evaluates it (glibc sqrt is correctly rounded)
// and the double nearest to its reciprocal.
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
constexpr double kInvSqrt2 = 0x1.6a09e667f3bccp-1;
// The D8 epsilon 0.000001 (apps/hydrology/flowdird8/d8Operator.cpp:3).
constexpr double kD8Eps = 0.000001;

// FP32 D8 filter (focal.hip): RN32(1/sqrt 2) and the thresholds that keep
// the filter's decisions provably those of the FP64 tests.
constexpr float kInvSqrt2f = 0x1.6a09e6p-1f;
constexpr float kD8SureF = 1.0001e-6f;        // K
constexpr float kD8Shrink = 0x1.ffffep-1f;    // 1 - 2^-20
constexpr float kD8FloorF = 0x1.0c6898p-20f;  // the float just below RN32(0.9999e-6)
constexpr float kD8BandMid = 1.5e-6f;         // |M - 1.5e-6| < 0.50025e-6: ambiguous
constexpr float kD8BandHalf = 0.50025e-6f;

// RN(x / kSqrt2) for finite x derived from float32 data (|x| <= 2^129,
// x = 0 or |x| >= 2^-149, so nothing below underflows).
// q0 = RN(x*y) with y = RN(1/s) is within 0.69 ulp of x/s because
// |y*s - 1| = 0.183 * 2^-53; by Markstein's theorem (y within half an ulp
// of 1/s, q0 within one ulp of x/s) r = x - q0*s is exact in one FMA and
// RN(q0 + r*y) is the correctly rounded quotient.  The identity is also
// checked on the host against IEEE division over every binade from 2^-149
// to 2^129 (tests/cpp/gpro_host_test.cpp).
__device__ __forceinline__ double div_sqrt2(double x)
{
    double q0 = x * kInvSqrt2;
    double r = __builtin_fma(-q0, kSqrt2, x);
    return __builtin_fma(r, kInvSqrt2, q0);
}

// Exact float interval test for NoData (see nodata_interval on the host).
__device__ __forceinline__ bool in_interval(float v, float lo, float hi) { return v >= lo && v <= hi; }

// ESRI code -> reference dirType 8 output (median tied window index).
// Built on the host from the reference's rule validDir[len >> 1]
// (d8Operator.cpp:78-88); 255 for code 0 (unused: flat cells are
// resolved before the lookup).
struct D8Lut8 {
    uint8_t v[256];
};

// Workgroups go to the 8 XCDs in turn by linear id.  Renumbers the
// workgroup of a (column block, row strip) grid so that each XCD takes a
// contiguous run of strips, column blocks first: neighbouring strips then
// share their halo rows in that XCD's L2 instead of reading them twice from
// HBM.  A bijection of the grid onto itself.
__device__ __forceinline__ void xcd_strip_order(unsigned &cb, unsigned &strip)
{
    const unsigned gx = gridDim.x, n = gx * gridDim.y, i = blockIdx.y * gx + blockIdx.x;
    const unsigned per = n / 8, rem = n % 8, x = i % 8;
    const unsigned w = x * per + (x < rem ? x : rem) + i / 8;
    cb = w % gx;
    strip = w / gx;
}

}  // namespace pargo
