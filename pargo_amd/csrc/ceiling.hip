// ceiling.hip -- measurement helper, not part of the product ABI
// (libpargo_ceiling.so, loaded by bench.py only).
//
// Streaming kernels with the HBM access pattern of the fused slope+D8 pass
// and no arithmetic: per cell 4 B of float32 read, 4 B of float32 and 1 B of
// code written, rows of 16384 cells in 16-row strips, 4 columns per lane,
// dwordx4 loads and stores.  What they reach is what this memory system
// delivers for that traffic mix, the practical ceiling of the fused pass.
//
// Variants (PARGO_CEIL_*):
//   STRIP       4 waves side by side per workgroup, one row strip each, like
//               the focal kernel; 4 rows loaded before their stores
//   STRIP_HALO  the same, also reading the row above and below each strip
//               (the focal kernel's halo rows)
//   FLAT        one grid-stride loop over the cells, 8 workgroups per CU
//   FLAT_NT     the same with non-temporal stores and 4 workgroups per CU:
//               the whole GPU sweeps one narrow window of the raster (the
//               fastest form measured for this traffic mix)
//   STRIP_NT    STRIP with non-temporal stores
//   READ        STRIP reading only (4 B/cell)
//   WRITE       STRIP writing only (5 B/cell)
//   ROW1        STRIP_NT with one-row strips and their two halo rows: every
//               workgroup reads rows r-1..r+1 and writes row r, so the GPU
//               sweeps a band of about 64 rows (the fastest form measured;
//               the halo rows are L2 hits)
//   STACK8      8 waves stacked on 256 columns walk a 64-row band 8 rows at
//               a time and pass their rows to each other through LDS, the
//               other narrow-band form a stencil could take
//
// Every access is in bounds for rows x cols with pitch >= cols and
// cols % 1024 == 0 (checked by the entry point): a wave covers 256 columns,
// a workgroup 1024.
#include <hip/hip_runtime.h>

#include <cstdint>

#define PARGO_CEIL_STRIP 0
#define PARGO_CEIL_STRIP_HALO 1
#define PARGO_CEIL_FLAT 2
#define PARGO_CEIL_STRIP_NT 3
#define PARGO_CEIL_READ 4
#define PARGO_CEIL_WRITE 5
#define PARGO_CEIL_FLAT_NT 6
#define PARGO_CEIL_ROW1 7
#define PARGO_CEIL_STACK8 8

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_codes(float4 v)
{
    // one byte per cell, taken from the float's exponent byte
    return __builtin_amdgcn_perm(__builtin_amdgcn_perm(__float_as_uint(v.w), __float_as_uint(v.z), 0x0c0c0703u),
                                 __builtin_amdgcn_perm(__float_as_uint(v.y), __float_as_uint(v.x), 0x0c0c0703u),
                                 0x05040100u);
}

template <bool READ, bool WRITE, bool NT, bool HALO>
__global__ __launch_bounds__(256) void strip_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                    uint8_t *__restrict__ dst8, int rows, long long pitch, int rps,
                                                    unsigned *__restrict__ sink)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned c0 = (blockIdx.x * 4u + (unsigned)wave) * 256u + (unsigned)lane * 4u;
    const int rbeg = (int)blockIdx.y * rps;
    const int rend = min(rbeg + rps, rows);
    const int rlo = HALO ? max(rbeg - 1, 0) : rbeg;
    const int rhi = HALO ? min(rend + 1, rows) : rend;
    float acc = 0.0f;
    for (int r = rlo; r < rhi; r += 4) {
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int rr = min(r + k, rhi - 1);
            if (READ) v[k] = *reinterpret_cast<const float4 *>(src + (long long)rr * pitch + c0);
            else v[k] = make_float4((float)rr, (float)c0, 1.0f, 2.0f);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int rr = r + k;
            // halo rows are read, never written
            if (rr >= rhi || rr < rbeg || rr >= rend) continue;
            if (WRITE) {
                float4 *d = reinterpret_cast<float4 *>(dst + (long long)rr * pitch + c0);
                unsigned *d8 = reinterpret_cast<unsigned *>(dst8 + (long long)rr * pitch + c0);
                if (NT) {
                    const f4v t = {v[k].x, v[k].y, v[k].z, v[k].w};
                    __builtin_nontemporal_store(t, reinterpret_cast<f4v *>(d));
                    __builtin_nontemporal_store(pack_codes(v[k]), d8);
                } else {
                    *d = v[k];
                    *d8 = pack_codes(v[k]);
                }
            } else {
                acc += v[k].x + v[k].y + v[k].z + v[k].w;
            }
        }
    }
    // keeps the read-only variant's loads alive; never true for finite data
    if (!WRITE && acc == -1.0f) sink[0] = 1u;
}

template <bool NT>
__global__ __launch_bounds__(256) void flat_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                   uint8_t *__restrict__ dst8, int rows, int cols, long long pitch)
{
    const unsigned per_row = (unsigned)cols / 4u;  // float4 groups per row
    const unsigned long long total = (unsigned long long)rows * per_row;
    for (unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (unsigned long long)gridDim.x * blockDim.x) {
        const long long r = (long long)(g / per_row);
        const unsigned c = (unsigned)(g % per_row) * 4u;
        const float4 v = *reinterpret_cast<const float4 *>(src + r * pitch + c);
        if (NT) {
            const f4v t = {v.x, v.y, v.z, v.w};
            __builtin_nontemporal_store(t, reinterpret_cast<f4v *>(dst + r * pitch + c));
            __builtin_nontemporal_store(pack_codes(v), reinterpret_cast<unsigned *>(dst8 + r * pitch + c));
        } else {
            *reinterpret_cast<float4 *>(dst + r * pitch + c) = v;
            *reinterpret_cast<unsigned *>(dst8 + r * pitch + c) = pack_codes(v);
        }
    }
}

// STACK8: workgroup = 8 waves on one 256-column segment; each step the
// waves load rows r0..r0+7, publish them in LDS, read the rows above and
// below after a barrier (the 3x3 window's rows; their sum is kept alive
// with a zero weight) and store their own row.
__global__ __launch_bounds__(512) void stack8_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                     uint8_t *__restrict__ dst8, int rows, long long pitch,
                                                     int band, int nseg)
{
    __shared__ float4 srow[8][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned seg = blockIdx.x % (unsigned)nseg, b = blockIdx.x / (unsigned)nseg;
    const unsigned c0 = seg * 256u + (unsigned)lane * 4u;
    const int rb = (int)b * band, re = min(rb + band, rows);
    for (int r0 = rb; r0 < re; r0 += 8) {
        const int r = min(r0 + w, re - 1);
        float4 v = *reinterpret_cast<const float4 *>(src + (long long)r * pitch + c0);
        srow[w][lane] = v;
        __syncthreads();
        const float4 a = srow[(w + 7) & 7][lane], c = srow[(w + 1) & 7][lane];
        v.x += a.x * 0.0f + c.x * 0.0f;
        v.y += a.y * 0.0f + c.y * 0.0f;
        __syncthreads();
        if (r0 + w < re) {
            const f4v t = {v.x, v.y, v.z, v.w};
            __builtin_nontemporal_store(t, reinterpret_cast<f4v *>(dst + (long long)r * pitch + c0));
            __builtin_nontemporal_store(pack_codes(v), reinterpret_cast<unsigned *>(dst8 + (long long)r * pitch + c0));
        }
    }
}

}  // namespace

extern "C" {

// Launches one variant on `stream` (a hipStream_t).  src, dst: rows x pitch
// float32; dst8: rows x pitch bytes; sink: one device uint32 (READ only).
// Returns 0, or -1 for a shape the kernels do not cover.
int pargo_ceiling_launch(int variant, const float *src, float *dst, uint8_t *dst8, unsigned *sink, int rows, int cols,
                         long long pitch, int rows_per_strip, void *stream)
{
    if (rows < 1 || cols < 1024 || cols % 1024 || pitch < cols || pitch % 4 || rows_per_strip < 1) return -1;
    if (!src || !dst || !dst8 || !sink) return -1;
    const hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)(cols / 1024), (unsigned)((rows + rows_per_strip - 1) / rows_per_strip));
    const dim3 block(256);
    switch (variant) {
    case PARGO_CEIL_STRIP:
        hipLaunchKernelGGL((strip_kernel<true, true, false, false>), grid, block, 0, s, src, dst, dst8, rows, pitch,
                           rows_per_strip, sink);
        break;
    case PARGO_CEIL_STRIP_HALO:
        hipLaunchKernelGGL((strip_kernel<true, true, false, true>), grid, block, 0, s, src, dst, dst8, rows, pitch,
                           rows_per_strip, sink);
        break;
    case PARGO_CEIL_STRIP_NT:
        hipLaunchKernelGGL((strip_kernel<true, true, true, false>), grid, block, 0, s, src, dst, dst8, rows, pitch,
                           rows_per_strip, sink);
        break;
    case PARGO_CEIL_READ:
        hipLaunchKernelGGL((strip_kernel<true, false, false, false>), grid, block, 0, s, src, dst, dst8, rows, pitch,
                           rows_per_strip, sink);
        break;
    case PARGO_CEIL_WRITE:
        hipLaunchKernelGGL((strip_kernel<false, true, false, false>), grid, block, 0, s, src, dst, dst8, rows, pitch,
                           rows_per_strip, sink);
        break;
    case PARGO_CEIL_ROW1:
        hipLaunchKernelGGL((strip_kernel<true, true, true, true>), dim3((unsigned)(cols / 1024), (unsigned)rows), block,
                           0, s, src, dst, dst8, rows, pitch, 1, sink);
        break;
    case PARGO_CEIL_STACK8: {
        const int nseg = cols / 256, band = 64;
        hipLaunchKernelGGL(stack8_kernel, dim3((unsigned)(nseg * ((rows + band - 1) / band))), dim3(512), 0, s, src,
                           dst, dst8, rows, pitch, band, nseg);
        break;
    }
    case PARGO_CEIL_FLAT:
    case PARGO_CEIL_FLAT_NT: {
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu < 1)
            cu = 256;
        if (variant == PARGO_CEIL_FLAT)
            hipLaunchKernelGGL(flat_kernel<false>, dim3((unsigned)(cu * 8)), block, 0, s, src, dst, dst8, rows, cols,
                               pitch);
        else
            hipLaunchKernelGGL(flat_kernel<true>, dim3((unsigned)(cu * 4)), block, 0, s, src, dst, dst8, rows, cols,
                               pitch);
        break;
    }
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
