// deviceRaster.cpp -- see deviceRaster.h.
#include "deviceRaster.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <set>

namespace GPRO {

unsigned long long DeviceTransfers::h2d = 0;
unsigned long long DeviceTransfers::d2h = 0;

namespace {

std::set<DeviceRaster *> &live()
{
    static std::set<DeviceRaster *> s;
    return s;
}

// One pinned staging buffer for the process, used in chunks.
constexpr size_t kChunkBytes = (size_t)32 << 20;
void *g_pinned = nullptr;

void *pinned()
{
    if (g_pinned) return g_pinned;
    pargo_ctx *ctx = Application::device();
    if (!ctx) return nullptr;
    if (!deviceOk(pargo_alloc_host(ctx, kChunkBytes, &g_pinned), "pargo_alloc_host(staging)")) g_pinned = nullptr;
    return g_pinned;
}

}  // namespace

bool DeviceRaster::allocate(DeviceForm f, int r, int c, bool withMask)
{
    current = newerThanHost = false;
    if (f == form && r == rows && c == cols && (mask != nullptr) == withMask) return true;
    release();
    pargo_ctx *ctx = Application::device();
    if (!ctx) return false;
    const size_t n = (size_t)r * (size_t)c;
    if (!deviceOk(pargo_alloc(ctx, n * elemBytes(f), &data), "pargo_alloc(device raster)")) return false;
    live().insert(this);
    if (withMask && !deviceOk(pargo_alloc(ctx, (size_t)r * maskWords(c) * sizeof(uint32_t), (void **)&mask),
                              "pargo_alloc(device NoData mask)")) {
        release();
        return false;
    }
    form = f;
    rows = r;
    cols = c;
    return true;
}

void DeviceRaster::release()
{
    pargo_ctx *ctx = Application::current();
    if (ctx && (data || mask)) {
        pargo_sync(ctx);  // queued work may still use the buffers
        if (data) pargo_free(ctx, data);
        if (mask) pargo_free(ctx, mask);
    }
    live().erase(this);
    data = nullptr;
    mask = nullptr;
    form = DeviceForm::None;
    rows = cols = 0;
    current = newerThanHost = false;
}

void DeviceRaster::releaseAll()
{
    const std::vector<DeviceRaster *> all(live().begin(), live().end());
    for (DeviceRaster *d : all) d->release();
    if (g_pinned) {
        if (pargo_ctx *ctx = Application::current()) pargo_free_host(ctx, g_pinned);
        g_pinned = nullptr;
    }
}

namespace staging {

bool uploadF32(const double *src, size_t n, int cols, float *dst, const char *what)
{
    pargo_ctx *ctx = Application::device();
    float *h = (float *)pinned();
    if (!ctx || !h) return false;
    const size_t per = kChunkBytes / sizeof(float);
    for (size_t off = 0; off < n; off += per) {
        const long long m = (long long)std::min(per, n - off);
        long long bad = LLONG_MAX;
#pragma omp parallel for reduction(min : bad) schedule(static)
        for (long long i = 0; i < m; i++) {
            const double v = src[off + i];
            const float f = (float)v;
            if ((double)f != v && !std::isnan(v)) bad = std::min(bad, (long long)off + i);
            h[i] = f;
        }
        if (bad != LLONG_MAX) {
            cerr << "[ERROR] " << what << ": value " << src[bad] << " at row " << bad / cols << ", column "
                 << bad % cols << " is not representable in float32, the device storage type of DEMs" << endl;
            return false;
        }
        if (!deviceOk(pargo_h2d(ctx, dst + off, h, (size_t)m * sizeof(float)), "pargo_h2d(raster)")) return false;
        DeviceTransfers::h2d += (unsigned long long)m * sizeof(float);
    }
    return true;
}

bool uploadF64(const double *src, size_t n, double *dst)
{
    pargo_ctx *ctx = Application::device();
    double *h = (double *)pinned();
    if (!ctx || !h) return false;
    const size_t per = kChunkBytes / sizeof(double);
    for (size_t off = 0; off < n; off += per) {
        const size_t m = std::min(per, n - off);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)m; i++) h[i] = src[off + i];
        if (!deviceOk(pargo_h2d(ctx, dst + off, h, m * sizeof(double)), "pargo_h2d(raster)")) return false;
        DeviceTransfers::h2d += (unsigned long long)m * sizeof(double);
    }
    return true;
}

bool download(const DeviceRaster &d, double nodata, double *dst)
{
    pargo_ctx *ctx = Application::device();
    unsigned char *h = (unsigned char *)pinned();
    if (!ctx || !h) return false;
    const size_t n = (size_t)d.rows * (size_t)d.cols, eb = DeviceRaster::elemBytes(d.form);
    const size_t wpr = DeviceRaster::maskWords(d.cols);
    std::vector<uint32_t> mask;
    if (d.mask) {
        mask.resize((size_t)d.rows * wpr);
        if (!deviceOk(pargo_d2h(ctx, mask.data(), d.mask, mask.size() * sizeof(uint32_t)), "pargo_d2h(mask)"))
            return false;
        DeviceTransfers::d2h += mask.size() * sizeof(uint32_t);
    }
    const unsigned flat = d.dirType == 8 ? PARGO_D8_FLAT_8 : PARGO_D8_FLAT_256;
    const size_t per = kChunkBytes / eb;
    const int cols = d.cols;
    for (size_t off = 0; off < n; off += per) {
        const size_t m = std::min(per, n - off);
        if (!deviceOk(pargo_d2h(ctx, h, (const unsigned char *)d.data + off * eb, m * eb), "pargo_d2h(raster)"))
            return false;
        DeviceTransfers::d2h += m * eb;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)m; i++) {
            const size_t k = off + (size_t)i;
            const bool nd = d.mask && ((mask[(k / cols) * wpr + (k % cols) / 32] >> (k % cols % 32)) & 1u);
            double v;
            if (d.form == DeviceForm::F32) v = (double)((const float *)h)[i];
            else if (d.form == DeviceForm::F64) v = ((const double *)h)[i];
            else v = h[i] == flat ? -1.0 : (double)h[i];
            dst[k] = nd ? nodata : v;
        }
    }
    return true;
}

}  // namespace staging

}  // namespace GPRO
