// ctx.hip -- context, error reporting and memory helpers of the C ABI.
//
// The context replaces what Application::START/END (gpro/application.cpp:5-31)
// set up for MPI: here it opens a HIP device and a stream.  Errors are kept
// per context as text, like the reference's cerr messages, and every entry
// point returns a status code instead of calling exit().
#include <cmath>
#include <cstring>
#include <new>

#include "faults.h"
#include "internal.h"

namespace pargo {

// Failures that happen before a context exists (pargo_create) or with a NULL
// context are reported through this per-thread buffer.
static thread_local char t_error[1024] = "";

int fail(pargo_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    else snprintf(t_error, sizeof(t_error), "%s", buf);
    return code;
}

int bind(pargo_ctx *ctx)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != ctx->device) {
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess)
            return fail(ctx, PARGO_E_DEVICE, "hipSetDevice(%d): %s", ctx->device, hipGetErrorString(e));
    }
    return PARGO_OK;
}

int check_launch(pargo_ctx *ctx, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, PARGO_E_DEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
    return PARGO_OK;
}

hipError_t dev_malloc(void **ptr, size_t bytes, const char *site)
{
    if (alloc_should_fail(site)) {
        *ptr = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(ptr, bytes);
}

const char *injected_note()
{
    const char *s = take_injected_site();
    static thread_local char buf[160];
    if (!*s) return "";
    snprintf(buf, sizeof(buf), " (injected at %s)", s);
    return buf;
}

int ensure_scratch(pargo_ctx *ctx, size_t bytes)
{
    if (ctx->scratch_bytes >= bytes) return PARGO_OK;
    if (ctx->scratch) {
        // The old buffer may still be in use by queued work.
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return fail(ctx, PARGO_E_DEVICE, "hipStreamSynchronize: %s", hipGetErrorString(e));
        hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = dev_malloc(&ctx->scratch, want, "context scratch");
    if (e != hipSuccess)
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc(%zu) for scratch: %s%s", want, hipGetErrorString(e), injected_note());
    ctx->scratch_bytes = want;
    return PARGO_OK;
}

// The reference tests NoData per window value as fabs(d - noData) < eps in
// double (slopeOperator.cpp:68, d8Operator.cpp:46).  d is a float widened to
// double, and |RN(d - noData)| is unimodal in d, so the floats passing the
// test form one interval.  Find its ends by binary search over the floats
// ordered as integers, starting from a passing float next to noData.
static uint32_t float_key(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static float key_float(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

void nodata_interval(double nodata, double eps, float *lo, float *hi, bool inclusive)
{
    auto pass = [&](float f) {
        const double d = std::fabs((double)f - nodata);
        return std::isfinite(f) && (inclusive ? d <= eps : d < eps);
    };
    *lo = 1.0f;
    *hi = 0.0f;  // empty
    if (!std::isfinite(nodata) || !(eps > 0)) return;
    float c = (float)nodata;
    float start = NAN;
    const float cand[3] = {c, std::nextafter(c, -INFINITY), std::nextafter(c, INFINITY)};
    for (float f : cand)
        if (pass(f)) { start = f; break; }
    if (std::isnan(start)) return;
    const uint32_t ks = float_key(start);
    const uint32_t kneg = float_key(-INFINITY), kpos = float_key(INFINITY);
    // Smallest passing key in [kneg, ks]: pass is false below it, true up to ks.
    uint32_t a = kneg, b = ks;
    while (a < b) {
        uint32_t m = a + (b - a) / 2;
        if (pass(key_float(m))) b = m; else a = m + 1;
    }
    *lo = key_float(a);
    // Largest passing key in [ks, kpos].
    a = ks;
    b = kpos;
    while (a < b) {
        uint32_t m = a + (b - a + 1) / 2;
        if (pass(key_float(m))) a = m; else b = m - 1;
    }
    *hi = key_float(a);
}

}  // namespace pargo

using pargo::fail;

extern "C" {

int pargo_abi_version(void) { return PARGO_ABI_VERSION; }

int pargo_create(int device, pargo_ctx **out)
{
    if (!out) return fail(nullptr, PARGO_E_INVALID, "pargo_create: null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return fail(nullptr, PARGO_E_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    if (n <= 0) return fail(nullptr, PARGO_E_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(nullptr, PARGO_E_INVALID, "device %d not in [0, %d)", device, n);
    pargo_ctx *ctx = pargo::alloc_should_fail("context") ? nullptr : new (std::nothrow) pargo_ctx();
    if (!ctx) return fail(nullptr, PARGO_E_NOMEM, "out of host memory%s", pargo::injected_note());
    ctx->device = device;
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, PARGO_E_DEVICE, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
    }
    // Default priority on purpose: no stream priorities anywhere.
    e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, PARGO_E_DEVICE, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
    }
    ctx->stream = ctx->own_stream;
    // the static fixup queue of pargo_slope_d8_noalloc on this device
    if (pargo::resolve_noalloc_scratch(ctx) != PARGO_OK) {
        fail(nullptr, PARGO_E_DEVICE, "%s", ctx->last_error.c_str());
        hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return PARGO_E_DEVICE;
    }
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess) ctx->num_cu = cu;
    *out = ctx;
    return PARGO_OK;
}

int pargo_destroy(pargo_ctx *ctx)
{
    if (!ctx) return PARGO_E_INVALID;
    pargo::bind(ctx);
    if (ctx->own_stream) hipStreamSynchronize(ctx->own_stream);
    if (ctx->stream && ctx->stream != ctx->own_stream) hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) hipFree(ctx->scratch);
    pargo::free_batch_plan(ctx);
    if (ctx->fix_mem) hipFree(ctx->fix_mem);
    if (ctx->switch_event) hipEventDestroy(ctx->switch_event);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return PARGO_OK;
}

const char *pargo_last_error(const pargo_ctx *ctx)
{
    if (!ctx) return pargo::t_error;
    return ctx->last_error.c_str();
}

int pargo_set_stream(pargo_ctx *ctx, void *hip_stream)
{
    if (!ctx) return PARGO_E_INVALID;
    const hipStream_t ns = (hipStream_t)hip_stream;  // NULL is HIP's default stream
    if (ns == ctx->stream) return PARGO_OK;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    // Work already queued on the old stream uses the context's device state
    // (fixup queue, scratch, batch queue): order the new stream after it.
    if (!ctx->switch_event) PARGO_HIP(ctx, hipEventCreateWithFlags(&ctx->switch_event, hipEventDisableTiming));
    PARGO_HIP(ctx, hipEventRecord(ctx->switch_event, ctx->stream));
    PARGO_HIP(ctx, hipStreamWaitEvent(ns, ctx->switch_event, 0));
    ctx->stream = ns;
    return PARGO_OK;
}

void *pargo_own_stream(const pargo_ctx *ctx) { return ctx ? (void *)ctx->own_stream : nullptr; }

int pargo_sync(pargo_ctx *ctx)
{
    if (!ctx) return PARGO_E_INVALID;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PARGO_OK;
}

int pargo_alloc(pargo_ctx *ctx, size_t bytes, void **dptr)
{
    if (!ctx || !dptr) return fail(ctx, PARGO_E_INVALID, "pargo_alloc: null argument");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    hipError_t e = pargo::dev_malloc(dptr, bytes ? bytes : 1, "pargo_alloc");
    if (e != hipSuccess)
        return fail(ctx, PARGO_E_NOMEM, "hipMalloc(%zu): %s%s", bytes, hipGetErrorString(e), pargo::injected_note());
    return PARGO_OK;
}

int pargo_free(pargo_ctx *ctx, void *dptr)
{
    if (!ctx) return PARGO_E_INVALID;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    if (dptr) PARGO_HIP(ctx, hipFree(dptr));
    return PARGO_OK;
}

int pargo_alloc_host(pargo_ctx *ctx, size_t bytes, void **hptr)
{
    if (!ctx || !hptr) return fail(ctx, PARGO_E_INVALID, "pargo_alloc_host: null argument");
    *hptr = nullptr;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    hipError_t e = pargo::alloc_should_fail("pargo_alloc_host") ? hipErrorOutOfMemory
                                                                : hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        *hptr = nullptr;
        return fail(ctx, PARGO_E_NOMEM, "hipHostMalloc(%zu): %s%s", bytes, hipGetErrorString(e), pargo::injected_note());
    }
    return PARGO_OK;
}

int pargo_free_host(pargo_ctx *ctx, void *hptr)
{
    if (!ctx) return PARGO_E_INVALID;
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    if (hptr) PARGO_HIP(ctx, hipHostFree(hptr));
    return PARGO_OK;
}

int pargo_h2d(pargo_ctx *ctx, void *dst, const void *host_src, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!host_src && bytes)) return fail(ctx, PARGO_E_INVALID, "pargo_h2d: null argument");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    PARGO_HIP(ctx, hipMemcpyAsync(dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PARGO_OK;
}

int pargo_d2h(pargo_ctx *ctx, void *host_dst, const void *src, size_t bytes)
{
    if (!ctx || (!src && bytes) || (!host_dst && bytes)) return fail(ctx, PARGO_E_INVALID, "pargo_d2h: null argument");
    int rc = pargo::bind(ctx);
    if (rc) return rc;
    PARGO_HIP(ctx, hipMemcpyAsync(host_dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PARGO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PARGO_OK;
}

}  // extern "C"
