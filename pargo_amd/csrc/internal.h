// internal.h -- shared definitions of libpargo (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pargo.h"
#include "faults.h"

namespace pargo {
struct BatchPlan;
}

struct pargo_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;  // where work goes (own or caller's)
    // recorded on the old stream at a stream switch; the new stream waits
    // for it, so the context's shared device state (fixup queue, scratch,
    // batch queue) is never used by two streams at once
    hipEvent_t switch_event = nullptr;
    std::string last_error;
    int num_cu = 0;
    // Small reusable device scratch owned by the context (grows on demand).
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *fix_mem = nullptr;  // focal fixup queue (counters + cell list)
    // device address of the static fixup queue of pargo_slope_d8_noalloc on
    // this context's device (resolved once by pargo_create)
    void *noalloc_scratch = nullptr;
    // batched slope/D8: the last plan (with its captured graph) and the
    // launch shape chosen by pargo_batch_autotune
    pargo::BatchPlan *batch = nullptr;
    int batch_cpt = 4;
    int batch_rows = 32;
    int batch_tuned = 0;
    // queue order of batch tiles: PARGO_ORDER_ESTIMATE (cells), _ROWS (as
    // given) or _MEASURED (batch_load, per-tile kernel times of one pass)
    int batch_order = 1;
    std::vector<double> batch_load;
};

namespace pargo {
void free_batch_plan(pargo_ctx *ctx);
// Resolves ctx->noalloc_scratch with ctx->device current.
int resolve_noalloc_scratch(pargo_ctx *ctx);
}

namespace pargo {

// Records a formatted message on the context and returns `code`.
int fail(pargo_ctx *ctx, int code, const char *fmt, ...);

// Makes ctx->device current for the calling thread.
int bind(pargo_ctx *ctx);

// Checks the HIP error state after a launch.
int check_launch(pargo_ctx *ctx, const char *what);

// hipMalloc behind an allocation point (PARGO_FAIL_ALLOC, faults.h): a
// selected point returns hipErrorOutOfMemory without allocating.
hipError_t dev_malloc(void **ptr, size_t bytes, const char *site);
// " (injected at <site>)" after an injected failure, else "" (clears it).
const char *injected_note();

// Grows the context scratch to at least `bytes`.
int ensure_scratch(pargo_ctx *ctx, size_t bytes);

// Float interval [lo, hi] of the float32 values v for which the reference's
// double test fabs((double)v - nodata) < eps (<= eps when inclusive) holds
// (empty when lo > hi).
void nodata_interval(double nodata, double eps, float *lo, float *hi, bool inclusive = false);

}  // namespace pargo

#define PARGO_HIP(ctx, call)                                                            \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess)                                                           \
            return pargo::fail((ctx), PARGO_E_DEVICE, "%s failed: %s", #call,           \
                               hipGetErrorString(_e));                                  \
    } while (0)
