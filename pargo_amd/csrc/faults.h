// faults.h -- allocation-failure injection and the exception guard of the
// C ABI (host code only; not part of the C ABI).
//
// PARGO_FAIL_ALLOC="<n>" (or "<n>:<tag>", the tag only making a new value)
// makes the n-th allocation point the library reaches, counted from 1 since
// the variable last changed value, fail: a host point throws std::bad_alloc,
// a device point returns hipErrorOutOfMemory.  Every C-ABI entry point turns
// either into PARGO_E_NOMEM with a message naming the point, so a test can
// walk n = 1, 2, ... over an entry point's allocations and check that each
// failure is reported and leaves the context usable.  Unset (the normal
// case), a point costs one getenv.
#pragma once

#include <cstddef>
#include <new>
#include <stdexcept>

#include "../../include/pargo.h"

namespace pargo {

int fail(pargo_ctx *ctx, int code, const char *fmt, ...);

// True when this allocation point is the one PARGO_FAIL_ALLOC selects.
bool alloc_should_fail(const char *site);

// A host allocation point: throws std::bad_alloc when selected.
inline void alloc_point(const char *site)
{
    if (alloc_should_fail(site)) throw std::bad_alloc();
}

// Name of the last point made to fail on this thread, cleared by the call
// ("" if none since the last call).
const char *take_injected_site();

// Runs an entry point's body; a std::bad_alloc becomes PARGO_E_NOMEM and
// any other exception PARGO_E_INVALID, each with a message on ctx (or on the
// thread's context-free message), so no exception crosses the C ABI.
template <class F>
int guard(pargo_ctx *ctx, const char *name, F &&body)
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        const char *site = take_injected_site();
        if (*site) return fail(ctx, PARGO_E_NOMEM, "%s: out of host memory (injected at %s)", name, site);
        return fail(ctx, PARGO_E_NOMEM, "%s: out of host memory", name);
    } catch (const std::exception &e) {
        return fail(ctx, PARGO_E_INVALID, "%s: %s", name, e.what());
    } catch (...) {
        return fail(ctx, PARGO_E_INVALID, "%s: unexpected error", name);
    }
}

}  // namespace pargo
