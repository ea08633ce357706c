// faults.cpp -- PARGO_FAIL_ALLOC allocation-failure injection (see faults.h).
#include "faults.h"

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

namespace pargo {

namespace {
std::mutex g_mu;
std::string g_seen;     // the variable's value the count refers to
long g_count = 0;       // allocation points reached since g_seen was set
thread_local std::string t_site;
}  // namespace

bool alloc_should_fail(const char *site)
{
    const char *v = std::getenv("PARGO_FAIL_ALLOC");
    if (!v || !*v) return false;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_seen != v) {
        g_seen = v;
        g_count = 0;
    }
    const long n = std::strtol(v, nullptr, 10);
    if (n < 1) return false;
    if (++g_count != n) return false;
    t_site = site;
    return true;
}

const char *take_injected_site()
{
    static thread_local std::string out;
    out.swap(t_site);
    t_site.clear();
    return out.c_str();
}

}  // namespace pargo
