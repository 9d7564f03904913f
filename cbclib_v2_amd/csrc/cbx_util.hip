// Host utilities of the C ABI: the per-thread error message, version,
// profile names and the device-wide scans (hipcub / rocPRIM).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <string>

#include "cbx_host.hpp"

namespace cbx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

template <typename V>
static hipError_t exclusive_scan(const V *in, V *out, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s);
    if (e != hipSuccess) return e;
    void *tmp = nullptr;
    e = hipMalloc(&tmp, bytes ? bytes : 1);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, s);
    hipError_t f = hipFree(tmp);
    return (e != hipSuccess) ? e : f;
}

hipError_t exclusive_scan_i32(const int *in, int *out, int64_t n, hipStream_t s) { return exclusive_scan(in, out, n, s); }
hipError_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s)
{
    return exclusive_scan(in, out, n, s);
}

} // namespace cbx

extern "C" {

const char *cbx_last_error(void) { return cbx::g_last_error.c_str(); }

const char *cbx_version(void) { return "cbx 0.1.0 (gfx950)"; }

int cbx_kernel_id(const char *name)
{
    static const char *names[5] = {"biweight", "gaussian", "parabolic", "rectangular", "triangular"};
    if (!name) return -1;
    for (int i = 0; i < 5; i++)
        if (std::strcmp(name, names[i]) == 0) return i;
    return -1;
}

} // extern "C"
