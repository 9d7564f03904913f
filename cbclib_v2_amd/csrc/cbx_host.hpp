// Host-side helpers shared by the C-ABI entry points: status codes, the
// thread-local error message, stream-ordered scratch buffers and the hipcub
// scans.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/cbx.h"

namespace cbx {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define CBX_HIP(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return ::cbx::fail(CBX_ERUNTIME, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Device allocation released on scope exit (hipFree waits for the device,
// so a buffer is never released under a kernel that still uses it).
template <typename V>
struct DevBuf {
    V *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    hipError_t alloc(size_t count, hipStream_t)
    {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&ptr, sizeof(V) * count);
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// Optional per-kernel timing with HIP events recorded on the launch stream
// (enabled by cbx_timing_enable); read back by cbx_timing_read.
bool timing_on();
void timing_record(const char *name, hipEvent_t begin, hipEvent_t end);

struct KernelTimer {
    const char *name;
    hipStream_t stream;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(const char *n, hipStream_t s) : name(n), stream(s)
    {
        if (!timing_on()) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        (void)hipEventRecord(a, stream);
    }
    void stop()
    {
        if (!a) return;
        (void)hipEventRecord(b, stream);
        timing_record(name, a, b);
        a = b = nullptr;
    }
    ~KernelTimer() { stop(); }
};

hipError_t exclusive_scan_i32(const int *in, int *out, int64_t n, hipStream_t s);
hipError_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

} // namespace cbx
