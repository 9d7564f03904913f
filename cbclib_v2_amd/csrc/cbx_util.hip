// Host utilities of the C ABI: the per-thread error message, version,
// profile names and the device-wide scans (hipcub / rocPRIM).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cbx_host.hpp"

namespace cbx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

struct TimingState {
    std::mutex mu;
    bool on = false;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::map<std::string, std::pair<double, int64_t>> total;
};

static TimingState &timing()
{
    static TimingState t;
    return t;
}

bool timing_on() { return timing().on; }

void timing_record(const char *name, hipEvent_t a, hipEvent_t b)
{
    TimingState &t = timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.pending.push_back({name, a, b});
}

static void timing_drain(TimingState &t)
{
    for (auto &p : t.pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &e = t.total[p.name];
            e.first += ms;
            e.second += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    t.pending.clear();
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

void cbx_timing_enable(int on)
{
    cbx::TimingState &t = cbx::timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.on = on != 0;
}

void cbx_timing_reset(void)
{
    cbx::TimingState &t = cbx::timing();
    std::lock_guard<std::mutex> g(t.mu);
    cbx::timing_drain(t);
    t.total.clear();
}

int cbx_timing_read(const char *name, double *total_ms, int64_t *count)
{
    cbx::TimingState &t = cbx::timing();
    std::lock_guard<std::mutex> g(t.mu);
    cbx::timing_drain(t);
    auto it = t.total.find(name ? name : "");
    *total_ms = (it == t.total.end()) ? 0.0 : it->second.first;
    *count = (it == t.total.end()) ? 0 : it->second.second;
    return CBX_OK;
}

int cbx_kernel_id(const char *name)
{
    static const char *names[5] = {"biweight", "gaussian", "parabolic", "rectangular", "triangular"};
    if (!name) return -1;
    for (int i = 0; i < 5; i++)
        if (std::strcmp(name, names[i]) == 0) return i;
    return -1;
}

} // extern "C"
