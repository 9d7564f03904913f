// Host utilities of the C ABI: the per-thread error message, version,
// profile names and the device-wide scans (hipcub / rocPRIM).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
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

// ---------------------------------------------------------------------------
// workspace caches
// ---------------------------------------------------------------------------
struct Cache {
    int device;
    hipStream_t stream;
    bool busy;
    std::vector<std::pair<void *, size_t>> blocks;
};

static std::mutex g_ws_mu;
static std::vector<Cache *> g_ws;
static thread_local Workspace *g_ws_current = nullptr;
static const size_t WS_ALIGN = 256;
static const size_t WS_MIN_BLOCK = size_t(64) << 20;

Workspace *Workspace::current() { return g_ws_current; }

Workspace::Workspace(hipStream_t s)
{
    int dev = 0;
    status_ = hipGetDevice(&dev);
    if (status_ != hipSuccess) return;
    {
        std::lock_guard<std::mutex> g(g_ws_mu);
        for (Cache *c : g_ws)
            if (c->device == dev && c->stream == s && !c->busy) {
                cache_ = c;
                break;
            }
        if (!cache_) {
            cache_ = new Cache{dev, s, false, {}};
            g_ws.push_back(cache_);
        }
        cache_->busy = true;
    }
    if (cache_->blocks.size() > 1) {
        // merge: wait for the stream's earlier work, then one block of the total
        size_t total = 0;
        for (auto &b : cache_->blocks) total += b.second;
        status_ = hipStreamSynchronize(s);
        for (auto &b : cache_->blocks) (void)hipFree(b.first);
        cache_->blocks.clear();
        if (status_ == hipSuccess) {
            void *p = nullptr;
            status_ = hipMalloc(&p, total);
            if (status_ == hipSuccess) cache_->blocks.push_back({p, total});
        }
    }
    outer_ = g_ws_current;
    g_ws_current = this;
}

Workspace::Workspace(hipStream_t, void *base, size_t capacity)
    : fixed_base_(static_cast<char *>(base)), fixed_cap_(capacity)
{
    if (!base) status_ = hipErrorInvalidValue;
    outer_ = g_ws_current;
    g_ws_current = this;
}

Workspace::~Workspace()
{
    if (fixed_base_) {
        g_ws_current = outer_;
        return;
    }
    if (!cache_) return;
    g_ws_current = outer_;
    std::lock_guard<std::mutex> g(g_ws_mu);
    cache_->busy = false;
}

hipError_t Workspace::take(void **p, size_t bytes)
{
    *p = nullptr;
    if (status_ != hipSuccess) return status_;
    bytes = (bytes + WS_ALIGN - 1) / WS_ALIGN * WS_ALIGN;
    if (bytes == 0) bytes = WS_ALIGN;
    if (fixed_base_) {
        requested_ += bytes;
        if (overflow_ || fixed_cap_ - off_ < bytes) {
            overflow_ = true;
            return hipErrorOutOfMemory;
        }
        *p = fixed_base_ + off_;
        off_ += bytes;
        return hipSuccess;
    }
    auto &bl = cache_->blocks;
    while (block_ < bl.size()) {
        if (bl[block_].second - off_ >= bytes) {
            *p = (char *)bl[block_].first + off_;
            off_ += bytes;
            return hipSuccess;
        }
        block_++;
        off_ = 0;
    }
    size_t have = 0;
    for (auto &b : bl) have += b.second;
    const size_t size = std::max(bytes, std::max(WS_MIN_BLOCK, have));
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, size);
    if (e != hipSuccess) return e;
    bl.push_back({q, size});
    block_ = bl.size() - 1;
    off_ = bytes;
    *p = q;
    return hipSuccess;
}

static size_t workspace_release()
{
    std::lock_guard<std::mutex> g(g_ws_mu);
    size_t freed = 0;
    int dev0 = 0;
    (void)hipGetDevice(&dev0);
    for (auto it = g_ws.begin(); it != g_ws.end();) {
        Cache *c = *it;
        if (c->busy) {
            ++it;
            continue;
        }
        (void)hipSetDevice(c->device);
        (void)hipDeviceSynchronize();
        for (auto &b : c->blocks) {
            (void)hipFree(b.first);
            freed += b.second;
        }
        delete c;
        it = g_ws.erase(it);
    }
    (void)hipSetDevice(dev0);
    return freed;
}

static size_t workspace_bytes()
{
    std::lock_guard<std::mutex> g(g_ws_mu);
    size_t n = 0;
    for (Cache *c : g_ws)
        for (auto &b : c->blocks) n += b.second;
    return n;
}

// ---------------------------------------------------------------------------
// small device-to-host reads
// ---------------------------------------------------------------------------
struct MailSrc {
    const void *p[4];
    int bytes[4];
    int n;
};

__global__ void k_mail(MailSrc s, uint64_t *dst)
{
    const int i = threadIdx.x;
    if (i < s.n)
        dst[i] = (s.bytes[i] == 8) ? *static_cast<const uint64_t *>(s.p[i])
                                   : (uint64_t) * static_cast<const uint32_t *>(s.p[i]);
}

// One pinned, mapped, coherent mailbox of 4 values per host thread and device
// (a host thread runs one call at a time; kept until the process ends).
struct Mailbox {
    uint64_t *host = nullptr;
    uint64_t *dev = nullptr;
};

hipError_t read_device_values(hipStream_t s, int n, const void *const *src, const int *bytes, uint64_t *out)
{
    if (n < 1 || n > 4) return hipErrorInvalidValue;
    hipError_t e;
    if (no_alloc_mode()) {
        // no pinned mailbox: one synchronous copy per value into host memory
        for (int i = 0; i < n; i++) {
            if (bytes[i] != 4 && bytes[i] != 8) return hipErrorInvalidValue;
            uint64_t v = 0;
            e = hipMemcpyAsync(&v, src[i], (size_t)bytes[i], hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            out[i] = v;
        }
        return hipSuccess;
    }
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    thread_local std::vector<Mailbox> boxes;
    if ((int)boxes.size() <= dev) boxes.resize(dev + 1);
    Mailbox &mb = boxes[dev];
    if (!mb.host) {
        void *h = nullptr;
        e = hipHostMalloc(&h, 4 * sizeof(uint64_t), hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return e;
        void *d = nullptr;
        e = hipHostGetDevicePointer(&d, h, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(h);
            return e;
        }
        mb.host = static_cast<uint64_t *>(h);
        mb.dev = static_cast<uint64_t *>(d);
    }
    MailSrc ms{};
    for (int i = 0; i < n; i++) {
        if (bytes[i] != 4 && bytes[i] != 8) return hipErrorInvalidValue;
        ms.p[i] = src[i];
        ms.bytes[i] = bytes[i];
    }
    ms.n = n;
    k_mail<<<1, 64, 0, s>>>(ms, mb.dev);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const volatile uint64_t *h = mb.host;
    for (int i = 0; i < n; i++) out[i] = h[i];
    return hipSuccess;
}

template <typename V>
static hipError_t exclusive_scan(const V *in, V *out, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s);
    if (e != hipSuccess) return e;
    DevBuf<uint8_t> tmp;
    e = tmp.alloc(bytes ? bytes : 1, s);
    if (e != hipSuccess) return e;
    return hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, in, out, (int)n, s);
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

int64_t cbx_workspace_release(void) { return (int64_t)cbx::workspace_release(); }

int64_t cbx_workspace_bytes(void) { return (int64_t)cbx::workspace_bytes(); }

int cbx_kernel_id(const char *name)
{
    static const char *names[5] = {"biweight", "gaussian", "parabolic", "rectangular", "triangular"};
    if (!name) return -1;
    for (int i = 0; i < 5; i++)
        if (std::strcmp(name, names[i]) == 0) return i;
    return -1;
}

} // extern "C"
