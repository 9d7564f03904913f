// Host-side helpers shared by the C-ABI entry points: status codes, the
// thread-local error message, stream-ordered scratch buffers and the
// device-wide scans.
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

// Intermediate device memory of one library call.  An entry point opens a
// Workspace on its stream; every DevBuf allocated on that host thread while it
// is open is carved (bump allocation, 256-byte aligned) out of device memory
// cached per (device, stream) and reused by the next call on that stream.
// Reuse is safe because work on one stream runs in order: the next call's
// kernels start after the previous call's have finished.  The cache only
// grows (a call that needed a second block has the blocks merged into one at
// the start of the next call) and is released by cbx_workspace_release().
// A cache already open on another host thread is not shared: that call opens
// a cache of its own.  Without an open Workspace, DevBuf falls back to
// hipMalloc / hipFree.
//
// A fixed Workspace (the second constructor) carves its buffers out of one
// caller-provided device region instead and never allocates: a request that
// does not fit fails with hipErrorOutOfMemory and sets overflow().  While a
// fixed Workspace is open, read_device_values and KernelTimer allocate
// nothing either (no pinned mailbox, no events).
struct ScanState {
    unsigned long long *state = nullptr;   // zero
    unsigned long long *clear = nullptr;   // for the scan to clear
    size_t n_clear = 0;
};

class Workspace {
  public:
    explicit Workspace(hipStream_t s);
    Workspace(hipStream_t s, void *base, size_t capacity);
    ~Workspace();
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    hipError_t take(void **p, size_t bytes);
    hipError_t status() const { return status_; }
    bool fixed() const { return fixed_base_ != nullptr; }
    bool overflow() const { return overflow_; }
    size_t requested() const { return requested_; }
    static Workspace *current();
    // zeroed look-back state of at least `words` words for one exclusive scan
    // on this workspace's stream, and the words the scan is to clear (two
    // states per cache, used in turn; carved and cleared in a fixed workspace)
    hipError_t scan_state(size_t words, struct ScanState *st);
    void scan_failed();

  private:
    hipStream_t stream_ = nullptr;
    struct Cache *cache_ = nullptr;
    Workspace *outer_ = nullptr;
    size_t block_ = 0, off_ = 0;
    hipError_t status_ = hipSuccess;
    char *fixed_base_ = nullptr;
    size_t fixed_cap_ = 0, requested_ = 0;
    bool overflow_ = false;
};

// True while a fixed Workspace is open on this host thread.
inline bool no_alloc_mode()
{
    const Workspace *ws = Workspace::current();
    return ws && ws->fixed();
}

template <typename V>
struct DevBuf {
    V *ptr = nullptr;
    size_t n = 0;
    bool owned = false;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    hipError_t alloc(size_t count, hipStream_t)
    {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        if (Workspace *ws = Workspace::current()) return ws->take((void **)&ptr, sizeof(V) * count);
        owned = true;
        return hipMalloc((void **)&ptr, sizeof(V) * count);
    }
    void release()
    {
        // hipFree waits for the device, so an owned buffer is never released
        // under a kernel that still uses it; workspace memory is not freed here
        if (ptr && owned) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
        owned = false;
    }
    ~DevBuf() { release(); }
};

// Optional per-kernel timing with HIP events recorded on the launch stream
// (enabled by cbx_timing_enable); read back by cbx_timing_read.
bool timing_on(const char *name);
void timing_record(const char *name, hipEvent_t begin, hipEvent_t end);

struct KernelTimer {
    const char *name;
    hipStream_t stream;
    hipEvent_t a = nullptr, b = nullptr;
    // (a null name times nothing: for calls that may be captured in a graph)
    KernelTimer(const char *n, hipStream_t s) : name(n), stream(s)
    {
        if (!n || !timing_on(n) || no_alloc_mode()) return;
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

// Reads up to 4 device values (4 or 8 bytes each) after the stream's earlier
// work: one small kernel writes them into pinned host memory, then one stream
// synchronisation.  (Replaces one staged device-to-host copy per value.)
// src[i]: device address, bytes[i]: 4 or 8; out[i]: the value, zero-extended.
hipError_t read_device_values(hipStream_t s, int n, const void *const *src, const int *bytes, uint64_t *out);

// Exclusive prefix sums (one kernel, single pass), ordered on stream s:
// out[i] = in[0] + ... + in[i - 1] for i < n.  The _total forms read n
// entries of a (+ b when given) and write n + 1 (out[n] = the total).
// int32 sums wrap (a negative total flags overflow); int64 totals must stay
// within +-2^61.
hipError_t exclusive_scan_i32(const int *in, int *out, int64_t n, hipStream_t s);
hipError_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);
hipError_t exclusive_scan_total_i32(const int *a, const int *b, int *out, int64_t n, hipStream_t s);
hipError_t exclusive_scan_total_i64(const int64_t *a, int64_t *out, int64_t n, hipStream_t s);
// The same scan on caller-held state: scan_state_words(n_out) words, zero
// when the scan starts (clear_scan_state or the caller's own kernel) and left
// set by it.
size_t scan_state_words(int64_t n_out);
hipError_t clear_scan_state(unsigned long long *state, size_t words, hipStream_t s);
hipError_t scan_i32_with_state(const int *a, const int *b, int *out, int64_t n_in, int64_t n_out,
                               unsigned long long *state, hipStream_t s);
hipError_t scan_i64_with_state(const int64_t *a, int64_t *out, int64_t n_in, int64_t n_out,
                               unsigned long long *state, hipStream_t s);

// Clears n (at most ZERO_MAX) buffers of words[i] 4-byte words with one
// kernel on stream s (a kernel, not memsets: usable under graph capture).
constexpr int ZERO_MAX = 6;
hipError_t zero_words(hipStream_t s, int n, void *const *p, const int64_t *words);

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

} // namespace cbx
