// Host utilities of the C ABI: the per-thread error message, version,
// profile names, workspace caches, device-to-host reads and the device-wide
// exclusive scans.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
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
    int on = 0;   // 0: off, 1: every timer, 2: the tile kernels' timers only
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

bool timing_on(const char *name)
{
    const int on = timing().on;
    return on == 1 || (on == 2 && std::strstr(name, "_tiles") != nullptr);
}

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
    unsigned long long *scan_base = nullptr;   // two look-back states of scan_words each
    size_t scan_words = 0;
    int scan_next = 0;        // the state the next scan uses (zero)
    size_t scan_dirty = 0;    // words the previous scan left set in the other
};

static std::mutex g_ws_mu;
static std::vector<Cache *> g_ws;
static thread_local Workspace *g_ws_current = nullptr;
static const size_t WS_ALIGN = 256;
static const size_t WS_MIN_BLOCK = size_t(64) << 20;

Workspace *Workspace::current() { return g_ws_current; }

Workspace::Workspace(hipStream_t s) : stream_(s)
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

Workspace::Workspace(hipStream_t s, void *base, size_t capacity)
    : stream_(s), fixed_base_(static_cast<char *>(base)), fixed_cap_(capacity)
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

hipError_t Workspace::scan_state(size_t words, ScanState *st)
{
    *st = ScanState{};
    if (status_ != hipSuccess) return status_;
    if (fixed_base_) {
        void *q = nullptr;
        hipError_t e = take(&q, sizeof(unsigned long long) * words);
        if (e != hipSuccess) return e;
        if ((e = clear_scan_state(static_cast<unsigned long long *>(q), words, stream_)) != hipSuccess) return e;
        st->state = static_cast<unsigned long long *>(q);
        return hipSuccess;
    }
    Cache *c = cache_;
    if (c->scan_words < words) {
        // hipFree waits for the device, so no scan still uses the old states
        if (c->scan_base) (void)hipFree(c->scan_base);
        c->scan_base = nullptr;
        c->scan_words = 0;
        const size_t w = std::max(words, size_t(4096));
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, 2 * sizeof(unsigned long long) * w);
        if (e != hipSuccess) return e;
        c->scan_base = static_cast<unsigned long long *>(q);
        if ((e = clear_scan_state(c->scan_base, 2 * w, stream_)) != hipSuccess) {
            (void)hipFree(q);
            c->scan_base = nullptr;
            return e;
        }
        c->scan_words = w;
        c->scan_next = 0;
        c->scan_dirty = 0;
    }
    // this scan uses state scan_next and clears the other, which the previous
    // scan on this stream (finished before this one starts) left dirty
    st->state = c->scan_base + c->scan_next * c->scan_words;
    st->clear = c->scan_base + (1 - c->scan_next) * c->scan_words;
    st->n_clear = c->scan_dirty;
    c->scan_dirty = words;
    c->scan_next = 1 - c->scan_next;
    return hipSuccess;
}

void Workspace::scan_failed()
{
    // the launch did not run: neither state is known to be clear any more
    if (cache_) cache_->scan_words = 0;
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
        if (c->scan_base) (void)hipFree(c->scan_base);
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

// Writes the values into dst[0 .. n), then the call's sequence number into
// dst[4] with a system-scope release, after every value is visible to the
// host.
__global__ void k_mail(MailSrc s, uint64_t *dst, uint64_t seq)
{
    const int i = threadIdx.x;
    if (i < s.n) {
        dst[i] = (s.bytes[i] == 8) ? *static_cast<const uint64_t *>(s.p[i])
                                   : (uint64_t) * static_cast<const uint32_t *>(s.p[i]);
        __threadfence_system();
    }
    __syncthreads();
    if (i == 0) __hip_atomic_store(dst + 4, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One pinned, mapped, coherent mailbox of 4 values and a sequence number per
// host thread and device (a host thread runs one call at a time; kept until
// the process ends).
struct Mailbox {
    uint64_t *host = nullptr;
    uint64_t *dev = nullptr;
    uint64_t seq = 0;
};

// How long read_device_values polls the mailbox before it waits for the
// stream instead (a long queue of work ahead of the read, or a failed launch).
constexpr auto MAIL_SPIN = std::chrono::milliseconds(20);

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
        e = hipHostMalloc(&h, 8 * sizeof(uint64_t), hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return e;
        void *d = nullptr;
        e = hipHostGetDevicePointer(&d, h, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(h);
            return e;
        }
        std::memset(h, 0, 8 * sizeof(uint64_t));   // no sequence number yet (they start at 1)
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
    const uint64_t seq = ++mb.seq;
    k_mail<<<1, 64, 0, s>>>(ms, mb.dev, seq);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // The values are final once the sequence number arrives (the stream runs
    // in order, so everything before the read has completed): polling the
    // coherent mailbox wakes the host sooner than a stream synchronisation.
    // Past MAIL_SPIN the stream is waited for, which also reports errors.
    const volatile uint64_t *h = mb.host;
    bool arrived = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t it = 0;; it++) {
        if (h[4] == seq) {
            arrived = true;
            break;
        }
        if ((it & 255u) == 255u && std::chrono::steady_clock::now() - t0 > MAIL_SPIN) break;
        __builtin_ia32_pause();
    }
    if (arrived) {
        std::atomic_thread_fence(std::memory_order_acquire);
    } else {
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    for (int i = 0; i < n; i++) out[i] = h[i];
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// clearing several small buffers with one launch
// ---------------------------------------------------------------------------
struct ZeroList {
    uint32_t *p[ZERO_MAX];
    int64_t end[ZERO_MAX];   // running totals of the 4-byte word counts
    int n;
};

__global__ void k_zero_words(ZeroList z)
{
    const int64_t total = z.end[z.n - 1];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int r = 0;
        while (i >= z.end[r]) r++;
        z.p[r][i - (r ? z.end[r - 1] : 0)] = 0u;
    }
}

hipError_t zero_words(hipStream_t s, int n, void *const *p, const int64_t *words)
{
    if (n < 1 || n > ZERO_MAX) return hipErrorInvalidValue;
    ZeroList z{};
    int64_t t = 0;
    for (int i = 0; i < n; i++) {
        if (words[i] < 0 || (words[i] > 0 && !p[i])) return hipErrorInvalidValue;
        z.p[i] = static_cast<uint32_t *>(p[i]);
        t += words[i];
        z.end[i] = t;
    }
    z.n = n;
    if (t == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((t + 255) / 256, 2048);
    k_zero_words<<<(unsigned)blocks, 256, 0, s>>>(z);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// exclusive prefix sums: one kernel, single pass with decoupled look-back
// ---------------------------------------------------------------------------
// Workgroups take tiles in the order they start (a ticket counter), so every
// tile a workgroup waits for belongs to a workgroup that is already running
// and waits only on tiles before its own: every wait ends.  Each tile
// publishes its sum (AGG) as soon as it has it and its inclusive prefix (INC)
// once its look-back is done; wave 0 reads 256 predecessors per step (four
// per lane, one round trip: the state words are device-coherent atomics that
// travel past the per-XCD L2 caches) and stops at the nearest INC.  Status
// and value share one 64-bit word (status in bits 62-63, the value's low 62
// bits, sign-extended when read), so no ordering between separate words is
// needed.  The state (ticket, one word per tile) must be zero when a scan
// starts; a scan may clear another, finished scan's state on the way
// (`clear`), which is how consecutive scans on one stream alternate between
// two states without a separate clearing launch.
// Tiles of 2048 entries (8 per thread) up to SCAN_BIG entries: the shortest
// workgroups for the small scans, whose time is a few dependent round trips;
// 4096 above, halving the look-back of large scans.
constexpr int SCAN_NT = 256;
constexpr int SCAN_TILE = SCAN_NT * 8;           // the state is sized for the smaller tile
constexpr int64_t SCAN_BIG = int64_t(1) << 19;
constexpr int SCAN_LB = 4;                       // predecessors per lane per look-back step
constexpr uint64_t SCAN_AGG = 1ull << 62;
constexpr uint64_t SCAN_INC = 2ull << 62;
constexpr uint64_t SCAN_VAL = SCAN_AGG - 1;
constexpr unsigned SCAN_MAX_POLLS = 1u << 20;   // a bound only: every wait ends

size_t scan_state_words(int64_t n_out) { return 1 + (size_t)((n_out + SCAN_TILE - 1) / SCAN_TILE); }

template <typename U>
__device__ __forceinline__ U scan_wave_incl(U v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const U o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

template <typename V>
__device__ __forceinline__ uint64_t scan_pack(V v)
{
    return (uint64_t)(int64_t)v & SCAN_VAL;
}

template <typename V>
__device__ __forceinline__ V scan_unpack(uint64_t w)
{
    return (V)(((int64_t)(w << 2)) >> 2);
}

// out[i] = sum of a[j] (+ b[j]) over j < i, for i < n_out; entries i >= n_in
// read as zero (n_out = n_in + 1 gives the total in out[n_in]).  Sums wrap
// like V's unsigned counterpart (callers detect int32 overflow as a negative
// total); int64 totals must stay within +-2^61.
template <typename V, int SCAN_IPT>
__global__ void __launch_bounds__(SCAN_NT)
k_scan(const V *__restrict__ a, const V *__restrict__ b, V *__restrict__ out, int64_t n_in, int64_t n_out,
       unsigned long long *__restrict__ state, unsigned long long *__restrict__ clear, int64_t n_clear)
{
    using U = typename std::make_unsigned<V>::type;
    constexpr int TILE = SCAN_NT * SCAN_IPT;
    __shared__ U buf[TILE];
    __shared__ U wsum[SCAN_NT / 64];
    __shared__ U s_prefix;
    __shared__ int s_tile;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = (int)atomicAdd(&state[0], 1ull);
    for (int64_t i = (int64_t)blockIdx.x * SCAN_NT + tid; i < n_clear; i += (int64_t)gridDim.x * SCAN_NT)
        clear[i] = 0ull;
    __syncthreads();
    const int tile = s_tile;
    const int64_t base = (int64_t)tile * TILE;
    // coalesced load, then each thread takes SCAN_IPT consecutive entries
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const int64_t i = base + k * SCAN_NT + tid;
        U v = 0;
        if (i < n_in) {
            v = (U)a[i];
            if (b) v += (U)b[i];
        }
        buf[k * SCAN_NT + tid] = v;
    }
    __syncthreads();
    U loc[SCAN_IPT], ts = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        loc[k] = buf[tid * SCAN_IPT + k];
        ts += loc[k];
    }
    const U incl = scan_wave_incl<U>(ts, lane);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    U wpre = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < SCAN_NT / 64; w++) {
        wpre += (w < wave) ? wsum[w] : U(0);
        agg += wsum[w];
    }
    if (wave == 0) {
        U prefix = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&state[1], SCAN_INC | scan_pack<V>((V)agg), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&state[1 + tile], SCAN_AGG | scan_pack<V>((V)agg), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            int j = tile - 1;   // the nearest predecessor not yet summed
            unsigned polls = 0;
            while (true) {
                // predecessor at distance d = lane * SCAN_LB + q from j
                uint64_t w[SCAN_LB];
#pragma unroll
                for (int q = 0; q < SCAN_LB; q++) {
                    const int k = j - (lane * SCAN_LB + q);
                    w[q] = (k >= 0) ? __hip_atomic_load(&state[1 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : SCAN_INC;   // before tile 0: an inclusive zero
                }
                int first = 64 * SCAN_LB;   // distance of the nearest INC
#pragma unroll
                for (int q = 0; q < SCAN_LB; q++) {
                    const uint64_t inc = __ballot((w[q] >> 62) == 2);
                    if (inc) first = min(first, __builtin_ctzll(inc) * SCAN_LB + q);
                }
                bool wait = false;   // a predecessor up to the nearest INC has not published
#pragma unroll
                for (int q = 0; q < SCAN_LB; q++) {
                    const uint64_t inv = __ballot((w[q] >> 62) == 0);
                    if (first >= q) {
                        const int lmax = (first - q) / SCAN_LB;
                        const uint64_t need = (lmax >= 63) ? ~0ull : ((2ull << lmax) - 1);
                        wait |= (inv & need) != 0;
                    }
                }
                if (wait && ++polls < SCAN_MAX_POLLS) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                U part = 0;
#pragma unroll
                for (int q = 0; q < SCAN_LB; q++)
                    if (lane * SCAN_LB + q <= first) part += (U)scan_unpack<V>(w[q]);
#pragma unroll
                for (int d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
                prefix += part;
                if (first < 64 * SCAN_LB || polls >= SCAN_MAX_POLLS) break;
                j -= 64 * SCAN_LB;
            }
            if (lane == 0) __hip_atomic_store(&state[1 + tile], SCAN_INC | scan_pack<V>((V)(prefix + agg)),
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    U run = s_prefix + wpre + incl - ts;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        buf[tid * SCAN_IPT + k] = run;
        run += loc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const int64_t i = base + k * SCAN_NT + tid;
        if (i < n_out) out[i] = (V)buf[k * SCAN_NT + tid];
    }
}

__global__ void k_scan_clear(unsigned long long *__restrict__ state, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) state[i] = 0ull;
}

template <typename V>
static hipError_t launch_scan(const V *a, const V *b, V *out, int64_t n_in, int64_t n_out,
                              unsigned long long *state, unsigned long long *clear, int64_t n_clear, hipStream_t s)
{
    if (n_out <= 0) return hipSuccess;
    if (n_in < 0 || n_in > n_out || n_out >= (int64_t(1) << 31)) return hipErrorInvalidValue;
    if (n_out <= SCAN_BIG) {
        const int64_t nblk = (n_out + SCAN_NT * 8 - 1) / (SCAN_NT * 8);
        k_scan<V, 8><<<(unsigned)nblk, SCAN_NT, 0, s>>>(a, b, out, n_in, n_out, state, clear, n_clear);
    } else {
        const int64_t nblk = (n_out + SCAN_NT * 16 - 1) / (SCAN_NT * 16);
        k_scan<V, 16><<<(unsigned)nblk, SCAN_NT, 0, s>>>(a, b, out, n_in, n_out, state, clear, n_clear);
    }
    return hipGetLastError();
}

hipError_t clear_scan_state(unsigned long long *state, size_t words, hipStream_t s)
{
    if (words == 0) return hipSuccess;
    k_scan_clear<<<grid_for((int64_t)words, 256), 256, 0, s>>>(state, (int64_t)words);
    return hipGetLastError();
}

hipError_t scan_i32_with_state(const int *a, const int *b, int *out, int64_t n_in, int64_t n_out,
                               unsigned long long *state, hipStream_t s)
{
    return launch_scan<int>(a, b, out, n_in, n_out, state, nullptr, 0, s);
}

hipError_t scan_i64_with_state(const int64_t *a, int64_t *out, int64_t n_in, int64_t n_out,
                               unsigned long long *state, hipStream_t s)
{
    return launch_scan<int64_t>(a, nullptr, out, n_in, n_out, state, nullptr, 0, s);
}

// The look-back state of scans on this thread's open workspace: two states
// per (device, stream) cache, used in turn, each scan clearing the one the
// previous scan used; in a fixed workspace (nothing may be allocated) or with
// no workspace open, carved or allocated for the call and cleared first.
template <typename V>
static hipError_t exclusive_scan(const V *a, const V *b, V *out, int64_t n_in, int64_t n_out, hipStream_t s)
{
    if (n_out <= 0) return hipSuccess;
    const size_t words = scan_state_words(n_out);
    Workspace *ws = Workspace::current();
    if (ws) {
        ScanState st;
        hipError_t e = ws->scan_state(words, &st);
        if (e != hipSuccess) return e;
        e = launch_scan<V>(a, b, out, n_in, n_out, st.state, st.clear, (int64_t)st.n_clear, s);
        if (e != hipSuccess) ws->scan_failed();
        return e;
    }
    DevBuf<unsigned long long> tmp;
    hipError_t e = tmp.alloc(words, s);
    if (e != hipSuccess) return e;
    if ((e = clear_scan_state(tmp.ptr, words, s)) != hipSuccess) return e;
    return launch_scan<V>(a, b, out, n_in, n_out, tmp.ptr, nullptr, 0, s);
}

hipError_t exclusive_scan_i32(const int *in, int *out, int64_t n, hipStream_t s)
{
    return exclusive_scan<int>(in, nullptr, out, n, n, s);
}

hipError_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s)
{
    return exclusive_scan<int64_t>(in, nullptr, out, n, n, s);
}

hipError_t exclusive_scan_total_i32(const int *a, const int *b, int *out, int64_t n, hipStream_t s)
{
    return exclusive_scan<int>(a, b, out, n, n + 1, s);
}

hipError_t exclusive_scan_total_i64(const int64_t *a, int64_t *out, int64_t n, hipStream_t s)
{
    return exclusive_scan<int64_t>(a, nullptr, out, n, n + 1, s);
}

} // namespace cbx

extern "C" {

const char *cbx_last_error(void) { return cbx::g_last_error.c_str(); }

const char *cbx_version(void) { return "cbx 0.1.0 (gfx950)"; }

void cbx_timing_enable(int on)
{
    cbx::TimingState &t = cbx::timing();
    std::lock_guard<std::mutex> g(t.mu);
    t.on = (on == 0 || on == 2) ? on : 1;
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

int cbx_exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, int total, void *stream)
{
    if (n < 0 || (!in && n > 0) || (!out && (n > 0 || total))) return cbx::fail(CBX_EVALUE, "invalid scan arguments");
    if (n + (total ? 1 : 0) >= (int64_t(1) << 31)) return cbx::fail(CBX_EVALUE, "scan longer than 2^31 - 1 entries");
    hipStream_t st = (hipStream_t)stream;
    cbx::Workspace ws(st);
    const hipError_t e = total ? cbx::exclusive_scan_total_i32(in, nullptr, out, n, st)
                               : cbx::exclusive_scan_i32(in, out, n, st);
    if (e != hipSuccess) return cbx::fail(CBX_ERUNTIME, std::string("exclusive scan: ") + hipGetErrorString(e));
    return CBX_OK;
}

int cbx_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, int total, void *stream)
{
    if (n < 0 || (!in && n > 0) || (!out && (n > 0 || total))) return cbx::fail(CBX_EVALUE, "invalid scan arguments");
    if (n + (total ? 1 : 0) >= (int64_t(1) << 31)) return cbx::fail(CBX_EVALUE, "scan longer than 2^31 - 1 entries");
    hipStream_t st = (hipStream_t)stream;
    cbx::Workspace ws(st);
    const hipError_t e = total ? cbx::exclusive_scan_total_i64(in, out, n, st) : cbx::exclusive_scan_i64(in, out, n, st);
    if (e != hipSuccess) return cbx::fail(CBX_ERUNTIME, std::string("exclusive scan: ") + hipGetErrorString(e));
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
