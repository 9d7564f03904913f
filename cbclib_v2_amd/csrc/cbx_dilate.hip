// binary_dilation on gfx950 (label.cpp:8-107; Region::dilate with n_iter,
// label.hpp:464-496).
//
// With one thread the reference grows, from every true input element, a
// region `iterations` steps wide by the structure's non-zero shifts (inside
// the array), a new element joining only where the mask is true.  The union
// of those breadth-first balls is `iterations` rounds of the masked dilation
//     X' = X | (M & OR_s X[. - s])
// started from X = input, which is what runs here.
//
// Layout: the array is viewed as rows along its last axis (R rows of W
// elements) and packed one bit per element into 64-bit words, Wq = ceil(W/64)
// words per row, bits beyond W zero.  A wave packs or unpacks one word: lane
// l owns element 64 j + l, so the byte loads and stores of a wave are 64
// consecutive bytes and the word is one ballot.  A dilation round is one
// thread per word: the shifts are grouped by their row part (all axes but
// the last); per group the source row's words j - 1, j, j + 1 are loaded once
// and every x shift of the group is a funnel shift of them.  HBM traffic of
// one call: the input and mask bytes once, the output bytes once, plus
// 3 x (P / 8) bytes of words per round (L2 absorbs the neighbour rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "cbx_host.hpp"

namespace cbx {

constexpr int DIL_MAX_DIM = 7;
constexpr int DIL_NT = 256;

struct DilGeom {
    int nlead;                  // leading axes (ndim - 1)
    int32_t lead[DIL_MAX_DIM];  // their sizes
    int64_t rows, W, Wq;
    int n_groups;
};

// group g: row shift lead_shift[g * nlead ..], flat row offset row_off[g],
// x shifts dx[dx_off[g] .. dx_off[g + 1]); near[g] = 1 when every |dx| < 64
struct DilGroups {
    const int32_t *lead_shift;
    const int64_t *row_off;
    const int32_t *dx_off;
    const int32_t *dx;
    const int32_t *near;
};

// one wave per word: word (r, j) of `bits` from the bytes of row r
__global__ void __launch_bounds__(DIL_NT)
k_pack_bits(const uint8_t *__restrict__ src, int64_t rows, int64_t W, int64_t Wq, uint64_t *__restrict__ bits)
{
    const int64_t word = (int64_t)blockIdx.x * (DIL_NT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (word >= rows * Wq) return;   // uniform per wave
    const int64_t r = word / Wq, j = word - r * Wq;
    const int64_t x = j * 64 + lane;
    const bool v = x < W && src[r * W + x] != 0;
    const uint64_t b = __ballot(v);
    if (lane == 0) bits[word] = b;
}

__global__ void __launch_bounds__(DIL_NT)
k_unpack_bits(const uint64_t *__restrict__ bits, int64_t rows, int64_t W, int64_t Wq, uint8_t *__restrict__ dst)
{
    const int64_t word = (int64_t)blockIdx.x * (DIL_NT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (word >= rows * Wq) return;
    const int64_t r = word / Wq, j = word - r * Wq;
    const int64_t x = j * 64 + lane;
    if (x >= W) return;
    dst[r * W + x] = (uint8_t)((bits[word] >> lane) & 1u);
}

__device__ __forceinline__ uint64_t word_at(const uint64_t *row, int64_t j, int64_t Wq)
{
    return (j >= 0 && j < Wq) ? row[j] : 0ull;
}

// the bits of X[x - dx] placed at x, for the 64 x of word j of `row`
__device__ __forceinline__ uint64_t shifted_word(const uint64_t *row, int64_t j, int64_t Wq, int dx)
{
    if (dx >= 0) {
        const int64_t a = dx >> 6;
        const int b = dx & 63;
        const uint64_t hi = word_at(row, j - a, Wq);
        if (b == 0) return hi;
        return (hi << b) | (word_at(row, j - a - 1, Wq) >> (64 - b));
    }
    const int64_t e = -(int64_t)dx;
    const int64_t a = e >> 6;
    const int b = (int)(e & 63);
    const uint64_t lo = word_at(row, j + a, Wq);
    if (b == 0) return lo;
    return (lo >> b) | (word_at(row, j + a + 1, Wq) << (64 - b));
}

__device__ __forceinline__ uint64_t near_shift(uint64_t wm, uint64_t w0, uint64_t wp, int dx)
{
    if (dx == 0) return w0;
    if (dx > 0) return (w0 << dx) | (wm >> (64 - dx));
    return (w0 >> -dx) | (wp << (64 + dx));
}

// one dilation round: Y = X | (M & OR_s X[. - s]); *changed |= (Y != X)
__global__ void __launch_bounds__(DIL_NT)
k_dilate_bits(const uint64_t *__restrict__ X, const uint64_t *__restrict__ M, DilGeom g, DilGroups grp,
              uint64_t *__restrict__ Y, int *__restrict__ changed)
{
    const int64_t word = (int64_t)blockIdx.x * DIL_NT + threadIdx.x;
    if (word >= g.rows * g.Wq) return;
    const int64_t r = word / g.Wq, j = word - r * g.Wq;
    // coordinates of row r on the leading axes (loops over the constant
    // maximum, so that c stays in registers)
    int32_t c[DIL_MAX_DIM - 1];
    int64_t rem = r;
#pragma unroll
    for (int d = DIL_MAX_DIM - 2; d >= 0; d--) {
        c[d] = 0;
        if (d < g.nlead) {
            c[d] = (int32_t)(rem % g.lead[d]);
            rem /= g.lead[d];
        }
    }
    const uint64_t x = X[word];
    const uint64_t m = M ? M[word] : ~0ull;
    uint64_t acc = 0;
    if ((m & ~x) != 0) {   // some element could still join
        for (int q = 0; q < g.n_groups; q++) {
            const int32_t *s = grp.lead_shift + (int64_t)q * g.nlead;
            bool in = true;
#pragma unroll
            for (int d = 0; d < DIL_MAX_DIM - 1; d++) {
                if (d < g.nlead) {
                    const int32_t cc = c[d] - s[d];
                    in &= (cc >= 0) & (cc < g.lead[d]);
                }
            }
            if (!in) continue;
            const uint64_t *row = X + (r - grp.row_off[q]) * g.Wq;
            const int k0 = grp.dx_off[q], k1 = grp.dx_off[q + 1];
            if (grp.near[q]) {
                const uint64_t wm = word_at(row, j - 1, g.Wq), w0 = row[j], wp = word_at(row, j + 1, g.Wq);
                for (int k = k0; k < k1; k++) acc |= near_shift(wm, w0, wp, grp.dx[k]);
            } else {
                for (int k = k0; k < k1; k++) acc |= shifted_word(row, j, g.Wq, grp.dx[k]);
            }
        }
    }
    uint64_t y = x | (acc & m);
    const int64_t tail = g.W - j * 64;
    if (tail < 64) y &= (1ull << tail) - 1ull;
    Y[word] = y;
    if (y != x) __hip_atomic_store(changed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_clear_flag(int *f)
{
    if (threadIdx.x == 0) *f = 0;
}

static int binary_dilation(const uint8_t *inp, const uint8_t *mask, int ndim, const int64_t *shape,
                           const int64_t *shifts, int64_t n_shifts, int64_t iterations, uint8_t *out, void *stream)
{
    if (ndim < 1 || ndim > DIL_MAX_DIM) return fail(CBX_EVALUE, "binary_dilation supports 1 to 7 dimensions");
    if (iterations < 0) return fail(CBX_EVALUE, "iterations must be non-negative");
    if (n_shifts < 0 || (n_shifts > 0 && !shifts)) return fail(CBX_EVALUE, "invalid shifts");
    if (n_shifts > (int64_t(1) << 16)) return fail(CBX_EVALUE, "structure has too many shifts");
    int64_t P = 1;
    for (int d = 0; d < ndim; d++) {
        if (shape[d] < 0) return fail(CBX_EVALUE, "negative shape");
        if (shape[d] >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "dimension too large");
        P *= shape[d];
    }
    if (P == 0) return CBX_OK;
    if (!inp || !out) return fail(CBX_EVALUE, "inp and out must not be NULL");

    DilGeom g{};
    g.nlead = ndim - 1;
    g.rows = 1;
    for (int d = 0; d < g.nlead; d++) {
        g.lead[d] = (int32_t)shape[d];
        g.rows *= shape[d];
    }
    g.W = shape[ndim - 1];
    g.Wq = (g.W + 63) / 64;
    // shifts grouped by their row part; shifts that leave every array are dropped
    std::map<std::vector<int32_t>, std::vector<int32_t>> groups;
    for (int64_t i = 0; i < n_shifts; i++) {
        const int64_t *s = shifts + i * ndim;
        bool zero = true, fits = true;
        for (int d = 0; d < ndim; d++) {
            zero &= s[d] == 0;
            fits &= s[d] > -shape[d] && s[d] < shape[d];
        }
        if (zero || !fits) continue;
        std::vector<int32_t> key(s, s + g.nlead);
        auto &v = groups[key];
        const int32_t dx = (int32_t)s[ndim - 1];
        if (std::find(v.begin(), v.end(), dx) == v.end()) v.push_back(dx);
    }
    std::vector<int32_t> h_lead, h_dx_off{0}, h_dx, h_near;
    std::vector<int64_t> h_row_off;
    for (auto &kv : groups) {
        int64_t off = 0, stride = 1;
        for (int d = g.nlead - 1; d >= 0; d--) {
            off += kv.first[d] * stride;
            stride *= shape[d];
        }
        h_lead.insert(h_lead.end(), kv.first.begin(), kv.first.end());
        h_row_off.push_back(off);
        bool near = true;
        for (int32_t dx : kv.second) near &= dx > -64 && dx < 64;
        h_near.push_back(near ? 1 : 0);
        h_dx.insert(h_dx.end(), kv.second.begin(), kv.second.end());
        h_dx_off.push_back((int32_t)h_dx.size());
    }
    g.n_groups = (int)h_row_off.size();

    hipStream_t st = (hipStream_t)stream;
    Workspace ws(st);
    const int64_t words = g.rows * g.Wq;
    const int64_t rounds = g.n_groups ? iterations : 0;
    DevBuf<uint64_t> a, b, mb;
    DevBuf<int32_t> d_lead, d_dx_off, d_dx, d_near;
    DevBuf<int64_t> d_row_off;
    DevBuf<int> flags;
    CBX_HIP(a.alloc(words, st));
    if (rounds > 0) {
        CBX_HIP(b.alloc(words, st));
        if (mask) CBX_HIP(mb.alloc(words, st));
        CBX_HIP(d_lead.alloc(std::max<size_t>(h_lead.size(), 1), st));
        CBX_HIP(d_row_off.alloc(h_row_off.size(), st));
        CBX_HIP(d_dx_off.alloc(h_dx_off.size(), st));
        CBX_HIP(d_dx.alloc(h_dx.size(), st));
        CBX_HIP(d_near.alloc(h_near.size(), st));
        CBX_HIP(flags.alloc(1, st));
        if (!h_lead.empty())
            CBX_HIP(hipMemcpyAsync(d_lead.ptr, h_lead.data(), sizeof(int32_t) * h_lead.size(), hipMemcpyHostToDevice, st));
        CBX_HIP(hipMemcpyAsync(d_row_off.ptr, h_row_off.data(), sizeof(int64_t) * h_row_off.size(),
                               hipMemcpyHostToDevice, st));
        CBX_HIP(hipMemcpyAsync(d_dx_off.ptr, h_dx_off.data(), sizeof(int32_t) * h_dx_off.size(),
                               hipMemcpyHostToDevice, st));
        CBX_HIP(hipMemcpyAsync(d_dx.ptr, h_dx.data(), sizeof(int32_t) * h_dx.size(), hipMemcpyHostToDevice, st));
        CBX_HIP(hipMemcpyAsync(d_near.ptr, h_near.data(), sizeof(int32_t) * h_near.size(), hipMemcpyHostToDevice, st));
    }
    const int wpb = DIL_NT / 64;
    const unsigned pack_grid = (unsigned)((words + wpb - 1) / wpb);
    KernelTimer timer("dilation", st);
    k_pack_bits<<<pack_grid, DIL_NT, 0, st>>>(inp, g.rows, g.W, g.Wq, a.ptr);
    if (rounds > 0 && mask) k_pack_bits<<<pack_grid, DIL_NT, 0, st>>>(mask, g.rows, g.W, g.Wq, mb.ptr);
    CBX_HIP(hipGetLastError());
    const DilGroups grp{d_lead.ptr, d_row_off.ptr, d_dx_off.ptr, d_dx.ptr, d_near.ptr};
    uint64_t *cur = a.ptr, *nxt = b.ptr;
    // the rounds in batches of 8; a batch whose last round changed nothing ends the dilation
    for (int64_t done = 0; done < rounds;) {
        const int64_t batch = std::min<int64_t>(8, rounds - done);
        for (int64_t k = 0; k < batch; k++) {
            if (k == batch - 1) k_clear_flag<<<1, 64, 0, st>>>(flags.ptr);
            k_dilate_bits<<<grid_for(words, DIL_NT), DIL_NT, 0, st>>>(cur, mask ? mb.ptr : nullptr, g, grp, nxt,
                                                                        flags.ptr);
            std::swap(cur, nxt);
        }
        CBX_HIP(hipGetLastError());
        done += batch;
        if (done < rounds) {
            uint64_t hv = 0;
            const void *src[1] = {flags.ptr};
            const int nb[1] = {4};
            CBX_HIP(read_device_values(st, 1, src, nb, &hv));
            if (hv == 0) break;
        }
    }
    k_unpack_bits<<<pack_grid, DIL_NT, 0, st>>>(cur, g.rows, g.W, g.Wq, out);
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

} // namespace cbx

extern "C" int cbx_binary_dilation(const uint8_t *inp, const uint8_t *mask, int ndim, const int64_t *shape,
                                   const int64_t *shifts, int64_t n_shifts, int64_t iterations, uint8_t *out,
                                   void *stream)
{
    return cbx::binary_dilation(inp, mask, ndim, shape, shifts, n_shifts, iterations, out, stream);
}
