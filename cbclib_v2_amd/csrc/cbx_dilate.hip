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
// words per row, bits beyond W zero.  Rows of a multiple of 64 elements at
// 16-byte alignment are packed and unpacked one word per thread with four
// 16-byte accesses (bytes to bits by a carry-free product); other rows one
// word per wave: lane l owns element 64 j + l, the word is one ballot.  A dilation round is one
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

// bytes != 0 of a 32-bit word (4 elements) as 4 bits: the high bit of each
// byte set when the byte is non-zero, then gathered by one product (the
// partial products land on distinct bits, so nothing carries into 28-31)
__device__ __forceinline__ uint32_t nz_nibble(uint32_t v)
{
    const uint32_t t = (((v & 0x7f7f7f7fu) + 0x7f7f7f7fu) | v) & 0x80808080u;
    return (t * 0x00204081u) >> 28;
}

// 4 bits as 4 bytes of 0 / 1 (bit k to bit 8 k, distinct partial products)
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t n)
{
    return (n * 0x00204081u) & 0x01010101u;
}

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// rows of a multiple of 64 elements, 16-byte aligned: one thread per word,
// four 16-byte loads
// one thread per 16 input bytes (loads of a wave 1 KiB contiguous); the four
// lanes of a word join their 16 bits with shuffles.  The grid covers whole
// words (4 lanes each), so every lane of a word is present.
__global__ void __launch_bounds__(DIL_NT)
k_pack_words(const uint8_t *__restrict__ src, uint32_t words, uint64_t *__restrict__ bits)
{
    const uint32_t c = blockIdx.x * DIL_NT + threadIdx.x;
    const bool in = c < 4u * words;
    uint32_t h = 0;
    if (in) {
        const v4u q = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(src) + c);
        h = nz_nibble(q.x) | (nz_nibble(q.y) << 4) | (nz_nibble(q.z) << 8) | (nz_nibble(q.w) << 12);
    }
    const uint32_t h1 = __shfl_down(h, 1, 4), h2 = __shfl_down(h, 2, 4), h3 = __shfl_down(h, 3, 4);
    if (in && (c & 3u) == 0)
        bits[c >> 2] = (uint64_t)h | ((uint64_t)h1 << 16) | ((uint64_t)h2 << 32) | ((uint64_t)h3 << 48);
}

// one thread per 16 output bytes (a quarter word), so that the stores of a
// wave are 1 KiB contiguous
__global__ void __launch_bounds__(DIL_NT)
k_unpack_words(const uint64_t *__restrict__ bits, uint32_t words, uint8_t *__restrict__ dst)
{
    const uint32_t c = blockIdx.x * DIL_NT + threadIdx.x;
    if (c >= 4u * words) return;
    const uint32_t h = (uint32_t)(bits[c >> 2] >> (16 * (c & 3u))) & 0xffffu;
    const v4u q = {nibble_bytes(h & 15u), nibble_bytes((h >> 4) & 15u), nibble_bytes((h >> 8) & 15u),
                   nibble_bytes(h >> 12)};
    __builtin_nontemporal_store(q, reinterpret_cast<v4u *>(dst) + c);
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

// one dilation round: Y = X | (M & OR_s X[. - s]); changed[workgroup] = 1
// where a word changed
__global__ void __launch_bounds__(DIL_NT)
k_dilate_bits(const uint64_t *__restrict__ X, const uint64_t *__restrict__ M, DilGeom g, DilGroups grp,
              uint64_t *__restrict__ Y, int *__restrict__ changed)   // changed: per-workgroup flags or NULL
{
    // 32-bit index arithmetic: the host keeps rows * Wq below 2^32
    const uint32_t word = blockIdx.x * DIL_NT + threadIdx.x;
    if (word >= (uint32_t)(g.rows * g.Wq)) return;
    const uint32_t Wq = (uint32_t)g.Wq;
    const uint32_t r = word / Wq, j = word - r * Wq;
    // coordinates of row r on the leading axes (loops over the constant
    // maximum, so that c stays in registers)
    int32_t c[DIL_MAX_DIM - 1];
    uint32_t rem = r;
#pragma unroll
    for (int d = DIL_MAX_DIM - 2; d >= 0; d--) {
        c[d] = 0;
        if (d < g.nlead) {
            const uint32_t q = rem / (uint32_t)g.lead[d];
            c[d] = (int32_t)(rem - q * (uint32_t)g.lead[d]);
            rem = q;
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
            const uint64_t *row = X + ((int64_t)r - grp.row_off[q]) * (int64_t)Wq;
            const int k0 = grp.dx_off[q], k1 = grp.dx_off[q + 1];
            if (grp.near[q]) {
                const uint64_t wm = word_at(row, (int64_t)j - 1, g.Wq), w0 = row[j];
                const uint64_t wp = word_at(row, (int64_t)j + 1, g.Wq);
                for (int k = k0; k < k1; k++) acc |= near_shift(wm, w0, wp, grp.dx[k]);
            } else {
                for (int k = k0; k < k1; k++) acc |= shifted_word(row, j, g.Wq, grp.dx[k]);
            }
        }
    }
    uint64_t y = x | (acc & m);
    const int64_t tail = g.W - (int64_t)j * 64;
    if (tail < 64) y &= (1ull << tail) - 1ull;
    Y[word] = y;
    // one flag per workgroup (stores to one address serialise), written
    // only when the host reads them
    if (changed && __ballot(y != x) != 0 && (threadIdx.x & 63) == 0) changed[blockIdx.x] = 1;
}

// any non-zero among the n per-workgroup flags -> *out (one workgroup)
__global__ void __launch_bounds__(DIL_NT)
k_any_flag(const int *__restrict__ f, uint32_t n, int *__restrict__ out)
{
    __shared__ int hit;
    if (threadIdx.x == 0) hit = 0;
    __syncthreads();
    int v = 0;
    for (uint32_t i = threadIdx.x; i < n; i += DIL_NT) v |= f[i];
    if (v) hit = 1;
    __syncthreads();
    if (threadIdx.x == 0) *out = hit;
}

__global__ void __launch_bounds__(DIL_NT) k_clear_flags(int *f, uint32_t n)
{
    const uint32_t i = blockIdx.x * DIL_NT + threadIdx.x;
    if (i < n) f[i] = 0;
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
    if (g.rows * g.Wq >= (int64_t(1) << 30)) return fail(CBX_EVALUE, "binary_dilation supports fewer than 2^36 elements");
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
        if (rounds > 8) CBX_HIP(flags.alloc((size_t)grid_for(words, DIL_NT) + 1, st));
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
    // whole words of 16-byte aligned rows: one thread per 16 bytes
    auto aligned = [&](const void *p) { return (g.W % 64) == 0 && ((uintptr_t)p & 15) == 0; };
    const unsigned quarter_grid = (unsigned)((4 * words + DIL_NT - 1) / DIL_NT);
    auto pack = [&](const uint8_t *src, uint64_t *dst) {
        if (aligned(src)) k_pack_words<<<quarter_grid, DIL_NT, 0, st>>>(src, (uint32_t)words, dst);
        else k_pack_bits<<<pack_grid, DIL_NT, 0, st>>>(src, g.rows, g.W, g.Wq, dst);
    };
    KernelTimer timer("dilation", st);
    pack(inp, a.ptr);
    if (rounds > 0 && mask) pack(mask, mb.ptr);
    CBX_HIP(hipGetLastError());
    const DilGroups grp{d_lead.ptr, d_row_off.ptr, d_dx_off.ptr, d_dx.ptr, d_near.ptr};
    uint64_t *cur = a.ptr, *nxt = b.ptr;
    // the rounds in batches of 8; a batch whose last round changed nothing
    // ends the dilation (the last round of a batch followed by another sets
    // per-workgroup flags, reduced into flags[n_blocks] and read back)
    const uint32_t n_blocks = (uint32_t)grid_for(words, DIL_NT);
    for (int64_t done = 0; done < rounds;) {
        const int64_t batch = std::min<int64_t>(8, rounds - done);
        const bool check = done + batch < rounds;
        for (int64_t k = 0; k < batch; k++) {
            int *fl = nullptr;
            if (check && k == batch - 1) {
                k_clear_flags<<<grid_for(n_blocks, DIL_NT), DIL_NT, 0, st>>>(flags.ptr, n_blocks);
                fl = flags.ptr;
            }
            k_dilate_bits<<<n_blocks, DIL_NT, 0, st>>>(cur, mask ? mb.ptr : nullptr, g, grp, nxt, fl);
            std::swap(cur, nxt);
        }
        CBX_HIP(hipGetLastError());
        done += batch;
        if (check) {
            k_any_flag<<<1, DIL_NT, 0, st>>>(flags.ptr, n_blocks, flags.ptr + n_blocks);
            CBX_HIP(hipGetLastError());
            uint64_t hv = 0;
            const void *src[1] = {flags.ptr + n_blocks};
            const int nb[1] = {4};
            CBX_HIP(read_device_values(st, 1, src, nb, &hv));
            if (hv == 0) break;
        }
    }
    if (aligned(out)) k_unpack_words<<<quarter_grid, DIL_NT, 0, st>>>(cur, (uint32_t)words, out);
    else k_unpack_bits<<<pack_grid, DIL_NT, 0, st>>>(cur, g.rows, g.W, g.Wq, out);
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
