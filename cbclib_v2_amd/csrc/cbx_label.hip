// label() on gfx950: connected components of equal-valued non-zero elements.
//
// Ids are the rank of each component's smallest flat index in raster order
// (what label<I> in cbclib_v2/_src/src/label.cpp:109-175 produces by keeping
// a region at its minimum index, and what cuda_label.cu:125-136 + 221-237
// produce by sorting union-find roots).  Here the union-find links the larger
// root under the smaller one, so every root IS its component's minimum index;
// ranks then come from one exclusive scan of the root flags, no sort.
//
// Passes (P elements):
//  1. k_init      parent[i] = i for non-zero elements, -1 otherwise (written
//                 into `out`, which doubles as the parent array)
//  2. k_union     lock-free union over the neighbour shifts with a raster-
//                 earlier offset (the structure is symmetric): CAS links the
//                 larger root under the smaller
//  3. k_flatten   parent[i] = root(i)
//  4. (npts > 1)  k_count: component sizes at the roots
//  5. k_keep      keep[i] = i is a root (and its size >= npts)
//  6. scan        rank = exclusive prefix of keep -> n_labels = rank[P]
//  7. k_final     out[i] = rank[root(i)] + 1, or 0
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cbx_host.hpp"

namespace cbx {

constexpr int LBL_MAX_DIM = 7;

// 0: fastest applicable path; 1: skip the run path; 2: general path only
static int g_label_path = 0;
constexpr int LBL_MAX_SHIFTS = 1 << 16;   // host-side sanity bound (radii 1 in 7D: 3^7 - 1 = 2186 shifts)

struct LabelGeom {
    int ndim;
    int n_shifts;
    int32_t shape[LBL_MAX_DIM];
};

// element access: truthiness and the equality predicate of label.cpp:141
template <int DT>
struct Elem;
template <>
struct Elem<CBX_BOOL> {
    using V = uint8_t;
};
template <>
struct Elem<CBX_INT32> {
    using V = int32_t;
};
template <>
struct Elem<CBX_INT64> {
    using V = int64_t;
};

struct ThresholdF32 {
    static constexpr bool binary = true;
    const float *img;
    float vmin;
    __device__ __forceinline__ bool on(int64_t i) const { return img[i] > vmin; }
    __device__ __forceinline__ bool same(int64_t, int64_t) const { return true; }
    using Raw = float;
    // (streamed once: non-temporal)
    __device__ __forceinline__ Raw load(int64_t i) const { return __builtin_nontemporal_load(&img[i]); }
    __device__ __forceinline__ bool test(Raw r) const { return r > vmin; }
};

template <typename V>
struct ValueIn {
    static constexpr bool binary = sizeof(V) == 1;
    const V *v;
    __device__ __forceinline__ bool on(int64_t i) const { return v[i] != V(0); }
    // bool inputs: every non-zero byte is the same value (true)
    __device__ __forceinline__ bool same(int64_t i, int64_t k) const { return binary || v[i] == v[k]; }
    using Raw = V;
    __device__ __forceinline__ Raw load(int64_t i) const { return v[i]; }
    __device__ __forceinline__ bool test(Raw r) const { return r != V(0); }
};

template <typename In>
__global__ void k_init(In in, int64_t P, int32_t *__restrict__ par)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P) par[i] = in.on(i) ? (int32_t)i : -1;
}

__device__ __forceinline__ int32_t find_root(const int32_t *par, int32_t i)
{
    int32_t p = par[i];
    while (p != i) {
        i = p;
        p = par[i];
    }
    return i;
}

template <typename In>
__global__ void k_union(In in, int64_t P, LabelGeom g, const int32_t *__restrict__ shifts,
                        const int64_t *__restrict__ offs, int32_t *par)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    if (par[i] < 0) return;
    int32_t coord[LBL_MAX_DIM];
    int64_t rem = i;
    for (int d = g.ndim - 1; d >= 0; d--) {
        coord[d] = (int32_t)(rem % g.shape[d]);
        rem /= g.shape[d];
    }
    for (int s = 0; s < g.n_shifts; s++) {
        const int32_t *sh = shifts + s * g.ndim;
        bool ok = true;
        for (int d = 0; d < g.ndim; d++) {
            int32_t c = coord[d] + sh[d];
            ok &= (c >= 0) & (c < g.shape[d]);
        }
        if (!ok) continue;
        const int64_t k = i + offs[s];
        if (__hip_atomic_load(&par[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) continue;
        if (!in.same(i, k)) continue;
        int32_t a = (int32_t)i, b = (int32_t)k;
        while (true) {
            // roots of both trees (parents only ever decrease)
            int32_t pa = __hip_atomic_load(&par[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (pa != a) {
                a = pa;
                pa = __hip_atomic_load(&par[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            int32_t pb = __hip_atomic_load(&par[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (pb != b) {
                b = pb;
                pb = __hip_atomic_load(&par[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (a == b) break;
            if (a < b) {
                const int32_t t = a;
                a = b;
                b = t;
            }
            // link the larger root a under the smaller root b
            int32_t old = atomicCAS(&par[a], a, b);
            if (old == a) break;
            a = old;
        }
    }
}

__global__ void k_flatten(int64_t P, int32_t *__restrict__ par)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    if (par[i] < 0) return;
    par[i] = find_root(par, (int32_t)i);
}

__global__ void k_count(int64_t P, const int32_t *__restrict__ par, int32_t *__restrict__ size)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int32_t r = par[i];
    if (r >= 0) atomicAdd(&size[r], 1);
}

__global__ void k_keep(int64_t P, const int32_t *__restrict__ par, const int32_t *__restrict__ size, int64_t npts,
                       int32_t *__restrict__ keep)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    keep[i] = (par[i] == (int32_t)i && (!size || size[i] >= npts)) ? 1 : 0;
}

__global__ void k_final(int64_t P, int32_t *__restrict__ par, const int32_t *__restrict__ keep,
                        const int32_t *__restrict__ rank)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int32_t r = par[i];
    par[i] = (r >= 0 && keep[r]) ? rank[r] + 1 : 0;
}

// ---------------------------------------------------------------------------
// Band path: 2D structures (or per-frame stacks, leading radii 0) whose
// shifts lie in the 3x3 neighbourhood.  A band is up to BAND_PX pixels of
// whole rows of one frame, so bands in order are pixels in raster order.
//   L1 k_band_local   union-find of the band in LDS; a band-local root is its
//                     component's smallest index; local ids k = rank of the
//                     local roots in the band.  Writes the local ids of the
//                     band's first and last rows, and one node per local
//                     root (node = band * cap + k) of a global union-find.
//   L2 k_band_merge   unions the nodes across each band boundary (the
//                     smaller node, i.e. the smaller raster index, is root).
//   L3 k_band_resolve / k_band_count / k_band_rank / k_node_label: global
//                     roots per band, their ranks (scan over bands), labels
//                     of every node.
//   L4 k_band_final   recomputes the band's local union-find (same result)
//                     and writes each pixel's label.
// HBM traffic: the input twice, the labels once, the band borders (int16).
// ---------------------------------------------------------------------------
constexpr int BAND_PX = 8192;
constexpr int BAND_WORDS = BAND_PX / 32;
constexpr int BAND_NT = 256;
static_assert(BAND_WORDS <= BAND_NT, "one root-mask word per thread");

struct BandGeom {
    int64_t F;
    int32_t H, W;
    int32_t B;               // rows per band
    int32_t bands_per_frame; // ceil(H / B)
    int64_t n_bands;
    int32_t cap;             // node slots per band
    int32_t prev;            // bit 0: (0,-1), 1: (-1,-1), 2: (-1,0), 3: (-1,+1)
};

// (relaxed workgroup-scope atomic loads: every step re-reads the entry,
// without the waits for all outstanding memory operations that volatile
// accesses get)
__device__ __forceinline__ int lds_find(int32_t *par, int i)
{
    int p = __hip_atomic_load(&par[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != i) {
        i = p;
        p = __hip_atomic_load(&par[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return i;
}

__device__ __forceinline__ void lds_union(int32_t *par, int a, int b)
{
    while (true) {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicCAS(&par[a], a, b);
        if (old == a) return;
        a = old;
    }
}

__device__ __forceinline__ void global_union(int32_t *par, int32_t a, int32_t b)
{
    while (true) {
        int32_t pa = __hip_atomic_load(&par[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (pa != a) {
            a = pa;
            pa = __hip_atomic_load(&par[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        int32_t pb = __hip_atomic_load(&par[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (pb != b) {
            b = pb;
            pb = __hip_atomic_load(&par[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (a == b) return;
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        const int32_t old = atomicCAS(&par[a], a, b);
        if (old == a) return;
        a = old;
    }
}

// exclusive scan of one int per thread over a 256-thread block
__device__ __forceinline__ int block_exclusive_scan(int v, int *tmp, int &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; w++) off += tmp[w];
    total = tmp[0] + tmp[1] + tmp[2] + tmp[3];
    __syncthreads();
    return off + x - v;
}

struct BandShared {
    int32_t par[BAND_PX];
    uint32_t words[BAND_WORDS];
    int32_t wpref[BAND_WORDS];
    int32_t tmp[4];
    int32_t nroots;
};

__device__ __forceinline__ int band_kid(const BandShared &S, int r)
{
    return S.wpref[r >> 5] + __popc(S.words[r >> 5] & ((1u << (r & 31)) - 1u));
}

// Local union-find of one band in LDS; leaves par[i] = local root (or -1),
// the root bitmask, its word prefix and the root count.
template <typename In>
__device__ void band_ccl(In in, const BandGeom &g, int64_t band, BandShared &S, int &npx, int &rows, int64_t &gbase)
{
    const int64_t f = band / g.bands_per_frame;
    const int bb = (int)(band - f * g.bands_per_frame);
    const int r0 = bb * g.B;
    rows = min(g.B, g.H - r0);
    npx = rows * g.W;
    gbase = (f * g.H + r0) * (int64_t)g.W;
    const int W = g.W, tid = threadIdx.x;
    for (int i = tid; i < npx; i += BAND_NT) S.par[i] = in.on(gbase + i) ? i : -1;
    __syncthreads();
    for (int i = tid; i < npx; i += BAND_NT) {
        if (S.par[i] < 0) continue;
        const int x = i % W, r = i / W;
        if ((g.prev & 1) && x > 0 && S.par[i - 1] >= 0 && in.same(gbase + i, gbase + i - 1)) lds_union(S.par, i, i - 1);
        if (r > 0) {
            if ((g.prev & 2) && x > 0 && S.par[i - W - 1] >= 0 && in.same(gbase + i, gbase + i - W - 1))
                lds_union(S.par, i, i - W - 1);
            if ((g.prev & 4) && S.par[i - W] >= 0 && in.same(gbase + i, gbase + i - W)) lds_union(S.par, i, i - W);
            if ((g.prev & 8) && x + 1 < W && S.par[i - W + 1] >= 0 && in.same(gbase + i, gbase + i - W + 1))
                lds_union(S.par, i, i - W + 1);
        }
    }
    __syncthreads();
    for (int i = tid; i < npx; i += BAND_NT)
        if (S.par[i] >= 0) S.par[i] = lds_find(S.par, i);
    __syncthreads();
    const int nwords = (npx + 31) >> 5;
    uint32_t bits = 0;
    if (tid < nwords) {
        for (int j = 0; j < 32; j++) {
            const int i = (tid << 5) + j;
            if (i < npx && S.par[i] == i) bits |= 1u << j;
        }
        S.words[tid] = bits;
    }
    int total = 0;
    const int pre = block_exclusive_scan(tid < nwords ? __popc(bits) : 0, S.tmp, total);
    if (tid < nwords) S.wpref[tid] = pre;
    if (tid == 0) S.nroots = total;
    __syncthreads();
}

template <typename In, bool SIZES>
__global__ void __launch_bounds__(BAND_NT)
k_band_local(In in, BandGeom g, int32_t *__restrict__ nroots, int16_t *__restrict__ top, int16_t *__restrict__ bot,
             int32_t *__restrict__ gpar, int32_t *__restrict__ lsize, int32_t *__restrict__ gsize)
{
    __shared__ BandShared S;
    __shared__ int32_t cnt[SIZES ? BAND_PX : 1];
    const int64_t band = blockIdx.x;
    int npx, rows;
    int64_t gbase;
    band_ccl(in, g, band, S, npx, rows, gbase);
    const int tid = threadIdx.x, W = g.W;
    const int64_t node0 = band * g.cap;
    if (tid == 0) nroots[band] = S.nroots;
    for (int i = tid; i < npx; i += BAND_NT) {
        if (S.par[i] != i) continue;
        const int64_t node = node0 + band_kid(S, i);
        gpar[node] = (int32_t)node;
        if (SIZES) gsize[node] = 0;
    }
    if (SIZES) {
        for (int k = tid; k < S.nroots; k += BAND_NT) cnt[k] = 0;
        __syncthreads();
        for (int i = tid; i < npx; i += BAND_NT)
            if (S.par[i] >= 0) atomicAdd(&cnt[band_kid(S, S.par[i])], 1);
        __syncthreads();
        for (int k = tid; k < S.nroots; k += BAND_NT) lsize[node0 + k] = cnt[k];
    }
    const int last = (rows - 1) * W;
    for (int x = tid; x < W; x += BAND_NT) {
        const int pt = S.par[x], pb = S.par[last + x];
        top[band * W + x] = (int16_t)(pt >= 0 ? band_kid(S, pt) : -1);
        bot[band * W + x] = (int16_t)(pb >= 0 ? band_kid(S, pb) : -1);
    }
}

template <typename In>
__global__ void __launch_bounds__(BAND_NT)
k_band_merge(In in, BandGeom g, const int16_t *__restrict__ top, const int16_t *__restrict__ bot, int32_t *gpar)
{
    const int64_t band = blockIdx.x;
    const int64_t f = band / g.bands_per_frame;
    const int bb = (int)(band - f * g.bands_per_frame);
    if (bb == 0) return;
    const int x = blockIdx.y * BAND_NT + threadIdx.x;
    const int W = g.W;
    if (x >= W) return;
    const int t = top[band * W + x];
    if (t < 0) return;
    const int32_t nt = (int32_t)(band * g.cap + t);
    const int64_t gt = (f * g.H + (int64_t)bb * g.B) * W + x;
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
        if (!(g.prev & (dx < 0 ? 2 : (dx == 0 ? 4 : 8)))) continue;
        const int xb = x + dx;
        if (xb < 0 || xb >= W) continue;
        const int v = bot[(band - 1) * W + xb];
        if (v < 0) continue;
        if (!in.same(gt, gt - W + dx)) continue;
        global_union(gpar, nt, (int32_t)((band - 1) * g.cap + v));
    }
}

template <bool SIZES>
__global__ void __launch_bounds__(BAND_NT)
k_band_resolve(BandGeom g, const int32_t *__restrict__ nroots, int32_t *gpar, const int32_t *__restrict__ lsize,
               int32_t *gsize)
{
    const int64_t band = blockIdx.x;
    const int n = nroots[band];
    for (int k = threadIdx.x; k < n; k += BAND_NT) {
        const int32_t node = (int32_t)(band * g.cap + k);
        int32_t r = node, p = gpar[r];
        while (p != r) {
            r = p;
            p = gpar[r];
        }
        gpar[node] = r;
        if (SIZES) atomicAdd(&gsize[r], lsize[node]);
    }
}

__global__ void __launch_bounds__(BAND_NT)
k_band_rank(BandGeom g, const int32_t *__restrict__ nroots, const int32_t *__restrict__ gpar,
            const int32_t *__restrict__ gsize, int64_t npts, const int32_t *__restrict__ band_off,
            int32_t *__restrict__ band_cnt, int32_t *__restrict__ lab)
{
    __shared__ int32_t tmp[4];
    const int64_t band = blockIdx.x;
    const int n = nroots[band];
    int base = band_off ? band_off[band] : 0;
    for (int k0 = 0; k0 < n; k0 += BAND_NT) {
        const int k = k0 + threadIdx.x;
        const int32_t node = (int32_t)(band * g.cap + k);
        const bool keep = k < n && gpar[node] == node && (!gsize || gsize[node] >= npts);
        int total = 0;
        const int pre = block_exclusive_scan(keep ? 1 : 0, tmp, total);
        if (band_off && k < n && gpar[node] == node) lab[node] = keep ? base + pre + 1 : 0;
        base += total;
    }
    if (!band_off && threadIdx.x == 0) band_cnt[band] = base;
}

__global__ void __launch_bounds__(BAND_NT)
k_node_label(BandGeom g, const int32_t *__restrict__ nroots, const int32_t *__restrict__ gpar, int32_t *lab)
{
    const int64_t band = blockIdx.x;
    const int n = nroots[band];
    for (int k = threadIdx.x; k < n; k += BAND_NT) {
        const int32_t node = (int32_t)(band * g.cap + k);
        const int32_t r = gpar[node];
        if (r != node) lab[node] = lab[r];
    }
}

template <typename In>
__global__ void __launch_bounds__(BAND_NT)
k_band_final(In in, BandGeom g, const int32_t *__restrict__ lab, int32_t *__restrict__ out)
{
    __shared__ BandShared S;
    const int64_t band = blockIdx.x;
    int npx, rows;
    int64_t gbase;
    band_ccl(in, g, band, S, npx, rows, gbase);
    const int64_t node0 = band * g.cap;
    for (int i = threadIdx.x; i < npx; i += BAND_NT) {
        const int p = S.par[i];
        out[gbase + i] = (p >= 0) ? lab[node0 + band_kid(S, p)] : 0;
    }
}

// ---------------------------------------------------------------------------
// Run path: binary inputs (bool arrays, thresholded images) with a 2D
// structure (or a per-frame stack, leading radii 0) that joins horizontal
// neighbours, (0,-1) and its mirror, and any of the three upper neighbours.
// A run is a maximal horizontal stretch of "on" pixels of one row.  With
// horizontal connectivity a run lies inside one component; runs are numbered
// in raster order of their first pixels, so the smallest run id of a
// component is the run holding its smallest flat index, and the union-find
// (smaller id is root) yields the reference's label order directly.
//   R1 k_run_pack     one wave per row: the row as 64-bit words (bit x%64 of
//                     word x/64 = pixel x is on) and its run count.
//   R2 scan           first run id of every row.
//   R3 k_run_band     union-find of the runs of a band of RUN_BAND rows in
//                     LDS (one workgroup per band); par[r] = band-local root.
//                     Contacts (a pixel of row y and an upper neighbour in
//                     row y-1, both on) are unioned where the run of row y or
//                     the run of row y-1 starts.  That covers every pair of
//                     touching runs: moving a contact one pixel to the left
//                     keeps it a contact until one of the two runs starts.
//   R4 k_run_border   the same for the first row of each band and the row
//                     above it, lock-free in global memory.
//                     (k_run_init: par[r] = r when no upper neighbour is in
//                     the structure; size[r] = 0 for npts > 1.)
//   R5 k_run_resolve  par[r] = root (read-only root walks, so every entry
//                     ends as its root); without a size limit also the number of
//                     roots per row.  (npts > 1: k_run_sizes adds each run's
//                     length to its root, k_run_count counts kept roots.)
//   R6 scan           first label of every row.
//   R7 k_run_rank     each root gets its label, stored as ~label in par.
//   R8 k_run_final    one wave per row: every pixel's label (0 if off):
//                     ~par[run] for a root, ~par[par[run]] for any other run.
// HBM traffic: the input once, the labels once; the bit rows (W/8 bytes per
// row) and the run arrays are small enough to stay in the caches.
// ---------------------------------------------------------------------------
struct RunGeom {
    int64_t rows; // frames * H
    int32_t H, W;
    int32_t nw;   // 64-bit words per row
    int32_t prev; // bit 1: (-1,-1), 2: (-1,0), 3: (-1,+1) neighbours
};

constexpr int RUN_NT = 256;
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int RUN_ROWS = RUN_NT / 64; // rows per workgroup, one per wave

__device__ __forceinline__ uint64_t le_mask(int b) { return (b >= 63) ? ~0ull : ((2ull << b) - 1ull); }

// sums and exclusive scans over groups of WIDTH lanes (64: the wave; 32:
// each half-wave on its own row); lane = the lane's index in its group
template <int WIDTH = 64>
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = WIDTH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WIDTH);
    return v;
}

template <int WIDTH = 64>
__device__ __forceinline__ int wave_excl_scan(int v, int lane)
{
    int x = v;
#pragma unroll
    for (int d = 1; d < WIDTH; d <<= 1) {
        const int y = __shfl_up(x, d, WIDTH);
        if (lane >= d) x += y;
    }
    return x - v;
}

// run starts of word w given word w-1 (pixel x starts a run if it is on and
// x - 1 is off or outside the row)
__device__ __forceinline__ uint64_t run_starts(uint64_t w, uint64_t left) { return w & ~((w << 1) | (left >> 63)); }

__device__ __forceinline__ int32_t run_find(int32_t *par, int32_t x)
{
    // path halving: a non-root's parent is only ever replaced by one of its
    // ancestors, and roots are only changed by the CAS in run_union
    while (true) {
        const int32_t p = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == x) return x;
        const int32_t gp = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gp == p) return p;
        __hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

__device__ __forceinline__ void run_union(int32_t *par, int32_t a, int32_t b)
{
    while (true) {
        a = run_find(par, a);
        b = run_find(par, b);
        if (a == b) return;
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        // link the larger root a under the smaller root b
        const int32_t old = atomicCAS(&par[a], a, b);
        if (old == a) return;
        a = old;
    }
}

constexpr int PACK_BATCH = 32;

template <typename In>
__global__ void __launch_bounds__(RUN_NT)
k_run_pack(In in, RunGeom g, uint64_t *__restrict__ bits, int32_t *__restrict__ cnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int64_t base = row * g.W;
    uint64_t *brow = bits + row * g.nw;
    int runs = 0;
    uint64_t carry = 0;
    for (int c0 = 0; c0 < g.nw; c0 += 64) {
        const int nwc = min(64, g.nw - c0);
        uint64_t mine = 0;
        // PACK_BATCH words per memory round trip (a 2048-pixel row in one)
        for (int k0 = 0; k0 < nwc; k0 += PACK_BATCH) {
            typename In::Raw v[PACK_BATCH];
#pragma unroll
            for (int k = 0; k < PACK_BATCH; k++) {
                const int64_t x = (int64_t)(c0 + k0 + k) * 64 + lane;
                v[k] = in.load(base + min(x, (int64_t)g.W - 1));
            }
#pragma unroll
            for (int k = 0; k < PACK_BATCH; k++) {
                const int64_t x = (int64_t)(c0 + k0 + k) * 64 + lane;
                const uint64_t m = __ballot(x < g.W && in.test(v[k]));
                if (lane == k0 + k) mine = m;
            }
        }
        uint64_t left = __shfl_up(mine, 1, 64);
        if (lane == 0) left = carry;
        runs += __popcll(run_starts(mine, left));
        if (lane < nwc) brow[c0 + lane] = mine;
        carry = __shfl(mine, nwc - 1, 64);
    }
    runs = wave_sum(runs);
    if (lane == 0) cnt[row] = runs;
}

template <bool SIZES>
__global__ void __launch_bounds__(RUN_NT)
k_run_init(RunGeom g, const int32_t *__restrict__ roff, int32_t *__restrict__ par, int32_t *__restrict__ size)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int32_t off = roff[row], n = roff[row + 1] - off;
    for (int k = lane; k < n; k += 64) {
        par[off + k] = off + k;
        if (SIZES) size[off + k] = 0;
    }
}

// Contacts between row y and row y-1 (same frame) for one wave: calls
// join(a, b) with the run ids of each contact pair that the event rule
// keeps (a run of row y, b of row y-1; b < a).
// The words a contact pass reads for one 64-word chunk of row y: row y's
// word w and w-1, row y-1's words w-1, w, w+1 (0 outside the row).
struct RowChunk {
    uint64_t cw, cwl, pw, pwl, pwr;
};

__device__ __forceinline__ RowChunk load_chunk(const uint64_t *__restrict__ cur, const uint64_t *__restrict__ up,
                                               int nw, int w)
{
    const bool act = w < nw;
    RowChunk r;
    r.cw = act ? cur[w] : 0;
    r.cwl = (act && w > 0) ? cur[w - 1] : 0;
    r.pw = act ? up[w] : 0;
    r.pwl = (act && w > 0) ? up[w - 1] : 0;
    r.pwr = (act && w + 1 < nw) ? up[w + 1] : 0;
    return r;
}

// Contacts of one chunk (see run_contacts); ccarry / pcarry count the runs of
// rows y and y-1 in the chunks before it.
template <int WIDTH, typename Join>
__device__ __forceinline__ void chunk_contacts(const RunGeom &g, const RowChunk &r, int32_t cbase, int32_t pbase,
                                               int &ccarry, int &pcarry, int lane, Join join)
{
    const uint64_t cw = r.cw, pw = r.pw, pwl = r.pwl, pwr = r.pwr;
    const uint64_t cst = run_starts(cw, r.cwl), pst = run_starts(pw, pwl);
    const int cn = __popcll(cst), pn = __popcll(pst);
    const int cpre = ccarry + wave_excl_scan<WIDTH>(cn, lane), ppre = pcarry + wave_excl_scan<WIDTH>(pn, lane);
#pragma unroll
    for (int d = -1; d <= 1; d++) {
        if (!(g.prev & (d < 0 ? 2 : (d == 0 ? 4 : 8)))) continue;
        // sp / sps: bit p = pixel / run start of row y-1 at p + d
        uint64_t sp, sps;
        if (d == 0) {
            sp = pw;
            sps = pst;
        } else if (d > 0) {
            sp = (pw >> 1) | (pwr << 63);
            sps = (pst >> 1) | ((pwr & ~(pw >> 63) & 1ull) << 63);
        } else {
            sp = (pw << 1) | (pwl >> 63);
            sps = (pst << 1) | ((pwl >> 63) & ~(pwl >> 62) & 1ull);
        }
        uint64_t ev = cw & sp & (cst | sps);
        while (ev) {
            const int p = __builtin_ctzll(ev);
            ev &= ev - 1;
            const uint64_t le = le_mask(p);
            const int32_t a = cbase + cpre + __popcll(cst & le) - 1;
            // runs of row y-1 starting at or before p + d
            int nb = ppre + __popcll(pst & le);
            if (d > 0) nb += (int)((sps >> p) & 1ull);
            if (d < 0) nb -= (int)((pst >> p) & 1ull);
            join(a, pbase + nb - 1);
        }
    }
    ccarry += wave_sum<WIDTH>(cn);
    pcarry += wave_sum<WIDTH>(pn);
}

// Contacts of row `row` with the row above, over a group of WIDTH lanes
// (lane: index in the group).
template <int WIDTH = 64, typename Join>
__device__ __forceinline__ void run_contacts(const RunGeom &g, const uint64_t *__restrict__ bits,
                                             const int32_t *__restrict__ roff, int64_t row, int lane, Join join)
{
    const int nw = g.nw;
    const uint64_t *cur = bits + row * nw, *up = cur - nw;
    const int32_t cbase = roff[row], pbase = roff[row - 1];
    int ccarry = 0, pcarry = 0;
    for (int c0 = 0; c0 < nw; c0 += WIDTH)
        chunk_contacts<WIDTH>(g, load_chunk(cur, up, nw, c0 + lane), cbase, pbase, ccarry, pcarry, lane, join);
}

// One workgroup per band of RUN_BAND rows of one frame.  The band's runs are
// the id range [roff[y0], roff[y0 + rows]); if it fits RUN_LCAP they are
// unioned in LDS and every run's band-local root is written to par (a
// band-local root is the smallest id of its band component, so the forest
// stays min-rooted); otherwise par is initialised and the band's contacts
// are unioned in global memory.  Contacts across band borders are left to
// k_run_border.
constexpr int RUN_BAND = 32;
constexpr int RUN_LCAP = 4096;

// WIDTH 32 (rows of at most 32 words): each half-wave takes a row of its own.
template <int WIDTH>
__global__ void __launch_bounds__(RUN_NT)
k_run_band(RunGeom g, const uint64_t *__restrict__ bits, const int32_t *__restrict__ roff, int32_t *par)
{
    __shared__ int32_t lpar[RUN_LCAP];
    constexpr int GROUPS = RUN_NT / WIDTH;   // rows in flight per workgroup
    const int lane = threadIdx.x % WIDTH, grp = threadIdx.x / WIDTH;
    const int bpf = (g.H + RUN_BAND - 1) / RUN_BAND;
    const int64_t f = blockIdx.x / bpf;
    const int y0 = (int)(blockIdx.x - f * bpf) * RUN_BAND;
    const int rows = min(RUN_BAND, g.H - y0);
    const int64_t row0 = f * g.H + y0;
    const int32_t r0 = roff[row0], nb = roff[row0 + rows] - r0;
    if (nb <= RUN_LCAP) {
        for (int k = threadIdx.x; k < nb; k += RUN_NT) lpar[k] = k;
        __syncthreads();
        for (int y = 1 + grp; y < rows; y += GROUPS)
            run_contacts<WIDTH>(g, bits, roff, row0 + y, lane,
                                [&](int32_t a, int32_t b) { lds_union(lpar, a - r0, b - r0); });
        __syncthreads();
        for (int k = threadIdx.x; k < nb; k += RUN_NT) par[r0 + k] = r0 + lds_find(lpar, k);
    } else {
        for (int k = threadIdx.x; k < nb; k += RUN_NT)
            __hip_atomic_store(&par[r0 + k], r0 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        __syncthreads();
        for (int y = 1 + grp; y < rows; y += GROUPS)
            run_contacts<WIDTH>(g, bits, roff, row0 + y, lane, [&](int32_t a, int32_t b) { run_union(par, a, b); });
    }
}

// Contacts between the first row of a band and the row above it (same frame).
template <int WIDTH>
__global__ void __launch_bounds__(RUN_NT)
k_run_border(RunGeom g, const uint64_t *__restrict__ bits, const int32_t *__restrict__ roff, int64_t n_bands,
             int32_t *par)
{
    const int lane = threadIdx.x % WIDTH;
    const int64_t band = (int64_t)blockIdx.x * (RUN_NT / WIDTH) + threadIdx.x / WIDTH;
    if (band >= n_bands) return;
    const int bpf = (g.H + RUN_BAND - 1) / RUN_BAND;
    const int64_t f = band / bpf;
    const int y0 = (int)(band - f * bpf) * RUN_BAND;
    if (y0 == 0) return;
    run_contacts<WIDTH>(g, bits, roff, f * g.H + y0, lane, [&](int32_t a, int32_t b) { run_union(par, a, b); });
}

// par[r] = root of run r, one lane per run (the root walks of a row's few
// runs would leave most lanes of a row's wave idle); a grid-stride loop over
// the *n_runs runs, so the launch needs no host read of the count
__global__ void __launch_bounds__(RUN_NT)
k_run_resolve_flat(const int32_t *__restrict__ n_runs, int32_t *par)
{
    const int64_t n = *n_runs;
    for (int64_t r = (int64_t)blockIdx.x * RUN_NT + threadIdx.x; r < n; r += (int64_t)gridDim.x * RUN_NT) {
        // read-only walk: a halving store here could replace another run's
        // freshly written root by a mere ancestor
        int32_t root = (int32_t)r, p = __hip_atomic_load(&par[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (p != root) {
            root = p;
            p = __hip_atomic_load(&par[root], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (root != (int32_t)r) __hip_atomic_store(&par[r], root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// roots per row (after the resolve: a run is a root iff par[r] == r)
__global__ void __launch_bounds__(RUN_NT)
k_run_roots(RunGeom g, const int32_t *__restrict__ roff, const int32_t *__restrict__ par, int32_t *__restrict__ kcnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int32_t off = roff[row], n = roff[row + 1] - off;
    int roots = 0;
    for (int k = lane; k < n; k += 64) roots += par[off + k] == off + k;
    roots = wave_sum(roots);
    if (lane == 0) kcnt[row] = roots;
}

template <bool COUNT>
__global__ void __launch_bounds__(RUN_NT)
k_run_resolve(RunGeom g, const int32_t *__restrict__ roff, int32_t *par, int32_t *__restrict__ kcnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int32_t off = roff[row], n = roff[row + 1] - off;
    int roots = 0;
    for (int k = lane; k < n; k += 64) {
        const int32_t r = off + k;
        // read-only walk: a halving store here could replace another run's
        // freshly written root by a mere ancestor
        int32_t root = r, p = __hip_atomic_load(&par[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (p != root) {
            root = p;
            p = __hip_atomic_load(&par[root], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (root == r) roots++;
        else __hip_atomic_store(&par[r], root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (COUNT) {
        roots = wave_sum(roots);
        if (lane == 0) kcnt[row] = roots;
    }
}

__global__ void __launch_bounds__(RUN_NT)
k_run_sizes(RunGeom g, const uint64_t *__restrict__ bits, const int32_t *__restrict__ roff,
            const int32_t *__restrict__ par, int32_t *size)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const uint64_t *cur = bits + row * g.nw;
    const int32_t cbase = roff[row];
    int ccarry = 0;
    for (int c0 = 0; c0 < g.nw; c0 += 64) {
        const int w = c0 + lane;
        const bool act = w < g.nw;
        const uint64_t cw = act ? cur[w] : 0, cwl = (act && w > 0) ? cur[w - 1] : 0;
        const uint64_t cst = run_starts(cw, cwl);
        const int cn = __popcll(cst);
        const int cpre = ccarry + wave_excl_scan(cn, lane);
        // pieces of runs in this word: the runs starting here, and the run
        // continuing from word w-1
        uint64_t piece = cst | (cw & 1ull);
        while (piece) {
            const int s = __builtin_ctzll(piece);
            piece &= piece - 1;
            const uint64_t rest = ~(cw >> s);
            const int len = rest ? __builtin_ctzll(rest) : 64;
            const int32_t r = cbase + cpre + __popcll(cst & le_mask(s)) - 1;
            atomicAdd(&size[par[r]], len);
        }
        ccarry += wave_sum(cn);
    }
}

__global__ void __launch_bounds__(RUN_NT)
k_run_count(RunGeom g, const int32_t *__restrict__ roff, const int32_t *__restrict__ par,
            const int32_t *__restrict__ size, int64_t npts, int32_t *__restrict__ kcnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int32_t off = roff[row], n = roff[row + 1] - off;
    int kept = 0;
    for (int k = lane; k < n; k += 64) {
        const int32_t r = off + k;
        if (par[r] == r && size[r] >= npts) kept++;
    }
    kept = wave_sum(kept);
    if (lane == 0) kcnt[row] = kept;
}

__global__ void __launch_bounds__(RUN_NT)
k_run_rank(RunGeom g, const int32_t *__restrict__ roff, const int32_t *__restrict__ koff, int32_t *par,
           const int32_t *__restrict__ size, int64_t npts)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int32_t off = roff[row], n = roff[row + 1] - off;
    int32_t base = koff[row];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const int32_t r = off + k;
        const bool root = k < n && par[r] == r;
        const bool keep = root && (!size || size[r] >= npts);
        const uint64_t m = __ballot(keep);
        if (root) par[r] = ~(keep ? base + __popcll(m & below) + 1 : 0);
        base += __popcll(m);
    }
}

__global__ void __launch_bounds__(RUN_NT)
k_run_final(RunGeom g, const uint64_t *__restrict__ bits, const int32_t *__restrict__ roff,
            const int32_t *__restrict__ par, int32_t *__restrict__ out, int vec)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * RUN_ROWS + (threadIdx.x >> 6);
    if (row >= g.rows) return;
    const int nw = g.nw;
    const uint64_t *cur = bits + row * nw;
    int32_t *orow = out + row * g.W;
    const int32_t cbase = roff[row];
    int ccarry = 0;
    for (int c0 = 0; c0 < nw; c0 += 64) {
        const int w = c0 + lane;
        const bool act = w < nw;
        const uint64_t cw = act ? cur[w] : 0, cwl = (act && w > 0) ? cur[w - 1] : 0;
        const uint64_t cst = run_starts(cw, cwl);
        const int cn = __popcll(cst);
        const int cpre = ccarry + wave_excl_scan(cn, lane);
        const int nwc = min(64, nw - c0);
        // 256 pixels (4 words) per step: lane l takes pixels 4(l%16)..+3 of
        // word l/16
        for (int j0 = 0; j0 < nwc; j0 += 4) {
            const int jj = j0 + (lane >> 4);
            const uint64_t wj = __shfl(cw, jj, 64);
            const uint64_t sj = __shfl(cst, jj, 64);
            const int pj = __shfl(cpre, jj, 64);
            const int b0 = (lane & 15) * 4;
            const int64_t x0 = (int64_t)(c0 + jj) * 64 + b0;
            const uint32_t nib = (uint32_t)(wj >> b0) & 0xFu;
            int v[4] = {0, 0, 0, 0};
            if (nib) {
                // a root holds ~label, any other run its root: two gathers,
                // each issued for the 4 pixels before any result is used
                int32_t q[4];
#pragma unroll
                for (int i = 0; i < 4; i++)
                    q[i] = ((nib >> i) & 1u) ? par[cbase + pj + __popcll(sj & le_mask(b0 + i)) - 1] : -1;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (q[i] >= 0) q[i] = par[q[i]];
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = ((nib >> i) & 1u) ? ~q[i] : 0;
            }
            if (vec && x0 + 3 < g.W) {
                // write-once output: non-temporal, not kept in the caches
                const v4i q = {v[0], v[1], v[2], v[3]};
                __builtin_nontemporal_store(q, reinterpret_cast<v4i *>(orow + x0));
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (x0 + i < g.W) orow[x0 + i] = v[i];
            }
        }
        ccarry += wave_sum(cn);
    }
}

template <typename In>
static int label_runs(In in, int ndim, const int64_t *shape, int prev, int64_t npts, int32_t *out, int64_t *n_labels,
                      hipStream_t st)
{
    RunGeom g;
    g.H = (int32_t)shape[ndim - 2];
    g.W = (int32_t)shape[ndim - 1];
    int64_t F = 1;
    for (int d = 0; d < ndim - 2; d++) F *= shape[d];
    g.rows = F * g.H;
    g.nw = (g.W + 63) / 64;
    g.prev = prev;
    const int64_t R = g.rows;
    // horizontal neighbours join: at most ceil(W / 2) runs per row
    const int64_t max_runs = R * ((g.W + 1) / 2);
    if (max_runs >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "label: too many elements for the run path");
    const bool sizes = npts > 1;

    DevBuf<uint64_t> bits;
    DevBuf<int32_t> rcnt, roff, par, kcnt, koff, size;
    CBX_HIP(bits.alloc(R * g.nw, st));
    CBX_HIP(rcnt.alloc(R, st));
    CBX_HIP(roff.alloc(R + 1, st));
    CBX_HIP(par.alloc(max_runs, st));
    CBX_HIP(kcnt.alloc(R, st));
    CBX_HIP(koff.alloc(R + 1, st));
    if (sizes) CBX_HIP(size.alloc(max_runs, st));
    const unsigned grid = (unsigned)((R + RUN_ROWS - 1) / RUN_ROWS);
    KernelTimer timer("label_total", st);
    k_run_pack<In><<<grid, RUN_NT, 0, st>>>(in, g, bits.ptr, rcnt.ptr);
    CBX_HIP(exclusive_scan_total_i32(rcnt.ptr, nullptr, roff.ptr, R, st));
    if (sizes) k_run_init<true><<<grid, RUN_NT, 0, st>>>(g, roff.ptr, par.ptr, size.ptr);
    if (prev & 14) {
        const int64_t n_bands = (int64_t)(R / g.H) * ((g.H + RUN_BAND - 1) / RUN_BAND);
        if (g.nw <= 32) {
            // rows of at most 2048 pixels: one row per half-wave
            k_run_band<32><<<(unsigned)n_bands, RUN_NT, 0, st>>>(g, bits.ptr, roff.ptr, par.ptr);
            k_run_border<32><<<(unsigned)((n_bands + 2 * RUN_ROWS - 1) / (2 * RUN_ROWS)), RUN_NT, 0, st>>>(
                g, bits.ptr, roff.ptr, n_bands, par.ptr);
        } else {
            k_run_band<64><<<(unsigned)n_bands, RUN_NT, 0, st>>>(g, bits.ptr, roff.ptr, par.ptr);
            k_run_border<64><<<(unsigned)((n_bands + RUN_ROWS - 1) / RUN_ROWS), RUN_NT, 0, st>>>(g, bits.ptr, roff.ptr,
                                                                                                 n_bands, par.ptr);
        }
    } else if (!sizes) {
        k_run_init<false><<<grid, RUN_NT, 0, st>>>(g, roff.ptr, par.ptr, nullptr);
    }
    if (sizes) {
        k_run_resolve<false><<<grid, RUN_NT, 0, st>>>(g, roff.ptr, par.ptr, nullptr);
        k_run_sizes<<<grid, RUN_NT, 0, st>>>(g, bits.ptr, roff.ptr, par.ptr, size.ptr);
        k_run_count<<<grid, RUN_NT, 0, st>>>(g, roff.ptr, par.ptr, size.ptr, npts, kcnt.ptr);
    } else {
        // resolve one lane per run (the run total is the scan's last entry),
        // 8 workgroups per CU of 256
        k_run_resolve_flat<<<2048, RUN_NT, 0, st>>>(roff.ptr + R, par.ptr);
        k_run_roots<<<grid, RUN_NT, 0, st>>>(g, roff.ptr, par.ptr, kcnt.ptr);
    }
    CBX_HIP(exclusive_scan_total_i32(kcnt.ptr, nullptr, koff.ptr, R, st));
    k_run_rank<<<grid, RUN_NT, 0, st>>>(g, roff.ptr, koff.ptr, par.ptr, sizes ? size.ptr : nullptr, npts);
    const int vec = (g.W % 4 == 0 && ((uintptr_t)out & 15) == 0) ? 1 : 0;
    k_run_final<<<grid, RUN_NT, 0, st>>>(g, bits.ptr, roff.ptr, par.ptr, out, vec);
    CBX_HIP(hipGetLastError());
    timer.stop();
    int32_t n = 0;
    {
        const void *src[1] = {koff.ptr + R};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        n = (int32_t)(uint32_t)hv;
    }
    *n_labels = n;
    return CBX_OK;
}

// Returns -1 if the structure is not handled by the band path.
static int band_prev_mask(int ndim, const int64_t *shape, const int64_t *shifts, int64_t n_shifts)
{
    if (ndim < 2) return -1;
    int mask = 0;
    for (int64_t j = 0; j < n_shifts; j++) {
        const int64_t *s = shifts + j * ndim;
        for (int d = 0; d < ndim - 2; d++)
            if (s[d] != 0) return -1;
        const int64_t dy = s[ndim - 2], dx = s[ndim - 1];
        if (dy < -1 || dy > 1 || dx < -1 || dx > 1) return -1;
        if (dy == 0 && dx == 0) continue;
        // keep the raster-earlier half; its mirror must be present too
        bool mirrored = false;
        for (int64_t m = 0; m < n_shifts; m++) {
            const int64_t *t = shifts + m * ndim;
            bool z = true;
            for (int d = 0; d < ndim - 2; d++) z &= t[d] == 0;
            if (z && t[ndim - 2] == -dy && t[ndim - 1] == -dx) mirrored = true;
        }
        if (!mirrored) return -1;
        if (dy == 0 && dx == -1) mask |= 1;
        if (dy == -1 && dx == -1) mask |= 2;
        if (dy == -1 && dx == 0) mask |= 4;
        if (dy == -1 && dx == 1) mask |= 8;
    }
    (void)shape;
    return mask;
}

template <typename In>
static int label_band(In in, int ndim, const int64_t *shape, int prev, int64_t npts, int32_t *out, int64_t *n_labels,
                      hipStream_t st)
{
    BandGeom g;
    g.H = (int32_t)shape[ndim - 2];
    g.W = (int32_t)shape[ndim - 1];
    g.F = 1;
    for (int d = 0; d < ndim - 2; d++) g.F *= shape[d];
    g.B = max(1, min(g.H, BAND_PX / g.W));
    g.bands_per_frame = (g.H + g.B - 1) / g.B;
    g.n_bands = g.F * g.bands_per_frame;
    // node slots per band = the most components a band can hold: for binary
    // inputs a left (up) neighbour merges horizontally (vertically) adjacent
    // pixels, so a row (column) holds at most ceil(W/2) (ceil(B/2)) of them;
    // equal-value inputs can have a component per pixel
    int cap = g.B * g.W;
    if (In::binary && (prev & 1)) cap = min(cap, g.B * ((g.W + 1) / 2));
    if (In::binary && (prev & 4)) cap = min(cap, g.W * ((g.B + 1) / 2));
    g.cap = cap;
    g.prev = prev;
    const int64_t n_nodes = g.n_bands * g.cap;
    if (n_nodes >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "label: too many bands");
    const bool sizes = npts > 1;

    DevBuf<int32_t> nroots, gpar, lab, lsize, gsize, band_cnt, band_off;
    DevBuf<int16_t> top, bot;
    CBX_HIP(nroots.alloc(g.n_bands, st));
    CBX_HIP(gpar.alloc(n_nodes, st));
    CBX_HIP(lab.alloc(n_nodes, st));
    CBX_HIP(top.alloc(g.n_bands * g.W, st));
    CBX_HIP(bot.alloc(g.n_bands * g.W, st));
    CBX_HIP(band_cnt.alloc(g.n_bands + 1, st));
    CBX_HIP(band_off.alloc(g.n_bands + 1, st));
    if (sizes) {
        CBX_HIP(lsize.alloc(n_nodes, st));
        CBX_HIP(gsize.alloc(n_nodes, st));
    }
    KernelTimer timer("label_total", st);
    const unsigned nb = (unsigned)g.n_bands;
    if (sizes)
        k_band_local<In, true><<<nb, BAND_NT, 0, st>>>(in, g, nroots.ptr, top.ptr, bot.ptr, gpar.ptr, lsize.ptr, gsize.ptr);
    else
        k_band_local<In, false><<<nb, BAND_NT, 0, st>>>(in, g, nroots.ptr, top.ptr, bot.ptr, gpar.ptr, nullptr, nullptr);
    k_band_merge<In><<<dim3(nb, (g.W + BAND_NT - 1) / BAND_NT), BAND_NT, 0, st>>>(in, g, top.ptr, bot.ptr, gpar.ptr);
    if (sizes) k_band_resolve<true><<<nb, BAND_NT, 0, st>>>(g, nroots.ptr, gpar.ptr, lsize.ptr, gsize.ptr);
    else k_band_resolve<false><<<nb, BAND_NT, 0, st>>>(g, nroots.ptr, gpar.ptr, nullptr, nullptr);
    CBX_HIP(hipMemsetAsync(band_cnt.ptr + g.n_bands, 0, sizeof(int32_t), st));
    k_band_rank<<<nb, BAND_NT, 0, st>>>(g, nroots.ptr, gpar.ptr, sizes ? gsize.ptr : nullptr, npts, nullptr,
                                        band_cnt.ptr, nullptr);
    CBX_HIP(exclusive_scan_i32(band_cnt.ptr, band_off.ptr, g.n_bands + 1, st));
    k_band_rank<<<nb, BAND_NT, 0, st>>>(g, nroots.ptr, gpar.ptr, sizes ? gsize.ptr : nullptr, npts, band_off.ptr,
                                        nullptr, lab.ptr);
    k_node_label<<<nb, BAND_NT, 0, st>>>(g, nroots.ptr, gpar.ptr, lab.ptr);
    k_band_final<In><<<nb, BAND_NT, 0, st>>>(in, g, lab.ptr, out);
    CBX_HIP(hipGetLastError());
    timer.stop();
    int32_t n = 0;
    {
        const void *src[1] = {band_off.ptr + g.n_bands};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        n = (int32_t)(uint32_t)hv;
    }
    *n_labels = n;
    return CBX_OK;
}

template <typename In>
static int label_dispatch(In in, int ndim, const int64_t *shape, const int64_t *shifts, int64_t n_shifts, int64_t npts,
                     int32_t *out, int64_t *n_labels, hipStream_t st)
{
    if (!n_labels) return fail(CBX_EVALUE, "n_labels must not be NULL");
    *n_labels = 0;
    if (ndim < 1 || ndim > LBL_MAX_DIM) return fail(CBX_EVALUE, "label supports 1 to 7 dimensions");
    int64_t P = 1;
    LabelGeom g;
    g.ndim = ndim;
    for (int d = 0; d < ndim; d++) {
        if (shape[d] < 0) return fail(CBX_EVALUE, "negative shape");
        if (shape[d] >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "dimension too large");
        g.shape[d] = (int32_t)shape[d];
        P *= shape[d];
    }
    if (P >= (int64_t(1) << 31) - 1) return fail(CBX_EVALUE, "label supports fewer than 2^31 - 1 elements");
    if (P == 0) return CBX_OK;
    {
        const int prev = band_prev_mask(ndim, shape, shifts, n_shifts);
        const int forced = g_label_path;
        if (prev >= 0 && In::binary && (prev & 1) && forced == 0)
            return label_runs(in, ndim, shape, prev, npts, out, n_labels, st);
        if (prev >= 0 && shape[ndim - 1] <= BAND_PX && shape[ndim - 1] >= 1 && forced <= 1)
            return label_band(in, ndim, shape, prev, npts, out, n_labels, st);
    }

    // keep the shifts that point to a raster-earlier neighbour when the set is
    // symmetric (every union is then made once); otherwise keep them all
    std::vector<int32_t> sh;
    std::vector<int64_t> off;
    std::vector<int64_t> strides(ndim, 1);
    for (int d = ndim - 2; d >= 0; d--) strides[d] = strides[d + 1] * shape[d + 1];
    auto offset_of = [&](const int64_t *s) {
        int64_t o = 0;
        for (int d = 0; d < ndim; d++) o += s[d] * strides[d];
        return o;
    };
    auto has = [&](const std::vector<int64_t> &v) {
        for (int64_t j = 0; j < n_shifts; j++) {
            bool eq = true;
            for (int d = 0; d < ndim; d++) eq &= shifts[j * ndim + d] == v[d];
            if (eq) return true;
        }
        return false;
    };
    bool symmetric = true;
    for (int64_t j = 0; j < n_shifts && symmetric; j++) {
        std::vector<int64_t> neg(ndim);
        for (int d = 0; d < ndim; d++) neg[d] = -shifts[j * ndim + d];
        symmetric = has(neg);
    }
    for (int64_t j = 0; j < n_shifts; j++) {
        const int64_t *s = shifts + j * ndim;
        bool zero = true;
        // lexicographic sign == sign of the raster offset for in-range shifts
        int sign = 0;
        for (int d = 0; d < ndim; d++) {
            zero &= s[d] == 0;
            if (sign == 0 && s[d] != 0) sign = (s[d] < 0) ? -1 : 1;
        }
        if (zero) continue;
        if (symmetric && sign > 0) continue;
        for (int d = 0; d < ndim; d++) sh.push_back((int32_t)s[d]);
        off.push_back(offset_of(s));
    }
    g.n_shifts = (int)off.size();
    if (g.n_shifts > LBL_MAX_SHIFTS) return fail(CBX_EVALUE, "structure has too many shifts");

    const int blk = 256;
    const int grid = grid_for(P, blk);
    // every buffer before the first kernel: a failed allocation leaves `out` untouched
    DevBuf<int32_t> dsh;
    DevBuf<int64_t> doff;
    DevBuf<int32_t> size, keep, rank;
    if (g.n_shifts) {
        CBX_HIP(dsh.alloc(sh.size(), st));
        CBX_HIP(doff.alloc(off.size(), st));
    }
    if (npts > 1) CBX_HIP(size.alloc(P, st));
    CBX_HIP(keep.alloc(P + 1, st));
    CBX_HIP(rank.alloc(P + 1, st));
    if (g.n_shifts) {
        CBX_HIP(hipMemcpyAsync(dsh.ptr, sh.data(), sizeof(int32_t) * sh.size(), hipMemcpyHostToDevice, st));
        CBX_HIP(hipMemcpyAsync(doff.ptr, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice, st));
    }
    KernelTimer timer("label_total", st);
    k_init<<<grid, blk, 0, st>>>(in, P, out);
    if (g.n_shifts) k_union<<<grid, blk, 0, st>>>(in, P, g, dsh.ptr, doff.ptr, out);
    k_flatten<<<grid, blk, 0, st>>>(P, out);
    CBX_HIP(hipGetLastError());

    if (npts > 1) {
        CBX_HIP(hipMemsetAsync(size.ptr, 0, sizeof(int32_t) * P, st));
        k_count<<<grid, blk, 0, st>>>(P, out, size.ptr);
    }
    CBX_HIP(hipMemsetAsync(keep.ptr + P, 0, sizeof(int32_t), st));
    k_keep<<<grid, blk, 0, st>>>(P, out, npts > 1 ? size.ptr : nullptr, npts, keep.ptr);
    CBX_HIP(exclusive_scan_i32(keep.ptr, rank.ptr, P + 1, st));
    k_final<<<grid, blk, 0, st>>>(P, out, keep.ptr, rank.ptr);
    CBX_HIP(hipGetLastError());
    timer.stop();
    int32_t n = 0;
    {
        const void *src[1] = {rank.ptr + P};
        const int nb[1] = {4};
        uint64_t hv = 0;
        CBX_HIP(read_device_values(st, 1, src, nb, &hv));
        n = (int32_t)(uint32_t)hv;
    }
    *n_labels = n;
    return CBX_OK;
}

// ---------------------------------------------------------------------------
// cbx_label_noalloc: the intermediate buffers come from `scratch`, a device
// array of NOALLOC_BYTES that every GPU holds from the library's load until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// the device array: its device address is looked up once per device with
// hipGetSymbolAddress and only that address is used.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];

// device address of this GPU's scratch array (the current device's)
static int scratch_address(void **p)
{
    *p = nullptr;
    int dev = 0;
    CBX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(CBX_ERUNTIME, "cbx_label_noalloc: device id out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q)
            return fail(CBX_ERUNTIME, std::string("cbx_label_noalloc: hipGetSymbolAddress(scratch): ") +
                                          hipGetErrorString(e));
        g_scratch_dev[dev] = q;
    }
    *p = g_scratch_dev[dev];
    return CBX_OK;
}

template <typename Run>
static int run_noalloc(hipStream_t st, Run &&run)
{
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    Workspace ws(st, base, NOALLOC_BYTES);
    rc = run();
    if (rc != CBX_OK && ws.overflow())
        return fail(CBX_EVALUE, "cbx_label_noalloc: the input needs more than the " +
                                    std::to_string(NOALLOC_BYTES >> 20) + " MiB of scratch (at least " +
                                    std::to_string(ws.requested()) + " bytes); use cbx_label");
    return rc;
}

} // namespace cbx

extern "C" {

int cbx_label_noalloc(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts,
                      int64_t n_shifts, int64_t npts, int32_t *out, int64_t *n_labels, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    return cbx::run_noalloc(st, [&]() {
        switch (dtype) {
        case CBX_BOOL:
            return cbx::label_dispatch(cbx::ValueIn<uint8_t>{(const uint8_t *)inp}, ndim, shape, shifts, n_shifts,
                                       npts, out, n_labels, st);
        case CBX_INT32:
            return cbx::label_dispatch(cbx::ValueIn<int32_t>{(const int32_t *)inp}, ndim, shape, shifts, n_shifts,
                                       npts, out, n_labels, st);
        case CBX_INT64:
            return cbx::label_dispatch(cbx::ValueIn<int64_t>{(const int64_t *)inp}, ndim, shape, shifts, n_shifts,
                                       npts, out, n_labels, st);
        default:
            return cbx::fail(CBX_EVALUE, "unsupported label input dtype");
        }
    });
}

int64_t cbx_label_noalloc_capacity(void) { return (int64_t)cbx::NOALLOC_BYTES; }

int cbx_label(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts, int64_t n_shifts,
              int64_t npts, int32_t *out, int64_t *n_labels, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    cbx::Workspace ws(st);
    switch (dtype) {
    case CBX_BOOL:
        return cbx::label_dispatch(cbx::ValueIn<uint8_t>{(const uint8_t *)inp}, ndim, shape, shifts, n_shifts, npts, out,
                              n_labels, st);
    case CBX_INT32:
        return cbx::label_dispatch(cbx::ValueIn<int32_t>{(const int32_t *)inp}, ndim, shape, shifts, n_shifts, npts, out,
                              n_labels, st);
    case CBX_INT64:
        return cbx::label_dispatch(cbx::ValueIn<int64_t>{(const int64_t *)inp}, ndim, shape, shifts, n_shifts, npts, out,
                              n_labels, st);
    default:
        return cbx::fail(CBX_EVALUE, "unsupported label input dtype");
    }
}

int cbx_label_threshold_f32(const float *image, float vmin, int ndim, const int64_t *shape, const int64_t *shifts,
                            int64_t n_shifts, int64_t npts, int32_t *out, int64_t *n_labels, void *stream)
{
    cbx::Workspace ws((hipStream_t)stream);
    return cbx::label_dispatch(cbx::ThresholdF32{image, vmin}, ndim, shape, shifts, n_shifts, npts, out, n_labels,
                          (hipStream_t)stream);
}

void cbx_label_force_path(int path) { cbx::g_label_path = (path >= 0 && path <= 2) ? path : 0; }

} // extern "C"
