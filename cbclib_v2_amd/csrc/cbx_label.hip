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
#include <string>
#include <vector>

#include "cbx_host.hpp"

namespace cbx {

constexpr int LBL_MAX_DIM = 7;
constexpr int LBL_MAX_SHIFTS = 256;

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
    const float *img;
    float vmin;
    __device__ __forceinline__ bool on(int64_t i) const { return img[i] > vmin; }
    __device__ __forceinline__ bool same(int64_t, int64_t) const { return true; }
};

template <typename V>
struct ValueIn {
    const V *v;
    __device__ __forceinline__ bool on(int64_t i) const { return v[i] != V(0); }
    __device__ __forceinline__ bool same(int64_t i, int64_t k) const { return v[i] == v[k]; }
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

template <typename In>
static int label_run(In in, int ndim, const int64_t *shape, const int64_t *shifts, int64_t n_shifts, int64_t npts,
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
    DevBuf<int32_t> dsh;
    DevBuf<int64_t> doff;
    if (g.n_shifts) {
        CBX_HIP(dsh.alloc(sh.size(), st));
        CBX_HIP(doff.alloc(off.size(), st));
        CBX_HIP(hipMemcpyAsync(dsh.ptr, sh.data(), sizeof(int32_t) * sh.size(), hipMemcpyHostToDevice, st));
        CBX_HIP(hipMemcpyAsync(doff.ptr, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice, st));
    }
    k_init<<<grid, blk, 0, st>>>(in, P, out);
    if (g.n_shifts) k_union<<<grid, blk, 0, st>>>(in, P, g, dsh.ptr, doff.ptr, out);
    k_flatten<<<grid, blk, 0, st>>>(P, out);
    CBX_HIP(hipGetLastError());

    DevBuf<int32_t> size, keep, rank;
    if (npts > 1) {
        CBX_HIP(size.alloc(P, st));
        CBX_HIP(hipMemsetAsync(size.ptr, 0, sizeof(int32_t) * P, st));
        k_count<<<grid, blk, 0, st>>>(P, out, size.ptr);
    }
    CBX_HIP(keep.alloc(P + 1, st));
    CBX_HIP(rank.alloc(P + 1, st));
    CBX_HIP(hipMemsetAsync(keep.ptr + P, 0, sizeof(int32_t), st));
    k_keep<<<grid, blk, 0, st>>>(P, out, npts > 1 ? size.ptr : nullptr, npts, keep.ptr);
    CBX_HIP(exclusive_scan_i32(keep.ptr, rank.ptr, P + 1, st));
    k_final<<<grid, blk, 0, st>>>(P, out, keep.ptr, rank.ptr);
    CBX_HIP(hipGetLastError());
    int32_t n = 0;
    CBX_HIP(hipMemcpyAsync(&n, rank.ptr + P, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CBX_HIP(hipStreamSynchronize(st));
    *n_labels = n;
    return CBX_OK;
}

} // namespace cbx

extern "C" {

int cbx_label(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts, int64_t n_shifts,
              int64_t npts, int32_t *out, int64_t *n_labels, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case CBX_BOOL:
        return cbx::label_run(cbx::ValueIn<uint8_t>{(const uint8_t *)inp}, ndim, shape, shifts, n_shifts, npts, out,
                              n_labels, st);
    case CBX_INT32:
        return cbx::label_run(cbx::ValueIn<int32_t>{(const int32_t *)inp}, ndim, shape, shifts, n_shifts, npts, out,
                              n_labels, st);
    case CBX_INT64:
        return cbx::label_run(cbx::ValueIn<int64_t>{(const int64_t *)inp}, ndim, shape, shifts, n_shifts, npts, out,
                              n_labels, st);
    default:
        return cbx::fail(CBX_EVALUE, "unsupported label input dtype");
    }
}

int cbx_label_threshold_f32(const float *image, float vmin, int ndim, const int64_t *shape, const int64_t *shifts,
                            int64_t n_shifts, int64_t npts, int32_t *out, int64_t *n_labels, void *stream)
{
    return cbx::label_run(cbx::ThresholdF32{image, vmin}, ndim, shape, shifts, n_shifts, npts, out, n_labels,
                          (hipStream_t)stream);
}

} // extern "C"
