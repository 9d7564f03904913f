// Region moments and p-values of a label map on gfx950: the consumers of
// label() in the reference's detection pipeline (cbclib_v2/label.py:2-4 ->
// _src/functions.py:401-641; CPU label.cpp:176-255, 534-583 on the moment
// classes of label.hpp:554-722; CUDA cuda_label.cu:380-829).
//
// Region j is the set of pixels whose label is index[j] (the first entry of
// index holding a label owns its accumulators; later duplicates read them).
//
// One pass over the label map (k_region_accumulate): one workgroup per patch
// of MOM_PR rows x MOM_PC chunks; each lane takes a chunk of MOM_CHUNK
// pixels of one row, folds consecutive pixels of equal label into run sums
// (sum v, sum v x, sum v x^2 along the row) and, per run, adds the region's
// raw moments in double (the other coordinates are constant along a row) and
// lowers the region's origin (its smallest flat index, integer atomicMin) in
// an LDS table of the patch's regions (a region that finds no entry goes to
// the global accumulators directly); the table then adds each region's sums
// to the global accumulators once.  Background chunks end after the label
// load.  k_region_finish turns the raw sums into the reference's
// quantities, in double, written in the data's type.  p_values counts
// (k_region_count) take the same patch table with 32-bit counts.
//
// The reference accumulates in the data's type relative to the region's
// first pixel (MomentsND::insert); here the sums are double and absolute, so
// f64 results agree with it to double rounding and f32 results are the
// correctly rounded double values (the reference's float sums carry their
// own rounding).  Double atomics make the order of additions vary between
// calls: last-bit differences in f64 outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>

#include "cbx_host.hpp"

namespace cbx {

constexpr int MOM_MAX_DIM = 7;
constexpr int MOM_CHUNK = 16;  // pixels per lane along a row
constexpr int MOM_NT = 256;
constexpr int32_t NO_SLOT = INT_MAX;

struct MomGeom {
    int64_t shape[MOM_MAX_DIM]; // zyx (the input's axis order)
    int64_t rows;               // product of every axis but the last
    int64_t W;                  // last axis
    int64_t chunks;             // chunks per row
    int32_t maxl;               // largest label in index (0: none)
    int32_t vec;                // rows start 16-byte aligned in the label map
    int32_t dvec;               // and in the data
    int32_t small;              // rows * chunks < 2^31: 32-bit index arithmetic
    int64_t n_pc;               // patch columns of the moment pass (chunks / MOM_PC, rounded up)
};

// accumulators per slot, xyz order (x = last axis):
// [mu, S_0 .. S_{N-1}, Q_0 .. Q_{N-1}, C_0 .. C_{NP-1}], S = sum v p, Q = sum v p^2,
// C over the pairs (i < j) in UniquePairs order (geometry.hpp:404-437)
template <int N>
struct MomLayout {
    static constexpr int NP = N * (N - 1) / 2;
    static constexpr int STRIDE = 1 + 2 * N + NP;
};

template <int N>
__device__ __forceinline__ void pair_of(int n, int &i, int &j)
{
    int a = 0, b = 0;
    for (int m = 0; m <= n; m++)
        if (++b == N) {
            ++a;
            b = a + 1;
        }
    i = a;
    j = b;
}

__global__ void k_index_max(const int64_t *__restrict__ index, int64_t n, unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long red[MOM_NT];
    unsigned long long m = 0;
    for (int64_t j = (int64_t)blockIdx.x * MOM_NT + threadIdx.x; j < n; j += (int64_t)gridDim.x * MOM_NT) {
        const int64_t v = index[j];
        if (v > 0 && (unsigned long long)v > m) m = (unsigned long long)v;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = MOM_NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicMax(out, red[0]);
}

__global__ void k_slot_fill(int32_t *__restrict__ slot, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * MOM_NT + threadIdx.x;
    if (i < n) slot[i] = NO_SLOT;
}

// slot[l] = first j with index[j] == l
__global__ void k_slot_map(const int64_t *__restrict__ index, int64_t n_index, int32_t maxl, int32_t *__restrict__ slot)
{
    const int64_t j = (int64_t)blockIdx.x * MOM_NT + threadIdx.x;
    if (j >= n_index) return;
    const int64_t l = index[j];
    if (l > 0 && l <= maxl) atomicMin(&slot[l], (int32_t)j);
}

// The labels of one lane's chunk (row `row`, columns [x0, x0 + n)).
__device__ __forceinline__ void load_chunk(const int32_t *__restrict__ labels, const MomGeom &g, int64_t base,
                                           int64_t x0, int n, int32_t *lab)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    if (g.vec && n == MOM_CHUNK) {
        const v4i *p = reinterpret_cast<const v4i *>(labels + base + x0);
#pragma unroll
        for (int q = 0; q < MOM_CHUNK / 4; q++) {
            const v4i v = __builtin_nontemporal_load(p + q);
            lab[4 * q] = v.x;
            lab[4 * q + 1] = v.y;
            lab[4 * q + 2] = v.z;
            lab[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < MOM_CHUNK; i++) lab[i] = (i < n) ? labels[base + x0 + i] : 0;
    }
}

// The data of one lane's chunk, every load issued before any use (the
// accumulation's atomics would otherwise order each load after them).
template <typename T>
__device__ __forceinline__ void load_data(const T *__restrict__ data, const MomGeom &g, int64_t base, int64_t x0,
                                          int n, T *dv)
{
    // 16-byte loads: 4 floats or 2 doubles
    constexpr int VW = 16 / (int)sizeof(T);
    typedef T vw __attribute__((ext_vector_type(VW)));
    if (g.dvec && n == MOM_CHUNK) {
        const vw *p = reinterpret_cast<const vw *>(data + base + x0);
#pragma unroll
        for (int q = 0; q < MOM_CHUNK / VW; q++) {
            const vw v = __builtin_nontemporal_load(p + q);
#pragma unroll
            for (int k = 0; k < VW; k++) dv[VW * q + k] = v[k];
        }
    } else {
#pragma unroll
        for (int i = 0; i < MOM_CHUNK; i++) dv[i] = (i < n) ? data[base + x0 + i] : T(0);
    }
}

// row and first column of lane t's chunk (32-bit division when it fits)
__device__ __forceinline__ void chunk_of(const MomGeom &g, int64_t t, int64_t &row, int64_t &x0)
{
    if (g.small) {
        const uint32_t tt = (uint32_t)t, ch = (uint32_t)g.chunks;
        const uint32_t r = tt / ch;
        row = r;
        x0 = (int64_t)(tt - r * ch) * MOM_CHUNK;
    } else {
        row = t / g.chunks;
        x0 = (t - row * g.chunks) * MOM_CHUNK;
    }
}

// xyz coordinates 1..N-1 of a row (coordinate 0 is the column)
template <int N>
__device__ __forceinline__ void row_coords(const MomGeom &g, int64_t row, int64_t *c)
{
    if (g.small) {
        uint32_t r = (uint32_t)row;
#pragma unroll
        for (int n = 1; n < N; n++) {
            const uint32_t d = (uint32_t)g.shape[N - 1 - n];
            const uint32_t q = r / d;
            c[n] = r - q * d;
            r = q;
        }
    } else {
        int64_t r = row;
#pragma unroll
        for (int n = 1; n < N; n++) {
            const int64_t d = g.shape[N - 1 - n];
            c[n] = r % d;
            r /= d;
        }
    }
    c[0] = 0;
}

// Patches of MOM_PR rows x (MOM_NT / MOM_PR) chunks: a workgroup sums the
// runs of its patch per region in an LDS table (MOM_TBL entries, open
// addressing on the slot) and adds each region's sums to the global
// accumulators once, so a region that spans several rows of the patch costs
// one set of global atomics instead of one per run (the global double
// atomics took three quarters of the pass).  A run whose region finds no
// table entry within MOM_PROBE probes goes straight to the global sums.
constexpr int MOM_PR = 8;
constexpr int MOM_PC = MOM_NT / MOM_PR;
constexpr int MOM_TBL = 128;
constexpr int MOM_PROBE = 16;

__device__ __forceinline__ int mom_entry(int32_t *tkey, int32_t s)
{
    const uint32_t h = ((uint32_t)s * 2654435761u) >> 25;   // 7 bits: MOM_TBL
    for (int p = 0; p < MOM_PROBE; p++) {
        const int e = (int)((h + p) & (MOM_TBL - 1));
        const int32_t k = atomicCAS(&tkey[e], -1, s);
        if (k == -1 || k == s) return e;
    }
    return -1;
}

template <int N, typename T>
__global__ void __launch_bounds__(MOM_NT)
k_region_accumulate(const int32_t *__restrict__ labels, const T *__restrict__ data, MomGeom g,
                    const int32_t *__restrict__ slot, double *__restrict__ acc, unsigned long long *__restrict__ origin)
{
    using Lay = MomLayout<N>;
    constexpr int S = Lay::STRIDE;
    __shared__ int32_t tkey[MOM_TBL];
    __shared__ unsigned long long torg[MOM_TBL];
    __shared__ double tval[MOM_TBL * S];
    const int tid = threadIdx.x;
    for (int e = tid; e < MOM_TBL; e += MOM_NT) {
        tkey[e] = -1;
        torg[e] = ~0ull;
    }
    for (int i = tid; i < MOM_TBL * S; i += MOM_NT) tval[i] = 0.0;
    __syncthreads();

    // this lane's chunk: patch (pr, pc), row pr * MOM_PR + tid / MOM_PC
    const int64_t pr = (int64_t)blockIdx.x / g.n_pc, pc = (int64_t)blockIdx.x - pr * g.n_pc;
    const int64_t row = pr * MOM_PR + tid / MOM_PC, ch = pc * MOM_PC + tid % MOM_PC;
    bool any = false;
    int32_t lab[MOM_CHUNK];
    T dv[MOM_CHUNK];
    int64_t x0 = 0, base = 0;
    int n = 0;
    if (row < g.rows && ch < g.chunks) {
        x0 = ch * MOM_CHUNK;
        n = (int)min((int64_t)MOM_CHUNK, g.W - x0);
        base = row * g.W;
        load_chunk(labels, g, base, x0, n, lab);
#pragma unroll
        for (int i = 0; i < MOM_CHUNK; i++) any |= lab[i] > 0;
    }
    if (any) {
        load_data(data, g, base, x0, n, dv);
        int64_t c[N];
        row_coords<N>(g, row, c);
        int32_t cur = 0;
        int64_t start = 0;
        double s0 = 0.0, sx = 0.0, sxx = 0.0;
        auto flush = [&]() {
            if (cur <= 0 || cur > g.maxl) return;
            const int32_t sl = slot[cur];
            if (sl == NO_SLOT) return;
            // the run's contributions to the region's components
            double v[S];
            v[0] = s0;
            v[1] = sx;
            v[1 + N] = sxx;
#pragma unroll
            for (int m = 1; m < N; m++) {
                const double cm = (double)c[m];
                v[1 + m] = cm * s0;
                v[1 + N + m] = cm * cm * s0;
            }
#pragma unroll
            for (int p = 0; p < Lay::NP; p++) {
                int i, j;
                pair_of<N>(p, i, j);
                v[1 + 2 * N + p] = (i == 0) ? (double)c[j] * sx : (double)c[i] * (double)c[j] * s0;
            }
            const int e = mom_entry(tkey, sl);
            if (e >= 0) {
                atomicMin(&torg[e], (unsigned long long)(base + start));
                if (s0 == 0.0) return;   // every value <= 0 (NaN sums go on)
#pragma unroll
                for (int j = 0; j < S; j++) atomicAdd(&tval[e * S + j], v[j]);
            } else {
                atomicMin(&origin[sl], (unsigned long long)(base + start));
                if (s0 == 0.0) return;
                double *a = acc + (int64_t)sl * S;
#pragma unroll
                for (int j = 0; j < S; j++) atomicAdd(&a[j], v[j]);
            }
        };
#pragma unroll
        for (int i = 0; i < MOM_CHUNK; i++) {
            if (i >= n) break;
            const int32_t l = lab[i];
            if (l != cur) {
                flush();
                cur = l;
                start = x0 + i;
                s0 = sx = sxx = 0.0;
            }
            if (l > 0) {
                double v = (double)dv[i];
                v = (v < 0.0) ? 0.0 : v;   // std::max(value, T()): NaN stays NaN
                const double x = (double)(x0 + i);
                s0 += v;
                sx += v * x;
                sxx += v * x * x;
            }
        }
        flush();
    }
    __syncthreads();
    // the table to the global accumulators (zero components skipped: adding
    // +0 changes nothing; NaN sums are added)
    for (int e = tid; e < MOM_TBL; e += MOM_NT) {
        const int32_t sl = tkey[e];
        if (sl < 0) continue;
        atomicMin(&origin[sl], torg[e]);
        double *a = acc + (int64_t)sl * S;
        for (int j = 0; j < S; j++) {
            const double v = tval[e * S + j];
            if (v != 0.0) atomicAdd(&a[j], v);
        }
    }
}

// The reference's outputs from the raw sums (label.hpp:554-722; what as
// CBX_TOTAL_MASS .. CBX_LINE_FIT), zyx order like the reference.
template <int N, typename T>
__global__ void k_region_finish(const double *__restrict__ acc, const unsigned long long *__restrict__ origin,
                                const int32_t *__restrict__ slot, const int64_t *__restrict__ index, int64_t n_index,
                                MomGeom g, int what, T *__restrict__ out)
{
    using Lay = MomLayout<N>;
    constexpr int NP = Lay::NP;
    const int64_t j = (int64_t)blockIdx.x * MOM_NT + threadIdx.x;
    if (j >= n_index) return;
    const int64_t l = index[j];
    const int32_t s = (l > 0 && l <= g.maxl) ? slot[l] : NO_SLOT;
    double mu = 0.0, S[N], Q[N], C[NP > 0 ? NP : 1], o[N];
    for (int n = 0; n < N; n++) S[n] = Q[n] = o[n] = 0.0;
    for (int p = 0; p < NP; p++) C[p] = 0.0;
    if (s != NO_SLOT && origin[s] != ~0ull) {
        const double *a = acc + (int64_t)s * Lay::STRIDE;
        mu = a[0];
        for (int n = 0; n < N; n++) {
            S[n] = a[1 + n];
            Q[n] = a[1 + N + n];
        }
        for (int p = 0; p < NP; p++) C[p] = a[1 + 2 * N + p];
        int64_t r = (int64_t)origin[s];
        for (int n = 0; n < N; n++) {
            const int64_t d = g.shape[N - 1 - n];
            o[n] = (double)(r % d);
            r /= d;
        }
    }
    const int M = (what == 0) ? 1 : (what <= 2) ? N : (what <= 4) ? N * N : 2 * N;
    T *q = out + j * M;
    if (what == 0) {            // total_mass
        q[0] = (T)mu;
        return;
    }
    if (what == 1) {            // mean: the raw first moment (MomentsND::first)
        for (int n = 0; n < N; n++) q[N - 1 - n] = (T)S[n];
        return;
    }
    if (what == 3) {            // moment_of_inertia: MomentsND::second
        for (int n = 0; n < N; n++) q[n + N * n] = (T)Q[N - 1 - n];
        // off-diagonals: the reference adds origin terms with the pair's
        // indices mapped to the input's axis order (label.hpp:699-706),
        // applied to moments relative to the origin stored in xyz order
        for (int p = 0; p < NP; p++) {
            int i, k;
            pair_of<N>(p, i, k);
            const int a = N - i - 1, b = N - k - 1;
            const double rel = C[p] - o[i] * S[k] - o[k] * S[i] + o[i] * o[k] * mu;   // sum v r_i r_k
            const double rx_a = S[a] - o[a] * mu, rx_b = S[b] - o[b] * mu;          // sum v r_a, sum v r_b
            const T m = (T)(rel + o[a] * rx_b + o[b] * rx_a + o[a] * o[b] * mu);
            q[a + N * b] = m;
            q[b + N * a] = m;
        }
        return;
    }
    // central moments (CentralMomentsND): zero when the mass is zero
    double cx[N], cxx[N], cxy[NP > 0 ? NP : 1];
    for (int n = 0; n < N; n++) cx[n] = cxx[n] = 0.0;
    for (int p = 0; p < NP; p++) cxy[p] = 0.0;
    if (mu != 0.0) {
        // Q / mu - cx^2 from absolute sums cancels to a few ulp of Q / mu; the
        // reference, summing coordinates relative to the region's first pixel,
        // gets exact zeros for zero spreads (one pixel, one row): differences
        // below 1e-13 of the raw moments (about 500 ulp) are those zeros
        for (int n = 0; n < N; n++) {
            cx[n] = S[n] / mu;
            const double q = Q[n] / mu;
            cxx[n] = q - cx[n] * cx[n];
            if (fabs(cxx[n]) <= 1e-13 * fabs(q)) cxx[n] = 0.0;
        }
        for (int p = 0; p < NP; p++) {
            int i, k;
            pair_of<N>(p, i, k);
            cxy[p] = C[p] / mu - cx[i] * cx[k];
            if (fabs(cxy[p]) <= 0.5e-13 * fabs((Q[i] + Q[k]) / mu)) cxy[p] = 0.0;
        }
    } else {
        for (int n = 0; n < N; n++) cx[n] = o[n];
    }
    if (what == 2) {            // center_of_mass
        for (int n = 0; n < N; n++) q[N - 1 - n] = (T)cx[n];
        return;
    }
    if (what == 4) {            // covariance_matrix
        for (int n = 0; n < N; n++) q[n + N * n] = (T)cxx[N - 1 - n];
        for (int p = 0; p < NP; p++) {
            int i, k;
            pair_of<N>(p, i, k);
            const int a = N - i - 1, b = N - k - 1;
            q[a + N * b] = (T)cxy[p];
            q[b + N * a] = (T)cxy[p];
        }
        return;
    }
    // line_fit: the major axis of the covariance ellipse (CentralMomentsND::line).
    // The reference sums small coordinates relative to the region's first
    // pixel, so a region with cxy = 0 or cxx = cyy gets exact zeros here and
    // atan2 its conventional direction.  The central moments from absolute
    // sums carry rounding of a few ulp of the raw second moments (~ the
    // squared centroid), which atan2 would turn into an arbitrary direction or
    // a swap of the end points: differences below 1e-13 of the raw moments
    // (about 500 ulp) are taken as those zeros.
    const double tr = cxx[0] + cxx[1];
    const double noise = (mu != 0.0) ? 1e-13 * fabs((Q[0] + Q[1]) / mu) : 0.0;
    double dxy = cxx[0] - cxx[1], xy2 = 2.0 * cxy[0];
    if (fabs(dxy) <= noise) dxy = 0.0;
    if (fabs(xy2) <= noise) xy2 = 0.0;
    const double angle = 0.5 * atan2(xy2, dxy);
    if (angle != angle) {
        for (int n = 0; n < N; n++) q[n] = q[N + n] = (T)cx[n];
        return;
    }
    double tau[N];
    for (int n = 0; n < N; n++) tau[n] = 0.0;
    tau[0] = cos(angle);
    tau[1] = sin(angle);
    const double dl = sqrt(xy2 * xy2 + dxy * dxy);
    const double hw = sqrt(fmax(0.0, 2.0 * 0.6931471805599453 * (cxx[0] + cxx[1] + dl)));
    for (int n = 0; n < N; n++) {
        q[n] = (T)(cx[n] + hw * tau[n]);
        q[N + n] = (T)(cx[n] - hw * tau[n]);
    }
}

// LineND<T, N>::distance of an integer point (geometry.hpp:319-339), in T
// and in the reference's order of operations.
template <int N, typename T>
__device__ __forceinline__ T seg_distance(const T *ln, const int64_t *p)
{
    T tau[N], mag = T(0);
    for (int n = 0; n < N; n++) tau[n] = ln[N + n] - ln[n];
    for (int n = 0; n < N; n++) mag += tau[n] * tau[n];
    T q[N];
    if (mag != T(0)) {
        T ctr[N], d = T(0);
        for (int n = 0; n < N; n++) ctr[n] = (ln[n] + ln[N + n]) / T(2);
        for (int n = 0; n < N; n++) d += tau[n] * ((T)p[n] - ctr[n]);
        T rt = d / mag;
        rt = (rt < T(-0.5)) ? T(-0.5) : ((rt > T(0.5)) ? T(0.5) : rt);
        for (int n = 0; n < N; n++) q[n] = rt * tau[n] + ctr[n];
    } else {
        for (int n = 0; n < N; n++) q[n] = ln[n];
    }
    T m2 = T(0);
    for (int n = 0; n < N; n++) m2 += ((T)p[n] - q[n]) * ((T)p[n] - q[n]);
    if constexpr (sizeof(T) == 4) return __builtin_sqrtf(m2);   // IEEE square root
    else return __builtin_sqrt(m2);
}

// p_values counts: pixels of region j within xtol of lines[j] (n_total) and
// those of them reaching vmin (n_signal), per run, pre-reduced per patch of
// MOM_PR rows x MOM_PC chunks in a workgroup table (as k_region_accumulate),
// then atomically per slot.
template <int N, typename T>
__global__ void __launch_bounds__(MOM_NT)
k_region_count(const int32_t *__restrict__ labels, const T *__restrict__ data, MomGeom g,
               const int32_t *__restrict__ slot, const T *__restrict__ lines, T vmin, T xtol,
               unsigned long long *__restrict__ counts)
{
    __shared__ int32_t tkey[MOM_TBL];
    __shared__ uint32_t tcnt[2 * MOM_TBL];
    const int tid = threadIdx.x;
    for (int e = tid; e < MOM_TBL; e += MOM_NT) {
        tkey[e] = -1;
        tcnt[2 * e] = tcnt[2 * e + 1] = 0;
    }
    __syncthreads();
    const int64_t pr = (int64_t)blockIdx.x / g.n_pc, pc = (int64_t)blockIdx.x - pr * g.n_pc;
    const int64_t row = pr * MOM_PR + tid / MOM_PC, ch = pc * MOM_PC + tid % MOM_PC;
    bool any = false;
    int32_t lab[MOM_CHUNK];
    int64_t x0 = 0, base = 0;
    int n = 0;
    if (row < g.rows && ch < g.chunks) {
        x0 = ch * MOM_CHUNK;
        n = (int)min((int64_t)MOM_CHUNK, g.W - x0);
        base = row * g.W;
        load_chunk(labels, g, base, x0, n, lab);
#pragma unroll
        for (int i = 0; i < MOM_CHUNK; i++) any |= lab[i] > 0;
    }
    if (any) {
        T dv[MOM_CHUNK];
        load_data(data, g, base, x0, n, dv);
        int64_t p[N];
        row_coords<N>(g, row, p);
        int32_t cur = 0, s = NO_SLOT;
        uint32_t nt = 0, ns = 0;
        T ln[2 * N];
        auto flush = [&]() {
            if (s == NO_SLOT || nt == 0) return;   // ns <= nt
            const int e = mom_entry(tkey, s);
            if (e >= 0) {
                atomicAdd(&tcnt[2 * e], nt);
                if (ns) atomicAdd(&tcnt[2 * e + 1], ns);
            } else {
                atomicAdd(&counts[2 * (int64_t)s], (unsigned long long)nt);
                if (ns) atomicAdd(&counts[2 * (int64_t)s + 1], (unsigned long long)ns);
            }
        };
#pragma unroll
        for (int i = 0; i < MOM_CHUNK; i++) {
            if (i >= n) break;
            const int32_t l = lab[i];
            if (l != cur) {
                flush();
                cur = l;
                nt = ns = 0;
                s = (l > 0 && l <= g.maxl) ? slot[l] : NO_SLOT;
                if (s != NO_SLOT)
                    for (int m = 0; m < 2 * N; m++) ln[m] = lines[2 * N * (int64_t)s + m];
            }
            if (s == NO_SLOT) continue;
            p[0] = x0 + i;
            if (seg_distance<N, T>(ln, p) < xtol) {
                nt++;
                if (dv[i] >= vmin) ns++;
            }
        }
        flush();
    }
    __syncthreads();
    for (int e = tid; e < MOM_TBL; e += MOM_NT) {
        const int32_t sl = tkey[e];
        if (sl < 0) continue;
        if (tcnt[2 * e]) atomicAdd(&counts[2 * (int64_t)sl], (unsigned long long)tcnt[2 * e]);
        if (tcnt[2 * e + 1]) atomicAdd(&counts[2 * (int64_t)sl + 1], (unsigned long long)tcnt[2 * e + 1]);
    }
}

__device__ __forceinline__ double dev_logaddexp(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const double m = (a > b) ? a : b;
    return m + log(exp(a - m) + exp(b - m));
}

__device__ __forceinline__ float dev_logaddexp(float a, float b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const float m = (a > b) ? a : b;
    return m + logf(expf(a - m) + expf(b - m));
}

__device__ __forceinline__ float dev_log(float v) { return logf(v); }
__device__ __forceinline__ double dev_log(double v) { return log(v); }
__device__ __forceinline__ float dev_log1p(float v) { return log1pf(v); }
__device__ __forceinline__ double dev_log1p(double v) { return log1p(v); }

// detail::logbinom (label.hpp:57-95): the log of the binomial upper tail,
// with the reference's types (integer-argument lgamma and log in double).
template <typename T>
__device__ T logbinom(int64_t n, int64_t k, T p)
{
    if (k <= 0) return T(0);
    if (k > n) return -INFINITY;
    if (p <= T(0)) return (k <= 0) ? T(0) : -INFINITY;
    if (p >= T(1)) return (k <= n) ? T(0) : -INFINITY;
    const T log_p = dev_log(p), log_q = dev_log1p(-p);
    T t = (T)(lgamma((double)(n + 1)) - lgamma((double)(k + 1)) - lgamma((double)(n - k + 1)) +
              (double)((T)k * log_p) + (double)((T)(n - k) * log_q));
    T tail = t;
    for (int64_t i = k + 1; i < n + 1; ++i) {
        t = (T)((double)t + (log((double)(n - i + 1)) - log((double)i) + (double)log_p - (double)log_q));
        tail = dev_logaddexp(tail, t);
    }
    return tail;
}

template <typename T>
__global__ void k_logbinom(const unsigned long long *__restrict__ counts, const int32_t *__restrict__ slot,
                           const int64_t *__restrict__ index, int64_t n_index, int32_t maxl, T p0,
                           T *__restrict__ out, int64_t *__restrict__ out_counts)
{
    const int64_t j = (int64_t)blockIdx.x * MOM_NT + threadIdx.x;
    if (j >= n_index) return;
    const int64_t l = index[j];
    const int32_t s = (l > 0 && l <= maxl) ? slot[l] : NO_SLOT;
    const int64_t nt = (s == NO_SLOT) ? 0 : (int64_t)counts[2 * (int64_t)s];
    const int64_t ns = (s == NO_SLOT) ? 0 : (int64_t)counts[2 * (int64_t)s + 1];
    out[j] = logbinom<T>(nt, ns, p0);
    if (out_counts) {
        out_counts[2 * j] = nt;
        out_counts[2 * j + 1] = ns;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int moment_geometry(int ndim, const int64_t *shape, const int32_t *labels, const void *data, MomGeom &g)
{
    if (ndim < 2 || ndim > MOM_MAX_DIM)
        return fail(CBX_EVALUE, "Unsupported number of dimensions: " + std::to_string(ndim));
    int64_t P = 1;
    for (int d = 0; d < ndim; d++) {
        if (shape[d] < 0) return fail(CBX_EVALUE, "negative shape");
        g.shape[d] = shape[d];
        P *= shape[d];
        if (P >= (int64_t(1) << 62)) return fail(CBX_EVALUE, "label map too large");
    }
    g.W = shape[ndim - 1];
    g.rows = g.W ? P / g.W : 0;
    g.chunks = (g.W + MOM_CHUNK - 1) / MOM_CHUNK;
    g.vec = (g.W % 4 == 0 && ((uintptr_t)labels & 15) == 0) ? 1 : 0;
    g.dvec = (g.W % 4 == 0 && ((uintptr_t)data & 15) == 0) ? 1 : 0;
    g.small = (g.rows * g.chunks < (int64_t(1) << 31)) ? 1 : 0;
    g.n_pc = 0;
    g.maxl = 0;
    return CBX_OK;
}

// the slot map of index (first occurrence of each label); sets g.maxl
static int build_slots(const int64_t *index, int64_t n_index, MomGeom &g, DevBuf<int32_t> &slot, hipStream_t st)
{
    if (n_index >= (int64_t)INT_MAX) return fail(CBX_EVALUE, "index has too many entries");
    DevBuf<unsigned long long> mx;
    CBX_HIP(mx.alloc(1, st));
    CBX_HIP(hipMemsetAsync(mx.ptr, 0, sizeof(unsigned long long), st));
    const unsigned nb = (unsigned)std::min<int64_t>(1024, (n_index + MOM_NT - 1) / MOM_NT);
    k_index_max<<<nb, MOM_NT, 0, st>>>(index, n_index, mx.ptr);
    CBX_HIP(hipGetLastError());
    uint64_t hv = 0;
    const void *src[1] = {mx.ptr};
    const int bytes[1] = {8};
    CBX_HIP(read_device_values(st, 1, src, bytes, &hv));
    if (hv > (uint64_t)INT_MAX - 1) return fail(CBX_EVALUE, "label ids must fit in int32");
    g.maxl = (int32_t)hv;
    CBX_HIP(slot.alloc((size_t)g.maxl + 1, st));
    k_slot_fill<<<grid_for(g.maxl + 1, MOM_NT), MOM_NT, 0, st>>>(slot.ptr, g.maxl + 1);
    k_slot_map<<<grid_for(n_index, MOM_NT), MOM_NT, 0, st>>>(index, n_index, g.maxl, slot.ptr);
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

template <int N, typename T>
static int moments_n(const int32_t *labels, const T *data, const MomGeom &g, const int32_t *slot,
                     const int64_t *index, int64_t n_index, int what, T *out, hipStream_t st)
{
    using Lay = MomLayout<N>;
    const int64_t n_slots = n_index;
    DevBuf<double> acc;
    DevBuf<unsigned long long> org;
    CBX_HIP(acc.alloc((size_t)n_slots * Lay::STRIDE, st));
    CBX_HIP(org.alloc((size_t)n_slots, st));
    CBX_HIP(hipMemsetAsync(acc.ptr, 0, sizeof(double) * n_slots * Lay::STRIDE, st));
    CBX_HIP(hipMemsetAsync(org.ptr, 0xFF, sizeof(unsigned long long) * n_slots, st));
    if (g.rows * g.chunks > 0) {
        // one workgroup per patch of MOM_PR rows x MOM_PC chunks
        MomGeom gp = g;
        gp.n_pc = (g.chunks + MOM_PC - 1) / MOM_PC;
        const int64_t patches = ((g.rows + MOM_PR - 1) / MOM_PR) * gp.n_pc;
        if (patches >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "label map too large");
        KernelTimer timer("region_moments", st);
        k_region_accumulate<N, T><<<(unsigned)patches, MOM_NT, 0, st>>>(labels, data, gp, slot, acc.ptr, org.ptr);
    }
    k_region_finish<N, T><<<grid_for(n_index, MOM_NT), MOM_NT, 0, st>>>(acc.ptr, org.ptr, slot, index, n_index, g,
                                                                        what, out);
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

template <int N, typename T>
static int count_n(const int32_t *labels, const T *data, const MomGeom &g, const int32_t *slot, const T *lines,
                   T vmin, T xtol, unsigned long long *counts, hipStream_t st)
{
    if (g.rows * g.chunks > 0) {
        MomGeom gp = g;
        gp.n_pc = (g.chunks + MOM_PC - 1) / MOM_PC;
        const int64_t patches = ((g.rows + MOM_PR - 1) / MOM_PR) * gp.n_pc;
        if (patches >= (int64_t(1) << 31)) return fail(CBX_EVALUE, "label map too large");
        KernelTimer timer("region_counts", st);
        k_region_count<N, T><<<(unsigned)patches, MOM_NT, 0, st>>>(labels, data, gp, slot, lines, vmin, xtol,
                                                                   counts);
    }
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

template <typename T>
static int region_moments(const int32_t *labels, const T *data, int ndim, const int64_t *shape, const int64_t *index,
                          int64_t n_index, int what, T *out, void *stream)
{
    if (what < 0 || what > 5) return fail(CBX_EVALUE, "unknown region moment");
    if (n_index < 0) return fail(CBX_EVALUE, "negative index size");
    MomGeom g;
    int rc = moment_geometry(ndim, shape, labels, data, g);
    if (rc) return rc;
    if (n_index == 0) return CBX_OK;
    hipStream_t st = (hipStream_t)stream;
    Workspace ws(st);
    DevBuf<int32_t> slot;
    if ((rc = build_slots(index, n_index, g, slot, st))) return rc;
    switch (ndim) {
    case 2: return moments_n<2, T>(labels, data, g, slot.ptr, index, n_index, what, out, st);
    case 3: return moments_n<3, T>(labels, data, g, slot.ptr, index, n_index, what, out, st);
    case 4: return moments_n<4, T>(labels, data, g, slot.ptr, index, n_index, what, out, st);
    case 5: return moments_n<5, T>(labels, data, g, slot.ptr, index, n_index, what, out, st);
    case 6: return moments_n<6, T>(labels, data, g, slot.ptr, index, n_index, what, out, st);
    default: return moments_n<7, T>(labels, data, g, slot.ptr, index, n_index, what, out, st);
    }
}

template <typename T>
static int p_values(const int32_t *labels, const T *data, int ndim, const int64_t *shape, const int64_t *index,
                    int64_t n_index, const T *lines, T p0, T vmin, T xtol, T *out, int64_t *counts_out, void *stream)
{
    if (n_index < 0) return fail(CBX_EVALUE, "negative index size");
    MomGeom g;
    int rc = moment_geometry(ndim, shape, labels, data, g);
    if (rc) return rc;
    if (n_index == 0) return CBX_OK;
    hipStream_t st = (hipStream_t)stream;
    Workspace ws(st);
    DevBuf<int32_t> slot;
    if ((rc = build_slots(index, n_index, g, slot, st))) return rc;
    DevBuf<unsigned long long> counts;
    CBX_HIP(counts.alloc((size_t)2 * n_index, st));
    CBX_HIP(hipMemsetAsync(counts.ptr, 0, sizeof(unsigned long long) * 2 * n_index, st));
    switch (ndim) {
    case 2: rc = count_n<2, T>(labels, data, g, slot.ptr, lines, vmin, xtol, counts.ptr, st); break;
    case 3: rc = count_n<3, T>(labels, data, g, slot.ptr, lines, vmin, xtol, counts.ptr, st); break;
    case 4: rc = count_n<4, T>(labels, data, g, slot.ptr, lines, vmin, xtol, counts.ptr, st); break;
    case 5: rc = count_n<5, T>(labels, data, g, slot.ptr, lines, vmin, xtol, counts.ptr, st); break;
    case 6: rc = count_n<6, T>(labels, data, g, slot.ptr, lines, vmin, xtol, counts.ptr, st); break;
    default: rc = count_n<7, T>(labels, data, g, slot.ptr, lines, vmin, xtol, counts.ptr, st); break;
    }
    if (rc) return rc;
    k_logbinom<T><<<grid_for(n_index, MOM_NT), MOM_NT, 0, st>>>(counts.ptr, slot.ptr, index, n_index, g.maxl, p0, out,
                                                                 counts_out);
    CBX_HIP(hipGetLastError());
    return CBX_OK;
}

} // namespace cbx

extern "C" {

int cbx_region_moments_f32(const int32_t *labels, const float *data, int ndim, const int64_t *shape,
                           const int64_t *index, int64_t n_index, int what, float *out, void *stream)
{
    return cbx::region_moments<float>(labels, data, ndim, shape, index, n_index, what, out, stream);
}

int cbx_region_moments_f64(const int32_t *labels, const double *data, int ndim, const int64_t *shape,
                           const int64_t *index, int64_t n_index, int what, double *out, void *stream)
{
    return cbx::region_moments<double>(labels, data, ndim, shape, index, n_index, what, out, stream);
}

int cbx_p_values_f32(const int32_t *labels, const float *data, int ndim, const int64_t *shape, const int64_t *index,
                     int64_t n_index, const float *lines, float p0, float vmin, float xtol, float *out,
                     int64_t *counts, void *stream)
{
    return cbx::p_values<float>(labels, data, ndim, shape, index, n_index, lines, p0, vmin, xtol, out, counts, stream);
}

int cbx_p_values_f64(const int32_t *labels, const double *data, int ndim, const int64_t *shape, const int64_t *index,
                     int64_t n_index, const double *lines, double p0, double vmin, double xtol, double *out,
                     int64_t *counts, void *stream)
{
    return cbx::p_values<double>(labels, data, ndim, shape, index, n_index, lines, p0, vmin, xtol, out, counts,
                                 stream);
}

} // extern "C"
