// 3D lines on gfx950: draw_lines and accumulate_lines for rows of 7
// (x0, y0, z0, x1, y1, z1, width) into volumes (..., D, H, W).
//
// The pixel walk is the reference's draw_line_3d (cbclib_v2/_src/src/
// bresenham.hpp:389-430) on BresenhamPlotter<T, 3> / LineIterator<T, 3>
// (118-348): a tangent error and one normal error per axis pair (0,1),
// (0,2), (1,2), the three-axis Bresenham step with the minimal-|error|
// choice per pair, and at every long-axis step the orthogonal plane filled
// in four quadrants by a 2D walk with single-axis rays.  Every value is
// accumulated with the reference's FP32 (or FP64) additions in its order, and
// divisions and square roots are IEEE (-fhip-fp32-correctly-rounded-divide-
// sqrt), so the emitted pixel set and the per-pixel values are the
// reference's exactly.
//
// Four lanes walk one line (each replays the main-axis walk and fills one
// quadrant of every orthogonal plane; the first also emits the main-axis
// pixels) and apply its pixels to `out` with global atomics (add for sum,
// compare-and-swap for max): sums land in no fixed order (within float
// rounding of the single-threaded reference), max is exact.  A validation pass runs first, so a call that fails leaves `out`
// untouched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>

#include "cbx_core.hpp"
#include "cbx_host.hpp"

namespace cbx {

constexpr int D3_NT = 128;
constexpr double MAX_VISITS3 = double(int64_t(1) << 30);
constexpr int64_t MAX_PLANES = int64_t(1) << 23;   // recorded planes of a two-phase draw (256 / 384 MiB)

template <typename T>
struct Iter3 {
    int64_t step[3], cur[3];
    bool next[3];
    T te, ne[3];
};

// pair q of UniquePairs<3>: (0,1), (0,2), (1,2)
__device__ __forceinline__ int pair_a(int q) { return q == 2 ? 1 : 0; }
__device__ __forceinline__ int pair_b(int q) { return q == 0 ? 1 : 2; }

// The constants of one line (BresenhamPlotter<T, 3>(line, ceil(w) + 1)).
template <typename T>
struct Line3 {
    T p0[3], t[3], nrm[3][3];
    T len, w2;
    int axis;
    int64_t step[3], m0[3], m1[3];
    int64_t ray_cap;   // steps of one plane ray: at most 2 sqrt(3) (w + 2) for a ray that leaves the ellipse
};

// 0: draws nothing; 1: ready; -1: the walk is longer than MAX_WALK or leaves
// the coordinate range (FLAG_LONG)
template <typename T>
__device__ int line3_setup(const T *ln, Line3<T> &L)
{
    T p1[3];
    for (int i = 0; i < 3; i++) {
        L.p0[i] = ln[i];
        p1[i] = ln[3 + i];
        L.t[i] = p1[i] - L.p0[i];
    }
    const T w = ln[6];
    for (int q = 0; q < 3; q++) {
        const int a = pair_a(q), b = pair_b(q);
        for (int i = 0; i < 3; i++) L.nrm[q][i] = T(0);
        L.nrm[q][a] = p1[b] - L.p0[b];
        L.nrm[q][b] = L.p0[a] - p1[a];
    }
    // error() is +inf for w <= 0, and a NaN never compares true
    if (!(w > T(0))) return 0;
    T mag = T(0);
    for (int i = 0; i < 3; i++) mag += L.t[i] * L.t[i];
    L.len = m_sqrt(mag);
    if (!(L.len > T(0)) || !__builtin_isfinite(L.len) || !__builtin_isfinite(w)) return 0;
    for (int i = 0; i < 3; i++)
        if (!__builtin_isfinite(L.p0[i]) || !__builtin_isfinite(p1[i])) return 0;
    // the first largest |t| (selects, not indexing: the arrays stay in registers)
    int axis = 0;
    T tmax = m_abs(L.t[0]);
    if (tmax < m_abs(L.t[1])) {
        axis = 1;
        tmax = m_abs(L.t[1]);
    }
    if (tmax < m_abs(L.t[2])) axis = 2;
    L.axis = axis;
    const long offset = (long)(m_ceil(w) + T(1));
    T tau[3];
    for (int i = 0; i < 3; i++) tau[i] = L.t[i] / L.len;
    const T tau_ax = (axis == 0) ? tau[0] : ((axis == 1) ? tau[1] : tau[2]);
    const T sc = m_abs((T)offset / tau_ax);
    for (int i = 0; i < 3; i++) {
        L.step[i] = (p1[i] >= L.p0[i]) ? 1 : -1;
        const T a0 = m_round(L.p0[i] - sc * tau[i]), a1 = m_round(p1[i] + sc * tau[i]);
        if (!(m_abs(a0) < T(MAX_COORD)) || !(m_abs(a1) < T(MAX_COORD))) return -1;
        L.m0[i] = (int64_t)a0;
        L.m1[i] = (int64_t)a1 + L.step[i];
    }
    const int64_t span = (axis == 0) ? L.m1[0] - L.m0[0] : ((axis == 1) ? L.m1[1] - L.m0[1] : L.m1[2] - L.m0[2]);
    if (span > MAX_WALK || span < -MAX_WALK) return -1;
    // one lane walks the whole line: at most MAX_VISITS3 pixel visits
    // (about 5.4 w^2 per plane: the ellipse of half-axes w and <= sqrt(3) w)
    const double wc = (double)m_ceil(w) + 2.0;
    if ((double)(span < 0 ? -span : span) * 8.0 * wc * wc > MAX_VISITS3) return -1;
    L.w2 = w * w;
    L.ray_cap = 8 * (int64_t)wc + 64;
    return 1;
}

// LineIterator::update: the next steps from the minimal-|error| choice of
// every pair of moving axes (NormalError::is_next)
template <typename T>
__device__ __forceinline__ void it3_update(const Line3<T> &L, Iter3<T> &it)
{
    for (int i = 0; i < 3; i++) it.next[i] = it.step[i] != 0;
    for (int q = 0; q < 3; q++) {
        const int a = pair_a(q), b = pair_b(q);
        if (!(it.step[a] && it.step[b])) continue;
        const T ex = it.ne[q] + (T)it.step[a] * L.nrm[q][a];
        const T ey = it.ne[q] + (T)it.step[b] * L.nrm[q][b];
        const T exy = ex + (T)it.step[b] * L.nrm[q][b];
        bool f, s;
        if (m_abs(ey) < m_abs(ex)) {
            f = m_abs(exy) < m_abs(ey);
            s = true;
        } else {
            f = true;
            s = m_abs(exy) < m_abs(ex);
        }
        it.next[a] = it.next[a] && f;
        it.next[b] = it.next[b] && s;
    }
}

// LineIterator::increment: the tangent error and the errors of the pairs
// containing the axis
template <typename T>
__device__ __forceinline__ void it3_inc(const Line3<T> &L, Iter3<T> &it, int64_t x, int ax)
{
    it.cur[ax] += x;
    it.te += (T)x * L.t[ax];
    for (int q = 0; q < 3; q++)
        if (pair_a(q) == ax || pair_b(q) == ax) it.ne[q] += (T)x * L.nrm[q][ax];
}

template <typename T>
__device__ __forceinline__ void it3_advance(const Line3<T> &L, Iter3<T> &it)
{
    for (int i = 0; i < 3; i++)
        if (it.next[i]) it3_inc(L, it, it.step[i], i);
    it3_update(L, it);
}

// BresenhamPlotter::normal_error: sum of the squared normal errors, each
// std::pow(T, 2) in double, accumulated into T
template <typename T>
__device__ __forceinline__ T it3_nerr(const Line3<T> &L, const Iter3<T> &it)
{
    T e = T(0);
    for (int q = 0; q < 3; q++) {
        const double v = (double)(it.ne[q] / L.len);
        e = (T)((double)e + v * v);
    }
    return e;
}

// BresenhamPlotter::error: (tangent term^2 + normal error) / w^2
template <typename T>
__device__ __forceinline__ T it3_err(const Line3<T> &L, const Iter3<T> &it)
{
    const T a = -it.te, b = it.te - L.len * L.len;
    const T m = (a < b) ? b : a;
    const T at = (m < T(0)) ? T(0) : m;
    const double qt = (double)(at / L.len);
    return (T)((qt * qt + (double)it3_nerr(L, it)) / (double)L.w2);
}

// draw_line_3d: emit(xyz, error) for every pixel the reference hands to its
// callback, in its order.  Every loop has a step cap that a walk in exact
// arithmetic never reaches (the main walk moves one of its axes towards m1 at
// every step; a plane ray moves monotonically through the ellipse of points
// within w of the line); a walk whose float errors stop changing hits it,
// sets FLAG_STUCK and stops.
//
// The long axis is a template argument, so that every array index below is a
// compile-time constant once the loops are unrolled and the iterator state
// stays in registers.

// the main iterator at m0 (LineIterator's constructor)
template <typename T>
__device__ __forceinline__ void it3_start(const Line3<T> &L, Iter3<T> &it)
{
    T d[3];
    for (int i = 0; i < 3; i++) {
        it.step[i] = L.step[i];
        it.cur[i] = L.m0[i];
        d[i] = (T)L.m0[i] - L.p0[i];
    }
    it.te = T(0);
    for (int i = 0; i < 3; i++) it.te += L.t[i] * d[i];
    for (int q = 0; q < 3; q++) {
        it.ne[q] = T(0);
        for (int i = 0; i < 3; i++) it.ne[q] += L.nrm[q][i] * d[i];
    }
    it3_update(L, it);
}

// The main loop: plane(it) for every iterator state whose long-axis step is
// taken (the planes draw_line_3d fills).  False if the step cap was hit.
template <int AXIS, typename T, typename Plane>
__device__ __forceinline__ bool walk3_main(const Line3<T> &L, int *flags, Plane &plane)
{
    Iter3<T> it;
    it3_start(L, it);
    int64_t main_cap = 64;
    for (int i = 0; i < 3; i++) main_cap += (L.m1[i] - L.m0[i]) * L.step[i];
    for (int64_t n = 0; !(it.cur[0] == L.m1[0] || it.cur[1] == L.m1[1] || it.cur[2] == L.m1[2]); n++) {
        if (n == main_cap) {
            atomicOr(flags, FLAG_STUCK);
            return false;
        }
        if (it.next[AXIS]) plane(it);
        it3_advance(L, it);
    }
    return true;
}

// Quadrant i of the plane through the main iterator state `it`: a 2D walk
// away from the line with single-axis rays from each of its pixels.
template <int AXIS, typename T, typename Emit>
__device__ __forceinline__ bool fill3_quadrant(const Line3<T> &L, const Iter3<T> &it, int i, int *flags, Emit &emit)
{
    constexpr int axis = AXIS, ax1 = (AXIS + 1) % 3, ax2 = (AXIS + 2) % 3;
    Iter3<T> xy = it;
    xy.step[axis] = 0;                                    // collapse(iter, axis)
    it3_update(L, xy);
    if (i & 1) xy.step[ax1] = -xy.step[ax1];              // flip
    if (i >> 1) xy.step[ax2] = -xy.step[ax2];
    it3_update(L, xy);
    const int mv = ((i & 1) ^ (i >> 1)) ? ax2 : ax1;     // move(axis(1 + ...))
    if (mv == ax1) it3_inc(L, xy, xy.step[ax1], ax1);
    else it3_inc(L, xy, xy.step[ax2], ax2);
    it3_update(L, xy);
    for (int64_t r = 0; it3_nerr(L, xy) < L.w2; it3_advance(L, xy), r++) {
        if (r == L.ray_cap) {
            atomicOr(flags, FLAG_STUCK);
            return false;
        }
        emit(xy.cur, it3_err(L, xy));
#pragma unroll
        for (int j = 1; j < 3; j++) {
            const int aj = (axis + j) % 3;
            if (!xy.next[aj]) continue;
            Iter3<T> x = xy;
            x.step[aj] = 0;                               // collapse(iter_xy, axis(j))
            it3_update(L, x);
            it3_advance(L, x);
            for (int64_t q = 0; it3_nerr(L, x) < L.w2; it3_advance(L, x), q++) {
                if (q == L.ray_cap) {
                    atomicOr(flags, FLAG_STUCK);
                    return false;
                }
                emit(x.cur, it3_err(L, x));
            }
        }
    }
    return true;
}

// the whole walk (quad = -1) or the main-axis pixels (quad 0 only) and the
// fills of one quadrant of every plane (quad = 0..3)
template <int AXIS, typename T, typename Emit>
__device__ __forceinline__ void walk3_axis(const Line3<T> &L, int quad, int *flags, Emit &emit)
{
    bool ok = true;
    auto plane = [&](const Iter3<T> &it) {
        if (!ok) return;
        if (quad <= 0) emit(it.cur, it3_err(L, it));
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (quad >= 0 && i != quad) continue;
            if (ok) ok = fill3_quadrant<AXIS>(L, it, i, flags, emit);
        }
    };
    walk3_main<AXIS>(L, flags, plane);
}

template <typename T, typename Emit>
__device__ void walk3(const Line3<T> &L, int quad, int *flags, Emit &&emit)
{
    if (L.axis == 0) walk3_axis<0>(L, quad, flags, emit);
    else if (L.axis == 1) walk3_axis<1>(L, quad, flags, emit);
    else walk3_axis<2>(L, quad, flags, emit);
}

// A recorded plane: the main iterator's errors and position (its steps are
// the line's, its next flags follow from them).  line < 0: unused slot.
template <typename T>
struct PlaneState {
    T te, ne[3];
    int32_t cur[3];
    int32_t line;
};

template <typename T>
__device__ __forceinline__ void it3_from(const Line3<T> &L, const PlaneState<T> &ps, Iter3<T> &it)
{
    for (int i = 0; i < 3; i++) {
        it.step[i] = L.step[i];
        it.cur[i] = ps.cur[i];
        it.ne[i] = ps.ne[i];
    }
    it.te = ps.te;
    it3_update(L, it);
}

// *p = std::max(*p, v) (replaced only if *p < v), or *p += v
template <int OV>
__device__ __forceinline__ void apply_atomic(float *p, float v)
{
    if (OV == 0) {
        atomicAdd(p, v);
    } else {
        unsigned int *q = (unsigned int *)p;
        unsigned int old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (__uint_as_float(old) < v) {
            const unsigned int prev = atomicCAS(q, old, __float_as_uint(v));
            if (prev == old) break;
            old = prev;
        }
    }
}

template <int OV>
__device__ __forceinline__ void apply_atomic(double *p, double v)
{
    if (OV == 0) {
        atomicAdd(p, v);
    } else {
        unsigned long long *q = (unsigned long long *)p;
        unsigned long long old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (__longlong_as_double((long long)old) < v) {
            const unsigned long long prev = atomicCAS(q, old, (unsigned long long)__double_as_longlong(v));
            if (prev == old) break;
            old = prev;
        }
    }
}

struct Vol3 {
    int64_t n_frames, D, H, W;
};

// validation: frame of every line in range, walks of bounded length
template <typename T, typename I>
__global__ void __launch_bounds__(D3_NT)
k_check3(const T *__restrict__ lines, int64_t n_lines, const I *__restrict__ idxs, Vol3 v, int *__restrict__ flags)
{
    const int64_t i = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (i >= n_lines) return;
    if (idxs && (idxs[i] < 0 || (int64_t)idxs[i] >= v.n_frames)) atomicOr(flags, FLAG_INDEX);
    Line3<T> L;
    if (line3_setup(lines + 7 * i, L) < 0) atomicOr(flags, FLAG_LONG);
}

template <typename T, typename I, int K, int OV>
__global__ void __launch_bounds__(D3_NT)
k_draw3(T *__restrict__ out, const T *__restrict__ lines, int64_t n_lines, const I *__restrict__ idxs, Vol3 v, T max_val,
        int *__restrict__ flags)
{
    const int64_t u = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (u >= 4 * n_lines) return;
    const int64_t i = u >> 2;
    Line3<T> L;
    if (line3_setup(lines + 7 * i, L) != 1) return;
    const int64_t f = idxs ? (int64_t)idxs[i] : ((v.n_frames == 1) ? 0 : i);
    T *fo = out + f * v.D * v.H * v.W;
    walk3(L, (int)(u & 3), flags, [&](const int64_t *p, T err) {
        if (!(err <= T(1)) || p[0] < 0 || p[0] >= v.W || p[1] < 0 || p[1] >= v.H || p[2] < 0 || p[2] >= v.D) return;
        apply_atomic<OV>(fo + (p[2] * v.H + p[1]) * v.W + p[0], pixel_value<K>(err, max_val));
    });
}

// Two-phase draw: planes of every line recorded by one lane per line
// (k_walk3_main), then one lane per (plane, quadrant) fills (k_fill3), so the
// lanes of a wave run consecutive planes of one line, whose fills take about
// as many steps each (one lane per line left most of a wave idle behind its
// longest and widest line).

// ub[i]: planes of line i at most (the long-axis span), 0 if it draws nothing;
// ub[n_lines] = 0 (the scan's total)
template <typename T>
__global__ void __launch_bounds__(D3_NT)
k_plane_bound(const T *__restrict__ lines, int64_t n_lines, int64_t *__restrict__ ub)
{
    const int64_t i = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (i > n_lines) return;
    int64_t n = 0;
    if (i < n_lines) {
        Line3<T> L;
        if (line3_setup(lines + 7 * i, L) == 1) {
            const int64_t sp = (L.axis == 0) ? L.m1[0] - L.m0[0] : ((L.axis == 1) ? L.m1[1] - L.m0[1] : L.m1[2] - L.m0[2]);
            n = sp < 0 ? -sp : sp;
        }
    }
    ub[i] = n;
}

template <typename T>
__global__ void __launch_bounds__(D3_NT)
k_walk3_main(const T *__restrict__ lines, int64_t n_lines, const int64_t *__restrict__ off,
             PlaneState<T> *__restrict__ states, int *__restrict__ flags)
{
    const int64_t i = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (i >= n_lines) return;
    const int64_t base = off[i], cap = off[i + 1] - base;
    if (cap == 0) return;
    Line3<T> L;
    int64_t k = 0;
    if (line3_setup(lines + 7 * i, L) == 1) {
        auto plane = [&](const Iter3<T> &it) {
            if (k == cap) {   // cannot happen: one plane per long-axis step
                atomicOr(flags, FLAG_STUCK);
                return;
            }
            PlaneState<T> ps;
            ps.te = it.te;
            for (int q = 0; q < 3; q++) {
                ps.ne[q] = it.ne[q];
                ps.cur[q] = (int32_t)it.cur[q];
            }
            ps.line = (int32_t)i;
            states[base + k++] = ps;
        };
        if (L.axis == 0) walk3_main<0>(L, flags, plane);
        else if (L.axis == 1) walk3_main<1>(L, flags, plane);
        else walk3_main<2>(L, flags, plane);
    }
    for (; k < cap; k++) states[base + k].line = -1;
}

template <int AXIS, typename T, typename Emit>
__device__ __forceinline__ void fill_plane(const Line3<T> &L, const Iter3<T> &it, int quad, int *flags, Emit &emit)
{
    if (quad == 0) emit(it.cur, it3_err(L, it));
    fill3_quadrant<AXIS>(L, it, quad, flags, emit);
}

template <typename T, typename I, int K, int OV>
__global__ void __launch_bounds__(D3_NT)
k_fill3(T *__restrict__ out, const T *__restrict__ lines, const PlaneState<T> *__restrict__ states, int64_t n_states,
        int64_t n_lines, const I *__restrict__ idxs, Vol3 v, T max_val, int *__restrict__ flags)
{
    const int64_t u = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (u >= 4 * n_states) return;
    const PlaneState<T> ps = states[u >> 2];
    if (ps.line < 0) return;
    const int64_t i = ps.line;
    Line3<T> L;
    if (line3_setup(lines + 7 * i, L) != 1) return;
    Iter3<T> it;
    it3_from(L, ps, it);
    const int64_t f = idxs ? (int64_t)idxs[i] : ((v.n_frames == 1) ? 0 : i);
    T *fo = out + f * v.D * v.H * v.W;
    auto emit = [&](const int64_t *p, T err) {
        if (!(err <= T(1)) || p[0] < 0 || p[0] >= v.W || p[1] < 0 || p[1] >= v.H || p[2] < 0 || p[2] >= v.D) return;
        apply_atomic<OV>(fo + (p[2] * v.H + p[1]) * v.W + p[0], pixel_value<K>(err, max_val));
    };
    const int quad = (int)(u & 3);
    if (L.axis == 0) fill_plane<0>(L, it, quad, flags, emit);
    else if (L.axis == 1) fill_plane<1>(L, it, quad, flags, emit);
    else fill_plane<2>(L, it, quad, flags, emit);
}

// accumulate_lines, mixed overlaps: the lines of terms [t0, t0 + nb) drawn
// into their term's zeroed buffer with in_overlap (the reference's term
// buffer: 0 + v, or std::max(0, v) through the compare-and-swap from zero),
// every drawn voxel marked in `touched`
template <typename T, typename I, int K, int IN_OV>
__global__ void __launch_bounds__(D3_NT)
k_acc3_draw(T *__restrict__ buf, uint32_t *__restrict__ touched, const T *__restrict__ lines, int64_t n_lines,
            const I *__restrict__ terms, int64_t t0, int64_t nb, Vol3 v, T max_val, int *__restrict__ flags)
{
    const int64_t u = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (u >= 4 * n_lines) return;
    const int64_t i = u >> 2;
    const int64_t b = (int64_t)terms[i] - t0;
    if (b < 0 || b >= nb) return;
    Line3<T> L;
    if (line3_setup(lines + 7 * i, L) != 1) return;
    const int64_t V = v.D * v.H * v.W;
    walk3(L, (int)(u & 3), flags, [&](const int64_t *p, T err) {
        if (!(err <= T(1)) || p[0] < 0 || p[0] >= v.W || p[1] < 0 || p[1] >= v.H || p[2] < 0 || p[2] >= v.D) return;
        const int64_t k = b * V + (p[2] * v.H + p[1]) * v.W + p[0];
        apply_atomic<IN_OV>(buf + k, pixel_value<K>(err, max_val));
        atomicOr(&touched[k >> 5], 1u << (k & 31));
    });
}

// every touched voxel of the batch combined into frame frames[term] with
// out_overlap (buffer + out, or std::max(out, buffer)); buffer and marks reset
template <typename T, typename I, int OUT_OV>
__global__ void __launch_bounds__(D3_NT)
k_acc3_fold(T *__restrict__ out, T *__restrict__ buf, uint32_t *__restrict__ touched, const I *__restrict__ frames,
            int64_t t0, int64_t nb, Vol3 v)
{
    const int64_t V = v.D * v.H * v.W;
    const int64_t k = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (k >= nb * V) return;
    const uint32_t bit = 1u << (k & 31);
    if (!(touched[k >> 5] & bit)) return;
    const int64_t b = k / V, r = k - b * V;
    apply_atomic<OUT_OV>(out + (int64_t)frames[t0 + b] * V + r, buf[k]);
    buf[k] = T(0);
}

__global__ void k_clear_words(uint32_t *__restrict__ w, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (i < n) w[i] = 0u;
}

// frame of every line (frames[terms[i]]) and the range checks of
// accumulate_lines (bresenham.cpp:171-187)
template <typename I>
__global__ void __launch_bounds__(D3_NT)
k_term_frames3(const I *__restrict__ terms, const I *__restrict__ frames, int64_t n_lines, int64_t n_terms,
               int64_t n_frames, I *__restrict__ comp, int *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * D3_NT + threadIdx.x;
    if (i < n_terms && (frames[i] < 0 || (int64_t)frames[i] >= n_frames)) atomicOr(&bad[1], 1);
    if (i >= n_lines) return;
    const I t = terms[i];
    if (t < 0 || (int64_t)t >= n_terms) {
        atomicOr(&bad[0], 1);
        return;
    }
    if (comp) comp[i] = frames[t];
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static int check_vol(int64_t n_frames, int64_t D, int64_t H, int64_t W, int64_t n_lines, int kernel, int overlap)
{
    if (n_frames < 0 || D < 0 || H < 0 || W < 0 || n_lines < 0) return fail(CBX_EVALUE, "negative size");
    if (n_lines >= (int64_t(1) << 34)) return fail(CBX_EVALUE, "too many lines");
    if (kernel < 0 || kernel > 4) return fail(CBX_EVALUE, "kernel is missing");
    if (overlap < 0 || overlap > 1) return fail(CBX_EVALUE, "Invalid overlap keyword");
    return CBX_OK;
}

static int read_flags(const int *flags, hipStream_t st, int *hv)
{
    uint64_t v = 0;
    const void *src[1] = {flags};
    const int bytes[1] = {4};
    CBX_HIP(read_device_values(st, 1, src, bytes, &v));
    *hv = (int)v;
    return CBX_OK;
}

static int flags3_status(int hv)
{
    if (hv & FLAG_INDEX) return fail(CBX_EINDEX, "idxs is outside of (0, n_frames)");
    if (hv & FLAG_LONG)
        return fail(CBX_EVALUE, "a line's 3D walk exceeds the supported length (2^26 px), coordinate range (1e9) "
                                "or pixel count (2^30)");
    if (hv & FLAG_STUCK) return fail(CBX_ERUNTIME, "a line walk did not terminate within its step budget");
    return CBX_OK;
}

// every line checked before anything is drawn; flags left zeroed for the draw
template <typename T, typename I>
static int validate3(const T *lines, int64_t n_lines, const I *idxs, const Vol3 &v, int *flags, hipStream_t st)
{
    CBX_HIP(hipMemsetAsync(flags, 0, sizeof(int), st));
    k_check3<T, I><<<grid_for(n_lines, D3_NT), D3_NT, 0, st>>>(lines, n_lines, idxs, v, flags);
    CBX_HIP(hipGetLastError());
    int hv = 0, rc = read_flags(flags, st, &hv);
    return rc ? rc : flags3_status(hv);
}

template <typename T, typename I, int OV>
static void launch_draw3(int kernel, T *out, const T *lines, int64_t n, const I *idxs, const Vol3 &v, T mv,
                         int *flags, hipStream_t st)
{
    const unsigned g = (unsigned)grid_for(4 * n, D3_NT);   // four lanes per line
    switch (kernel) {
    case K_BIWEIGHT: k_draw3<T, I, K_BIWEIGHT, OV><<<g, D3_NT, 0, st>>>(out, lines, n, idxs, v, mv, flags); break;
    case K_GAUSSIAN: k_draw3<T, I, K_GAUSSIAN, OV><<<g, D3_NT, 0, st>>>(out, lines, n, idxs, v, mv, flags); break;
    case K_PARABOLIC: k_draw3<T, I, K_PARABOLIC, OV><<<g, D3_NT, 0, st>>>(out, lines, n, idxs, v, mv, flags); break;
    case K_RECTANGULAR:
        k_draw3<T, I, K_RECTANGULAR, OV><<<g, D3_NT, 0, st>>>(out, lines, n, idxs, v, mv, flags);
        break;
    default: k_draw3<T, I, K_TRIANGULAR, OV><<<g, D3_NT, 0, st>>>(out, lines, n, idxs, v, mv, flags); break;
    }
}

template <typename T, typename I, int OV>
static void launch_fill3(int kernel, T *out, const T *lines, const PlaneState<T> *states, int64_t n_states,
                         int64_t n_lines, const I *idxs, const Vol3 &v, T mv, int *flags, hipStream_t st)
{
    const unsigned g = (unsigned)grid_for(4 * n_states, D3_NT);   // four lanes per plane
#define CBX_FILL3(KK) k_fill3<T, I, KK, OV><<<g, D3_NT, 0, st>>>(out, lines, states, n_states, n_lines, idxs, v, mv, flags)
    switch (kernel) {
    case K_BIWEIGHT: CBX_FILL3(K_BIWEIGHT); break;
    case K_GAUSSIAN: CBX_FILL3(K_GAUSSIAN); break;
    case K_PARABOLIC: CBX_FILL3(K_PARABOLIC); break;
    case K_RECTANGULAR: CBX_FILL3(K_RECTANGULAR); break;
    default: CBX_FILL3(K_TRIANGULAR); break;
    }
#undef CBX_FILL3
}

// validation, the draw, then its step-budget flag (inside an open Workspace).
// Above MAX_PLANES recorded planes the single-kernel walk (four lanes per
// line) runs instead of the two phases.
template <typename T, typename I>
static int draw3_run(T *out, const Vol3 &v, const T *lines, int64_t n_lines, const I *idxs, T max_val, int kernel,
                     int overlap, hipStream_t st)
{
    DevBuf<int> flags;
    CBX_HIP(flags.alloc(1, st));
    int rc = validate3(lines, n_lines, idxs, v, flags.ptr, st);
    if (rc) return rc;
    // planes of every line: offsets by a scan of their bounds
    DevBuf<int64_t> ub, off;
    CBX_HIP(ub.alloc((size_t)n_lines + 1, st));
    CBX_HIP(off.alloc((size_t)n_lines + 1, st));
    k_plane_bound<T><<<grid_for(n_lines + 1, D3_NT), D3_NT, 0, st>>>(lines, n_lines, ub.ptr);
    CBX_HIP(hipGetLastError());
    CBX_HIP(exclusive_scan_i64(ub.ptr, off.ptr, n_lines + 1, st));
    uint64_t total = 0;
    {
        const void *src[1] = {off.ptr + n_lines};
        const int nb[1] = {8};
        CBX_HIP(read_device_values(st, 1, src, nb, &total));
    }
    DevBuf<PlaneState<T>> states;
    const bool two_phase = total <= (uint64_t)MAX_PLANES;
    if (two_phase && total > 0) CBX_HIP(states.alloc((size_t)total, st));
    {
        KernelTimer timer("draw3", st);
        if (two_phase) {
            if (total > 0) {
                k_walk3_main<T><<<grid_for(n_lines, D3_NT), D3_NT, 0, st>>>(lines, n_lines, off.ptr, states.ptr,
                                                                           flags.ptr);
                if (overlap == 0)
                    launch_fill3<T, I, 0>(kernel, out, lines, states.ptr, (int64_t)total, n_lines, idxs, v, max_val,
                                          flags.ptr, st);
                else
                    launch_fill3<T, I, 1>(kernel, out, lines, states.ptr, (int64_t)total, n_lines, idxs, v, max_val,
                                          flags.ptr, st);
            }
        } else if (overlap == 0) {
            launch_draw3<T, I, 0>(kernel, out, lines, n_lines, idxs, v, max_val, flags.ptr, st);
        } else {
            launch_draw3<T, I, 1>(kernel, out, lines, n_lines, idxs, v, max_val, flags.ptr, st);
        }
        CBX_HIP(hipGetLastError());
    }
    int hv = 0;
    if ((rc = read_flags(flags.ptr, st, &hv))) return rc;
    return flags3_status(hv);
}

template <typename T, typename I>
static int draw_lines3(T *out, int64_t n_frames, int64_t D, int64_t H, int64_t W, const T *lines, int64_t n_lines,
                       const I *idxs, T max_val, int kernel, int overlap, void *stream)
{
    int rc = check_vol(n_frames, D, H, W, n_lines, kernel, overlap);
    if (rc) return rc;
    // idxs = None: one frame, or one frame per line (bresenham.cpp:5-20)
    if (!idxs && n_lines > 0 && n_frames != 1 && n_frames != n_lines)
        return fail(CBX_EVALUE, "idxs has an incompatible size (" + std::to_string(n_lines) + " != " +
                                    std::to_string(n_frames) + ")");
    if (n_lines == 0 || n_frames * D * H * W == 0) return CBX_OK;
    hipStream_t st = (hipStream_t)stream;
    Workspace ws(st);
    return draw3_run(out, Vol3{n_frames, D, H, W}, lines, n_lines, idxs, max_val, kernel, overlap, st);
}

// mixed overlaps: terms in batches of at most 256 MiB of term buffers
template <typename T, typename I, int K, int IN_OV, int OUT_OV>
static int accumulate3_batches(T *out, const T *lines, int64_t n_lines, const I *terms, const I *frames,
                               int64_t n_terms, const Vol3 &v, T max_val, int *flags, hipStream_t st)
{
    const int64_t V = v.D * v.H * v.W;
    const int64_t budget = int64_t(256) << 20;
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(n_terms, budget / (V * (int64_t)sizeof(T))));
    const int64_t words = (nb * V + 31) / 32;
    DevBuf<T> buf;
    DevBuf<uint32_t> touched;
    CBX_HIP(buf.alloc((size_t)(nb * V), st));
    CBX_HIP(touched.alloc((size_t)words, st));
    CBX_HIP(hipMemsetAsync(buf.ptr, 0, sizeof(T) * nb * V, st));
    CBX_HIP(hipMemsetAsync(touched.ptr, 0, sizeof(uint32_t) * words, st));
    for (int64_t t0 = 0; t0 < n_terms; t0 += nb) {
        const int64_t n = std::min(nb, n_terms - t0);
        k_acc3_draw<T, I, K, IN_OV><<<grid_for(4 * n_lines, D3_NT), D3_NT, 0, st>>>(buf.ptr, touched.ptr, lines, n_lines,
                                                                               terms, t0, n, v, max_val, flags);
        k_acc3_fold<T, I, OUT_OV><<<grid_for(n * V, D3_NT), D3_NT, 0, st>>>(out, buf.ptr, touched.ptr, frames, t0, n, v);
        k_clear_words<<<grid_for((n * V + 31) / 32, D3_NT), D3_NT, 0, st>>>(touched.ptr, (n * V + 31) / 32);
        CBX_HIP(hipGetLastError());
    }
    return CBX_OK;
}

template <typename T, typename I, int IN_OV, int OUT_OV>
static int accumulate3_k(int kernel, T *out, const T *lines, int64_t n_lines, const I *terms, const I *frames,
                         int64_t n_terms, const Vol3 &v, T max_val, int *flags, hipStream_t st)
{
#define CBX_ACC3(KK)                                                                                          \
    accumulate3_batches<T, I, KK, IN_OV, OUT_OV>(out, lines, n_lines, terms, frames, n_terms, v, max_val, flags, st)
    switch (kernel) {
    case K_BIWEIGHT: return CBX_ACC3(K_BIWEIGHT);
    case K_GAUSSIAN: return CBX_ACC3(K_GAUSSIAN);
    case K_PARABOLIC: return CBX_ACC3(K_PARABOLIC);
    case K_RECTANGULAR: return CBX_ACC3(K_RECTANGULAR);
    default: return CBX_ACC3(K_TRIANGULAR);
    }
#undef CBX_ACC3
}

template <typename T, typename I>
static int accumulate_lines3(T *out, int64_t n_frames, int64_t D, int64_t H, int64_t W, const T *lines,
                             int64_t n_lines, const I *terms, const I *frames, int64_t n_terms, T max_val, int kernel,
                             int in_overlap, int out_overlap, void *stream)
{
    if (in_overlap < 0 || in_overlap > 1 || out_overlap < 0 || out_overlap > 1)
        return fail(CBX_EVALUE, "Invalid in_overlap and out_overlap keywords");
    int rc = check_vol(n_frames, D, H, W, n_lines, kernel, 0);
    if (rc) return rc;
    if (n_terms < 0) return fail(CBX_EVALUE, "negative size");
    if (n_lines > 0 && !terms) return fail(CBX_EVALUE, "terms must not be NULL");
    if (n_terms > 0 && !frames) return fail(CBX_EVALUE, "frames must not be NULL");
    hipStream_t st = (hipStream_t)stream;
    Workspace ws(st);
    const bool same = in_overlap == out_overlap;
    DevBuf<I> comp;
    const int64_t n_chk = std::max(n_lines, n_terms);
    if (n_chk > 0) {
        DevBuf<int> bad;
        CBX_HIP(bad.alloc(2, st));
        CBX_HIP(hipMemsetAsync(bad.ptr, 0, 2 * sizeof(int), st));
        if (same && n_lines > 0) CBX_HIP(comp.alloc(n_lines, st));
        k_term_frames3<I><<<grid_for(n_chk, D3_NT), D3_NT, 0, st>>>(terms, frames, n_lines, n_terms, n_frames,
                                                                   comp.ptr, bad.ptr);
        CBX_HIP(hipGetLastError());
        uint64_t hv[2] = {0, 0};
        const void *src[2] = {bad.ptr, bad.ptr + 1};
        const int bytes[2] = {4, 4};
        CBX_HIP(read_device_values(st, 2, src, bytes, hv));
        if (hv[1]) return fail(CBX_EINDEX, "frames is outside of (0, n_frames)");
        if (hv[0]) return fail(CBX_EINDEX, "a term index is outside of (0, len(frames))");
    }
    if (n_lines == 0 || n_frames * D * H * W == 0) return CBX_OK;
    if (same) {
        // max/max: the term buffer starts at zero, so with max_val < 0 every
        // drawn voxel holds 0 when its term ends
        const T mv = (in_overlap == 1 && max_val < T(0)) ? T(0) : max_val;
        return draw3_run<T, I>(out, Vol3{n_frames, D, H, W}, lines, n_lines, comp.ptr, mv, kernel, in_overlap, st);
    }
    const Vol3 v{n_frames, D, H, W};
    DevBuf<int> flags;
    CBX_HIP(flags.alloc(1, st));
    if ((rc = validate3<T, I>(lines, n_lines, nullptr, v, flags.ptr, st))) return rc;
    {
        KernelTimer timer("accumulate3", st);
        if (in_overlap == 0)
            rc = accumulate3_k<T, I, 0, 1>(kernel, out, lines, n_lines, terms, frames, n_terms, v, max_val, flags.ptr, st);
        else
            rc = accumulate3_k<T, I, 1, 0>(kernel, out, lines, n_lines, terms, frames, n_terms, v, max_val, flags.ptr, st);
    }
    if (rc) return rc;
    int hv = 0;
    if ((rc = read_flags(flags.ptr, st, &hv))) return rc;
    return flags3_status(hv);
}

} // namespace cbx

extern "C" {

int cbx_draw_lines3_f32(float *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                        const float *lines, int64_t n_lines, const int32_t *idxs, float max_val, int kernel,
                        int overlap, void *stream)
{
    return cbx::draw_lines3<float, int32_t>(out, n_frames, depth, height, width, lines, n_lines, idxs, max_val, kernel,
                                            overlap, stream);
}

int cbx_draw_lines3_f64(double *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                        const double *lines, int64_t n_lines, const int64_t *idxs, double max_val, int kernel,
                        int overlap, void *stream)
{
    return cbx::draw_lines3<double, int64_t>(out, n_frames, depth, height, width, lines, n_lines, idxs, max_val,
                                             kernel, overlap, stream);
}

int cbx_accumulate_lines3_f32(float *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                              const float *lines, int64_t n_lines, const int32_t *terms, const int32_t *frames,
                              int64_t n_terms, float max_val, int kernel, int in_overlap, int out_overlap, void *stream)
{
    return cbx::accumulate_lines3<float, int32_t>(out, n_frames, depth, height, width, lines, n_lines, terms, frames,
                                                  n_terms, max_val, kernel, in_overlap, out_overlap, stream);
}

int cbx_accumulate_lines3_f64(double *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                              const double *lines, int64_t n_lines, const int64_t *terms, const int64_t *frames,
                              int64_t n_terms, double max_val, int kernel, int in_overlap, int out_overlap,
                              void *stream)
{
    return cbx::accumulate_lines3<double, int64_t>(out, n_frames, depth, height, width, lines, n_lines, terms, frames,
                                                   n_terms, max_val, kernel, in_overlap, out_overlap, stream);
}

} // extern "C"
