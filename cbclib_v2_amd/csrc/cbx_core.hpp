// Device core shared by the draw_lines kernels: per-line setup, the exact
// walk, the pixel error and the five profiles.
//
// Exactness contract.  A pixel belongs to a segment's footprint iff the
// reference's float walk emits it with error <= 1
// (cbclib_v2/_src/src/bresenham.hpp:352-383, bresenham.cpp:80-86).  That set
// is decided by float values accumulated one rounded addition at a time, so
// the kernels here replay the same additions in the same order (no FMA
// contraction: the library is built with -ffp-contract=off) and evaluate the
// same predicates.  Divisions by the segment length and by w^2 use a
// precomputed correctly-rounded reciprocal and one FMA correction step, which
// returns the correctly-rounded quotient (Markstein's theorem), i.e. exactly
// what the reference's IEEE division returns.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbx {

enum Profile { K_BIWEIGHT = 0, K_GAUSSIAN = 1, K_PARABOLIC = 2, K_RECTANGULAR = 3, K_TRIANGULAR = 4 };
// template value for "each line carries its own profile" (LineSetup::kid)
constexpr int K_MIXED = 5;

// Longest walk accepted (steps along the long axis) and the largest |pixel
// coordinate| a walk end point may have.
constexpr int64_t MAX_WALK = int64_t(1) << 26;
constexpr double MAX_COORD = 1e9;

// error flags raised by the device passes
constexpr int FLAG_INDEX = 1;     // idxs outside [0, n_frames)
constexpr int FLAG_LONG = 2;      // walk longer than MAX_WALK or end point beyond MAX_COORD
constexpr int FLAG_STUCK = 4;     // a walk did not leave its segment within its step budget
constexpr int FLAG_KERNEL = 16;   // a per-line profile id outside K_BIWEIGHT .. K_TRIANGULAR
constexpr int FLAG_OVERFLOW = 32; // a plan's segment or item capacity was exceeded

// ---------------------------------------------------------------------------
// type-exact math (float stays float: roundf/ceilf/sqrtf of the reference)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float m_abs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double m_abs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float m_round(float x) { return __builtin_roundf(x); }
__device__ __forceinline__ double m_round(double x) { return __builtin_round(x); }
__device__ __forceinline__ float m_ceil(float x) { return __builtin_ceilf(x); }
__device__ __forceinline__ double m_ceil(double x) { return __builtin_ceil(x); }
// IEEE square roots.  Note: HIP's __fsqrt_rn is the native v_sqrt_f32 (1 ulp,
// not correctly rounded) unless OCML_BASIC_ROUNDED_OPERATIONS is defined;
// __builtin_sqrtf under -fhip-fp32-correctly-rounded-divide-sqrt is exact.
__device__ __forceinline__ float m_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return __builtin_sqrt(x); }

// element i (0 or 1) of a two-element array without dynamic indexing, which
// would move the array to scratch memory
template <typename V>
__device__ __forceinline__ V pick(const V (&v)[2], int i)
{
    return i ? v[1] : v[0];
}

// ---------------------------------------------------------------------------
// correctly rounded quotient from a correctly rounded reciprocal
// ---------------------------------------------------------------------------
__device__ __forceinline__ float div_rn(float a, float b, float rb)
{
    float q = a * rb;
    float r = __builtin_fmaf(-q, b, a);
    float q1 = __builtin_fmaf(r, rb, q);
    // the correction needs a normal quotient; tiny numerators take the IEEE path
    return (__builtin_fabsf(q) < 1e-30f) ? a / b : q1;
}

__device__ __forceinline__ double div_rn(double a, double b, double rb)
{
    double q = a * rb;
    double r = __builtin_fma(-q, b, a);
    double q1 = __builtin_fma(r, rb, q);
    return (__builtin_fabs(q) < 1e-290) ? a / b : q1;
}

// Same without the guard, for arguments known to give a normal quotient.
__device__ __forceinline__ float div_rn_nz(float a, float b, float rb)
{
    const float q = a * rb;
    const float r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, rb, q);
}

// ---------------------------------------------------------------------------
// per-line constants (everything the walk, the fills and the error need)
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(16) LineSetup {
    T t[2];        // tangent p1 - p0 (geometry.hpp:316)
    T nv[2];       // normal derror (y1 - y0, x0 - x1) (bresenham.hpp:310-316)
    T len, rlen;   // |t| (float sqrt of the float magnitude) and RN(1/len)
    T LL;          // RN(len * len) (TangentError::error_at)
    T w, w2;       // width and RN(w * w)
    float nlim;    // float lines: |Ne| < nlim exactly when RN(RN(Ne / len)^2) < w2 (band_limit)
    double rw2;    // RN64(1 / (double)w2)
    float rw2f;    // RN32(1 / w2) for the fast error estimate
    float ea;      // float lines: RN32(1 / (len^2 w2)), the fast path's error scale (err_fast)
    int32_t m1[2]; // walk end (exclusive, "any coordinate equal" test)
    int32_t m0[2]; // walk start
    int32_t frame;
    int8_t axis;   // long axis (0 = x, 1 = y)
    int8_t st[2];  // step per axis (+1 / -1)
    int8_t valid;  // 0: the line emits nothing
    int8_t safe;   // 1: quotients of its errors by len cannot be subnormal (see line_setup)
    int8_t kid;    // profile of the line (K_BIWEIGHT .. K_TRIANGULAR)
    T e0[2];       // tangent and normal errors at m0
    T p0[2];       // first end point (for the adjoint's closed-form derivatives)
};

// Smallest float x >= 0 with !(RN(RN(x / len)^2) < w2) (the normal-error test
// of nsq_of).  That square is non-decreasing in |x| and RN is symmetric, so
// |Ne| < nlim decides the fill loops' band test exactly.  Found by stepping
// or bisecting over the bit patterns of the non-negative floats, which are
// ordered like their values.
__device__ __forceinline__ float band_limit(const LineSetup<float> &L)
{
    auto out = [&](uint32_t bits) {
        const float q = div_rn(__uint_as_float(bits), L.len, L.rlen);
        return !(q * q < L.w2);
    };
    if (out(0u)) return 0.0f;
    // the threshold lies within a few ulps of w len: step from there (x is
    // the answer when out(x) and not out(x - 1)), and bisect only if that
    // has not settled within 16 steps
    uint32_t x = __float_as_uint(L.w * L.len);
    if (x > 0u && x < 0x7f800000u) {
        for (int i = 0; i < 16; i++) {
            if (!out(x)) x++;
            else if (out(x - 1u)) x--;
            else return __uint_as_float(x);
        }
    }
    uint32_t lo = 0u, hi = 0x7f800000u;   // out(lo) false; out(+inf) true (a NaN or +inf square)
    while (hi - lo > 1u) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (out(mid)) hi = mid;
        else lo = mid;
    }
    return __uint_as_float(hi);
}

// Builds the setup exactly as BresenhamPlotter's constructor does
// (bresenham.hpp:243-253, 303-347).  Returns a FLAG_* on a rejected line.
template <typename T>
__device__ __forceinline__ int line_setup(const T *ln, LineSetup<T> &L)
{
    T x0 = ln[0], y0 = ln[1], x1 = ln[2], y1 = ln[3], w = ln[4];
    T p0[2] = {x0, y0}, p1[2] = {x1, y1};
    L.t[0] = x1 - x0;
    L.t[1] = y1 - y0;
    L.nv[0] = y1 - y0;
    L.nv[1] = x0 - x1;
    L.w = w;
    L.p0[0] = x0;
    L.p0[1] = y0;
    L.valid = 0;
    // error() is +inf for w <= 0 and NaN never compares true: nothing emitted
    if (!(w > T(0))) return 0;
    T mag = T(0);
    mag += L.t[0] * L.t[0];
    mag += L.t[1] * L.t[1];
    T len = m_sqrt(mag);
    if (!(len > T(0)) || !__builtin_isfinite(len) || !__builtin_isfinite(w)) return 0;
    if (!__builtin_isfinite(x0) || !__builtin_isfinite(y0) || !__builtin_isfinite(x1) || !__builtin_isfinite(y1))
        return 0;
    L.len = len;
    L.rlen = T(1) / len;
    L.LL = len * len;
    L.w2 = w * w;
    L.rw2 = 1.0 / (double)L.w2;
    L.rw2f = (float)(1.0 / (double)L.w2);
    int axis = (m_abs(L.t[0]) < m_abs(L.t[1])) ? 1 : 0;
    L.axis = (int8_t)axis;
    long offset = (long)(m_ceil(w) + T(1));
    T tau[2] = {L.t[0] / len, L.t[1] / len};
    T sc = m_abs((T)offset / tau[axis]);
    for (int i = 0; i < 2; i++) {
        int st = (p1[i] >= p0[i]) ? 1 : -1;
        T a0 = m_round(p0[i] - tau[i] * sc);
        T a1 = m_round(p1[i] + tau[i] * sc);
        if (!(m_abs(a0) < T(MAX_COORD)) || !(m_abs(a1) < T(MAX_COORD))) return FLAG_LONG;
        L.st[i] = (int8_t)st;
        L.m0[i] = (int32_t)a0;
        L.m1[i] = (int32_t)a1 + st;
    }
    int64_t span = (int64_t)L.m1[axis] - (int64_t)L.m0[axis];
    if (span > MAX_WALK || span < -MAX_WALK) return FLAG_LONG;
    T d0 = (T)L.m0[0] - x0, d1 = (T)L.m0[1] - y0;
    T Te = T(0);
    Te += L.t[0] * d0;
    Te += L.t[1] * d1;
    T Ne = T(0);
    Ne += L.nv[0] * d0;
    Ne += L.nv[1] * d1;
    L.e0[0] = Te;
    L.e0[1] = Ne;
    L.valid = 1;
    // With every end point coordinate 0 or at least 2^-20 in magnitude (and
    // w >= 2^-20), each walk error is 0 or far above the subnormal range, so
    // the reciprocal-based division needs no tiny-quotient guard.
    // Widths up to 2^20 also keep every square in err_fast finite and normal
    // near the threshold.
    const T lim = T(1.0 / 1048576.0);
    auto ok = [&](T v) { return v == T(0) || m_abs(v) >= lim; };
    L.safe = (int8_t)(ok(x0) && ok(y0) && ok(x1) && ok(y1) && w >= lim && w <= T(1048576.0));
    if constexpr (sizeof(T) == 4) {
        L.ea = (float)(1.0 / ((double)len * (double)len * (double)L.w2));
        L.nlim = band_limit(L);
    } else {
        L.ea = 0.0f;
        L.nlim = 0.0f;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// walk state and one step of LineIterator (bresenham.hpp:99-112, 149-154,
// 198-222)
// ---------------------------------------------------------------------------
// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt left at their
// maxima).  Placed before a loop whose first iteration reads loaded values, it
// settles them there; otherwise the compiler's wait for them sits in the loop
// and, being a count, also waits on every iteration for memory operations the
// loop itself issued (returning atomics, stores).
__device__ __forceinline__ void wait_vm_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// s_waitcnt lgkmcnt(0) (vmcnt and expcnt left at their maxima): the same for
// LDS and scalar loads, before a loop that issues LDS atomics, whose pending
// count would otherwise make the loop wait for its own previous atomic.
__device__ __forceinline__ void wait_lds_loads() { __builtin_amdgcn_s_waitcnt(0xC07F); }

template <typename T>
struct WalkState {
    int32_t c[2];
    T Te, Ne;
};

// next[] decision of NormalError::is_next for the pair (x, y).
template <typename T>
__device__ __forceinline__ void walk_next(const LineSetup<T> &L, T Ne, bool &nx, bool &ny)
{
    T dx = (L.st[0] > 0) ? L.nv[0] : -L.nv[0];
    T dy = (L.st[1] > 0) ? L.nv[1] : -L.nv[1];
    T ex = Ne + dx;
    T ey = Ne + dy;
    T exy = ex + dy;
    T aex = m_abs(ex), aey = m_abs(ey), aexy = m_abs(exy);
    if (aey < aex) {
        nx = aexy < aey;
        ny = true;
    } else {
        nx = true;
        ny = aexy < aex;
    }
}

template <typename T>
__device__ __forceinline__ void walk_step(const LineSetup<T> &L, WalkState<T> &s, bool nx, bool ny)
{
    if (nx) {
        s.c[0] += L.st[0];
        s.Te += (L.st[0] > 0) ? L.t[0] : -L.t[0];
        s.Ne += (L.st[0] > 0) ? L.nv[0] : -L.nv[0];
    }
    if (ny) {
        s.c[1] += L.st[1];
        s.Te += (L.st[1] > 0) ? L.t[1] : -L.t[1];
        s.Ne += (L.st[1] > 0) ? L.nv[1] : -L.nv[1];
    }
}

template <typename T>
__device__ __forceinline__ bool walk_done(const LineSetup<T> &L, const WalkState<T> &s)
{
    return s.c[0] == L.m1[0] || s.c[1] == L.m1[1];
}

// ---------------------------------------------------------------------------
// pixel error (BresenhamPlotter::error / normal_error, bresenham.hpp:279-291)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float nsq_of(const LineSetup<float> &L, float Ne)
{
    float qn = div_rn(Ne, L.len, L.rlen);
    return qn * qn;   // == (float)((double)qn * qn): the double square is exact
}

__device__ __forceinline__ double nsq_of(const LineSetup<double> &L, double Ne)
{
    double qn = div_rn(Ne, L.len, L.rlen);
    return qn * qn;
}

// tangential overshoot max(-T, T - len^2, 0) (TangentError::error_at)
template <typename T>
__device__ __forceinline__ T tangent_at(const LineSetup<T> &L, T Te)
{
    const T a = -Te, b = Te - L.LL;
    const T m = (a < b) ? b : a;
    return (m < T(0)) ? T(0) : m;
}

// exact error from the tangential overshoot and the normal error's square
__device__ __forceinline__ float err_exact(const LineSetup<float> &L, float at, float nsq)
{
    const float qt = div_rn(at, L.len, L.rlen);
    // RN64(qt^2 + nsq): the square of a float is exact in double, so one FMA
    // rounds exactly like the reference's double addition
    const double e = __builtin_fma((double)qt, (double)qt, (double)nsq);
    return (float)div_rn(e, (double)L.w2, L.rw2);
}

__device__ __forceinline__ double err_exact(const LineSetup<double> &L, double at, double nsq)
{
    const double qt = div_rn(at, L.len, L.rlen);
    const double qt2 = qt * qt;
    const double e = qt2 + nsq;
    return div_rn(e, L.w2, L.rw2);
}

// Error of a pixel in the float fast path (L.safe lines), for the decision
// e <= 1 and the fast profiles.  Estimate: (at^2 + Ne^2) * ea, i.e. the exact
// error's (qt^2 + nsq) / w2 with the divisions by len^2 and w2 folded into one
// product: within ~10 ulp (< 1e-6 relative) of the exact error, whose
// roundings it skips.  Within 4e-6 of the threshold the exact error is
// recomputed, so e <= 1 decides as the reference does.  tangent_at as
// max(-Te, Te - LL, 0): the errors are finite, and a -0 overshoot squares to +0.
__device__ __forceinline__ float err_fast(const LineSetup<float> &L, float LL, float ea, float Te, float Ne)
{
    const float at = __builtin_fmaxf(__builtin_fmaxf(-Te, Te - LL), 0.0f);
    const float e = __builtin_fmaf(at, at, Ne * Ne) * ea;
    if (__builtin_fabsf(e - 1.0f) <= 4e-6f) return err_exact(L, at, nsq_of(L, Ne));
    return e;
}

template <typename T>
__device__ __forceinline__ T err_of(const LineSetup<T> &L, T Te, T Ne)
{
    return err_exact(L, tangent_at(L, Te), nsq_of(L, Ne));
}

// ---------------------------------------------------------------------------
// profiles (include.hpp:446-468): the float versions keep the reference's
// promotions (std::pow(float, int) is a double square, std::max<T> rounds
// back to T)
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ float profile(float x)
{
    if constexpr (K == K_BIWEIGHT) {
        float m = (float)(1.0 - (double)x * (double)x);
        m = (m < 0.0f) ? 0.0f : m;
        return (float)(0.9375 * ((double)m * (double)m));
    } else if constexpr (K == K_GAUSSIAN) {
        if (__builtin_fabsf(x) <= 1.0f) {
            float x3 = 3.0f * x;
            double e = ::exp(-((double)x3 * (double)x3) / 2.0);
            return (float)(0.3989422804014327 * e);
        }
        return 0.0f;
    } else if constexpr (K == K_PARABOLIC) {
        float m = (float)(1.0 - (double)x * (double)x);
        m = (m < 0.0f) ? 0.0f : m;
        return 0.75f * m;
    } else if constexpr (K == K_RECTANGULAR) {
        return (__builtin_fabsf(x) <= 1.0f) ? 1.0f : 0.0f;
    } else {
        float m = 1.0f - __builtin_fabsf(x);
        return (m < 0.0f) ? 0.0f : m;
    }
}

template <int K>
__device__ __forceinline__ double profile(double x)
{
    if constexpr (K == K_BIWEIGHT) {
        double m = 1.0 - x * x;
        m = (m < 0.0) ? 0.0 : m;
        return 0.9375 * (m * m);
    } else if constexpr (K == K_GAUSSIAN) {
        if (__builtin_fabs(x) <= 1.0) {
            double x3 = 3.0 * x;
            return 0.3989422804014327 * ::exp(-(x3 * x3) / 2.0);
        }
        return 0.0;
    } else if constexpr (K == K_PARABOLIC) {
        double m = 1.0 - x * x;
        m = (m < 0.0) ? 0.0 : m;
        return 0.75 * m;
    } else if constexpr (K == K_RECTANGULAR) {
        return (__builtin_fabs(x) <= 1.0) ? 1.0 : 0.0;
    } else {
        double m = 1.0 - __builtin_fabs(x);
        return (m < 0.0) ? 0.0 : m;
    }
}

// value of an emitted pixel: max_val * K(sqrt(err)) (bresenham.cpp:84)
template <int K>
__device__ __forceinline__ float pixel_value(float err, float max_val)
{
    return max_val * profile<K>(m_sqrt(err));
}

template <int K>
__device__ __forceinline__ double pixel_value(double err, double max_val)
{
    return max_val * profile<K>(m_sqrt(err));
}

// Fast profile values from u = err = x^2 (float arithmetic, a few ulp from
// the reference's values; used by the tile path, never by the exact modes).
template <int K>
__device__ __forceinline__ float profile_fast(float u)
{
    if constexpr (K == K_BIWEIGHT) {
        const float m = 1.0f - u;
        return 0.9375f * m * m;
    } else if constexpr (K == K_GAUSSIAN) {
        return 0.3989422804014327f * __expf(-4.5f * u);
    } else if constexpr (K == K_PARABOLIC) {
        return 0.75f * (1.0f - u);
    } else if constexpr (K == K_RECTANGULAR) {
        return 1.0f;
    } else {
        const float m = 1.0f - __builtin_amdgcn_sqrtf(u);
        return (m < 0.0f) ? 0.0f : m;
    }
}

// The fast profile as c * shape(u) with c = s * (the profile's constant
// factor) computed once per column (fast_scale) and shape evaluated per pixel
// (fast_value): s * K(sqrt(u)) to a few ulp.  K_MIXED picks the profile kid.
template <int K>
__device__ __forceinline__ float fast_scale(int kid, float s)
{
    const int k = (K == K_MIXED) ? kid : K;
    return (k == K_BIWEIGHT) ? s * 0.9375f
         : (k == K_GAUSSIAN) ? s * 0.3989422804014327f
         : (k == K_PARABOLIC) ? s * 0.75f
                              : s;
}

template <int K>
__device__ __forceinline__ float fast_value(int kid, float u, float c)
{
    if constexpr (K == K_MIXED) {
        switch (kid) {
        case K_BIWEIGHT: return fast_value<K_BIWEIGHT>(kid, u, c);
        case K_GAUSSIAN: return fast_value<K_GAUSSIAN>(kid, u, c);
        case K_PARABOLIC: return fast_value<K_PARABOLIC>(kid, u, c);
        case K_RECTANGULAR: return c;
        default: return fast_value<K_TRIANGULAR>(kid, u, c);
        }
    } else if constexpr (K == K_BIWEIGHT) {
        const float m = 1.0f - u;
        return c * (m * m);
    } else if constexpr (K == K_GAUSSIAN) {
        return c * __builtin_amdgcn_exp2f(u * -6.4921276840f);   // exp(-4.5 u) = 2^(-4.5 log2(e) u)
    } else if constexpr (K == K_PARABOLIC) {
        return c * (1.0f - u);
    } else if constexpr (K == K_RECTANGULAR) {
        return c;
    } else {
        const float m = 1.0f - __builtin_amdgcn_sqrtf(u);
        return (m < 0.0f) ? 0.0f : c * m;
    }
}

// Profiles whose fast value is positive on 0 <= u <= 1 for s > 0.
template <int K>
constexpr bool fast_positive() { return K == K_GAUSSIAN || K == K_RECTANGULAR; }

// Profile chosen at run time (per-line profiles).
__device__ __forceinline__ float profile_fast_dyn(int k, float u)
{
    switch (k) {
    case K_BIWEIGHT: return profile_fast<K_BIWEIGHT>(u);
    case K_GAUSSIAN: return profile_fast<K_GAUSSIAN>(u);
    case K_PARABOLIC: return profile_fast<K_PARABOLIC>(u);
    case K_RECTANGULAR: return profile_fast<K_RECTANGULAR>(u);
    default: return profile_fast<K_TRIANGULAR>(u);
    }
}

template <typename T>
__device__ __forceinline__ T pixel_value_dyn(int k, T err, T max_val)
{
    switch (k) {
    case K_BIWEIGHT: return pixel_value<K_BIWEIGHT>(err, max_val);
    case K_GAUSSIAN: return pixel_value<K_GAUSSIAN>(err, max_val);
    case K_PARABOLIC: return pixel_value<K_PARABOLIC>(err, max_val);
    case K_RECTANGULAR: return pixel_value<K_RECTANGULAR>(err, max_val);
    default: return pixel_value<K_TRIANGULAR>(err, max_val);
    }
}

// profile value of u = err for profile K, or the line's own profile
template <int K>
__device__ __forceinline__ float profile_fast_of(int kid, float u)
{
    if constexpr (K == K_MIXED) return profile_fast_dyn(kid, u);
    else return profile_fast<K>(u);
}

template <int K, typename T>
__device__ __forceinline__ T pixel_value_of(int kid, T err, T max_val)
{
    if constexpr (K == K_MIXED) return pixel_value_dyn(kid, err, max_val);
    else return pixel_value<K>(err, max_val);
}

// Decides whether a pixel is emitted (exactly as the reference: error <= 1)
// and computes its value.  FAST: the error is estimated with reciprocal
// products and recomputed exactly only within 4e-6 of the threshold (the
// estimate is within 1e-6 of the exact value); values from profile_fast.
template <bool FAST, int K>
__device__ __forceinline__ bool pixel_eval(const LineSetup<float> &L, float Te, float nsq, float max_val, float &v)
{
    const float at = tangent_at(L, Te);
    if constexpr (FAST) {
        const float qt = at * L.rlen;
        float e = __builtin_fmaf(qt, qt, nsq) * L.rw2f;
        if (__builtin_fabsf(e - 1.0f) <= 4e-6f) e = err_exact(L, at, nsq);
        if (!(e <= 1.0f)) return false;
        v = max_val * profile_fast_of<K>(L.kid, e);
    } else {
        const float e = err_exact(L, at, nsq);
        if (!(e <= 1.0f)) return false;
        v = pixel_value_of<K>(L.kid, e, max_val);
    }
    return true;
}

template <bool FAST, int K>
__device__ __forceinline__ bool pixel_eval(const LineSetup<double> &L, double Te, double nsq, double max_val, double &v)
{
    const double e = err_exact(L, tangent_at(L, Te), nsq);
    if (!(e <= 1.0)) return false;
    v = pixel_value_of<K>(L.kid, e, max_val);
    return true;
}

// dK/du at u = err = (dist / w)^2 for the adjoint; the triangular profile is
// handled through dist in the caller.
template <int K>
__device__ __forceinline__ float profile_du(float u, float kval)
{
    if constexpr (K == K_BIWEIGHT) return -1.875f * (1.0f - u);
    else if constexpr (K == K_PARABOLIC) return -0.75f;
    else if constexpr (K == K_GAUSSIAN) return -4.5f * kval;
    else return 0.0f;
}

// order-preserving encodings for max accumulation in LDS integer atomics
__device__ __forceinline__ uint32_t ord_enc(float v)
{
    uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t e)
{
    uint32_t b = (e & 0x80000000u) ? (e & 0x7fffffffu) : ~e;
    return __uint_as_float(b);
}
__device__ __forceinline__ unsigned long long ord_enc(double v)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord_dec(unsigned long long e)
{
    unsigned long long b = (e & 0x8000000000000000ull) ? (e & 0x7fffffffffffffffull) : ~e;
    return __longlong_as_double((long long)b);
}

} // namespace cbx
