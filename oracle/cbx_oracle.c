/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of cbclib_v2's streak rasterizer (draw_lines and
 * accumulate_lines, 2D and 3D lines), of its connected-component labelling
 * (label), region moments, p-values and binary_dilation, written from the
 * reference's behaviour, used as the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library; the product (cbclib_v2_amd) never does.
 *
 * Arithmetic contract (what makes pixel indices bit-exact):
 *   - float walk state is accumulated incrementally in T (float or double),
 *     one rounding per addition, exactly as BaseError::increment
 *     (cbclib_v2/_src/src/bresenham.hpp:59-63) does;
 *   - std::pow(float, 2) in the reference promotes to double (C++11 overload
 *     rules), so squares of float quotients are taken in double here too
 *     (bresenham.hpp:282, 289);
 *   - the file is compiled with -ffp-contract=off so that no product is
 *     fused into an FMA (the reference is built with plain -O3 on x86-64,
 *     which has no FMA in its baseline ISA).
 *
 * Every function cites the reference lines it follows.  Pinning (what checks
 * this file against something independent) is listed per function.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CBXO_OK 0
#define CBXO_EVALUE 1   /* ValueError  (std::invalid_argument)   */
#define CBXO_EINDEX 2   /* IndexError  (std::out_of_range)       */
#define CBXO_ERUNTIME 3 /* RuntimeError                          */

enum { K_BIWEIGHT = 0, K_GAUSSIAN = 1, K_PARABOLIC = 2, K_RECTANGULAR = 3, K_TRIANGULAR = 4 };

/* Longest walk (in steps) accepted; the reference has no limit and simply
 * runs for as long as the segment is. */
#define CBXO_MAX_WALK (1L << 26)

/* Planted mistakes, for the tests that show each pin catches a plausible
 * error in this file (0 = none, the default; tests only):
 *   1 drop the tangent term of the pixel error
 *   2 wrong sign of one normal component
 *   3 transposed operands (x and y of the end points swapped)
 *   4 label ignores neighbour shifts of L1 norm > 1 (the connectivity norm
 *     confused)
 *   5 profile evaluated at the error instead of its square root */
static int g_plant = 0;
void cbxo_plant(int mistake) { g_plant = mistake; }

/* ------------------------------------------------------------------------ */
/* Profiles: include.hpp:446-468 (kernel_traits).  Float versions reproduce  */
/* the promotions of the C++ expressions (std::pow(float, int) -> double,    */
/* std::max<T> -> conversion back to T).                                     */
/* Pinned by: closed forms (tests/test_oracle_cpu.py) and by the compiled    */
/* reference through draw_lines images and the golden fixtures.              */
/* ------------------------------------------------------------------------ */

static float prof_f32(int k, float x)
{
    switch (k) {
    case K_BIWEIGHT: {
        /* T(0.9375) * pow(max<T>(1 - pow(x, 2), 0), 2) */
        float m = (float)(1.0 - (double)x * (double)x);
        m = (m < 0.0f) ? 0.0f : m;
        return (float)((double)0.9375f * ((double)m * (double)m));
    }
    case K_GAUSSIAN: {
        /* if |x| <= 1: M_1_SQRT2PI * exp(-pow(3 * x, 2) / 2) */
        if (fabsf(x) <= 1.0f) {
            float x3 = 3.0f * x;
            double e = exp(-((double)x3 * (double)x3) / 2.0);
            return (float)(0.3989422804014327 * e);
        }
        return 0.0f;
    }
    case K_PARABOLIC: {
        /* T(0.75) * max<T>(1 - pow(x, 2), 0) */
        float m = (float)(1.0 - (double)x * (double)x);
        m = (m < 0.0f) ? 0.0f : m;
        return 0.75f * m;
    }
    case K_RECTANGULAR:
        return (fabsf(x) <= 1.0f) ? 1.0f : 0.0f;
    case K_TRIANGULAR: {
        float m = 1.0f - fabsf(x);
        return (m < 0.0f) ? 0.0f : m;
    }
    }
    return 0.0f;
}

static double prof_f64(int k, double x)
{
    switch (k) {
    case K_BIWEIGHT: {
        double m = 1.0 - x * x;
        m = (m < 0.0) ? 0.0 : m;
        return 0.9375 * (m * m);
    }
    case K_GAUSSIAN:
        if (fabs(x) <= 1.0) {
            double x3 = 3.0 * x;
            return 0.3989422804014327 * exp(-(x3 * x3) / 2.0);
        }
        return 0.0;
    case K_PARABOLIC: {
        double m = 1.0 - x * x;
        m = (m < 0.0) ? 0.0 : m;
        return 0.75 * m;
    }
    case K_RECTANGULAR:
        return (fabs(x) <= 1.0) ? 1.0 : 0.0;
    case K_TRIANGULAR: {
        double m = 1.0 - fabs(x);
        return (m < 0.0) ? 0.0 : m;
    }
    }
    return 0.0;
}

double cbxo_profile_f64(int k, double x) { return prof_f64(k, x); }
float cbxo_profile_f32(int k, float x) { return prof_f32(k, x); }

/* ------------------------------------------------------------------------ */
/* The walk.  Follows BresenhamPlotter (bresenham.hpp:233-348), LineIterator */
/* (118-231), the error classes (38-113) and draw_line_2d (352-383).         */
/* Emits every (x, y, err) the reference hands to its draw_pixel callback,   */
/* in the same order (bresenham.cpp:80-88).                                  */
/* Pinned by: the compiled reference module (oracle/_ref) on random inputs   */
/* and the committed golden fixtures (tests/test_oracle_cpu.py); and by a    */
/* brute-force {dist <= w} footprint away from the boundary.                 */
/* ------------------------------------------------------------------------ */

typedef void (*emit_f32_t)(void *ctx, long x, long y, float err);
typedef void (*emit_f64_t)(void *ctx, long x, long y, double err);

/* error() of BresenhamPlotter for T = float (bresenham.hpp:279-291). */
static float err_f32(float T, float N, float len, float w)
{
    if (w <= 0.0f) return INFINITY;
    float a = -T, b = T - len * len;
    float m = (a < b) ? b : a;          /* std::max(a, b) */
    float at = (m < 0.0f) ? 0.0f : m;   /* std::max(., T()) */
    float qt = (g_plant == 1) ? 0.0f : at / len;
    float qn = N / len;
    float nsq = (float)((double)qn * (double)qn);
    double e = (double)qt * (double)qt + (double)nsq;
    return (float)(e / (double)(w * w));
}

static float nerr_f32(float N, float len)
{
    float qn = N / len;
    return (float)((double)qn * (double)qn);
}

static double err_f64(double T, double N, double len, double w)
{
    if (w <= 0.0) return INFINITY;
    double a = -T, b = T - len * len;
    double m = (a < b) ? b : a;
    double at = (m < 0.0) ? 0.0 : m;
    double qt = (g_plant == 1) ? 0.0 : at / len;
    double qn = N / len;
    double e = qt * qt + qn * qn;
    return e / (w * w);
}

static double nerr_f64(double N, double len)
{
    double qn = N / len;
    return qn * qn;
}

#define WALK_BODY(REAL, ROUND, CEIL, FABS, SQRT, ERRF, NERRF, EMIT_T)                           \
    REAL x0 = ln[0], y0 = ln[1], x1 = ln[2], y1 = ln[3], w = ln[4];                             \
    if (g_plant == 3) { REAL s_ = x0; x0 = y0; y0 = s_; s_ = x1; x1 = y1; y1 = s_; }            \
    REAL p0[2] = {x0, y0}, p1[2] = {x1, y1};                                                    \
    /* tangent (geometry.hpp:316) and normal (bresenham.hpp:310-316) */                          \
    REAL t[2] = {x1 - x0, y1 - y0};                                                             \
    REAL nv[2] = {y1 - y0, (g_plant == 2) ? (x1 - x0) : (x0 - x1)};                             \
    if (!(w > (REAL)0)) return CBXO_OK; /* error() == inf for w <= 0; NaN never emits */        \
    REAL mag = (REAL)0; mag += t[0] * t[0]; mag += t[1] * t[1];                                 \
    REAL len = SQRT(mag);                                                                       \
    if (!(len > (REAL)0) || !isfinite(len) || !isfinite(w)) return CBXO_OK;                     \
    if (!isfinite(x0) || !isfinite(y0) || !isfinite(x1) || !isfinite(y1)) return CBXO_OK;       \
    int axis = (FABS(t[0]) < FABS(t[1])) ? 1 : 0; /* first max |tau| */                         \
    int sax = 1 - axis;                                                                         \
    long offset = (long)(CEIL(w) + (REAL)1);                                                    \
    REAL tau[2] = {t[0] / len, t[1] / len};                                                     \
    REAL sc = FABS((REAL)offset / tau[axis]);                                                   \
    long step[2], m0[2], m1[2];                                                                 \
    for (int i = 0; i < 2; i++) {                                                               \
        step[i] = (p1[i] >= p0[i]) ? 1 : -1;                                                    \
        REAL a0 = ROUND(p0[i] - tau[i] * sc);                                                   \
        REAL a1 = ROUND(p1[i] + tau[i] * sc);                                                   \
        if (!(FABS(a0) < (REAL)1e9) || !(FABS(a1) < (REAL)1e9)) return CBXO_EVALUE;             \
        m0[i] = (long)a0;                                                                       \
        m1[i] = (long)a1 + step[i];                                                             \
    }                                                                                           \
    if (labs(m1[axis] - m0[axis]) > CBXO_MAX_WALK) return CBXO_EVALUE;                          \
    /* BaseError(derr, point, origin): error = dot(derr, point - origin) */                     \
    REAL d0 = (REAL)m0[0] - p0[0], d1 = (REAL)m0[1] - p0[1];                                    \
    REAL T = (REAL)0; T += t[0] * d0; T += t[1] * d1;                                           \
    REAL N = (REAL)0; N += nv[0] * d0; N += nv[1] * d1;                                         \
    REAL w2 = w * w;                                                                            \
    long cur[2] = {m0[0], m0[1]};                                                               \
    /* for (iter = begin; iter != end; ++iter): end test is "any coordinate equal" */           \
    while (!(cur[0] == m1[0] || cur[1] == m1[1])) {                                             \
        /* update() with NormalError::is_next (bresenham.hpp:99-112, 208-222) */               \
        REAL ex = N + (REAL)step[0] * nv[0];                                                    \
        REAL ey = N + (REAL)step[1] * nv[1];                                                    \
        REAL exy = ex + (REAL)step[1] * nv[1];                                                  \
        int nx, ny;                                                                             \
        if (FABS(ey) < FABS(ex)) { nx = (FABS(exy) < FABS(ey)); ny = 1; }                       \
        else { nx = 1; ny = (FABS(exy) < FABS(ex)); }                                           \
        int nxt[2] = {nx, ny};                                                                  \
        if (nxt[axis]) {                                                                        \
            emit(ctx, cur[0], cur[1], ERRF(T, N, len, w));                                      \
            for (int half = 0; half < 2; half++) {                                              \
                long dir = half ? -step[sax] : step[sax];                                       \
                long c[2] = {cur[0], cur[1]};                                                   \
                REAL Tf = T, Nf = N;                                                            \
                for (;;) {                                                                      \
                    c[sax] += dir;                                                              \
                    Tf += (REAL)dir * t[sax];                                                   \
                    Nf += (REAL)dir * nv[sax];                                                  \
                    if (!(NERRF(Nf, len) < w2)) break;                                          \
                    emit(ctx, c[0], c[1], ERRF(Tf, Nf, len, w));                                \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        /* operator++: increment axis 0 then axis 1 (bresenham.hpp:149-154, 198-206) */         \
        for (int i = 0; i < 2; i++) if (nxt[i]) {                                               \
            cur[i] += step[i];                                                                  \
            T += (REAL)step[i] * t[i];                                                          \
            N += (REAL)step[i] * nv[i];                                                         \
        }                                                                                       \
    }                                                                                           \
    return CBXO_OK;

static int walk_f32(const float *ln, emit_f32_t emit, void *ctx)
{
    WALK_BODY(float, roundf, ceilf, fabsf, sqrtf, err_f32, nerr_f32, float)
}

static int walk_f64(const double *ln, emit_f64_t emit, void *ctx)
{
    WALK_BODY(double, round, ceil, fabs, sqrt, err_f64, nerr_f64, double)
}

/* ------------------------------------------------------------------------ */
/* idxs rules: fill_indices / check_indices (bresenham.cpp:5-36, 54-55).     */
/* ------------------------------------------------------------------------ */

static int check_idxs_i32(const int32_t *idxs, long n_lines, long n_frames)
{
    if (!idxs) return (n_lines == 0 || n_frames == 1 || n_frames == n_lines) ? CBXO_OK : CBXO_EVALUE;
    for (long i = 0; i < n_lines; i++)
        if (idxs[i] < 0 || idxs[i] >= n_frames) return CBXO_EINDEX;
    return CBXO_OK;
}

static int check_idxs_i64(const int64_t *idxs, long n_lines, long n_frames)
{
    if (!idxs) return (n_lines == 0 || n_frames == 1 || n_frames == n_lines) ? CBXO_OK : CBXO_EVALUE;
    for (long i = 0; i < n_lines; i++)
        if (idxs[i] < 0 || idxs[i] >= n_frames) return CBXO_EINDEX;
    return CBXO_OK;
}

static long frame_of(long i, long n_frames, const void *idxs, int wide)
{
    if (idxs) return wide ? (long)((const int64_t *)idxs)[i] : (long)((const int32_t *)idxs)[i];
    return (n_frames == 1) ? 0 : i;
}

/* ------------------------------------------------------------------------ */
/* Image accumulation: draw_lines_nd_impl with one thread                    */
/* (bresenham.cpp:38-101): keep error <= 1 && inbound, value =               */
/* max_val * K(sqrt(error)), then out[idx + HW*frame] += value or max.       */
/* ------------------------------------------------------------------------ */

typedef struct {
    void *out; long H, W, frame; int kernel, overlap; double max_val;
    /* pixel-list mode */
    long count, cap; long line;
    int32_t *pl_line, *pl_frame, *pl_y, *pl_x; void *pl_val; void *pl_err;
} draw_ctx;

static void draw_px_f32(void *vctx, long x, long y, float err)
{
    draw_ctx *c = (draw_ctx *)vctx;
    if (!(err <= 1.0f) || y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    float v = (float)c->max_val * prof_f32(c->kernel, (g_plant == 5) ? err : sqrtf(err));
    float *o = (float *)c->out + (c->frame * c->H + y) * c->W + x;
    if (c->overlap) *o = (*o < v) ? v : *o;   /* std::max(out, value) */
    else *o = *o + v;
}

static void draw_px_f64(void *vctx, long x, long y, double err)
{
    draw_ctx *c = (draw_ctx *)vctx;
    if (!(err <= 1.0) || y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    double v = c->max_val * prof_f64(c->kernel, (g_plant == 5) ? err : sqrt(err));
    double *o = (double *)c->out + (c->frame * c->H + y) * c->W + x;
    if (c->overlap) *o = (*o < v) ? v : *o;
    else *o = *o + v;
}

int cbxo_draw_lines_f32(float *out, long n_frames, long H, long W, const float *lines, long n_lines,
                        const int32_t *idxs, float max_val, int kernel, int overlap)
{
    if (kernel < 0 || kernel > 4 || overlap < 0 || overlap > 1) return CBXO_EVALUE;
    int rc = check_idxs_i32(idxs, n_lines, n_frames);
    if (rc) return rc;
    draw_ctx c = {out, H, W, 0, kernel, overlap, max_val};
    for (long i = 0; i < n_lines; i++) {
        c.frame = frame_of(i, n_frames, idxs, 0);
        rc = walk_f32(lines + 5 * i, draw_px_f32, &c);
        if (rc) return rc;
    }
    return CBXO_OK;
}

int cbxo_draw_lines_f64(double *out, long n_frames, long H, long W, const double *lines, long n_lines,
                        const int64_t *idxs, double max_val, int kernel, int overlap)
{
    if (kernel < 0 || kernel > 4 || overlap < 0 || overlap > 1) return CBXO_EVALUE;
    int rc = check_idxs_i64(idxs, n_lines, n_frames);
    if (rc) return rc;
    draw_ctx c = {out, H, W, 0, kernel, overlap, max_val};
    for (long i = 0; i < n_lines; i++) {
        c.frame = frame_of(i, n_frames, idxs, 1);
        rc = walk_f64(lines + 5 * i, draw_px_f64, &c);
        if (rc) return rc;
    }
    return CBXO_OK;
}

/* ------------------------------------------------------------------------ */
/* Raw pixel list: the entries the reference pushes into its per-thread      */
/* ImageBuffer (bresenham.cpp:80-86, bresenham.hpp:10-34), in emission order */
/* for one thread: (line, frame, y, x, value, err).  Returns the number of   */
/* entries; arrays are written only up to cap.                               */
/* ------------------------------------------------------------------------ */

static void list_px_f32(void *vctx, long x, long y, float err)
{
    draw_ctx *c = (draw_ctx *)vctx;
    if (!(err <= 1.0f) || y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    long k = c->count++;
    if (k >= c->cap) return;
    c->pl_line[k] = (int32_t)c->line; c->pl_frame[k] = (int32_t)c->frame;
    c->pl_y[k] = (int32_t)y; c->pl_x[k] = (int32_t)x;
    ((float *)c->pl_val)[k] = (float)c->max_val * prof_f32(c->kernel, (g_plant == 5) ? err : sqrtf(err));
    if (c->pl_err) ((float *)c->pl_err)[k] = err;
}

static void list_px_f64(void *vctx, long x, long y, double err)
{
    draw_ctx *c = (draw_ctx *)vctx;
    if (!(err <= 1.0) || y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    long k = c->count++;
    if (k >= c->cap) return;
    c->pl_line[k] = (int32_t)c->line; c->pl_frame[k] = (int32_t)c->frame;
    c->pl_y[k] = (int32_t)y; c->pl_x[k] = (int32_t)x;
    ((double *)c->pl_val)[k] = c->max_val * prof_f64(c->kernel, (g_plant == 5) ? err : sqrt(err));
    if (c->pl_err) ((double *)c->pl_err)[k] = err;
}

long cbxo_pixel_list_f32(long n_frames, long H, long W, const float *lines, long n_lines, const int32_t *idxs,
                         float max_val, int kernel, int32_t *line, int32_t *frame, int32_t *y, int32_t *x,
                         float *val, float *err, long cap)
{
    if (kernel < 0 || kernel > 4) return -CBXO_EVALUE;
    int rc = check_idxs_i32(idxs, n_lines, n_frames);
    if (rc) return -rc;
    draw_ctx c = {NULL, H, W, 0, kernel, 0, max_val, 0, cap, 0, line, frame, y, x, val, err};
    for (long i = 0; i < n_lines; i++) {
        c.frame = frame_of(i, n_frames, idxs, 0);
        c.line = i;
        rc = walk_f32(lines + 5 * i, list_px_f32, &c);
        if (rc) return -rc;
    }
    return c.count;
}

long cbxo_pixel_list_f64(long n_frames, long H, long W, const double *lines, long n_lines, const int64_t *idxs,
                         double max_val, int kernel, int32_t *line, int32_t *frame, int32_t *y, int32_t *x,
                         double *val, double *err, long cap)
{
    if (kernel < 0 || kernel > 4) return -CBXO_EVALUE;
    int rc = check_idxs_i64(idxs, n_lines, n_frames);
    if (rc) return -rc;
    draw_ctx c = {NULL, H, W, 0, kernel, 0, max_val, 0, cap, 0, line, frame, y, x, val, err};
    for (long i = 0; i < n_lines; i++) {
        c.frame = frame_of(i, n_frames, idxs, 1);
        c.line = i;
        rc = walk_f64(lines + 5 * i, list_px_f64, &c);
        if (rc) return -rc;
    }
    return c.count;
}

/* ------------------------------------------------------------------------ */
/* FP64 closed form (dense brute force): value = max_val * K(dist / w) for   */
/* dist(p, segment) <= w, dist from the clamped projection of                */
/* LineND::project_to_streak (geometry.hpp:319-339); the same formula is the */
/* reference test's image_numpy (tests/draw_line_test.py:154-167).           */
/* Pinned by: analytic special cases (tests/test_oracle_cpu.py).             */
/* ------------------------------------------------------------------------ */

static double seg_dist2(double px, double py, double x0, double y0, double x1, double y1)
{
    double tx = x1 - x0, ty = y1 - y0, mag = tx * tx + ty * ty;
    double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
    double qx = x0, qy = y0;
    if (mag > 0.0) {
        double r = ((px - cx) * tx + (py - cy) * ty) / mag;
        r = (r < -0.5) ? -0.5 : ((r > 0.5) ? 0.5 : r);
        qx = r * tx + cx; qy = r * ty + cy;
    }
    return (px - qx) * (px - qx) + (py - qy) * (py - qy);
}

int cbxo_render_closed_f64(double *out, long n_frames, long H, long W, const double *lines, long n_lines,
                           const int64_t *idxs, double max_val, int kernel)
{
    if (kernel < 0 || kernel > 4) return CBXO_EVALUE;
    int rc = check_idxs_i64(idxs, n_lines, n_frames);
    if (rc) return rc;
    for (long i = 0; i < n_lines; i++) {
        const double *l = lines + 5 * i;
        double w = l[4];
        if (!(w > 0.0)) continue;
        long f = frame_of(i, n_frames, idxs, 1);
        long xa = (long)floor(fmin(l[0], l[2]) - w) - 1, xb = (long)ceil(fmax(l[0], l[2]) + w) + 1;
        long ya = (long)floor(fmin(l[1], l[3]) - w) - 1, yb = (long)ceil(fmax(l[1], l[3]) + w) + 1;
        if (xa < 0) xa = 0;
        if (ya < 0) ya = 0;
        if (xb > W - 1) xb = W - 1;
        if (yb > H - 1) yb = H - 1;
        for (long y = ya; y <= yb; y++)
            for (long x = xa; x <= xb; x++) {
                double d2 = seg_dist2((double)x, (double)y, l[0], l[1], l[2], l[3]);
                double u = d2 / (w * w);
                if (u <= 1.0) out[(f * H + y) * W + x] += max_val * prof_f64(kernel, sqrt(u));
            }
    }
    return CBXO_OK;
}

/* ------------------------------------------------------------------------ */
/* FP64 adjoint (no reference counterpart: the reference has no gradient).   */
/* dL/dtheta_l = sum over the pixels line l emits of                         */
/*   g[p] * max_val * dK/du * du/dtheta,   u = dist^2 / w^2,                 */
/* with d(dist^2)/dp0 = -2 (1 - t*) (p - q), d(dist^2)/dp1 = -2 t* (p - q),  */
/* q the clamped projection and t* in [0, 1] its parameter (envelope         */
/* theorem), du/dw = -2u/w.  The pixel set is the float walk's (f32 lines)   */
/* or the double walk's (f64 lines), so it is the set the forward uses.      */
/* Pinned by: FP64 central finite differences of cbxo_render_closed_f64 and  */
/* analytic special cases (tests/test_oracle_cpu.py).                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    const double *g; long H, W, frame; int kernel; double max_val;
    double x0, y0, x1, y1, w; double acc[5];
} adj_ctx;

static void adj_px(adj_ctx *c, long x, long y)
{
    if (y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    double gv = c->g[(c->frame * c->H + y) * c->W + x];
    if (gv == 0.0) return;
    double tx = c->x1 - c->x0, ty = c->y1 - c->y0, mag = tx * tx + ty * ty;
    double rx = (double)x - c->x0, ry = (double)y - c->y0;
    double ts = (rx * tx + ry * ty) / mag;
    ts = (ts < 0.0) ? 0.0 : ((ts > 1.0) ? 1.0 : ts);
    double dx = rx - ts * tx, dy = ry - ts * ty;
    double d2 = dx * dx + dy * dy, w = c->w, u = d2 / (w * w);
    double dist = sqrt(d2);
    /* dK/d(dist^2) and dK/dw */
    double dk_dd2, dk_dw;
    switch (c->kernel) {
    case K_BIWEIGHT: { double k1 = -1.875 * (1.0 - u); dk_dd2 = k1 / (w * w); dk_dw = k1 * (-2.0 * u / w); break; }
    case K_PARABOLIC: { double k1 = -0.75; dk_dd2 = k1 / (w * w); dk_dw = k1 * (-2.0 * u / w); break; }
    case K_GAUSSIAN: { double k1 = -4.5 * prof_f64(K_GAUSSIAN, sqrt(u)); dk_dd2 = k1 / (w * w); dk_dw = k1 * (-2.0 * u / w); break; }
    case K_TRIANGULAR: dk_dd2 = (dist > 0.0) ? -1.0 / (2.0 * w * dist) : 0.0; dk_dw = dist / (w * w); break;
    default: dk_dd2 = 0.0; dk_dw = 0.0; break;
    }
    double s = gv * c->max_val;
    c->acc[0] += s * dk_dd2 * (-2.0 * (1.0 - ts) * dx);
    c->acc[1] += s * dk_dd2 * (-2.0 * (1.0 - ts) * dy);
    c->acc[2] += s * dk_dd2 * (-2.0 * ts * dx);
    c->acc[3] += s * dk_dd2 * (-2.0 * ts * dy);
    c->acc[4] += s * dk_dw;
}

static void adj_px_f32(void *vctx, long x, long y, float err)
{
    if (!(err <= 1.0f)) return;
    adj_px((adj_ctx *)vctx, x, y);
}

static void adj_px_f64(void *vctx, long x, long y, double err)
{
    if (!(err <= 1.0)) return;
    adj_px((adj_ctx *)vctx, x, y);
}

/* lines_f32 may be NULL: then the pixel set comes from the double walk of lines. */
int cbxo_adjoint_f64(const double *grad_img, long n_frames, long H, long W, const double *lines,
                     const float *lines_f32, long n_lines, const int64_t *idxs, double max_val, int kernel,
                     double *grad_lines)
{
    if (kernel < 0 || kernel > 4) return CBXO_EVALUE;
    int rc = check_idxs_i64(idxs, n_lines, n_frames);
    if (rc) return rc;
    for (long i = 0; i < n_lines; i++) {
        const double *l = lines + 5 * i;
        adj_ctx c = {grad_img, H, W, frame_of(i, n_frames, idxs, 1), kernel, max_val,
                     l[0], l[1], l[2], l[3], l[4], {0, 0, 0, 0, 0}};
        if (lines_f32) rc = walk_f32(lines_f32 + 5 * i, adj_px_f32, &c);
        else rc = walk_f64(l, adj_px_f64, &c);
        if (rc) return rc;
        for (int j = 0; j < 5; j++) grad_lines[5 * i + j] = c.acc[j];
    }
    return CBXO_OK;
}

/* ------------------------------------------------------------------------ */
/* label: connected components of equal-valued truthy pixels under the      */
/* structure's shifts, ids in raster order of each component's minimum flat */
/* index, components with fewer than npts pixels dropped and the rest       */
/* renumbered consecutively.  Same partition and numbering as label<I>      */
/* (label.cpp:109-175) + LabelResult.to_array (label.cpp:459-493), which    */
/* keeps a region at its minimum index and concatenates chunks in order.    */
/* Union-find with the smaller index as the root.                           */
/* Pinned by: scipy.ndimage.label (identical numbering), brute-force BFS on */
/* tiny masks, and the compiled reference label module.                     */
/* dtype: 0 = bool/uint8, 1 = int32, 2 = int64.                             */
/* ------------------------------------------------------------------------ */

static long uf_find(long *p, long i)
{
    long r = i;
    while (p[r] != r) r = p[r];
    while (p[i] != r) { long n = p[i]; p[i] = r; i = n; }
    return r;
}

static long val_at(const void *inp, int dtype, long i)
{
    if (dtype == 0) return ((const uint8_t *)inp)[i];
    if (dtype == 1) return ((const int32_t *)inp)[i];
    return (long)((const int64_t *)inp)[i];
}

long cbxo_label(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts, long n_shifts,
                long npts, int32_t *out)
{
    if (ndim < 1 || ndim > 16) return -CBXO_EVALUE;
    long size = 1;
    for (int d = 0; d < ndim; d++) size *= (long)shape[d];
    if (size == 0) return 0;
    long *par = (long *)malloc(sizeof(long) * (size_t)size);
    long *cnt = (long *)calloc((size_t)size, sizeof(long));
    if (!par || !cnt) { free(par); free(cnt); return -CBXO_ERUNTIME; }
    for (long i = 0; i < size; i++) par[i] = val_at(inp, dtype, i) ? i : -1;
    long coord[16];
    for (long i = 0; i < size; i++) {
        if (par[i] < 0) continue;
        long rem = i;
        for (int d = ndim - 1; d >= 0; d--) { coord[d] = rem % shape[d]; rem /= shape[d]; }
        long vi = val_at(inp, dtype, i);
        for (long s = 0; s < n_shifts; s++) {
            const int64_t *sh = shifts + s * ndim;
            if (g_plant == 4) {
                long l1 = 0;
                for (int d = 0; d < ndim; d++) l1 += labs((long)sh[d]);
                if (l1 > 1) continue;
            }
            long k = 0, ok = 1;
            for (int d = 0; d < ndim; d++) {
                long c = coord[d] + (long)sh[d];
                if (c < 0 || c >= shape[d]) { ok = 0; break; }
                k = k * (long)shape[d] + c;
            }
            if (!ok || par[k] < 0 || val_at(inp, dtype, k) != vi) continue;
            long a = uf_find(par, i), b = uf_find(par, k);
            if (a == b) continue;
            if (a < b) par[b] = a; else par[a] = b;
        }
    }
    for (long i = 0; i < size; i++) if (par[i] >= 0) cnt[uf_find(par, i)]++;
    long n = 0;
    for (long i = 0; i < size; i++) {
        if (par[i] == i) cnt[i] = (cnt[i] >= npts) ? ++n : 0;
    }
    for (long i = 0; i < size; i++) out[i] = (par[i] >= 0) ? (int32_t)cnt[uf_find(par, i)] : 0;
    free(par); free(cnt);
    return n;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: the reference's multi-threaded CPU path restated with the  */
/* same parallel structure (timed beside the GPU; not a parity checker).    */
/*                                                                          */
/* draw_lines (bresenham.cpp:62-94): an OpenMP team; each thread walks its  */
/* share of the lines (`omp for nowait`, static schedule) into a private    */
/* buffer of (flat index, frame, value) entries (the 24-byte tuples of      */
/* ImageBuffer, bresenham.hpp:10-34), then merges the buffer into `out`     */
/* inside one `omp critical` section.  With one thread the result is        */
/* cbxo_draw_lines_*'s exactly (same order); with more, sums are in the     */
/* reference's nondeterministic merge order.                                */
/* ------------------------------------------------------------------------ */
#include <omp.h>

typedef struct { uint64_t idx, frame; float v; } ent_f32;   /* std::tuple<size_t, size_t, float> */
typedef struct { uint64_t idx, frame; double v; } ent_f64;

typedef struct {
    void *a; long n, cap; size_t esz;
    long H, W, frame; int kernel; double max_val;
} omp_buf;

static int buf_grow(omp_buf *b)
{
    long cap = b->cap ? 2 * b->cap : 4096;
    void *p = realloc(b->a, (size_t)cap * b->esz);
    if (!p) return 0;
    b->a = p;
    b->cap = cap;
    return 1;
}

static void omp_px_f32(void *vctx, long x, long y, float err)
{
    omp_buf *b = (omp_buf *)vctx;
    if (!(err <= 1.0f) || y < 0 || y >= b->H || x < 0 || x >= b->W) return;
    if (b->n == b->cap && !buf_grow(b)) return;
    ent_f32 *e = (ent_f32 *)b->a + b->n++;
    e->idx = (uint64_t)(y * b->W + x);
    e->frame = (uint64_t)b->frame;
    e->v = (float)b->max_val * prof_f32(b->kernel, sqrtf(err));
}

static void omp_px_f64(void *vctx, long x, long y, double err)
{
    omp_buf *b = (omp_buf *)vctx;
    if (!(err <= 1.0) || y < 0 || y >= b->H || x < 0 || x >= b->W) return;
    if (b->n == b->cap && !buf_grow(b)) return;
    ent_f64 *e = (ent_f64 *)b->a + b->n++;
    e->idx = (uint64_t)(y * b->W + x);
    e->frame = (uint64_t)b->frame;
    e->v = b->max_val * prof_f64(b->kernel, sqrt(err));
}

int cbxo_draw_lines_omp_f32(float *out, long n_frames, long H, long W, const float *lines, long n_lines,
                            const int32_t *idxs, float max_val, int kernel, int overlap, int threads)
{
    if (kernel < 0 || kernel > 4 || overlap < 0 || overlap > 1 || threads < 1) return CBXO_EVALUE;
    int rc = check_idxs_i32(idxs, n_lines, n_frames);
    if (rc) return rc;
    int rc_all = CBXO_OK;
    const long HW = H * W;
#pragma omp parallel num_threads(threads)
    {
        omp_buf b = {NULL, 0, 0, sizeof(ent_f32), H, W, 0, kernel, max_val};
#pragma omp for nowait
        for (long i = 0; i < n_lines; i++) {
            b.frame = frame_of(i, n_frames, idxs, 0);
            int r = walk_f32(lines + 5 * i, omp_px_f32, &b);
            if (r) {
#pragma omp atomic write
                rc_all = r;
            }
        }
#pragma omp critical
        {
            const ent_f32 *e = (const ent_f32 *)b.a;
            for (long k = 0; k < b.n; k++) {
                float *o = out + e[k].idx + HW * e[k].frame;
                if (overlap) *o = (*o < e[k].v) ? e[k].v : *o;
                else *o = *o + e[k].v;
            }
        }
        free(b.a);
    }
    return rc_all;
}

int cbxo_draw_lines_omp_f64(double *out, long n_frames, long H, long W, const double *lines, long n_lines,
                            const int64_t *idxs, double max_val, int kernel, int overlap, int threads)
{
    if (kernel < 0 || kernel > 4 || overlap < 0 || overlap > 1 || threads < 1) return CBXO_EVALUE;
    int rc = check_idxs_i64(idxs, n_lines, n_frames);
    if (rc) return rc;
    int rc_all = CBXO_OK;
    const long HW = H * W;
#pragma omp parallel num_threads(threads)
    {
        omp_buf b = {NULL, 0, 0, sizeof(ent_f64), H, W, 0, kernel, max_val};
#pragma omp for nowait
        for (long i = 0; i < n_lines; i++) {
            b.frame = frame_of(i, n_frames, idxs, 1);
            int r = walk_f64(lines + 5 * i, omp_px_f64, &b);
            if (r) {
#pragma omp atomic write
                rc_all = r;
            }
        }
#pragma omp critical
        {
            const ent_f64 *e = (const ent_f64 *)b.a;
            for (long k = 0; k < b.n; k++) {
                double *o = out + e[k].idx + HW * e[k].frame;
                if (overlap) *o = (*o < e[k].v) ? e[k].v : *o;
                else *o = *o + e[k].v;
            }
        }
        free(b.a);
    }
    return rc_all;
}

/* ------------------------------------------------------------------------ */
/* label (label.cpp:109-175): the flat input is cut into one chunk per      */
/* thread; each thread grows a region (Region::dilate, label.hpp:396-425:   */
/* breadth-first over the structure's shifts, keeping neighbours equal to   */
/* the seed's value) from every unvisited truthy pixel of its chunk, marks  */
/* the region's pixels of its chunk visited, and keeps the region when it   */
/* has >= npts pixels and its smallest index lies in the chunk.  Regions    */
/* are concatenated thread by thread, so ids follow the raster order of the */
/* smallest index; the map is then written as LabelResult.to_array does     */
/* (label.cpp:459-493).  Region membership uses a per-thread stamp array    */
/* instead of the reference's std::set.                                     */
/* ------------------------------------------------------------------------ */

typedef struct { int32_t *px; long n, cap; long *roff; long nr, rcap; } region_list;

static int push_long(long **a, long *n, long *cap, long v)
{
    if (*n == *cap) {
        long c = *cap ? 2 * *cap : 1024;
        long *p = (long *)realloc(*a, (size_t)c * sizeof(long));
        if (!p) return 0;
        *a = p;
        *cap = c;
    }
    (*a)[(*n)++] = v;
    return 1;
}

static int push_i32(int32_t **a, long *n, long *cap, int32_t v)
{
    if (*n == *cap) {
        long c = *cap ? 2 * *cap : 4096;
        int32_t *p = (int32_t *)realloc(*a, (size_t)c * sizeof(int32_t));
        if (!p) return 0;
        *a = p;
        *cap = c;
    }
    (*a)[(*n)++] = v;
    return 1;
}

long cbxo_label_omp(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts, long n_shifts,
                    long npts, int32_t *out, int threads)
{
    if (ndim < 1 || ndim > 16 || threads < 1) return -CBXO_EVALUE;
    long size = 1;
    for (int d = 0; d < ndim; d++) size *= (long)shape[d];
    if (size >= ((long)1 << 31)) return -CBXO_EVALUE;
    memset(out, 0, sizeof(int32_t) * (size_t)size);
    if (size == 0) return 0;
    region_list *rl = (region_list *)calloc((size_t)threads, sizeof(region_list));
    if (!rl) return -CBXO_ERUNTIME;
    int failed = 0;
#pragma omp parallel num_threads(threads)
    {
        const int tid = omp_get_thread_num();
        const long chunk = (size + threads - 1) / threads;
        const long t0 = tid * chunk, t1 = (t0 + chunk < size) ? t0 + chunk : size;
        region_list *me = &rl[tid];
        uint8_t *visited = (t1 > t0) ? (uint8_t *)calloc((size_t)(t1 - t0), 1) : NULL;
        uint32_t *stamp = (uint32_t *)calloc((size_t)size, sizeof(uint32_t));
        int32_t *reg = NULL;
        long rn = 0, rcap = 0;
        uint32_t cur = 0;
        long coord[16];
        if ((t1 > t0 && !visited) || !stamp) {
#pragma omp atomic write
            failed = 1;
        } else {
            for (long index = t0; index < t1; index++) {
                if (visited[index - t0] || !val_at(inp, dtype, index)) continue;
                const long root_v = val_at(inp, dtype, index);
                cur++;
                rn = 0;
                stamp[index] = cur;
                push_i32(&reg, &rn, &rcap, (int32_t)index);
                for (long q = 0; q < rn; q++) {
                    long rem = reg[q];
                    for (int d = ndim - 1; d >= 0; d--) { coord[d] = rem % shape[d]; rem /= shape[d]; }
                    for (long s = 0; s < n_shifts; s++) {
                        const int64_t *sh = shifts + s * ndim;
                        long k = 0, ok = 1;
                        for (int d = 0; d < ndim; d++) {
                            long c = coord[d] + (long)sh[d];
                            if (c < 0 || c >= shape[d]) { ok = 0; break; }
                            k = k * (long)shape[d] + c;
                        }
                        if (!ok || stamp[k] == cur || val_at(inp, dtype, k) != root_v) continue;
                        stamp[k] = cur;
                        push_i32(&reg, &rn, &rcap, (int32_t)k);
                    }
                }
                long mn = reg[0];
                for (long q = 0; q < rn; q++) {
                    const long p = reg[q];
                    if (p >= t0 && p < t1) visited[p - t0] = 1;
                    if (p < mn) mn = p;
                }
                if (rn >= npts && mn >= t0 && mn < t1) {
                    push_long(&me->roff, &me->nr, &me->rcap, me->n);
                    for (long q = 0; q < rn; q++) push_i32(&me->px, &me->n, &me->cap, reg[q]);
                }
            }
        }
        free(reg);
        free(stamp);
        free(visited);
    }
    long n = 0;
    if (!failed) {
        long *first = (long *)malloc(sizeof(long) * (size_t)threads);
        for (int t = 0; t < threads; t++) { first[t] = n; n += rl[t].nr; }
#pragma omp parallel for num_threads(threads) schedule(static, 1)
        for (int t = 0; t < threads; t++) {
            const region_list *r = &rl[t];
            for (long j = 0; j < r->nr; j++) {
                const long a = r->roff[j], b = (j + 1 < r->nr) ? r->roff[j + 1] : r->n;
                for (long q = a; q < b; q++) out[r->px[q]] = (int32_t)(first[t] + j + 1);
            }
        }
        free(first);
    }
    for (int t = 0; t < threads; t++) { free(rl[t].px); free(rl[t].roff); }
    free(rl);
    return failed ? -CBXO_ERUNTIME : n;
}

/* ------------------------------------------------------------------------ */
/* accumulate_lines with one thread (bresenham.cpp:151-252): lines sorted by */
/* term (stably here: ascending line order inside a term, one of the orders */
/* the reference's unstable std::sort may give), terms visited in ascending */
/* order.  Each term is drawn into a zeroed buffer; a pixel's index is      */
/* recorded whenever it is drawn while its buffer value is still zero       */
/* (possibly more than once), the buffer is updated with in_overlap         */
/* (buffer + v, or std::max(buffer, v)), and after the term every recorded  */
/* index is combined into frame frames[term] with out_overlap (buffer + out */
/* or std::max(out, buffer)) and its buffer value reset to zero.            */
/* Pinned by: the compiled reference (tests/test_accumulate.py) and the     */
/* independent Python formulation oracle.accumulate_lines_py (a touched     */
/* mask instead of the index list).                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    void *buf; long H, W; int kernel, in_ov; double max_val;
    long *idx; long n, cap; int fail;
    long D;   /* depth of 3D frames (1 for 2D) */
} acc_ctx;

static void acc_push(acc_ctx *c, long i)
{
    if (c->n == c->cap) {
        long cap = c->cap ? 2 * c->cap : 4096;
        long *p = (long *)realloc(c->idx, (size_t)cap * sizeof(long));
        if (!p) { c->fail = 1; return; }
        c->idx = p;
        c->cap = cap;
    }
    c->idx[c->n++] = i;
}

static void acc_px_f32(void *vctx, long x, long y, float err)
{
    acc_ctx *c = (acc_ctx *)vctx;
    if (!(err <= 1.0f) || y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    const long i = y * c->W + x;
    float *b = (float *)c->buf + i;
    if (*b == 0.0f) acc_push(c, i);
    const float v = (float)c->max_val * prof_f32(c->kernel, sqrtf(err));
    if (c->in_ov) *b = (*b < v) ? v : *b;
    else *b = *b + v;
}

static void acc_px_f64(void *vctx, long x, long y, double err)
{
    acc_ctx *c = (acc_ctx *)vctx;
    if (!(err <= 1.0) || y < 0 || y >= c->H || x < 0 || x >= c->W) return;
    const long i = y * c->W + x;
    double *b = (double *)c->buf + i;
    if (*b == 0.0) acc_push(c, i);
    const double v = c->max_val * prof_f64(c->kernel, sqrt(err));
    if (c->in_ov) *b = (*b < v) ? v : *b;
    else *b = *b + v;
}

/* order[]: the lines stably sorted by term; toff[t]..toff[t+1] the lines of term t. */
static int term_order(const void *terms, int wide, long n_lines, long n_terms, long **order, long **toff)
{
    *order = (long *)malloc(sizeof(long) * (size_t)(n_lines ? n_lines : 1));
    *toff = (long *)calloc((size_t)n_terms + 2, sizeof(long));
    if (!*order || !*toff) return CBXO_ERUNTIME;
    for (long i = 0; i < n_lines; i++) {
        long t = wide ? (long)((const int64_t *)terms)[i] : (long)((const int32_t *)terms)[i];
        if (t < 0 || t >= n_terms) return CBXO_EINDEX;
        (*toff)[t + 2]++;
    }
    for (long t = 0; t < n_terms; t++) (*toff)[t + 2] += (*toff)[t + 1];
    for (long i = 0; i < n_lines; i++) {
        long t = wide ? (long)((const int64_t *)terms)[i] : (long)((const int32_t *)terms)[i];
        (*order)[(*toff)[t + 1]++] = i;
    }
    return CBXO_OK;
}

#define ACCUMULATE_BODY(REAL, WALK, PX, FR_T) ACCUMULATE_ND(REAL, WALK, PX, FR_T, 5, 1)

/* ROW: values per line (5 or 7); DEPTH: frame depth (1 for 2D frames) */
#define ACCUMULATE_ND(REAL, WALK, PX, FR_T, ROW, DEPTH)                                             \
    if (kernel < 0 || kernel > 4 || in_ov < 0 || in_ov > 1 || out_ov < 0 || out_ov > 1)             \
        return CBXO_EVALUE;                                                                         \
    for (long t = 0; t < n_terms; t++)                                                              \
        if (frames[t] < 0 || frames[t] >= n_frames) return CBXO_EINDEX;                             \
    long *order = NULL, *toff = NULL;                                                               \
    int rc = term_order(terms, sizeof(FR_T) == 8, n_lines, n_terms, &order, &toff);                 \
    const long fsz = (long)(DEPTH) * H * W;                                                         \
    REAL *buf = (rc == CBXO_OK) ? (REAL *)calloc((size_t)(fsz ? fsz : 1), sizeof(REAL)) : NULL;     \
    if (rc == CBXO_OK && !buf) rc = CBXO_ERUNTIME;                                                  \
    acc_ctx c = {buf, H, W, kernel, in_ov, max_val, NULL, 0, 0, 0, (long)(DEPTH)};                  \
    for (long t = 0; rc == CBXO_OK && t < n_terms; t++) {                                           \
        REAL *dst = out + (long)frames[t] * fsz;                                                    \
        for (long j = toff[t]; j < toff[t + 1] && rc == CBXO_OK; j++)                               \
            rc = WALK(lines + (ROW) * order[j], PX, &c);                                            \
        if (c.fail) rc = CBXO_ERUNTIME;                                                             \
        for (long k = 0; rc == CBXO_OK && k < c.n; k++) {                                           \
            const long i = c.idx[k];                                                                \
            if (out_ov) dst[i] = (dst[i] < buf[i]) ? buf[i] : dst[i];                               \
            else dst[i] = buf[i] + dst[i];                                                          \
            buf[i] = (REAL)0;                                                                       \
        }                                                                                           \
        c.n = 0;                                                                                    \
    }                                                                                               \
    free(c.idx); free(buf); free(order); free(toff);                                                \
    return rc;

int cbxo_accumulate_lines_f32(float *out, long n_frames, long H, long W, const float *lines, long n_lines,
                              const int32_t *terms, const int32_t *frames, long n_terms, float max_val, int kernel,
                              int in_ov, int out_ov)
{
    ACCUMULATE_BODY(float, walk_f32, acc_px_f32, int32_t)
}

int cbxo_accumulate_lines_f64(double *out, long n_frames, long H, long W, const double *lines, long n_lines,
                              const int64_t *terms, const int64_t *frames, long n_terms, double max_val, int kernel,
                              int in_ov, int out_ov)
{
    ACCUMULATE_BODY(double, walk_f64, acc_px_f64, int64_t)
}

/* ------------------------------------------------------------------------ */
/* Region moments (label.hpp:554-722 MomentsND / CentralMomentsND,          */
/* label.cpp:534-583 total_mass, mean, center_of_mass, moment_of_inertia,   */
/* covariance_matrix, line_fit) and p_values (label.cpp:228-255 with        */
/* LineND::distance, geometry.hpp:319-339, and detail::logbinom,            */
/* label.hpp:57-95).  A region is the set of pixels of a label map holding  */
/* index[j], visited in ascending flat order (the reference's Region is a   */
/* sorted set).  Arithmetic in T, in the reference's order: coordinates in  */
/* xyz order (x = last axis) relative to the region's first pixel, value =  */
/* std::max(value, 0), sums in T; outputs in zyx order like the reference.  */
/* what: 0 total_mass (1), 1 mean (N: raw first moment), 2 center_of_mass   */
/* (N), 3 moment_of_inertia (N*N raw), 4 covariance_matrix (N*N),           */
/* 5 line_fit (2N, xyz end points).                                         */
/* Pinned by: the compiled reference label module (golden fixtures and      */
/* random inputs, tests/test_moments.py).                                   */
/* ------------------------------------------------------------------------ */

#define MOM_MAXN 7

/* regions[j]: pixels of label index[j], ascending (CSR over a counting sort).
 * own[j]: the first entry of index with the same label, which holds the
 * pixels (later duplicates have empty ranges and read the owner's). */
static int region_lists(const int32_t *labels, long size, const int64_t *index, long n_index, long **off, long **pix,
                        long **own)
{
    int64_t maxl = 0;
    for (long j = 0; j < n_index; j++) if (index[j] > maxl) maxl = index[j];
    long *slot = (long *)malloc(sizeof(long) * (size_t)(maxl + 1));
    *off = (long *)calloc((size_t)n_index + 1, sizeof(long));
    *own = (long *)malloc(sizeof(long) * (size_t)(n_index ? n_index : 1));
    if (!slot || !*off || !*own) { free(slot); return CBXO_ERUNTIME; }
    for (long l = 0; l <= maxl; l++) slot[l] = -1;
    for (long j = n_index - 1; j >= 0; j--) if (index[j] > 0) slot[index[j]] = j;  /* first occurrence wins */
    for (long j = 0; j < n_index; j++) (*own)[j] = (index[j] > 0) ? slot[index[j]] : j;
    for (long i = 0; i < size; i++) {
        long l = labels[i];
        if (l > 0 && l <= maxl && slot[l] >= 0) (*off)[slot[l] + 1]++;
    }
    for (long j = 0; j < n_index; j++) (*off)[j + 1] += (*off)[j];
    *pix = (long *)malloc(sizeof(long) * (size_t)((*off)[n_index] ? (*off)[n_index] : 1));
    long *cur = (long *)malloc(sizeof(long) * (size_t)(n_index ? n_index : 1));
    if (!*pix || !cur) { free(slot); free(cur); return CBXO_ERUNTIME; }
    for (long j = 0; j < n_index; j++) cur[j] = (*off)[j];
    for (long i = 0; i < size; i++) {
        long l = labels[i];
        if (l > 0 && l <= maxl && slot[l] >= 0) (*pix)[cur[slot[l]]++] = i;
    }
    free(slot); free(cur);
    return CBXO_OK;
}

static void point_of(long idx, int ndim, const int64_t *shape, long *xyz)
{
    for (int n = 0; n < ndim; n++) {           /* xyz[0] = last axis */
        long d = shape[ndim - 1 - n];
        xyz[n] = idx % d;
        idx /= d;
    }
}

#define MOMENTS_BODY(REAL, SQRT, ATAN2, COS, SIN)                                                       \
    const int N = ndim;                                                                                 \
    if (N < 1 || N > MOM_MAXN || what < 0 || what > 5 || (what == 5 && N < 2)) return CBXO_EVALUE;    \
    long size = 1;                                                                                      \
    for (int d = 0; d < N; d++) size *= (long)shape[d];                                                 \
    long *off = NULL, *pix = NULL, *own = NULL;                                                         \
    int rc = region_lists(labels, size, index, n_index, &off, &pix, &own);                             \
    if (rc) { free(off); free(pix); free(own); return rc; }                                             \
    const int NP = N * (N - 1) / 2;                                                                     \
    int pi[MOM_MAXN * MOM_MAXN], pj[MOM_MAXN * MOM_MAXN];                                               \
    { int a = 0, b = 0; for (int n = 0; n < NP; n++) { if (++b == N) { ++a; b = a + 1; } pi[n] = a; pj[n] = b; } } \
    const int M = (what == 0) ? 1 : (what <= 2) ? N : (what <= 4) ? N * N : 2 * N;                   \
    for (long j = 0; j < n_index; j++) {                                                                \
        const long s = own[j];                                                                          \
        const long a0 = off[s], a1 = off[s + 1];                                                        \
        REAL org[MOM_MAXN] = {0}, mu = 0, mx[MOM_MAXN] = {0}, mxx[MOM_MAXN] = {0}, mxy[MOM_MAXN * MOM_MAXN] = {0}; \
        long p[MOM_MAXN];                                                                               \
        if (a1 > a0) {                                                                                  \
            point_of(pix[a0], N, shape, p);                                                             \
            for (int n = 0; n < N; n++) org[n] = (REAL)p[n];                                            \
        }                                                                                               \
        for (long q = a0; q < a1; q++) {                                                                \
            point_of(pix[q], N, shape, p);                                                              \
            REAL r[MOM_MAXN];                                                                           \
            for (int n = 0; n < N; n++) r[n] = (REAL)p[n] - org[n];                                     \
            REAL v = data[pix[q]];                                                                      \
            v = (v < (REAL)0) ? (REAL)0 : v;                /* std::max(value, T()) */                 \
            mu += v;                                                                                    \
            for (int n = 0; n < N; n++) mx[n] += r[n] * v;                                              \
            for (int n = 0; n < N; n++) mxx[n] += r[n] * r[n] * v;                                      \
            for (int n = 0; n < NP; n++) mxy[n] += r[pi[n]] * r[pj[n]] * v;                             \
        }                                                                                               \
        REAL *o = out + j * M;                                                                          \
        if (what == 0) { o[0] = mu; continue; }                                                         \
        if (what == 1) { for (int n = 0; n < N; n++) o[N - 1 - n] = mx[n] + org[n] * mu; continue; }   \
        if (what == 3) {                                                                                \
            for (int n = 0; n < N; n++) { int z = N - n - 1;                                            \
                o[n + N * n] = mxx[z] + 2 * org[z] * mx[z] + org[z] * org[z] * mu; }                    \
            for (int n = 0; n < NP; n++) { int a = N - pi[n] - 1, b = N - pj[n] - 1;                    \
                REAL m = mxy[n] + org[a] * mx[b] + org[b] * mx[a] + org[a] * org[b] * mu;               \
                o[a + N * b] = m; o[b + N * a] = m; }                                                   \
            continue;                                                                                   \
        }                                                                                               \
        /* central moments (CentralMomentsND): zero when the mass is zero */                          \
        REAL cx[MOM_MAXN] = {0}, cxx[MOM_MAXN] = {0}, cxy[MOM_MAXN * MOM_MAXN] = {0};                    \
        if (mu != (REAL)0) {                                                                            \
            for (int n = 0; n < N; n++) cx[n] = mx[n] / mu;                                             \
            for (int n = 0; n < N; n++) cxx[n] = mxx[n] / mu - cx[n] * cx[n];                           \
            for (int n = 0; n < NP; n++) cxy[n] = mxy[n] / mu - cx[pi[n]] * cx[pj[n]];                  \
        }                                                                                               \
        if (what == 2) { for (int n = 0; n < N; n++) o[N - 1 - n] = cx[n] + org[n]; continue; }        \
        if (what == 4) {                                                                                \
            for (int n = 0; n < N; n++) o[n + N * n] = cxx[N - n - 1];                                  \
            for (int n = 0; n < NP; n++) { int a = N - pi[n] - 1, b = N - pj[n] - 1;                    \
                o[a + N * b] = cxy[n]; o[b + N * a] = cxy[n]; }                                         \
            continue;                                                                                   \
        }                                                                                               \
        /* line(): theta = 0.5 atan2(2 mu_xy, mu_xx - mu_yy) (a double product) */                     \
        REAL angle = (REAL)(0.5 * (double)ATAN2(2 * cxy[0], cxx[0] - cxx[1]));                          \
        REAL c0[MOM_MAXN];                                                                              \
        for (int n = 0; n < N; n++) c0[n] = cx[n] + org[n];                                             \
        if (angle != angle) { for (int n = 0; n < N; n++) { o[n] = c0[n]; o[N + n] = c0[n]; } continue; } \
        REAL tau[MOM_MAXN] = {0};                                                                       \
        tau[0] = COS(angle); tau[1] = SIN(angle);                                                       \
        REAL dl = SQRT(4 * cxy[0] * cxy[0] + (cxx[0] - cxx[1]) * (cxx[0] - cxx[1]));                    \
        REAL hw = (REAL)sqrt(2 * log(2.0) * (double)(cxx[0] + cxx[1] + dl));                            \
        for (int n = 0; n < N; n++) { o[n] = c0[n] + hw * tau[n]; o[N + n] = c0[n] - hw * tau[n]; }     \
    }                                                                                                   \
    free(off); free(pix); free(own);                                                                    \
    return CBXO_OK;

int cbxo_region_moments_f32(const int32_t *labels, const float *data, int ndim, const int64_t *shape,
                            const int64_t *index, long n_index, int what, float *out)
{
    MOMENTS_BODY(float, sqrtf, atan2f, cosf, sinf)
}

int cbxo_region_moments_f64(const int32_t *labels, const double *data, int ndim, const int64_t *shape,
                            const int64_t *index, long n_index, int what, double *out)
{
    MOMENTS_BODY(double, sqrt, atan2, cos, sin)
}

/* detail::logbinom (label.hpp:57-95): log of the binomial upper tail in T. */
static double logaddexp_f64(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    double m = (a > b) ? a : b;
    return m + log(exp(a - m) + exp(b - m));
}

static float logaddexp_f32(float a, float b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    float m = (a > b) ? a : b;
    return m + logf(expf(a - m) + expf(b - m));
}

/* The integer-argument calls (std::lgamma(n + 1), std::log(i)) take the
 * double overloads whatever T is; k * log_p is computed in T; sums of a
 * double and a T are double, assigned back to T. */
#define LOGBINOM_BODY(REAL, LOG, LOG1P, ADDEXP)                                                        \
    if (k <= 0) return (REAL)0;                                                                         \
    if (k > n) return -INFINITY;                                                                        \
    if (p <= (REAL)0) return (k <= 0) ? (REAL)0 : -INFINITY;                                            \
    if (p >= (REAL)1) return (k <= n) ? (REAL)0 : -INFINITY;                                            \
    REAL log_p = LOG(p), log_q = LOG1P(-p);                                                             \
    REAL t = (REAL)(lgamma((double)(n + 1)) - lgamma((double)(k + 1)) - lgamma((double)(n - k + 1)) +  \
                    (double)((REAL)k * log_p) + (double)((REAL)(n - k) * log_q));                       \
    REAL tail = t;                                                                                      \
    for (long i = k + 1; i < n + 1; ++i) {                                                              \
        t = (REAL)((double)t + (log((double)(n - i + 1)) - log((double)i) + (double)log_p - (double)log_q)); \
        tail = ADDEXP(tail, t);                                                                         \
    }                                                                                                   \
    return tail;

static double logbinom_f64(long n, long k, double p) { LOGBINOM_BODY(double, log, log1p, logaddexp_f64) }
static float logbinom_f32(long n, long k, float p) { LOGBINOM_BODY(float, logf, log1pf, logaddexp_f32) }

double cbxo_logbinom_f64(long n, long k, double p) { return logbinom_f64(n, k, p); }
float cbxo_logbinom_f32(long n, long k, float p) { return logbinom_f32(n, k, p); }

/* LineND<T, N>::distance(point) (geometry.hpp:319-339) with an integer point. */
#define SEGDIST_BODY(REAL, SQRT)                                                                        \
    REAL tau[MOM_MAXN], ctr[MOM_MAXN], mag = 0;                                                         \
    for (int n = 0; n < N; n++) tau[n] = ln[N + n] - ln[n];                                             \
    for (int n = 0; n < N; n++) mag += tau[n] * tau[n];                                                 \
    REAL q[MOM_MAXN];                                                                                   \
    if (mag != (REAL)0) {                                                                               \
        for (int n = 0; n < N; n++) ctr[n] = (ln[n] + ln[N + n]) / (REAL)2;                             \
        REAL d = 0;                                                                                     \
        for (int n = 0; n < N; n++) d += tau[n] * ((REAL)p[n] - ctr[n]);                                \
        REAL rt = d / mag;                                                                              \
        rt = (rt < (REAL)-0.5) ? (REAL)-0.5 : ((rt > (REAL)0.5) ? (REAL)0.5 : rt);                      \
        for (int n = 0; n < N; n++) q[n] = rt * tau[n] + ctr[n];                                        \
    } else {                                                                                            \
        for (int n = 0; n < N; n++) q[n] = ln[n];                                                       \
    }                                                                                                   \
    REAL m2 = 0;                                                                                        \
    for (int n = 0; n < N; n++) m2 += ((REAL)p[n] - q[n]) * ((REAL)p[n] - q[n]);                        \
    return SQRT(m2);

static double segdist_f64(const double *ln, const long *p, int N) { SEGDIST_BODY(double, sqrt) }
static float segdist_f32(const float *ln, const long *p, int N) { SEGDIST_BODY(float, sqrtf) }

#define PVALUES_BODY(REAL, DIST, LOGBINOM)                                                              \
    const int N = ndim;                                                                                 \
    if (N < 1 || N > MOM_MAXN) return CBXO_EVALUE;                                                      \
    long size = 1;                                                                                      \
    for (int d = 0; d < N; d++) size *= (long)shape[d];                                                 \
    long *off = NULL, *pix = NULL, *own = NULL;                                                         \
    int rc = region_lists(labels, size, index, n_index, &off, &pix, &own);                             \
    if (rc) { free(off); free(pix); free(own); return rc; }                                             \
    for (long j = 0; j < n_index; j++) {                                                                \
        const long s = own[j];                                                                          \
        long n = 0, k = 0, p[MOM_MAXN];                                                                 \
        for (long q = off[s]; q < off[s + 1]; q++) {                                                    \
            point_of(pix[q], N, shape, p);                                                              \
            if (DIST(lines + 2 * N * j, p, N) < xtol) {                                                 \
                n++;                                                                                    \
                if (data[pix[q]] >= vmin) k++;                                                          \
            }                                                                                           \
        }                                                                                               \
        out[j] = LOGBINOM(n, k, p0);                                                                    \
        if (counts) { counts[2 * j] = n; counts[2 * j + 1] = k; }                                       \
    }                                                                                                   \
    free(off); free(pix); free(own);                                                                    \
    return CBXO_OK;

int cbxo_p_values_f32(const int32_t *labels, const float *data, int ndim, const int64_t *shape, const int64_t *index,
                      long n_index, const float *lines, float p0, float vmin, float xtol, float *out, long *counts)
{
    PVALUES_BODY(float, segdist_f32, logbinom_f32)
}

int cbxo_p_values_f64(const int32_t *labels, const double *data, int ndim, const int64_t *shape,
                      const int64_t *index, long n_index, const double *lines, double p0, double vmin, double xtol,
                      double *out, long *counts)
{
    PVALUES_BODY(double, segdist_f64, logbinom_f64)
}

/* ------------------------------------------------------------------------ */
/* 3D lines (rows of 7: x0 y0 z0 x1 y1 z1 w): draw_line_3d                  */
/* (bresenham.hpp:389-430) on BresenhamPlotter<T, 3> and LineIterator<T, 3>  */
/* (118-348): a tangent error and one normal error per axis pair            */
/* ((0,1), (0,2), (1,2): UniquePairs, geometry.hpp:404-437), the 3-axis      */
/* Bresenham step, and at every long-axis step the orthogonal plane filled  */
/* in four quadrants by a 2D walk with single-axis rays.  Same arithmetic   */
/* contract as the 2D walk (T increments, pow squares in double).           */
/* Pinned by: the compiled reference bresenham module on 3D inputs (golden  */
/* fixtures tests/golden/draw3_cases.npz and random inputs) and the brute-   */
/* force distance-to-segment image of the reference's own test               */
/* (tests/test_draw3.py).                                                    */
/* ------------------------------------------------------------------------ */

#define W3_BODY(REAL, ROUND, CEIL, FABS, SQRT)                                                          \
    typedef struct { long step[3], cur[3]; int next[3]; REAL te, ne[3]; } it3;                          \
    static const int PA[3] = {0, 0, 1}, PB[3] = {1, 2, 2};                                               \
    REAL p0[3] = {ln[0], ln[1], ln[2]}, p1[3] = {ln[3], ln[4], ln[5]}, w = ln[6];                       \
    REAL t[3], nrm[3][3];                                                                               \
    for (int i = 0; i < 3; i++) t[i] = p1[i] - p0[i];                                                   \
    for (int q = 0; q < 3; q++) {                                                                       \
        for (int i = 0; i < 3; i++) nrm[q][i] = (REAL)0;                                                \
        nrm[q][PA[q]] = p1[PB[q]] - p0[PB[q]];                                                          \
        nrm[q][PB[q]] = p0[PA[q]] - p1[PA[q]];                                                          \
    }                                                                                                   \
    if (!(w > (REAL)0)) return CBXO_OK;                                                                 \
    REAL mag = (REAL)0;                                                                                 \
    for (int i = 0; i < 3; i++) mag += t[i] * t[i];                                                     \
    const REAL len = SQRT(mag);                                                                         \
    if (!(len > (REAL)0) || !isfinite(len) || !isfinite(w)) return CBXO_OK;                            \
    for (int i = 0; i < 3; i++) if (!isfinite(p0[i]) || !isfinite(p1[i])) return CBXO_OK;              \
    int axis = 0;                                                                                       \
    for (int i = 1; i < 3; i++) if (FABS(t[axis]) < FABS(t[i])) axis = i;   /* first max |t| */         \
    const long offset = (long)(CEIL(w) + (REAL)1);                                                      \
    REAL tau[3];                                                                                        \
    for (int i = 0; i < 3; i++) tau[i] = t[i] / len;                                                    \
    const REAL sc = FABS((REAL)offset / tau[axis]);                                                     \
    long step[3], m0[3], m1[3];                                                                         \
    for (int i = 0; i < 3; i++) {                                                                       \
        step[i] = (p1[i] >= p0[i]) ? 1 : -1;                                                            \
        REAL a0 = ROUND(p0[i] - sc * tau[i]), a1 = ROUND(p1[i] + sc * tau[i]);                          \
        if (!(FABS(a0) < (REAL)1e9) || !(FABS(a1) < (REAL)1e9)) return CBXO_EVALUE;                     \
        m0[i] = (long)a0;                                                                               \
        m1[i] = (long)a1 + step[i];                                                                     \
    }                                                                                                   \
    if (labs(m1[axis] - m0[axis]) > CBXO_MAX_WALK) return CBXO_EVALUE;                                  \
    const REAL w2 = w * w;                                                                              \
    /* iterator operations */                                                                           \
    void upd(it3 *it) {                                                                                 \
        for (int i = 0; i < 3; i++) it->next[i] = it->step[i] != 0;                                     \
        for (int q = 0; q < 3; q++) {                                                                   \
            const int a = PA[q], b = PB[q];                                                             \
            if (!(it->step[a] && it->step[b])) continue;                                                \
            REAL ex = it->ne[q] + (REAL)it->step[a] * nrm[q][a];                                        \
            REAL ey = it->ne[q] + (REAL)it->step[b] * nrm[q][b];                                        \
            REAL exy = ex + (REAL)it->step[b] * nrm[q][b];                                              \
            int f, s;                                                                                   \
            if (FABS(ey) < FABS(ex)) { if (FABS(exy) < FABS(ey)) { f = 1; s = 1; } else { f = 0; s = 1; } } \
            else if (FABS(exy) < FABS(ex)) { f = 1; s = 1; }                                            \
            else { f = 1; s = 0; }                                                                      \
            it->next[a] &= f;                                                                           \
            it->next[b] &= s;                                                                           \
        }                                                                                               \
    }                                                                                                   \
    void inc(it3 *it, long x, int ax) {                                                                 \
        it->cur[ax] += x;                                                                               \
        it->te += (REAL)x * t[ax];                                                                      \
        for (int q = 0; q < 3; q++) if (PA[q] == ax || PB[q] == ax) it->ne[q] += (REAL)x * nrm[q][ax];  \
    }                                                                                                   \
    void adv(it3 *it) {                                                                                 \
        for (int i = 0; i < 3; i++) if (it->next[i]) inc(it, it->step[i], i);                           \
        upd(it);                                                                                        \
    }                                                                                                   \
    REAL nerr(const it3 *it) {                                                                          \
        REAL e = (REAL)0;                                                                               \
        for (int q = 0; q < 3; q++) { double v = (double)(it->ne[q] / len); e = (REAL)((double)e + v * v); } \
        return e;                                                                                       \
    }                                                                                                   \
    REAL errf(const it3 *it) {                                                                          \
        REAL a = -it->te, b = it->te - len * len;                                                       \
        REAL m = (a < b) ? b : a;                                                                       \
        REAL at = (m < (REAL)0) ? (REAL)0 : m;                                                          \
        double qt = (double)(at / len);                                                                 \
        double e = qt * qt + (double)nerr(it);                                                          \
        return (REAL)(e / (double)w2);                                                                  \
    }                                                                                                   \
    it3 it;                                                                                             \
    for (int i = 0; i < 3; i++) { it.step[i] = step[i]; it.cur[i] = m0[i]; }                           \
    {                                                                                                   \
        REAL d[3];                                                                                      \
        for (int i = 0; i < 3; i++) d[i] = (REAL)m0[i] - p0[i];                                         \
        it.te = (REAL)0;                                                                                \
        for (int i = 0; i < 3; i++) it.te += t[i] * d[i];                                               \
        for (int q = 0; q < 3; q++) {                                                                   \
            it.ne[q] = (REAL)0;                                                                         \
            for (int i = 0; i < 3; i++) it.ne[q] += nrm[q][i] * d[i];                                   \
        }                                                                                               \
    }                                                                                                   \
    upd(&it);                                                                                           \
    const int ax1 = (axis + 1) % 3, ax2 = (axis + 2) % 3;                                               \
    while (!(it.cur[0] == m1[0] || it.cur[1] == m1[1] || it.cur[2] == m1[2])) {                        \
        if (it.next[axis]) {                                                                            \
            emit(ctx, it.cur, errf(&it));                                                               \
            for (int i = 0; i < 4; i++) {                                                               \
                it3 xy = it;                                                                            \
                xy.step[axis] = 0; upd(&xy);                          /* collapse(iter, axis) */        \
                if (i & 1) xy.step[ax1] *= -1;                                                          \
                if (i >> 1) xy.step[ax2] *= -1;                                                         \
                upd(&xy);                                             /* flip */                        \
                const int mv = (1 + ((i & 1) ^ (i >> 1)) == 1) ? ax1 : ax2;                             \
                inc(&xy, xy.step[mv], mv); upd(&xy);                  /* move */                        \
                for (; nerr(&xy) < w2; adv(&xy)) {                                                      \
                    emit(ctx, xy.cur, errf(&xy));                                                       \
                    for (int j = 1; j < 3; j++) {                                                       \
                        const int aj = (axis + j) % 3;                                                  \
                        if (!xy.next[aj]) continue;                                                     \
                        it3 x = xy;                                                                     \
                        x.step[aj] = 0; upd(&x);                      /* collapse(iter_xy, axis(j)) */  \
                        for (adv(&x); nerr(&x) < w2; adv(&x)) emit(ctx, x.cur, errf(&x));              \
                    }                                                                                   \
                }                                                                                       \
            }                                                                                           \
        }                                                                                               \
        adv(&it);                                                                                       \
    }                                                                                                   \
    return CBXO_OK;

typedef void (*emit3_f32_t)(void *ctx, const long *xyz, float err);
typedef void (*emit3_f64_t)(void *ctx, const long *xyz, double err);

static int walk3_f32(const float *ln, emit3_f32_t emit, void *ctx) { W3_BODY(float, roundf, ceilf, fabsf, sqrtf) }
static int walk3_f64(const double *ln, emit3_f64_t emit, void *ctx) { W3_BODY(double, round, ceil, fabs, sqrt) }

typedef struct { void *out; long D, H, W, frame; int kernel, overlap; double max_val; } draw3_ctx;

static void draw3_px_f32(void *vctx, const long *p, float err)
{
    draw3_ctx *c = (draw3_ctx *)vctx;
    if (!(err <= 1.0f) || p[0] < 0 || p[0] >= c->W || p[1] < 0 || p[1] >= c->H || p[2] < 0 || p[2] >= c->D) return;
    float v = (float)c->max_val * prof_f32(c->kernel, sqrtf(err));
    float *o = (float *)c->out + ((c->frame * c->D + p[2]) * c->H + p[1]) * c->W + p[0];
    if (c->overlap) *o = (*o < v) ? v : *o;
    else *o = *o + v;
}

static void draw3_px_f64(void *vctx, const long *p, double err)
{
    draw3_ctx *c = (draw3_ctx *)vctx;
    if (!(err <= 1.0) || p[0] < 0 || p[0] >= c->W || p[1] < 0 || p[1] >= c->H || p[2] < 0 || p[2] >= c->D) return;
    double v = c->max_val * prof_f64(c->kernel, sqrt(err));
    double *o = (double *)c->out + ((c->frame * c->D + p[2]) * c->H + p[1]) * c->W + p[0];
    if (c->overlap) *o = (*o < v) ? v : *o;
    else *o = *o + v;
}

/* draw_lines_nd_impl<T, I, 3> with one thread (bresenham.cpp:38-101): out is
 * n_frames volumes of D x H x W (z, y, x), lines n_lines rows of 7. */
int cbxo_draw_lines3_f32(float *out, long n_frames, long D, long H, long W, const float *lines, long n_lines,
                         const int32_t *idxs, float max_val, int kernel, int overlap)
{
    if (kernel < 0 || kernel > 4 || overlap < 0 || overlap > 1) return CBXO_EVALUE;
    int rc = check_idxs_i32(idxs, n_lines, n_frames);
    if (rc) return rc;
    draw3_ctx c = {out, D, H, W, 0, kernel, overlap, max_val};
    for (long i = 0; i < n_lines; i++) {
        c.frame = frame_of(i, n_frames, idxs, 0);
        if ((rc = walk3_f32(lines + 7 * i, draw3_px_f32, &c))) return rc;
    }
    return CBXO_OK;
}

int cbxo_draw_lines3_f64(double *out, long n_frames, long D, long H, long W, const double *lines, long n_lines,
                         const int64_t *idxs, double max_val, int kernel, int overlap)
{
    if (kernel < 0 || kernel > 4 || overlap < 0 || overlap > 1) return CBXO_EVALUE;
    int rc = check_idxs_i64(idxs, n_lines, n_frames);
    if (rc) return rc;
    draw3_ctx c = {out, D, H, W, 0, kernel, overlap, max_val};
    for (long i = 0; i < n_lines; i++) {
        c.frame = frame_of(i, n_frames, idxs, 1);
        if ((rc = walk3_f64(lines + 7 * i, draw3_px_f64, &c))) return rc;
    }
    return CBXO_OK;
}

/* accumulate_lines_nd_impl<T, I, 3, ...> with one thread (bresenham.cpp:151-252):
 * the 2D term-buffer procedure above on D x H x W frames and 3D walks.
 * Pinned by: the compiled reference on 3D inputs (tests/test_draw3.py). */
static void acc3_px_f32(void *vctx, const long *p, float err)
{
    acc_ctx *c = (acc_ctx *)vctx;
    if (!(err <= 1.0f) || p[0] < 0 || p[0] >= c->W || p[1] < 0 || p[1] >= c->H || p[2] < 0 || p[2] >= c->D) return;
    const long i = (p[2] * c->H + p[1]) * c->W + p[0];
    float *b = (float *)c->buf + i;
    if (*b == 0.0f) acc_push(c, i);
    const float v = (float)c->max_val * prof_f32(c->kernel, sqrtf(err));
    if (c->in_ov) *b = (*b < v) ? v : *b;
    else *b = *b + v;
}

static void acc3_px_f64(void *vctx, const long *p, double err)
{
    acc_ctx *c = (acc_ctx *)vctx;
    if (!(err <= 1.0) || p[0] < 0 || p[0] >= c->W || p[1] < 0 || p[1] >= c->H || p[2] < 0 || p[2] >= c->D) return;
    const long i = (p[2] * c->H + p[1]) * c->W + p[0];
    double *b = (double *)c->buf + i;
    if (*b == 0.0) acc_push(c, i);
    const double v = c->max_val * prof_f64(c->kernel, sqrt(err));
    if (c->in_ov) *b = (*b < v) ? v : *b;
    else *b = *b + v;
}

int cbxo_accumulate_lines3_f32(float *out, long n_frames, long D, long H, long W, const float *lines, long n_lines,
                               const int32_t *terms, const int32_t *frames, long n_terms, float max_val, int kernel,
                               int in_ov, int out_ov)
{
    ACCUMULATE_ND(float, walk3_f32, acc3_px_f32, int32_t, 7, D)
}

int cbxo_accumulate_lines3_f64(double *out, long n_frames, long D, long H, long W, const double *lines,
                               long n_lines, const int64_t *terms, const int64_t *frames, long n_terms,
                               double max_val, int kernel, int in_ov, int out_ov)
{
    ACCUMULATE_ND(double, walk3_f64, acc3_px_f64, int64_t, 7, D)
}

/* ------------------------------------------------------------------------ */
/* binary_dilation with one thread (dilate / dilate_with_mask,              */
/* label.cpp:8-107, Region::dilate with n_iter, label.hpp:464-496, and      */
/* detail::shift_index, label.hpp:28-45): the output starts as a copy of    */
/* the input; every true input element seeds a region that grows           */
/* `iterations` times by the structure's non-zero shifts (staying inside    */
/* the array), a new element joining only where the mask is true (always   */
/* without a mask); every element of every region is set.  One breadth-    */
/* first search per seed, elements stamped with the seed's id.              */
/* Pinned by: scipy.ndimage.binary_dilation applied one masked step per     */
/* iteration (its own multi-iteration masked mode is another operation),   */
/* golden fixtures tests/golden/dilation_cases.npz, and the compiled        */
/* reference label module (num_threads = 1) (tests/test_dilation.py).      */
/* ------------------------------------------------------------------------ */

static long shift_flat(long index, const int64_t *shift, int ndim, const int64_t *shape)
{
    long out = 0, stride = 1;
    for (int d = ndim - 1; d >= 0; d--) {
        long c = index % shape[d] + shift[d];
        if (c < 0 || c >= shape[d]) return -1;
        out += c * stride;
        index /= shape[d];
        stride *= shape[d];
    }
    return out;
}

int cbxo_binary_dilation(const uint8_t *inp, const uint8_t *mask, int ndim, const int64_t *shape,
                         const int64_t *shifts, long n_shifts, long iterations, uint8_t *out)
{
    if (ndim < 1 || iterations < 0) return CBXO_EVALUE;
    long P = 1;
    for (int d = 0; d < ndim; d++) P *= shape[d];
    for (long i = 0; i < P; i++) out[i] = inp[i] != 0;
    if (P == 0 || iterations == 0) return CBXO_OK;
    long *stamp = (long *)malloc(sizeof(long) * (size_t)P);
    long *cur = (long *)malloc(sizeof(long) * (size_t)P);
    long *nxt = (long *)malloc(sizeof(long) * (size_t)P);
    if (!stamp || !cur || !nxt) { free(stamp); free(cur); free(nxt); return CBXO_ERUNTIME; }
    for (long i = 0; i < P; i++) stamp[i] = -1;
    for (long seed = 0; seed < P; seed++) {
        if (!inp[seed]) continue;
        long nc = 0;
        cur[nc++] = seed;
        stamp[seed] = seed;
        for (long it = 0; it < iterations && nc; it++) {
            long nn = 0;
            for (long k = 0; k < nc; k++)
                for (long s = 0; s < n_shifts; s++) {
                    long q = shift_flat(cur[k], shifts + s * ndim, ndim, shape);
                    if (q < 0 || stamp[q] == seed) continue;
                    if (mask && !mask[q]) continue;
                    stamp[q] = seed;
                    out[q] = 1;
                    nxt[nn++] = q;
                }
            long *t = cur; cur = nxt; nxt = t;
            nc = nn;
        }
    }
    free(stamp); free(cur); free(nxt);
    return CBXO_OK;
}
