/*
 * cbx — MI355X (gfx950) streak rasterization and connected-component
 * labelling, C ABI.
 *
 * Drop-in boundary for the hot path of cbclib_v2: ndimage.draw_lines and
 * label.label.  Every pointer argument below is a DEVICE pointer (HBM) unless
 * stated otherwise; every call is stream-ordered on `stream` (a hipStream_t,
 * NULL = the default stream).  Arrays are C-contiguous.  Calls return a
 * status code; on failure cbx_last_error() holds a message for the calling
 * thread.  Status codes map onto the exceptions the reference raises:
 *   CBX_EVALUE   -> ValueError   (std::invalid_argument in the reference)
 *   CBX_EINDEX   -> IndexError   (std::out_of_range)
 *   CBX_ERUNTIME -> RuntimeError (std::runtime_error / CUDA errors)
 * A call that returns an error has not modified its outputs.
 */
#ifndef CBX_H_
#define CBX_H_

#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define CBX_API __attribute__((visibility("default")))
#else
#define CBX_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define CBX_OK 0
#define CBX_EVALUE 1
#define CBX_EINDEX 2
#define CBX_ERUNTIME 3
#define CBX_ECAPACITY 4 /* a plan's capacity was exceeded: create a larger plan (RuntimeError) */

/* Profiles, numbered as kernels::type (cbclib_v2/_src/src/include.hpp:473-480). */
#define CBX_BIWEIGHT 0
#define CBX_GAUSSIAN 1
#define CBX_PARABOLIC 2
#define CBX_RECTANGULAR 3
#define CBX_TRIANGULAR 4

/* Overlap modes (bresenham.cpp:142-149). */
#define CBX_SUM 0
#define CBX_MAX 1

/* label() input element types. */
#define CBX_BOOL 0
#define CBX_INT32 1
#define CBX_INT64 2

/* Message of the last failed call on this thread ("" if none). */
CBX_API const char *cbx_last_error(void);

/* Library version string. */
CBX_API const char *cbx_version(void);

/* Profile name -> CBX_* id, -1 if unknown (kernels::get_type, include.hpp:495-503). */
CBX_API int cbx_kernel_id(const char *name);

/*
 * Launch options of draw_lines (NULL = automatic).  `tile` is the side of the
 * square pixel tile one workgroup accumulates in LDS (32 or 64), `seg` the
 * number of long-axis columns per walk checkpoint (8 or 16).  `deterministic`
 * = 1 accumulates every pixel in ascending line order starting from the value
 * already in `out`, which is the single-threaded reference's order: the
 * result is then bit-identical to bresenham.draw_lines(..., num_threads=1).
 * Otherwise (the tile path) the pixel set is still the reference's exactly;
 * values are computed in float (a few ulp from the reference's) unless
 * `exact_values` = 1, and summed in LDS in a different order.
 * The replacement of the reference's `grid` argument
 * (cuda_draw_lines.pyi:8-10).
 */
typedef struct cbx_draw_opts {
    int32_t tile;
    int32_t seg;
    int32_t deterministic;
    int32_t threads;      /* workgroup size of the tile kernel: 64, 128 or 256 (0 = 256) */
    int32_t exact_values; /* 1: the tile path computes values with the reference's arithmetic */
    int32_t chunk;        /* items per phase of the tile kernel: 64 or the workgroup size (0 = 64) */
    const uint8_t *kernels; /* NULL, or a DEVICE array of n_lines profile ids (CBX_BIWEIGHT ..): each line
                               drawn with its own profile (tile path only; the `kernel` argument is then
                               ignored).  No reference counterpart: the reference takes one profile per
                               call, so mixed-profile frames need one call per profile there. */
    void *out_ready;        /* NULL, or a hipEvent_t recorded where `out` becomes ready (e.g. a fill on
                               another stream): draw_lines waits for it just before its first access to
                               `out`, so the line setup, walk and binning overlap that work.  `lines` and
                               `idxs` must be ready on `stream` as usual.  No reference counterpart. */
} cbx_draw_opts;

/*
 * draw_lines: render thick profiled segments into `out` (accumulated in
 * place).  out: n_frames x height x width; lines: n_lines x 5 rows
 * [x0, y0, x1, y1, width] (x = column); idxs: frame of each line or NULL
 * (then n_frames must be 1 or n_lines, bresenham.cpp:5-20).  Out-of-range
 * idxs -> CBX_EINDEX (bresenham.cpp:22-36).  overlap = CBX_SUM | CBX_MAX.
 * Replaces cuda_draw_lines.draw_lines (cuda_draw_lines.cu:940-946, bound at
 * 1098-1113) and bresenham.draw_lines (bresenham.cpp:142-149, bound at
 * 329-330); the f32 entry takes int32 idxs and the f64 entry int64, the two
 * overloads the reference binds.
 */
CBX_API int cbx_draw_lines_f32(float *out, int64_t n_frames, int64_t height, int64_t width, const float *lines,
                       int64_t n_lines, const int32_t *idxs, float max_val, int kernel, int overlap,
                       const cbx_draw_opts *opts, void *stream);
CBX_API int cbx_draw_lines_f64(double *out, int64_t n_frames, int64_t height, int64_t width, const double *lines,
                       int64_t n_lines, const int64_t *idxs, double max_val, int kernel, int overlap,
                       const cbx_draw_opts *opts, void *stream);

/*
 * Pixel-list mode: the (frame, y, x, value) entries of the render instead of
 * a dense image.  merge = 0 returns the raw per-line entries in the order the
 * reference emits them into its per-thread ImageBuffer (bresenham.cpp:80-86,
 * bresenham.hpp:10-34), with `line` = the segment index; merge = 1 returns one
 * entry per touched pixel, sorted by (frame, y, x), whose value is the
 * overlap-reduction of its entries in ascending line order (the image value
 * on a zero frame for CBX_SUM; `line` is then the first contributing line).
 * `*count` receives the number of entries; outputs are written only if
 * *count <= capacity (call again with a larger buffer otherwise).
 */
CBX_API int cbx_draw_pixels_f32(int64_t n_frames, int64_t height, int64_t width, const float *lines, int64_t n_lines,
                        const int32_t *idxs, float max_val, int kernel, int overlap, int merge, int32_t *frame,
                        int32_t *y, int32_t *x, float *value, int32_t *line, int64_t capacity, int64_t *count,
                        void *stream);
CBX_API int cbx_draw_pixels_f64(int64_t n_frames, int64_t height, int64_t width, const double *lines, int64_t n_lines,
                        const int64_t *idxs, double max_val, int kernel, int overlap, int merge, int32_t *frame,
                        int32_t *y, int32_t *x, double *value, int32_t *line, int64_t capacity, int64_t *count,
                        void *stream);

/*
 * accumulate_lines: lines grouped into terms (terms[i] in [0, n_terms) for
 * each of the n_lines lines), each term drawn with `in_overlap` over its lines
 * starting from zero, then every term combined into frame frames[t] of `out`
 * (n_frames x height x width, in place) with `out_overlap`
 * (CBX_SUM | CBX_MAX each).  A term or frame index out of range ->
 * CBX_EINDEX.  Equal overlaps are draw_lines with idxs = frames[terms]; mixed
 * overlaps reduce each term per pixel tile in the reference's line order and
 * add (or max) the tile into its frame with atomics, so terms sharing a frame
 * pixel combine in no fixed order.  opts as for draw_lines (the tile path of
 * equal overlaps);
 * opts->kernels is not accepted.  Replaces bresenham.accumulate_lines
 * (bresenham.cpp:151-309, bound at 326-327) and
 * cuda_draw_lines.accumulate_lines (cuda_draw_lines.cu:948-1057); the f32
 * entry takes int32 terms/frames and the f64 entry int64.
 */
CBX_API int cbx_accumulate_lines_f32(float *out, int64_t n_frames, int64_t height, int64_t width, const float *lines,
                             int64_t n_lines, const int32_t *terms, const int32_t *frames, int64_t n_terms,
                             float max_val, int kernel, int in_overlap, int out_overlap, const cbx_draw_opts *opts,
                             void *stream);
CBX_API int cbx_accumulate_lines_f64(double *out, int64_t n_frames, int64_t height, int64_t width, const double *lines,
                             int64_t n_lines, const int64_t *terms, const int64_t *frames, int64_t n_terms,
                             double max_val, int kernel, int in_overlap, int out_overlap, const cbx_draw_opts *opts,
                             void *stream);

/*
 * Adjoint of draw_lines (overlap = sum): grad_lines[i] = d<grad_out, image>/d(x0, y0, x1, y1, width)
 * of line i, over the pixels line i draws.  No reference counterpart (the
 * reference has no gradient); the forward it differentiates is
 * cbx_draw_lines_*.
 */
CBX_API int cbx_draw_lines_adjoint_f32(const float *grad_out, int64_t n_frames, int64_t height, int64_t width,
                               const float *lines, int64_t n_lines, const int32_t *idxs, float max_val, int kernel,
                               float *grad_lines, const cbx_draw_opts *opts, void *stream);
CBX_API int cbx_draw_lines_adjoint_f64(const double *grad_out, int64_t n_frames, int64_t height, int64_t width,
                               const double *lines, int64_t n_lines, const int64_t *idxs, double max_val,
                               int kernel, double *grad_lines, const cbx_draw_opts *opts, void *stream);

/*
 * 3D lines: draw_lines on volumes.  out: n_frames x depth x height x width
 * (z, y, x); lines: n_lines x 7 rows [x0, y0, z0, x1, y1, z1, width].  Same
 * idxs rules, error codes and overlap values as cbx_draw_lines_*; the pixels
 * and values are draw_line_3d's (bresenham.hpp:389-430).  CBX_MAX is exact;
 * CBX_SUM adds with atomics (no fixed order).  A line whose walk would visit
 * more than 2^30 pixels -> CBX_EVALUE.  Replaces bresenham.draw_lines on
 * rows of 7 (draw_lines_2d_3d, bresenham.cpp:126-149, the L = 7 case).
 */
CBX_API int cbx_draw_lines3_f32(float *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                                const float *lines, int64_t n_lines, const int32_t *idxs, float max_val, int kernel,
                                int overlap, void *stream);
CBX_API int cbx_draw_lines3_f64(double *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                                const double *lines, int64_t n_lines, const int64_t *idxs, double max_val, int kernel,
                                int overlap, void *stream);

/*
 * 3D lines: accumulate_lines on volumes (arguments as cbx_accumulate_lines_*
 * with depth and rows of 7).  Equal overlaps are cbx_draw_lines3 with
 * idxs = frames[terms]; mixed overlaps draw the terms into zeroed term
 * volumes (batches of at most 256 MiB) and combine every touched voxel into
 * its frame.  Replaces bresenham.accumulate_lines on rows of 7
 * (accumulate_lines_2d_3d, bresenham.cpp:284-309, the L = 7 case).
 */
CBX_API int cbx_accumulate_lines3_f32(float *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                                      const float *lines, int64_t n_lines, const int32_t *terms,
                                      const int32_t *frames, int64_t n_terms, float max_val, int kernel,
                                      int in_overlap, int out_overlap, void *stream);
CBX_API int cbx_accumulate_lines3_f64(double *out, int64_t n_frames, int64_t depth, int64_t height, int64_t width,
                                      const double *lines, int64_t n_lines, const int64_t *terms,
                                      const int64_t *frames, int64_t n_terms, double max_val, int kernel,
                                      int in_overlap, int out_overlap, void *stream);

/*
 * label: connected components of equal-valued non-zero elements of `inp`
 * (dtype CBX_BOOL / CBX_INT32 / CBX_INT64, ndim dims of `shape`) under the
 * neighbour `shifts` (n_shifts x ndim, HOST array, the structure's non-zero
 * shifts).  out (int32, same shape) receives 0 for background and ids
 * 1..n in raster order of each component's smallest flat index, after
 * dropping components with fewer than npts elements.  *n_labels (HOST)
 * receives n.  Replaces cuda_label.label (cuda_label.cu:294-378, bound at
 * 847-849) and label.label (label.cpp:109-175, bound at 530-532).
 */
CBX_API int cbx_label(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts, int64_t n_shifts,
              int64_t npts, int32_t *out, int64_t *n_labels, void *stream);

/*
 * cbx_label without allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors as cbx_label; its intermediate buffers come from a device array
 * of cbx_label_noalloc_capacity() bytes (64 MiB) that every GPU holds from
 * the library's load until the process ends, so the call makes no hipMalloc /
 * hipFree / pinned-host allocation.  An input that needs more scratch than
 * that (about 2.2 bytes per element for binary inputs with a left neighbour,
 * up to 12 for the general path) -> CBX_EVALUE before any output is written.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams or host threads) race and must be serialised by the
 * caller.  Still waits on `stream` for *n_labels, like cbx_label.
 */
CBX_API int cbx_label_noalloc(const void *inp, int dtype, int ndim, const int64_t *shape, const int64_t *shifts,
                              int64_t n_shifts, int64_t npts, int32_t *out, int64_t *n_labels, void *stream);
CBX_API int64_t cbx_label_noalloc_capacity(void);

/*
 * binary_dilation: out = inp dilated `iterations` times by the structure's
 * `shifts` (n_shifts x ndim, HOST array, the non-zero shifts), an element
 * joining only where `mask` is non-zero (mask NULL: everywhere); inp, mask
 * and out are DEVICE arrays of ndim dims of `shape` with one byte per element
 * (non-zero = true; out receives 0 / 1; out may alias inp).  iterations = 0
 * copies the input.  Equals the reference with num_threads = 1 (its
 * multi-threaded result keeps only the pixels each thread's seeds reach
 * inside that thread's chunk).  Replaces label.binary_dilation
 * (label.cpp:8-107, bound at 527-528).
 */
CBX_API int cbx_binary_dilation(const uint8_t *inp, const uint8_t *mask, int ndim, const int64_t *shape,
                                const int64_t *shifts, int64_t n_shifts, int64_t iterations, uint8_t *out,
                                void *stream);

/*
 * label of (image > vmin) without materialising the mask: the threshold of
 * the detection pipeline (RegionDetector.detect_regions,
 * data_processing.py:630-631) fused into the first labelling pass.
 */
CBX_API int cbx_label_threshold_f32(const float *image, float vmin, int ndim, const int64_t *shape, const int64_t *shifts,
                            int64_t n_shifts, int64_t npts, int32_t *out, int64_t *n_labels, void *stream);

/* Region quantities of cbx_region_moments (the reference's per-region functions, label.cpp:534-583). */
#define CBX_TOTAL_MASS 0        /* (n)          sum of max(v, 0) */
#define CBX_MEAN 1              /* (n, N)       raw first moment sum v p (MomentsND::first) */
#define CBX_CENTER_OF_MASS 2    /* (n, N)       centroid, the region's first pixel when the mass is 0 */
#define CBX_MOMENT_OF_INERTIA 3 /* (n, N, N)    raw second moment (MomentsND::second) */
#define CBX_COVARIANCE_MATRIX 4 /* (n, N, N)    central second moment, zero when the mass is 0 */
#define CBX_LINE_FIT 5          /* (n, 2N)      major axis end points (x0, y0, .., x1, y1, ..), xyz order */

/*
 * Region moments of a label map: for each of the n_index entries of `index`
 * (DEVICE, int64 label ids), the quantity `what` of the region of pixels of
 * `labels` (DEVICE int32, ndim = 2..7 dims of `shape`, HOST) holding that id,
 * over the values max(data, 0) (DEVICE, same shape).  Coordinates and
 * matrices in the input's axis order (zyx), line end points in xyz order,
 * as the reference returns them.  Sums in double (the reference sums in the
 * data's type).  Unsupported ndim -> CBX_EVALUE.  Replaces
 * cuda_label.center_of_mass / covariance_matrix (cuda_label.cu:380-724) and
 * label.{total_mass, mean, center_of_mass, moment_of_inertia,
 * covariance_matrix, line_fit} (label.cpp:534-583); line_fit is
 * functions.py:582-620 (to_line on the GPU).
 */
CBX_API int cbx_region_moments_f32(const int32_t *labels, const float *data, int ndim, const int64_t *shape,
                                   const int64_t *index, int64_t n_index, int what, float *out, void *stream);
CBX_API int cbx_region_moments_f64(const int32_t *labels, const double *data, int ndim, const int64_t *shape,
                                   const int64_t *index, int64_t n_index, int what, double *out, void *stream);

/*
 * p_values of the regions (label.cpp:228-255, cuda_label.cu:726-829): for
 * region j, n = its pixels within xtol of lines[j] (LineND::distance, the
 * clamped projection; lines DEVICE, n_index x 2N, xyz end points) and k = those
 * of them with data >= vmin; out[j] = log of the binomial upper tail
 * P(X >= k), X ~ Bin(n, p0) (detail::logbinom, label.hpp:57-95).  counts
 * (DEVICE, n_index x 2 int64: n, k) may be NULL.
 */
CBX_API int cbx_p_values_f32(const int32_t *labels, const float *data, int ndim, const int64_t *shape,
                             const int64_t *index, int64_t n_index, const float *lines, float p0, float vmin,
                             float xtol, float *out, int64_t *counts, void *stream);
CBX_API int cbx_p_values_f64(const int32_t *labels, const double *data, int ndim, const int64_t *shape,
                             const int64_t *index, int64_t n_index, const double *lines, double p0, double vmin,
                             double xtol, double *out, int64_t *counts, void *stream);

/*
 * Plans (no reference counterpart): the tile path of cbx_draw_lines_f32 split
 * into sizing and drawing, for repeated draws of similar line sets (batched
 * frames, refinement loops) and for HIP-graph capture.
 *
 * cbx_plan_create_f32 sizes the intermediate buffers (segment checkpoints and
 * per-tile item lists) on the given lines, times `headroom` (>= 1), and
 * allocates them once; the frame shape, the maximum line count and the launch
 * options (tile, seg, threads, chunk, exact_values) are fixed by the plan.
 * cbx_plan_draw_f32 draws `n_lines` <= the plan's line count into `out` with
 * no host synchronisation, no allocation and no memset (the whole call can be
 * captured in a HIP graph and replayed); `kernels` (DEVICE, n_lines profile
 * ids, or NULL for `kernel`) gives per-line profiles.  Errors of a draw (bad
 * idxs or profile ids, capacity exceeded) leave `out` untouched and are
 * reported by the next cbx_plan_check, which waits for the stream;
 * CBX_ECAPACITY means: create a plan with more headroom and draw again.
 */
typedef struct cbx_plan cbx_plan;
CBX_API int cbx_plan_create_f32(cbx_plan **plan, int64_t n_frames, int64_t height, int64_t width, const float *lines,
                                int64_t n_lines, const int32_t *idxs, const uint8_t *kernels, int kernel,
                                const cbx_draw_opts *opts, double headroom, void *stream);
CBX_API int cbx_plan_draw_f32(cbx_plan *plan, float *out, const float *lines, int64_t n_lines, const int32_t *idxs,
                              const uint8_t *kernels, float max_val, int kernel, int overlap, void *stream);
/*
 * cbx_plan_adjoint_f32: the adjoint of the plan's last draw (no reference
 * counterpart; the gradient of cbx_draw_lines_adjoint_f32, see there), on
 * that draw's segments and tile lists instead of a second walk, with no host
 * synchronisation.  `lines`, `idxs`, `n_lines` and `kernel` must be those of
 * the last cbx_plan_draw_f32 on this plan (a different count or profile, or a
 * draw with per-line profiles -> CBX_EVALUE; other differences are the
 * caller's to avoid); the draw's overlap and max_val do not matter.  After a
 * failed draw grad_lines stays zero and cbx_plan_check reports the error.
 * The first call allocates the adjoint's buffers (about 40 bytes per item of
 * the plan's capacity).
 */
CBX_API int cbx_plan_adjoint_f32(cbx_plan *plan, const float *grad_out, const float *lines, int64_t n_lines,
                                 const int32_t *idxs, float max_val, int kernel, float *grad_lines, void *stream);
CBX_API int cbx_plan_check(cbx_plan *plan, void *stream);
CBX_API int64_t cbx_plan_bytes(const cbx_plan *plan);
CBX_API void cbx_plan_destroy(cbx_plan *plan);

/*
 * Path selection for tests and measurements (no reference counterpart):
 * 0 (default) picks the fastest applicable label algorithm (run-length
 * union-find for binary inputs with horizontal connectivity, band union-find
 * for 3x3 2D structures, element union-find otherwise); 1 never uses the
 * run-length path; 2 uses the element union-find only.  All give the same
 * labels.  Process-wide; not for concurrent use with running label calls.
 */
CBX_API void cbx_label_force_path(int path);

/*
 * Instrumentation (no reference counterpart; the reference has no timers):
 * per-kernel durations measured with HIP events recorded on the launch
 * stream around the library's main kernels ("draw_tiles",
 * "adjoint_tiles", "segments_pass1", "label_union", ...).  cbx_timing_read
 * waits for the recorded events and returns the accumulated milliseconds and
 * the number of launches since the last reset.  cbx_timing_enable: 0 off,
 * 2 only the tile kernels' timers (names ending in "_tiles": the fewest
 * events in a timed loop), any other value every timer.
 */
CBX_API void cbx_timing_enable(int on);
CBX_API void cbx_timing_reset(void);
CBX_API int cbx_timing_read(const char *name, double *total_ms, int64_t *count);

/*
 * Workspace (no reference counterpart; the reference allocates per call on
 * the host).  Intermediate device buffers of every entry point come from a
 * cache kept per (device, stream) and reused by later calls on the same
 * stream, so steady-state calls make no hipMalloc / hipFree.  The cache only
 * grows.  cbx_workspace_release waits for the devices and frees every cache
 * not in use by a running call, returning the bytes freed;
 * cbx_workspace_bytes returns the bytes held.
 */
CBX_API int64_t cbx_workspace_release(void);
CBX_API int64_t cbx_workspace_bytes(void);

/*
 * The device-wide exclusive prefix sum the entry points use (no reference
 * counterpart; exported for tests and callers): out[i] = in[0] + ... +
 * in[i-1] for i < n, one kernel (single pass, decoupled look-back) on
 * `stream`; total != 0 also writes out[n] = the sum of all n (out holds
 * n + 1 entries).  int32 sums wrap; int64 sums must stay within +-2^61.
 * At most 2^31 - 1 outputs.
 */
CBX_API int cbx_exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, int total, void *stream);
CBX_API int cbx_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, int total, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CBX_H_ */
