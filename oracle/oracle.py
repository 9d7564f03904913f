"""ORACLE — test infrastructure only.

ctypes front end of ``liboracle.so`` (the C restatement in ``cbx_oracle.c``)
and a loader for the reference's own compiled CPU modules in ``oracle/_ref``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module; the product package never does.

The functions mirror the reference call signatures they restate:

* :func:`draw_lines` — ``bresenham.draw_lines`` with ``num_threads=1``
  (cbclib_v2/_src/src/bresenham.cpp:38-149, 329-330);
* :func:`pixel_list` — the per-line ImageBuffer entries the reference
  accumulates (bresenham.cpp:80-86), in emission order;
* :func:`accumulate_lines` — ``bresenham.accumulate_lines`` with
  ``num_threads=1`` (bresenham.cpp:151-309), on top of :func:`pixel_list`;
* :func:`label` — ``label.label`` + ``LabelResult.to_array``
  (label.cpp:109-175, 459-493) returned as an int32 map and a count;
* :func:`render_closed_f64` / :func:`adjoint_f64` — the FP64 closed form and
  its gradient (no reference counterpart for the gradient);
* :func:`draw_lines3` / :func:`accumulate_lines3` — the same two on 3D lines
  (rows of 7, draw_line_3d, bresenham.hpp:389-430);
* :func:`region_moments` / :func:`p_values` — the per-region functions and
  p_values of the label module (label.cpp:176-255, 534-583);
* :func:`binary_dilation` — ``label.binary_dilation`` with ``num_threads=1``
  (label.cpp:8-107);
* :func:`draw_lines_omp` / :func:`label_omp` — the reference's multi-threaded
  CPU procedures restated (the CPU baseline).
"""
from __future__ import annotations

import ctypes
import glob
import importlib.util
import os
import subprocess
import sys
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = {'biweight': 0, 'gaussian': 1, 'parabolic': 2, 'rectangular': 3, 'triangular': 4}
OVERLAPS = {'sum': 0, 'max': 1}

_lib = None


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources exist)."""
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        _lib = ctypes.CDLL(path)
        L, P, F, D, I = ctypes.c_long, ctypes.c_void_p, ctypes.c_float, ctypes.c_double, ctypes.c_int
        _lib.cbxo_draw_lines_f32.argtypes = [P, L, L, L, P, L, P, F, I, I]
        _lib.cbxo_draw_lines_f64.argtypes = [P, L, L, L, P, L, P, D, I, I]
        _lib.cbxo_pixel_list_f32.argtypes = [L, L, L, P, L, P, F, I, P, P, P, P, P, P, L]
        _lib.cbxo_pixel_list_f32.restype = L
        _lib.cbxo_pixel_list_f64.argtypes = [L, L, L, P, L, P, D, I, P, P, P, P, P, P, L]
        _lib.cbxo_pixel_list_f64.restype = L
        _lib.cbxo_render_closed_f64.argtypes = [P, L, L, L, P, L, P, D, I]
        _lib.cbxo_adjoint_f64.argtypes = [P, L, L, L, P, P, L, P, D, I, P]
        _lib.cbxo_label.argtypes = [P, I, I, P, P, L, L, P]
        _lib.cbxo_label.restype = L
        _lib.cbxo_draw_lines_omp_f32.argtypes = [P, L, L, L, P, L, P, F, I, I, I]
        _lib.cbxo_draw_lines_omp_f64.argtypes = [P, L, L, L, P, L, P, D, I, I, I]
        _lib.cbxo_label_omp.argtypes = [P, I, I, P, P, L, L, P, I]
        _lib.cbxo_label_omp.restype = L
        _lib.cbxo_draw_lines3_f32.argtypes = [P, L, L, L, L, P, L, P, F, I, I]
        _lib.cbxo_draw_lines3_f64.argtypes = [P, L, L, L, L, P, L, P, D, I, I]
        _lib.cbxo_binary_dilation.argtypes = [P, P, I, P, P, L, L, P]
        _lib.cbxo_accumulate_lines3_f32.argtypes = [P, L, L, L, L, P, L, P, P, L, F, I, I, I]
        _lib.cbxo_accumulate_lines3_f64.argtypes = [P, L, L, L, L, P, L, P, P, L, D, I, I, I]
        _lib.cbxo_region_moments_f32.argtypes = [P, P, I, P, P, L, I, P]
        _lib.cbxo_region_moments_f64.argtypes = [P, P, I, P, P, L, I, P]
        _lib.cbxo_p_values_f32.argtypes = [P, P, I, P, P, L, P, F, F, F, P, P]
        _lib.cbxo_p_values_f64.argtypes = [P, P, I, P, P, L, P, D, D, D, P, P]
        _lib.cbxo_logbinom_f64.argtypes = [L, L, D]
        _lib.cbxo_logbinom_f64.restype = D
        _lib.cbxo_logbinom_f32.argtypes = [L, L, F]
        _lib.cbxo_logbinom_f32.restype = F
        _lib.cbxo_accumulate_lines_f32.argtypes = [P, L, L, L, P, L, P, P, L, F, I, I, I]
        _lib.cbxo_accumulate_lines_f64.argtypes = [P, L, L, L, P, L, P, P, L, D, I, I, I]
        _lib.cbxo_plant.argtypes = [I]
        _lib.cbxo_plant.restype = None
        _lib.cbxo_profile_f64.argtypes = [I, D]
        _lib.cbxo_profile_f64.restype = D
        _lib.cbxo_profile_f32.argtypes = [I, F]
        _lib.cbxo_profile_f32.restype = F
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _raise(rc: int, what: str):
    if rc == 1:
        raise ValueError(what)
    if rc == 2:
        raise IndexError(what)
    if rc:
        raise RuntimeError(what)


def _frames(out_shape) -> Tuple[int, int, int]:
    H, W = out_shape[-2], out_shape[-1]
    return int(np.prod(out_shape[:-2], dtype=np.int64)), int(H), int(W)


def draw_lines(out: np.ndarray, lines: np.ndarray, idxs: Optional[np.ndarray] = None,
               max_val: float = 1.0, kernel: str = 'rectangular', overlap: str = 'sum') -> np.ndarray:
    """Single-threaded restatement of bresenham.draw_lines; ``out`` is updated in place."""
    if kernel not in KERNELS:
        raise ValueError(f"kernel is missing for {kernel}")
    if overlap not in OVERLAPS:
        raise ValueError(f"Invalid overlap keyword: {overlap}")
    assert out.flags.c_contiguous
    F, H, W = _frames(out.shape)
    if out.dtype == np.float32:
        ln = np.ascontiguousarray(lines, dtype=np.float32).reshape(-1, 5)
        ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int32).ravel()
        if ix is not None and ix.size != ln.shape[0]:
            raise ValueError("idxs has an invalid size")
        rc = lib().cbxo_draw_lines_f32(_ptr(out), F, H, W, _ptr(ln), ln.shape[0], _ptr(ix),
                                       float(max_val), KERNELS[kernel], OVERLAPS[overlap])
    elif out.dtype == np.float64:
        ln = np.ascontiguousarray(lines, dtype=np.float64).reshape(-1, 5)
        ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int64).ravel()
        if ix is not None and ix.size != ln.shape[0]:
            raise ValueError("idxs has an invalid size")
        rc = lib().cbxo_draw_lines_f64(_ptr(out), F, H, W, _ptr(ln), ln.shape[0], _ptr(ix),
                                       float(max_val), KERNELS[kernel], OVERLAPS[overlap])
    else:
        raise TypeError(out.dtype)
    _raise(rc, "oracle draw_lines failed")
    return out


def draw_lines3(out: np.ndarray, lines: np.ndarray, idxs: Optional[np.ndarray] = None, max_val: float = 1.0,
                kernel: str = 'rectangular', overlap: str = 'sum') -> np.ndarray:
    """Single-threaded restatement of bresenham.draw_lines for 3D lines (rows
    of 7, draw_line_3d, bresenham.hpp:389-430) into out (..., D, H, W); in place."""
    if kernel not in KERNELS:
        raise ValueError(f"kernel is missing for {kernel}")
    if overlap not in OVERLAPS:
        raise ValueError(f"Invalid overlap keyword: {overlap}")
    assert out.flags.c_contiguous and out.ndim >= 3
    D, H, W = out.shape[-3:]
    F = int(np.prod(out.shape[:-3], dtype=np.int64))
    f64 = out.dtype == np.float64
    ln = np.ascontiguousarray(lines, dtype=np.float64 if f64 else np.float32).reshape(-1, 7)
    ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int64 if f64 else np.int32).ravel()
    if ix is not None and ix.size != ln.shape[0]:
        raise ValueError("idxs has an invalid size")
    fn = lib().cbxo_draw_lines3_f64 if f64 else lib().cbxo_draw_lines3_f32
    rc = fn(_ptr(out), F, D, H, W, _ptr(ln), ln.shape[0], _ptr(ix), float(max_val), KERNELS[kernel],
            OVERLAPS[overlap])
    _raise(rc, "oracle draw_lines3 failed")
    return out


def pixel_list(shape, lines: np.ndarray, idxs: Optional[np.ndarray] = None, max_val: float = 1.0,
               kernel: str = 'rectangular', dtype=np.float32):
    """Raw emission list (line, frame, y, x, value, err), one entry per emitted pixel."""
    F, H, W = _frames(shape)
    f64 = np.dtype(dtype) == np.float64
    ln = np.ascontiguousarray(lines, dtype=np.float64 if f64 else np.float32).reshape(-1, 5)
    ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int64 if f64 else np.int32).ravel()
    fn = lib().cbxo_pixel_list_f64 if f64 else lib().cbxo_pixel_list_f32
    cap = 0
    while True:
        arrs = [np.empty(cap, np.int32) for _ in range(4)]
        val = np.empty(cap, ln.dtype)
        err = np.empty(cap, ln.dtype)
        n = fn(F, H, W, _ptr(ln), ln.shape[0], _ptr(ix), float(max_val), KERNELS[kernel],
               *[_ptr(a) for a in arrs], _ptr(val), _ptr(err), cap)
        if n < 0:
            _raise(-n, "oracle pixel_list failed")
        if n <= cap:
            return tuple(a[:n] for a in arrs) + (val[:n], err[:n])
        cap = n


def accumulate_lines(out: np.ndarray, lines: np.ndarray, terms: np.ndarray, frames: np.ndarray,
                     max_val: float = 1.0, kernel: str = 'rectangular', in_overlap: str = 'sum',
                     out_overlap: str = 'sum') -> np.ndarray:
    """Single-threaded restatement of bresenham.accumulate_lines (bresenham.cpp:151-252) in C
    (cbxo_accumulate_lines_*: the reference's recorded-index term buffer); ``out`` in place."""
    if kernel not in KERNELS:
        raise ValueError(f"kernel is missing for {kernel}")
    if in_overlap not in OVERLAPS or out_overlap not in OVERLAPS:
        raise ValueError(f"Invalid in_overlap and out_overlap keywords: {in_overlap} and {out_overlap}")
    assert out.flags.c_contiguous and out.dtype in (np.float32, np.float64)
    F, H, W = _frames(out.shape)
    f64 = out.dtype == np.float64
    it = np.int64 if f64 else np.int32
    ln = np.ascontiguousarray(lines, dtype=out.dtype).reshape(-1, 5)
    tm = np.ascontiguousarray(terms, dtype=it).ravel()
    fr = np.ascontiguousarray(frames, dtype=it).ravel()
    if np.any((fr < 0) | (fr >= F)):
        raise IndexError("frames is outside of (0, n_frames)")
    if tm.size != ln.shape[0]:
        raise ValueError("Term indices is incompatible with number of lines")
    fn = lib().cbxo_accumulate_lines_f64 if f64 else lib().cbxo_accumulate_lines_f32
    rc = fn(_ptr(out), F, H, W, _ptr(ln), ln.shape[0], _ptr(tm), _ptr(fr), fr.size, float(max_val),
            KERNELS[kernel], OVERLAPS[in_overlap], OVERLAPS[out_overlap])
    _raise(rc, "Term index is out of range")
    return out


def accumulate_lines3(out: np.ndarray, lines: np.ndarray, terms: np.ndarray, frames: np.ndarray,
                      max_val: float = 1.0, kernel: str = 'rectangular', in_overlap: str = 'sum',
                      out_overlap: str = 'sum') -> np.ndarray:
    """:func:`accumulate_lines` for 3D lines (rows of 7) into out (..., D, H, W)
    (accumulate_lines_nd_impl with N = 3, bresenham.cpp:151-252); in place."""
    if kernel not in KERNELS:
        raise ValueError(f"kernel is missing for {kernel}")
    if in_overlap not in OVERLAPS or out_overlap not in OVERLAPS:
        raise ValueError(f"Invalid in_overlap and out_overlap keywords: {in_overlap} and {out_overlap}")
    assert out.flags.c_contiguous and out.dtype in (np.float32, np.float64) and out.ndim >= 3
    D, H, W = out.shape[-3:]
    F = int(np.prod(out.shape[:-3], dtype=np.int64))
    f64 = out.dtype == np.float64
    it = np.int64 if f64 else np.int32
    ln = np.ascontiguousarray(lines, dtype=out.dtype).reshape(-1, 7)
    tm = np.ascontiguousarray(terms, dtype=it).ravel()
    fr = np.ascontiguousarray(frames, dtype=it).ravel()
    if np.any((fr < 0) | (fr >= F)):
        raise IndexError("frames is outside of (0, n_frames)")
    if tm.size != ln.shape[0]:
        raise ValueError("Term indices is incompatible with number of lines")
    fn = lib().cbxo_accumulate_lines3_f64 if f64 else lib().cbxo_accumulate_lines3_f32
    rc = fn(_ptr(out), F, D, H, W, _ptr(ln), ln.shape[0], _ptr(tm), _ptr(fr), fr.size, float(max_val),
            KERNELS[kernel], OVERLAPS[in_overlap], OVERLAPS[out_overlap])
    _raise(rc, "Term index is out of range")
    return out


def accumulate_lines_py(out: np.ndarray, lines: np.ndarray, terms: np.ndarray, frames: np.ndarray,
                        max_val: float = 1.0, kernel: str = 'rectangular', in_overlap: str = 'sum',
                        out_overlap: str = 'sum') -> np.ndarray:
    """Single-threaded restatement of bresenham.accumulate_lines (bresenham.cpp:151-252); ``out`` in place.

    Terms are visited in ascending order (one order the reference's OpenMP
    loop over terms may take).  Each term's pixels, the :func:`pixel_list`
    entries of its lines in ascending line order, are reduced with
    ``in_overlap`` into a buffer that starts at zero (bresenham.cpp:215-223);
    every touched pixel of the buffer is then combined into frame
    ``frames[term]`` with ``out_overlap`` (bresenham.cpp:230-248).
    """
    if kernel not in KERNELS:
        raise ValueError(f"kernel is missing for {kernel}")
    if in_overlap not in OVERLAPS or out_overlap not in OVERLAPS:
        raise ValueError(f"Invalid in_overlap and out_overlap keywords: {in_overlap} and {out_overlap}")
    assert out.flags.c_contiguous and out.dtype in (np.float32, np.float64)
    F, H, W = _frames(out.shape)
    ln = np.asarray(lines, dtype=out.dtype).reshape(-1, 5)
    terms = np.asarray(terms, dtype=np.int64).ravel()
    frames = np.asarray(frames, dtype=np.int64).ravel()
    if np.any((frames < 0) | (frames >= F)):
        raise IndexError("frames is outside of (0, n_frames)")
    if terms.size != ln.shape[0]:
        raise ValueError("Term indices is incompatible with number of lines")
    if np.any((terms < 0) | (terms >= frames.size)):
        raise IndexError("Term index is out of range")
    o = out.reshape(F, H * W)
    for t in np.unique(terms):
        sel = np.flatnonzero(terms == t)
        _, _, y, x, val, _ = pixel_list((1, H, W), ln[sel], None, max_val, kernel, dtype=out.dtype)
        pix = y.astype(np.int64) * W + x
        buf = np.zeros(H * W, out.dtype)
        touched = np.zeros(H * W, bool)
        for k in range(pix.size):
            p = pix[k]
            touched[p] = True
            if in_overlap == 'max':
                buf[p] = val[k] if buf[p] < val[k] else buf[p]
            else:
                buf[p] = buf[p] + val[k]
        dst = o[frames[t]]
        for p in np.flatnonzero(touched):
            if out_overlap == 'max':
                dst[p] = buf[p] if dst[p] < buf[p] else dst[p]
            else:
                dst[p] = buf[p] + dst[p]
    return out


def render_closed_f64(shape, lines: np.ndarray, idxs: Optional[np.ndarray] = None, max_val: float = 1.0,
                      kernel: str = 'rectangular') -> np.ndarray:
    F, H, W = _frames(shape)
    out = np.zeros((F, H, W), np.float64)
    ln = np.ascontiguousarray(lines, dtype=np.float64).reshape(-1, 5)
    ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int64).ravel()
    rc = lib().cbxo_render_closed_f64(_ptr(out), F, H, W, _ptr(ln), ln.shape[0], _ptr(ix),
                                      float(max_val), KERNELS[kernel])
    _raise(rc, "oracle render_closed failed")
    return out.reshape(shape)


def _profile_f64(kernel: str, x: np.ndarray) -> np.ndarray:
    """K(x) in FP64 on |x| <= 1 (include.hpp:446-459), vectorised."""
    ax = np.abs(x)
    if kernel == 'biweight':
        return 0.9375 * np.maximum(1.0 - x * x, 0.0) ** 2
    if kernel == 'gaussian':
        return np.where(ax <= 1.0, 0.3989422804014327 * np.exp(-((3.0 * x) ** 2) / 2.0), 0.0)
    if kernel == 'parabolic':
        return 0.75 * np.maximum(1.0 - x * x, 0.0)
    if kernel == 'rectangular':
        return np.where(ax <= 1.0, 1.0, 0.0)
    if kernel == 'triangular':
        return np.maximum(1.0 - ax, 0.0)
    raise ValueError(f"kernel is missing for {kernel}")


def segment_window_f64(line, box, max_val: float = 1.0, kernel: str = 'biweight') -> np.ndarray:
    """FP64 closed form max_val K(dist / w) of one line over the pixel box
    (y0, y1, x0, x1) (inclusive), 0 where dist > w: the per-line term of
    :func:`render_closed_f64` (the clamped projection of geometry.hpp:319-339)."""
    x0, y0, x1, y1, w = (float(v) for v in line)
    ya, yb, xa, xb = box
    out = np.zeros((yb - ya + 1, xb - xa + 1))
    if not w > 0.0 or out.size == 0:
        return out
    py, px = np.mgrid[ya:yb + 1, xa:xb + 1].astype(np.float64)
    tx, ty = x1 - x0, y1 - y0
    mag = tx * tx + ty * ty
    cx, cy = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    if mag > 0.0:
        r = np.clip(((px - cx) * tx + (py - cy) * ty) / mag, -0.5, 0.5)
        qx, qy = r * tx + cx, r * ty + cy
    else:
        qx, qy = x0, y0
    u = ((px - qx) ** 2 + (py - qy) ** 2) / (w * w)
    inside = u <= 1.0
    out[inside] = max_val * _profile_f64(kernel, np.sqrt(u[inside]))
    return out


def _line_box(line, H, W, pad=1.0):
    x0, y0, x1, y1, w = (float(v) for v in line)
    w = max(w, 0.0) + pad
    xa = max(0, int(np.floor(min(x0, x1) - w)) - 1)
    xb = min(W - 1, int(np.ceil(max(x0, x1) + w)) + 1)
    ya = max(0, int(np.floor(min(y0, y1) - w)) - 1)
    yb = min(H - 1, int(np.ceil(max(y0, y1) + w)) + 1)
    return ya, yb, xa, xb


def loss_fd_f64(images: dict, targets: dict, lines: np.ndarray, idxs: np.ndarray, pick, h: float = 1e-3,
                max_val: float = 1.0, kernel: str = 'biweight') -> np.ndarray:
    """FP64 central differences of loss = sum over frames of sum((I - T)^2)
    w.r.t. (x0, y0, x1, y1, w) of the lines in ``pick``.

    ``images[f]`` / ``targets[f]``: FP64 closed-form renders (H, W) of the
    frames the picked lines lie in (:func:`render_closed_f64`).  Only the
    line's own term changes, so each difference is evaluated on the line's
    window: (L+ - L-) / 2h = sum (c+ - c-) (2 R + c+ + c-) / 2h, with R the
    render minus the line's term minus the target.
    """
    out = np.zeros((len(pick), 5))
    for n, i in enumerate(pick):
        f = int(idxs[i])
        img, tgt = images[f], targets[f]
        H, W = img.shape
        ln = np.asarray(lines[i], np.float64)
        box = _line_box(ln, H, W, pad=1.0 + 4 * h)
        ya, yb, xa, xb = box
        if yb < ya or xb < xa:
            continue
        c0 = segment_window_f64(ln, box, max_val, kernel)
        R = img[ya:yb + 1, xa:xb + 1] - c0 - tgt[ya:yb + 1, xa:xb + 1]
        for j in range(5):
            p = ln.copy()
            p[j] += h
            m = ln.copy()
            m[j] -= h
            cp = segment_window_f64(p, box, max_val, kernel)
            cm = segment_window_f64(m, box, max_val, kernel)
            out[n, j] = float(np.sum((cp - cm) * (2.0 * R + cp + cm))) / (2.0 * h)
    return out


def adjoint_f64(grad_img: np.ndarray, lines: np.ndarray, idxs: Optional[np.ndarray] = None,
                max_val: float = 1.0, kernel: str = 'rectangular', walk: str = 'f32') -> np.ndarray:
    """FP64 gradient of sum(grad_img * image) w.r.t. (x0, y0, x1, y1, w) of every line.

    ``walk`` picks the pixel set: 'f32' uses the float walk (what an f32 forward
    draws), 'f64' the double walk.
    """
    F, H, W = _frames(grad_img.shape)
    g = np.ascontiguousarray(grad_img, dtype=np.float64)
    ln = np.ascontiguousarray(lines, dtype=np.float64).reshape(-1, 5)
    l32 = np.ascontiguousarray(lines, dtype=np.float32).reshape(-1, 5) if walk == 'f32' else None
    ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int64).ravel()
    out = np.zeros_like(ln)
    rc = lib().cbxo_adjoint_f64(_ptr(g), F, H, W, _ptr(ln), _ptr(l32), ln.shape[0], _ptr(ix),
                                float(max_val), KERNELS[kernel], _ptr(out))
    _raise(rc, "oracle adjoint failed")
    return out


def label(inp: np.ndarray, shifts: np.ndarray, npts: int = 1) -> Tuple[np.ndarray, int]:
    """Connected components; ``shifts`` are the structure's non-zero shifts (n, ndim)."""
    a = np.ascontiguousarray(inp)
    if a.dtype == np.bool_:
        a, dt = a.view(np.uint8), 0
    elif a.dtype == np.int32:
        dt = 1
    elif a.dtype == np.int64:
        dt = 2
    else:
        raise TypeError(a.dtype)
    shape = np.asarray(a.shape, np.int64)
    sh = np.ascontiguousarray(shifts, dtype=np.int64).reshape(-1, a.ndim)
    out = np.zeros(a.shape, np.int32)
    n = lib().cbxo_label(_ptr(a), dt, a.ndim, _ptr(shape), _ptr(sh), sh.shape[0], int(npts), _ptr(out))
    if n < 0:
        _raise(-n, "oracle label failed")
    return out, int(n)


def draw_lines_omp(out: np.ndarray, lines: np.ndarray, idxs: Optional[np.ndarray] = None, max_val: float = 1.0,
                   kernel: str = 'rectangular', overlap: str = 'sum', threads: int = 1) -> np.ndarray:
    """The reference's multi-threaded CPU draw_lines restated (bresenham.cpp:62-94): per-thread
    pixel buffers merged in a critical section.  The CPU baseline; ``out`` in place."""
    F, H, W = _frames(out.shape)
    assert out.flags.c_contiguous
    f64 = out.dtype == np.float64
    ln = np.ascontiguousarray(lines, dtype=np.float64 if f64 else np.float32).reshape(-1, 5)
    ix = None if idxs is None else np.ascontiguousarray(idxs, dtype=np.int64 if f64 else np.int32).ravel()
    fn = lib().cbxo_draw_lines_omp_f64 if f64 else lib().cbxo_draw_lines_omp_f32
    rc = fn(_ptr(out), F, H, W, _ptr(ln), ln.shape[0], _ptr(ix), float(max_val), KERNELS[kernel],
            OVERLAPS[overlap], int(threads))
    _raise(rc, "oracle draw_lines_omp failed")
    return out


def label_omp(inp: np.ndarray, shifts: np.ndarray, npts: int = 1, threads: int = 1) -> Tuple[np.ndarray, int]:
    """The reference's multi-threaded CPU label restated (label.cpp:109-175): per-thread chunks,
    region growing, regions kept by the thread holding their smallest index.  The CPU baseline."""
    a = np.ascontiguousarray(inp)
    if a.dtype == np.bool_:
        a, dt = a.view(np.uint8), 0
    elif a.dtype == np.int32:
        dt = 1
    elif a.dtype == np.int64:
        dt = 2
    else:
        raise TypeError(a.dtype)
    shape = np.asarray(a.shape, np.int64)
    sh = np.ascontiguousarray(shifts, dtype=np.int64).reshape(-1, a.ndim)
    out = np.empty(a.shape, np.int32)
    n = lib().cbxo_label_omp(_ptr(a), dt, a.ndim, _ptr(shape), _ptr(sh), sh.shape[0], int(npts), _ptr(out),
                             int(threads))
    if n < 0:
        _raise(-n, "oracle label_omp failed")
    return out, int(n)


MOMENTS = {'total_mass': 0, 'mean': 1, 'center_of_mass': 2, 'moment_of_inertia': 3, 'covariance_matrix': 4,
           'line_fit': 5}


def region_moments(what: str, labels: np.ndarray, index, data: np.ndarray) -> np.ndarray:
    """The reference's per-region functions (label.cpp:534-583) on the regions
    of a label map: region j = the pixels holding index[j], ascending.
    Arithmetic in data's type, in the reference's order (MomentsND,
    label.hpp:554-722).  Shapes: total_mass (n,), mean / center_of_mass
    (n, ndim), moment_of_inertia / covariance_matrix (n, ndim, ndim),
    line_fit (n, 2 ndim)."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    d = np.ascontiguousarray(data)
    if d.dtype not in (np.float32, np.float64):
        d = d.astype(np.float64)
    if lab.shape != d.shape:
        raise ValueError("labels and data shapes differ")
    idx = np.ascontiguousarray(index, dtype=np.int64).ravel()
    N = lab.ndim
    w = MOMENTS[what]
    M = {0: 1, 1: N, 2: N, 3: N * N, 4: N * N, 5: 2 * N}[w]
    out = np.zeros((idx.size, M), d.dtype)
    shape = np.asarray(lab.shape, np.int64)
    fn = lib().cbxo_region_moments_f64 if d.dtype == np.float64 else lib().cbxo_region_moments_f32
    _raise(fn(_ptr(lab), _ptr(d), N, _ptr(shape), _ptr(idx), idx.size, w, _ptr(out)), "oracle moments failed")
    if w == 0:
        return out.reshape(-1)
    if w in (3, 4):
        return out.reshape(-1, N, N)
    return out


def p_values(labels: np.ndarray, index, lines: np.ndarray, data: np.ndarray, p0: float, vmin: float,
             xtol: float, counts: bool = False):
    """label.p_values (label.cpp:228-255): log binomial tail of the pixels of
    each region within xtol of its line that reach vmin, in data's type;
    ``counts=True`` also returns the (n, k) pairs."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    d = np.ascontiguousarray(data)
    idx = np.ascontiguousarray(index, dtype=np.int64).ravel()
    N = lab.ndim
    ln = np.ascontiguousarray(lines, dtype=d.dtype).reshape(-1, 2 * N)
    if ln.shape[0] < idx.size:
        raise ValueError("lines array must have shape (n_lines, data.ndim() * 2)")
    out = np.zeros(idx.size, d.dtype)
    cnt = np.zeros((idx.size, 2), np.int64)
    shape = np.asarray(lab.shape, np.int64)
    fn = lib().cbxo_p_values_f64 if d.dtype == np.float64 else lib().cbxo_p_values_f32
    _raise(fn(_ptr(lab), _ptr(d), N, _ptr(shape), _ptr(idx), idx.size, _ptr(ln), float(p0), float(vmin),
              float(xtol), _ptr(out), _ptr(cnt)), "oracle p_values failed")
    return (out, cnt) if counts else out


class planted:
    """Context manager that plants mistake ``m`` in the oracle (tests only; see
    cbxo_plant in cbx_oracle.c for the list)."""

    def __init__(self, m: int):
        self.m = int(m)

    def __enter__(self):
        lib().cbxo_plant(self.m)
        return self

    def __exit__(self, *exc):
        lib().cbxo_plant(0)
        return False


def profile(kernel: str, x: float, dtype=np.float64) -> float:
    if np.dtype(dtype) == np.float32:
        return lib().cbxo_profile_f32(KERNELS[kernel], float(np.float32(x)))
    return lib().cbxo_profile_f64(KERNELS[kernel], float(x))


# ---------------------------------------------------------------------------
# the reference's compiled CPU modules (built by oracle/Makefile from the
# reference sources; absent when the reference tree was not available)
# ---------------------------------------------------------------------------

def ref_module(name: str):
    """Import oracle/_ref/<name>*.so (``bresenham`` or ``label``) or return None."""
    hits = glob.glob(os.path.join(HERE, '_ref', name + '*.so'))
    if not hits:
        return None
    key = 'cbx_oracle_ref_' + name
    if key in sys.modules:
        return sys.modules[key]
    spec = importlib.util.spec_from_file_location(name, hits[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules[key] = mod
    return mod


def binary_dilation(inp: np.ndarray, shifts: np.ndarray, iterations: int = 1,
                    mask: Optional[np.ndarray] = None) -> np.ndarray:
    """Single-threaded restatement of label.binary_dilation (label.cpp:8-107);
    ``shifts`` are the structure's non-zero shifts (n, ndim).  Returns a new
    bool array."""
    a = np.ascontiguousarray(inp, dtype=np.bool_)
    if iterations < 0:
        raise ValueError("iterations must be non-negative")
    m = None
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.bool_)
        if m.shape != a.shape:
            raise ValueError("input and mask must have the same shape")
    shape = np.asarray(a.shape, np.int64)
    sh = np.ascontiguousarray(shifts, dtype=np.int64).reshape(-1, max(a.ndim, 1))
    out = np.zeros(a.shape, np.uint8)
    rc = lib().cbxo_binary_dilation(_ptr(a.view(np.uint8)), None if m is None else _ptr(m.view(np.uint8)), a.ndim,
                                    _ptr(shape), _ptr(sh), sh.shape[0], int(iterations), _ptr(out))
    _raise(rc, "oracle binary_dilation failed")
    return out.view(np.bool_)
