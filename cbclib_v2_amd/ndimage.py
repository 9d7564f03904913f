"""``ndimage.draw_lines`` on MI355X, a drop-in for cbclib_v2.ndimage.draw_lines
(cbclib_v2/ndimage.py:1 -> cbclib_v2/_src/functions.py:275-302), plus the
pixel-list mode :func:`draw_pixels`.

Same arguments, defaults, in-place semantics and exceptions as the reference:

* ``out`` (..., H, W) float32 or float64 is accumulated in place and returned;
* ``lines`` (..., 5) rows ``[x0, y0, x1, y1, width]`` (x = column), or
  (..., 7) rows ``[x0, y0, z0, x1, y1, z1, width]`` of 3D lines drawn into
  volumes (..., D, H, W) (draw_line_3d, bresenham.hpp:389-430);
* ``idxs`` frame of each line in the flattened leading dims of ``out``, or
  None (then there must be one frame or one line per frame,
  bresenham.cpp:5-20); wrong size -> ValueError, out of range -> IndexError
  (bresenham.cpp:22-36);
* ``kernel`` one of biweight, gaussian, parabolic, rectangular, triangular
  (unknown -> ValueError, include.hpp:495-503); ``overlap`` 'sum' or 'max'
  (else ValueError, bresenham.cpp:146-148).

Pixel sets are bit-identical to the reference CPU walk (bresenham.hpp:352-383);
values are the reference's per-pixel values, accumulated per LDS tile (sum
order differs from the reference's, within float rounding) or, with
``deterministic=True``, in the single-threaded reference's exact order (then
the image is bit-identical to ``bresenham.draw_lines(..., num_threads=1)``).
"""
from __future__ import annotations

import ctypes
from math import prod
from typing import Any, NamedTuple, Optional

import numpy as np
import torch

from . import _native, device
from ._arrays import NUMPY, dtype_of, inout, namespace, ptr, result, shape_of, stream_ptr, to_device

__all__ = ['accumulate_lines', 'draw_lines', 'draw_pixels', 'PixelList']


def _line_rows(lines: Any) -> tuple:
    """(number of lines, spatial dimensions): rows of 5 are 2D lines, rows of 7
    3D lines (draw_lines_2d_3d, bresenham.cpp:126-139)."""
    shape = shape_of(lines)
    L = shape[-1] if len(shape) else 0
    if L not in (5, 7):
        # bresenham.cpp:137 / cuda_draw_lines.cu:936
        raise RuntimeError(f"Invalid lines size ({L}) at axis {max(len(shape) - 1, 0)}")
    return prod(shape[:-1]), (L - 1) // 2


def _line_count(lines: Any) -> int:
    n_lines, ndim = _line_rows(lines)
    if ndim == 3:
        raise NotImplementedError("pixel lists of 3D lines (rows of 7) are not provided; use draw_lines")
    return n_lines


def _frames_of(oshape: tuple, ndim: int) -> tuple:
    """(n_frames, frame shape) of an output whose last ``ndim`` axes are a frame."""
    if len(oshape) < ndim:
        raise ValueError(f"Output array has insufficient number of dimensions ({len(oshape)} < {ndim})")
    return prod(oshape[:-ndim]), tuple(oshape[-ndim:])


def _check_idxs(idxs: Any, n_lines: int, n_frames: int) -> None:
    if idxs is None:
        if n_lines and n_frames not in (1, n_lines):
            raise ValueError(f"idxs has an icnompatible size ({n_lines} != {n_frames})")
        return
    n = prod(shape_of(idxs))
    if n != n_lines:
        raise ValueError(f"idxs has an invalid size ({n} != {n_lines})")


def _event_handle(ev: Any) -> Optional[int]:
    """hipEvent_t of a torch.cuda.Event (or a raw handle), None for None."""
    if ev is None:
        return None
    if isinstance(ev, int):
        return ev
    handle = getattr(ev, 'cuda_event', None)
    if not handle:
        raise ValueError("out_ready must be a recorded torch.cuda.Event or a hipEvent_t handle")
    return int(handle)


def _opts(tile: Optional[int], seg: Optional[int], deterministic: bool, threads: Optional[int] = None,
          exact_values: bool = False, chunk: Optional[int] = None, kernels_ptr: Optional[int] = None,
          out_ready: Any = None):
    return _native.DrawOpts(int(tile or 0), int(seg or 0), 1 if deterministic else 0, int(threads or 0),
                            1 if exact_values else 0, int(chunk or 0), kernels_ptr, _event_handle(out_ready))


def draw_lines(out: Any, lines: Any, idxs: Any = None, max_val: float = 1.0, kernel: str = 'rectangular',
               overlap: str = 'sum', *, tile: Optional[int] = None, seg: Optional[int] = None,
               deterministic: bool = False, threads: Optional[int] = None, exact_values: bool = False,
               chunk: Optional[int] = None, kernels: Any = None, out_ready: Any = None):
    """Draw thick lines with a profile into ``out`` (in place) and return ``out``.

    Extra keyword-only arguments (no reference counterpart): ``tile`` (32 or
    64, pixel tile per workgroup), ``seg`` (8 or 16, walk checkpoint spacing),
    ``threads`` (64, 128 or 256, tile workgroup size), ``deterministic``
    (reference summation order: bit-exact images), ``exact_values``
    (reference value arithmetic in the tile path), ``kernels`` (one profile
    per line, names or ids; ``kernel`` is then ignored; tile path only) and
    ``out_ready`` (a torch.cuda.Event recorded where ``out`` becomes ready on
    another stream: the draw waits for it only before it touches ``out``, so
    its line setup, walk and binning overlap that work; 2D lines, GPU
    tensors).
    """
    device.require_rocm('draw_lines')
    kid = _native.kernel_id(kernel)
    oid = _native.overlap_id(overlap)
    oshape = shape_of(out)
    odt = dtype_of(out)
    if odt not in (np.float32, np.float64):
        raise TypeError(f"out must be float32 or float64, got {odt}")
    n_lines, ndim = _line_rows(lines)
    n_frames, fshape = _frames_of(oshape, ndim)
    _check_idxs(idxs, n_lines, n_frames)
    f64 = odt == np.float64
    t_out, finish = inout(out, torch.float64 if f64 else torch.float32)
    if n_lines == 0 or t_out.numel() == 0:
        return finish()
    t_lines = to_device(lines, torch.float64 if f64 else torch.float32)
    t_idxs = None if idxs is None else to_device(idxs, torch.int64 if f64 else torch.int32)
    L = _native.lib()
    if ndim == 3:
        if kernels is not None:
            raise ValueError("kernels (one profile per line) is available for 2D lines only")
        fn = L.cbx_draw_lines3_f64 if f64 else L.cbx_draw_lines3_f32
        _native.check(fn(t_out.data_ptr(), n_frames, *fshape, t_lines.data_ptr(), n_lines, ptr(t_idxs),
                         float(max_val), kid, oid, stream_ptr()))
        return finish()
    H, W = fshape
    fn = L.cbx_draw_lines_f64 if f64 else L.cbx_draw_lines_f32
    t_kern = None
    if kernels is not None:
        t_kern = to_device(_native.kernel_ids(kernels), torch.uint8).reshape(-1)
        if t_kern.numel() != n_lines:
            raise ValueError(f"kernels has an invalid size ({t_kern.numel()} != {n_lines})")
    opts = _opts(tile, seg, deterministic, threads, exact_values, chunk, ptr(t_kern), out_ready)
    _native.check(fn(t_out.data_ptr(), n_frames, H, W, t_lines.data_ptr(), n_lines, ptr(t_idxs), float(max_val),
                     kid, oid, ctypes.byref(opts), stream_ptr()))
    return finish()


def accumulate_lines(out: Any, lines: Any, terms: Any, frames: Any, max_val: float = 1.0,
                     kernel: str = 'rectangular', in_overlap: str = 'sum', out_overlap: str = 'sum', *,
                     tile: Optional[int] = None, seg: Optional[int] = None, threads: Optional[int] = None,
                     chunk: Optional[int] = None):
    """Accumulate thick lines grouped into terms across frames (in place); return ``out``.

    Line ``i`` belongs to term ``terms[i]``; each term is drawn over its lines
    with ``in_overlap`` ('sum' or 'max', starting from zero) and then combined
    into frame ``frames[term]`` of ``out`` with ``out_overlap``.  Same
    arguments and errors as cbclib_v2.ndimage.accumulate_lines
    (functions.py:191-240, bresenham.cpp:151-309): a term outside
    ``[0, len(frames))`` or a frame outside ``[0, n_frames)`` raises
    IndexError, ``terms`` of a size other than the number of lines ValueError.
    ``tile``, ``seg``, ``threads`` and ``chunk`` are the draw_lines launch
    options (equal overlaps only).
    """
    device.require_rocm('accumulate_lines')
    kid = _native.kernel_id(kernel)
    if in_overlap not in ('sum', 'max') or out_overlap not in ('sum', 'max'):
        raise ValueError(f"Invalid in_overlap and out_overlap keywords: {in_overlap} and {out_overlap}")
    iid, oid = _native.overlap_id(in_overlap), _native.overlap_id(out_overlap)
    oshape = shape_of(out)
    odt = dtype_of(out)
    if odt not in (np.float32, np.float64):
        raise TypeError(f"out must be float32 or float64, got {odt}")
    n_lines, ndim = _line_rows(lines)
    n_frames, fshape = _frames_of(oshape, ndim)
    n_terms = prod(shape_of(frames))
    if prod(shape_of(terms)) != n_lines:
        raise ValueError(f"Term indices ({prod(shape_of(terms))}) is incompatible with number of lines "
                         f"({n_lines})")
    f64 = odt == np.float64
    it = torch.int64 if f64 else torch.int32
    t_out, finish = inout(out, torch.float64 if f64 else torch.float32)
    t_lines = to_device(lines, torch.float64 if f64 else torch.float32)
    t_terms = to_device(terms, it).reshape(-1)
    t_frames = to_device(frames, it).reshape(-1)
    L = _native.lib()
    if ndim == 3:
        fn = L.cbx_accumulate_lines3_f64 if f64 else L.cbx_accumulate_lines3_f32
        _native.check(fn(t_out.data_ptr(), n_frames, *fshape, t_lines.data_ptr(), n_lines, ptr(t_terms),
                         ptr(t_frames), n_terms, float(max_val), kid, iid, oid, stream_ptr()))
        return finish()
    fn = L.cbx_accumulate_lines_f64 if f64 else L.cbx_accumulate_lines_f32
    opts = _opts(tile, seg, False, threads, False, chunk)
    _native.check(fn(t_out.data_ptr(), n_frames, *fshape, t_lines.data_ptr(), n_lines, ptr(t_terms), ptr(t_frames),
                     n_terms, float(max_val), kid, iid, oid, ctypes.byref(opts), stream_ptr()))
    return finish()


class PixelList(NamedTuple):
    """Sparse render: entry i is pixel (frame[i], y[i], x[i]) with value[i].

    ``line[i]`` is the line that emitted the entry (raw lists) or the first
    line that touched the pixel (merged lists).
    """
    frame: Any
    y: Any
    x: Any
    value: Any
    line: Any


def draw_pixels(shape, lines: Any, idxs: Any = None, max_val: float = 1.0, kernel: str = 'rectangular',
                overlap: str = 'sum', merge: bool = True, dtype=np.float32) -> PixelList:
    """Pixel-list mode of :func:`draw_lines`.

    ``shape`` is the shape ``out`` would have.  ``merge=False`` returns every
    (line, pixel) entry in the order the reference emits them
    (bresenham.cpp:80-86); ``merge=True`` one entry per touched pixel sorted by
    (frame, y, x), reduced in ascending line order.  Indices are bit-identical
    to the reference walk.  Arrays are torch tensors on the GPU when ``lines``
    is a GPU array, numpy arrays otherwise.
    """
    device.require_rocm('draw_pixels')
    kid = _native.kernel_id(kernel)
    oid = _native.overlap_id(overlap)
    shape = tuple(int(s) for s in shape)
    if len(shape) < 2:
        raise ValueError("shape must have at least two dimensions")
    n_lines = _line_count(lines)
    n_frames, H, W = prod(shape[:-2]), shape[-2], shape[-1]
    _check_idxs(idxs, n_lines, n_frames)
    f64 = np.dtype(dtype) == np.float64
    tdt = torch.float64 if f64 else torch.float32
    like = namespace(lines)
    t_lines = to_device(lines, tdt)
    t_idxs = None if idxs is None else to_device(idxs, torch.int64 if f64 else torch.int32)
    L = _native.lib()
    fn = L.cbx_draw_pixels_f64 if f64 else L.cbx_draw_pixels_f32
    dev = t_lines.device
    count = ctypes.c_int64(0)
    cap = 0
    bufs = None
    for _ in range(3):
        bufs = [torch.empty(cap, dtype=torch.int32, device=dev) for _ in range(3)] + \
               [torch.empty(cap, dtype=tdt, device=dev), torch.empty(cap, dtype=torch.int32, device=dev)]
        _native.check(fn(n_frames, H, W, t_lines.data_ptr(), n_lines, ptr(t_idxs), float(max_val), kid, oid,
                         1 if merge else 0, *[b.data_ptr() for b in bufs], cap, ctypes.byref(count), stream_ptr()))
        if count.value <= cap:
            break
        cap = count.value
    n = count.value
    return PixelList(*[result(b[:n], like) for b in bufs])
