"""Batched drawing: plans, HIP-graph replay and the launch-shape autotuner.

No reference counterpart (the reference draws with one call per profile and
synchronises after every call).  Many frames with uneven line counts and
per-line profiles are drawn by ONE tile-path launch sequence over all of
their tiles:

* :class:`Plan` sizes the intermediate buffers once on a set of lines
  (``cbx_plan_create_f32``) and then draws with no host synchronisation,
  allocation or memset (``cbx_plan_draw_f32``), so a draw can be captured;
* :class:`GraphDraw` captures ``out.zero_()`` + one plan draw in a HIP graph
  (``torch.cuda.CUDAGraph``) and replays it;
* :func:`tune` times the tile-path launch shapes on the caller's lines and
  returns (and caches) the fastest.

Float32 images only.  Errors of a draw (bad ``idxs`` or profile ids, capacity
exceeded) leave ``out`` untouched and are raised by :meth:`Plan.check`.
"""
from __future__ import annotations

import ctypes
from math import prod
from typing import Any, Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native, device
from ._arrays import ptr, shape_of, stream_ptr, to_device

__all__ = ['Plan', 'GraphDraw', 'tune', 'CANDIDATES']

# launch shapes the autotuner tries: (tile, seg, threads, chunk)
CANDIDATES: Tuple[Tuple[int, int, int, int], ...] = (
    (64, 8, 256, 64), (64, 16, 256, 64), (64, 8, 128, 64), (32, 8, 256, 64), (32, 16, 256, 64), (32, 8, 128, 64),
    (64, 8, 256, 256), (64, 16, 256, 256),
)
_tuned: Dict[tuple, Dict[str, Any]] = {}


def _frames(shape: Sequence[int]) -> Tuple[int, int, int]:
    shape = tuple(int(s) for s in shape)
    if len(shape) < 2:
        raise ValueError("shape must have at least two dimensions")
    return prod(shape[:-2]), shape[-2], shape[-1]


def _lines_args(lines: Any, idxs: Any, kernels: Any):
    t_lines = to_device(lines, torch.float32).reshape(-1, 5)
    t_idxs = None if idxs is None else to_device(idxs, torch.int32).reshape(-1)
    t_kern = None if kernels is None else to_device(_native.kernel_ids(kernels), torch.uint8).reshape(-1)
    n = t_lines.shape[0]
    if t_idxs is not None and t_idxs.numel() != n:
        raise ValueError(f"idxs has an invalid size ({t_idxs.numel()} != {n})")
    if t_kern is not None and t_kern.numel() != n:
        raise ValueError(f"kernels has an invalid size ({t_kern.numel()} != {n})")
    return t_lines, t_idxs, t_kern


class Plan:
    """Capacities for drawing up to ``len(lines)`` lines into frames of ``shape``.

    Args:
        shape: shape of the images drawn into, (..., H, W).
        lines, idxs, kernels: a representative line set (end points, frames,
            per-line profiles) on which the buffers are sized.
        kernel: profile of every line when ``kernels`` is None.
        headroom: capacity factor over the representative set (>= 1); draws of
            other line sets may need up to this much more.
        tile, seg, threads, chunk: tile-path launch shape (see ``tune``).
    """

    def __init__(self, shape: Sequence[int], lines: Any, idxs: Any = None, kernels: Any = None,
                 kernel: str = 'rectangular', headroom: float = 1.25, tile: Optional[int] = None,
                 seg: Optional[int] = None, threads: Optional[int] = None, chunk: Optional[int] = None,
                 exact_values: bool = False):
        device.require_rocm('Plan')
        self.shape = tuple(int(s) for s in shape)
        F, H, W = _frames(self.shape)
        t_lines, t_idxs, t_kern = _lines_args(lines, idxs, kernels)
        self.n_lines = t_lines.shape[0]
        self.kernel = kernel
        self.draws = 0   # draws so far (render_plan's backward checks that none came between)
        self.opts = _native.DrawOpts(int(tile or 0), int(seg or 0), 0, int(threads or 0), 1 if exact_values else 0,
                                     int(chunk or 0), None)
        self._ptr = ctypes.c_void_p()
        L = _native.lib()
        _native.check(L.cbx_plan_create_f32(ctypes.byref(self._ptr), F, H, W, t_lines.data_ptr(), self.n_lines,
                                            ptr(t_idxs), ptr(t_kern), _native.kernel_id(kernel),
                                            ctypes.byref(self.opts), float(headroom), stream_ptr()))

    @property
    def nbytes(self) -> int:
        """Device bytes the plan holds."""
        return int(_native.lib().cbx_plan_bytes(self._ptr)) if self._ptr else 0

    def draw(self, out: torch.Tensor, lines: Any, idxs: Any = None, kernels: Any = None, max_val: float = 1.0,
             kernel: Optional[str] = None, overlap: str = 'sum') -> torch.Tensor:
        """Accumulate the lines into ``out`` (float32 CUDA tensor of the plan's
        shape) on the current stream; returns ``out`` without waiting."""
        if not self._ptr:
            raise RuntimeError("the plan is closed")
        if not isinstance(out, torch.Tensor) or out.dtype != torch.float32 or not out.is_cuda:
            raise TypeError("out must be a float32 CUDA tensor")
        if tuple(out.shape) != self.shape or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous tensor of shape {self.shape}")
        t_lines, t_idxs, t_kern = _lines_args(lines, idxs, kernels)
        kid = _native.kernel_id(kernel or self.kernel)
        self.draws += 1
        _native.check(_native.lib().cbx_plan_draw_f32(self._ptr, out.data_ptr(), t_lines.data_ptr(),
                                                      t_lines.shape[0], ptr(t_idxs), ptr(t_kern), float(max_val),
                                                      kid, _native.overlap_id(overlap), stream_ptr()))
        return out

    def adjoint(self, grad_out: torch.Tensor, lines: Any, idxs: Any = None, max_val: float = 1.0,
                kernel: Optional[str] = None) -> torch.Tensor:
        """Gradient of ``<grad_out, draw(lines)>`` with respect to the lines'
        (x0, y0, x1, y1, width), shape (N, 5), as ``autograd.adjoint`` but on
        the segments and tile lists of the plan's last draw, which must have
        drawn the same ``lines``, ``idxs`` and profile (one profile for every
        line).  On the current stream, without waiting; errors of that draw
        leave the gradient zero and are raised by :meth:`check`."""
        if not self._ptr:
            raise RuntimeError("the plan is closed")
        if not isinstance(grad_out, torch.Tensor) or grad_out.dtype != torch.float32 or not grad_out.is_cuda:
            raise TypeError("grad_out must be a float32 CUDA tensor")
        if tuple(grad_out.shape) != self.shape:
            raise ValueError(f"grad_out must have the plan's shape {self.shape}")
        g = grad_out.contiguous()
        t_lines, t_idxs, _ = _lines_args(lines, idxs, None)
        res = torch.empty((t_lines.shape[0], 5), dtype=torch.float32, device=t_lines.device)
        kid = _native.kernel_id(kernel or self.kernel)
        _native.check(_native.lib().cbx_plan_adjoint_f32(self._ptr, g.data_ptr(), t_lines.data_ptr(),
                                                         t_lines.shape[0], ptr(t_idxs), float(max_val), kid,
                                                         res.data_ptr(), stream_ptr()))
        return res

    def check(self) -> None:
        """Wait for the stream and raise the error of the last draw, if any
        (:class:`_native.CapacityError` when the plan was too small)."""
        if self._ptr:
            _native.check(_native.lib().cbx_plan_check(self._ptr, stream_ptr()))

    def close(self) -> None:
        if self._ptr:
            _native.lib().cbx_plan_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GraphDraw:
    """``out.zero_()`` (optional) + one plan draw, captured in a HIP graph.

    The tensors passed here are the graph's inputs: update them in place
    (``lines.copy_(...)``) between replays.  Call :meth:`Plan.check` after
    replays to see draw errors.
    """

    def __init__(self, plan: Plan, out: torch.Tensor, lines: torch.Tensor, idxs: Optional[torch.Tensor] = None,
                 kernels: Optional[torch.Tensor] = None, max_val: float = 1.0, kernel: Optional[str] = None,
                 overlap: str = 'sum', zero: bool = True):
        self.plan, self.out = plan, out
        self.args = (lines, idxs, kernels)   # the graph reads these tensors
        _native.lib()                        # load the library before capture

        def body():
            if zero:
                out.zero_()
            plan.draw(out, lines, idxs, kernels, max_val, kernel, overlap)

        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            body()                       # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        plan.check()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            body()

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.out


def _signature(shape, n_lines, kernels, kernel) -> tuple:
    mix = 'mixed' if kernels is not None else kernel
    return (tuple(int(s) for s in shape), int(np.ceil(np.log2(max(n_lines, 1)))), mix)


def tune(shape: Sequence[int], lines: Any, idxs: Any = None, kernels: Any = None, kernel: str = 'rectangular',
         candidates: Sequence[Tuple[int, int, int, int]] = CANDIDATES, repeats: int = 3,
         use_cache: bool = True) -> Dict[str, Any]:
    """Fastest tile-path launch shape for drawing these lines into ``shape``.

    Each candidate (tile, seg, threads, chunk) gets a plan and one warm-up
    draw, then ``repeats`` timed draws (HIP events on the current stream).
    Returns ``{'tile', 'seg', 'threads', 'chunk', 'ms', 'table'}`` and caches
    it per (shape, power-of-two line count, profile).
    """
    device.require_rocm('tune')
    key = _signature(shape, prod(shape_of(lines)[:-1]), kernels, kernel)
    if use_cache and key in _tuned:
        return _tuned[key]
    t_lines, t_idxs, t_kern = _lines_args(lines, idxs, kernels)
    out = torch.zeros(tuple(shape), dtype=torch.float32, device=t_lines.device)
    table = []
    for tile, seg, threads, chunk in candidates:
        plan = Plan(shape, t_lines, t_idxs, t_kern, kernel, 1.0, tile, seg, threads, chunk)
        plan.draw(out, t_lines, t_idxs, t_kern, kernel=kernel)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeats):
            plan.draw(out, t_lines, t_idxs, t_kern, kernel=kernel)
        stop.record()
        plan.check()
        table.append({'tile': tile, 'seg': seg, 'threads': threads, 'chunk': chunk,
                      'ms': start.elapsed_time(stop) / repeats})
        plan.close()
    best = dict(min(table, key=lambda r: r['ms']))
    best['table'] = table
    _tuned[key] = best
    return best
