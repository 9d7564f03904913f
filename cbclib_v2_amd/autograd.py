"""Differentiable rendering: ``torch.autograd.Function`` over draw_lines and
its hand-written adjoint kernel.

``render(lines, shape, idxs, ...)`` draws the lines (sum overlap) into a fresh
zero image; its backward scatters dL/dI back to each line's end points and
width through the profile derivative:

    dL/dtheta_l = sum_{p drawn by l} dL/dI[p] * max_val * dK/du * du/dtheta_l,
    u = dist(p, segment)^2 / w^2,

with d(dist^2)/dp0 = -2 (1 - t*) (p - q), d(dist^2)/dp1 = -2 t* (p - q)
(q = the clamped projection of p, t* its segment parameter) and
du/dw = -2 u / w.  The pixel set is the forward's (the walk is recomputed
from the line end points, not stored).  The reference has no gradient of
draw_lines; the forward is cbclib_v2's (bresenham.hpp:352-383).
"""
from __future__ import annotations

import ctypes
from math import prod
from typing import Optional, Sequence

import torch

from . import _native, device
from ._arrays import stream_ptr


def _call_forward(out: torch.Tensor, lines: torch.Tensor, idxs: Optional[torch.Tensor], max_val: float,
                  kid: int, opts) -> None:
    L = _native.lib()
    f64 = out.dtype == torch.float64
    fn = L.cbx_draw_lines_f64 if f64 else L.cbx_draw_lines_f32
    H, W = out.shape[-2], out.shape[-1]
    F = prod(out.shape[:-2])
    _native.check(fn(out.data_ptr(), F, H, W, lines.data_ptr(), lines.shape[0],
                     None if idxs is None else idxs.data_ptr(), float(max_val), kid, 0, ctypes.byref(opts),
                     stream_ptr()))


def adjoint(grad_image: torch.Tensor, lines: torch.Tensor, idxs: Optional[torch.Tensor] = None,
            max_val: float = 1.0, kernel: str = 'biweight', tile: Optional[int] = None,
            seg: Optional[int] = None) -> torch.Tensor:
    """d<grad_image, draw_lines(lines)>/d lines, shape (N, 5)."""
    device.require_rocm('adjoint')
    kid = _native.kernel_id(kernel)
    g = grad_image.contiguous()
    f64 = g.dtype == torch.float64
    ln = lines.detach().to(dtype=g.dtype).contiguous().reshape(-1, 5)
    ix = None if idxs is None else idxs.to(dtype=torch.int64 if f64 else torch.int32).contiguous()
    out = torch.empty_like(ln)
    L = _native.lib()
    fn = L.cbx_draw_lines_adjoint_f64 if f64 else L.cbx_draw_lines_adjoint_f32
    opts = _native.DrawOpts(int(tile or 0), int(seg or 0), 0, 0)
    H, W = g.shape[-2], g.shape[-1]
    F = prod(g.shape[:-2])
    _native.check(fn(g.data_ptr(), F, H, W, ln.data_ptr(), ln.shape[0], None if ix is None else ix.data_ptr(),
                     float(max_val), kid, out.data_ptr(), ctypes.byref(opts), stream_ptr()))
    return out


class DrawLinesFunction(torch.autograd.Function):
    """image = draw_lines(zeros(shape), lines, idxs, max_val, kernel, 'sum')."""

    @staticmethod
    def forward(ctx, lines: torch.Tensor, idxs: Optional[torch.Tensor], shape: Sequence[int], max_val: float,
                kernel: str, tile: Optional[int], seg: Optional[int]):
        kid = _native.kernel_id(kernel)
        ln = lines.detach().contiguous().reshape(-1, 5)
        f64 = ln.dtype == torch.float64
        ix = None if idxs is None else idxs.to(dtype=torch.int64 if f64 else torch.int32).contiguous()
        out = torch.zeros(tuple(shape), dtype=ln.dtype, device=ln.device)
        _call_forward(out, ln, ix, max_val, kid, _native.DrawOpts(int(tile or 0), int(seg or 0), 0, 0))
        ctx.save_for_backward(ln, ix if ix is not None else torch.empty(0))
        ctx.has_idxs = ix is not None
        ctx.max_val, ctx.kernel, ctx.tile, ctx.seg = max_val, kernel, tile, seg
        ctx.lines_shape = lines.shape
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        ln, ix = ctx.saved_tensors
        g = adjoint(grad_out, ln, ix if ctx.has_idxs else None, ctx.max_val, ctx.kernel, ctx.tile, ctx.seg)
        return g.reshape(ctx.lines_shape), None, None, None, None, None, None


class PlanDrawFunction(torch.autograd.Function):
    """image = plan.draw(zeros(plan.shape), lines, idxs, max_val, kernel);
    the backward is ``plan.adjoint`` on the forward's segments."""

    @staticmethod
    def forward(ctx, lines: torch.Tensor, idxs: Optional[torch.Tensor], plan, max_val: float, kernel: str):
        ln = lines.detach().to(torch.float32).contiguous().reshape(-1, 5)
        ix = None if idxs is None else idxs.to(torch.int32).contiguous()
        out = torch.zeros(plan.shape, dtype=torch.float32, device=ln.device)
        plan.draw(out, ln, ix, max_val=max_val, kernel=kernel)
        ctx.save_for_backward(ln, ix if ix is not None else torch.empty(0))
        ctx.plan, ctx.draw_no, ctx.has_idxs = plan, plan.draws, ix is not None
        ctx.max_val, ctx.kernel, ctx.lines_shape = max_val, kernel, lines.shape
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        ln, ix = ctx.saved_tensors
        if ctx.plan.draws != ctx.draw_no:
            raise RuntimeError("the plan drew again between this forward and its backward")
        g = ctx.plan.adjoint(grad_out.to(torch.float32), ln, ix if ctx.has_idxs else None, ctx.max_val,
                             ctx.kernel)
        return g.reshape(ctx.lines_shape).to(grad_out.dtype), None, None, None, None


def render_plan(plan, lines: torch.Tensor, idxs: Optional[torch.Tensor] = None, max_val: float = 1.0,
                kernel: str = 'biweight') -> torch.Tensor:
    """Differentiable draw through a :class:`batch.Plan` (float32, sum
    overlap): no host synchronisation, and the backward reuses the forward's
    segments and tile lists instead of walking the lines again.  The
    backward must run before the plan draws again; a draw that exceeded the
    plan's capacity is raised by ``plan.check()``."""
    device.require_rocm('render_plan')
    return PlanDrawFunction.apply(lines, idxs, plan, float(max_val), kernel)


def render(lines: torch.Tensor, shape: Sequence[int], idxs: Optional[torch.Tensor] = None, max_val: float = 1.0,
           kernel: str = 'biweight', tile: Optional[int] = None, seg: Optional[int] = None) -> torch.Tensor:
    """Differentiable draw_lines into a zero image of ``shape`` (sum overlap)."""
    device.require_rocm('render')
    return DrawLinesFunction.apply(lines, idxs, tuple(shape), float(max_val), kernel, tile, seg)
