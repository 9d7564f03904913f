"""Array plumbing between the caller's arrays and HBM.

The reference dispatches on the array namespace of its argument
(``array_dispatch``, cbclib_v2/_src/functions.py:25-80): NumPy arrays go to
the CPU modules, CuPy arrays to CUDA.  Here every namespace goes to the HIP
library:

* ``torch.Tensor`` on the GPU and objects exposing ``__cuda_array_interface__``
  are used in place (zero copy) when they already have the dtype and layout
  the kernel needs;
* NumPy arrays (and CPU tensors) are staged into HBM, processed there and, for
  in-place outputs, copied back into the caller's array, so ``out`` is still
  modified in place and returned, as in the reference.
"""
from __future__ import annotations

from typing import Any, Callable, Optional, Tuple

import numpy as np
import torch

NUMPY, TORCH = 'numpy', 'torch'


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("cbclib_v2_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device('cuda', torch.cuda.current_device())


def namespace(x: Any) -> str:
    if isinstance(x, torch.Tensor) or hasattr(x, '__cuda_array_interface__'):
        return TORCH
    return NUMPY


def to_device(x: Any, dtype: torch.dtype) -> torch.Tensor:
    """Contiguous GPU tensor of ``dtype`` holding ``x`` (a view when possible)."""
    dev = _device()
    if isinstance(x, torch.Tensor):
        t = x
    elif hasattr(x, '__cuda_array_interface__'):
        t = torch.as_tensor(x, device=dev)
    else:
        a = np.asarray(x)
        t = torch.from_numpy(np.ascontiguousarray(a))
    if t.device != dev or t.dtype != dtype:
        t = t.to(device=dev, dtype=dtype)
    return t.contiguous()


def inout(x: Any, dtype: torch.dtype) -> Tuple[torch.Tensor, Callable[[], Any]]:
    """GPU tensor for an in-place output and a ``finish()`` that writes it back.

    ``finish()`` returns the caller's object (``out`` is returned as in the
    reference's bindings, cupy_array.hpp:758-763).
    """
    dev = _device()
    if isinstance(x, torch.Tensor):
        if x.device == dev and x.dtype == dtype and x.is_contiguous():
            return x, lambda: x
        t = x.detach().to(device=dev, dtype=dtype).contiguous()

        def finish_t():
            with torch.no_grad():
                x.copy_(t)
            return x
        return t, finish_t
    if hasattr(x, '__cuda_array_interface__'):
        view = torch.as_tensor(x, device=dev)
        if view.dtype == dtype and view.is_contiguous():
            return view, lambda: x
        t = view.to(dtype=dtype).contiguous()

        def finish_c():
            view.copy_(t)
            return x
        return t, finish_c
    if not isinstance(x, np.ndarray):
        raise TypeError(f"out must be an array, got {type(x).__name__}")
    if not x.flags.writeable:
        raise ValueError("out is read-only")
    t = torch.from_numpy(np.ascontiguousarray(x)).to(device=dev, dtype=dtype)

    def finish_n():
        x[...] = t.cpu().numpy()
        return x
    return t, finish_n


def result(t: torch.Tensor, like: str):
    """Return a library-made array in the caller's namespace."""
    if like == NUMPY:
        return t.cpu().numpy()
    return t


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: Optional[torch.Tensor]):
    return None if t is None else t.data_ptr()


def shape_of(x: Any) -> Tuple[int, ...]:
    if isinstance(x, torch.Tensor):
        return tuple(x.shape)
    if hasattr(x, '__cuda_array_interface__'):
        return tuple(x.__cuda_array_interface__['shape'])
    return tuple(np.shape(x))


def dtype_of(x: Any):
    """The element type of ``x`` as a numpy dtype."""
    if isinstance(x, torch.Tensor):
        return torch.empty((), dtype=x.dtype).numpy().dtype
    if hasattr(x, '__cuda_array_interface__'):
        return np.dtype(x.__cuda_array_interface__['typestr'])
    return np.asarray(x).dtype if not isinstance(x, np.ndarray) else x.dtype


TORCH_DTYPES = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
                np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64, np.dtype(np.bool_): torch.bool,
                np.dtype(np.uint8): torch.uint8}
