"""Device selection: the ``device`` module the reference documents
(cbclib_v2 README.md:46-65; functions.py:1-8) with a ROCm device.

``'rocm'`` (aliases ``'cuda'`` and ``'gpu'``, the names the reference and
PyTorch use for the GPU) runs every operation through the HIP library.  This
package has no CPU implementation: selecting ``'cpu'`` is allowed (so code
written for the reference keeps working) but the operations then raise
``RuntimeError`` instead of silently falling back to another path.

The selection is thread-local, like the reference's ``CPUConfig``
(cbclib_v2/_src/config.py:38-44).
"""
from __future__ import annotations

import threading
from contextlib import contextmanager
from typing import Iterator

_ALIASES = {'rocm': 'rocm', 'cuda': 'rocm', 'gpu': 'rocm', 'hip': 'rocm', 'cpu': 'cpu'}
_local = threading.local()


def _normalise(name: str) -> str:
    key = str(name).lower()
    if key not in _ALIASES:
        raise ValueError(f"Unknown device '{name}'. Supported devices are 'rocm' (alias 'cuda') and 'cpu'.")
    return _ALIASES[key]


def is_rocm_available() -> bool:
    """True when a GPU is visible to PyTorch-ROCm and the HIP library is built."""
    try:
        import torch
        from . import _native
        return bool(torch.cuda.is_available()) and _native.lib() is not None
    except Exception:
        return False


# the reference's spelling (README.md:53)
is_cuda_available = is_rocm_available


def get_device() -> str:
    return getattr(_local, 'device', 'rocm')


def set_device(name: str) -> None:
    _local.device = _normalise(name)


@contextmanager
def context(name: str) -> Iterator[str]:
    """Run the enclosed operations on ``name`` and restore the previous device."""
    old = get_device()
    set_device(name)
    try:
        yield get_device()
    finally:
        _local.device = old


def require_rocm(op: str) -> None:
    if get_device() != 'rocm':
        raise RuntimeError(f"{op}: the selected device is '{get_device()}', but cbclib_v2_amd implements {op} "
                           "only on ROCm (select it with device.set_device('rocm')).")
