"""ctypes binding of ``libcbx.so`` (the C ABI declared in ``include/cbx.h``).

The library is the only compute path of this package: there is no CPU
fallback.  :func:`lib` raises ``RuntimeError`` when the shared library is
missing, so a broken build fails loudly instead of silently running
something else.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libcbx.so')

CBX_OK, CBX_EVALUE, CBX_EINDEX, CBX_ERUNTIME, CBX_ECAPACITY = 0, 1, 2, 3, 4
KERNELS = {'biweight': 0, 'gaussian': 1, 'parabolic': 2, 'rectangular': 3, 'triangular': 4}
OVERLAPS = {'sum': 0, 'max': 1}
DTYPES = {'bool': 0, 'int32': 1, 'int64': 2}

EXPORTS = ('cbx_last_error', 'cbx_version', 'cbx_kernel_id', 'cbx_draw_lines_f32', 'cbx_draw_lines_f64',
           'cbx_draw_pixels_f32', 'cbx_draw_pixels_f64', 'cbx_accumulate_lines_f32', 'cbx_accumulate_lines_f64',
           'cbx_draw_lines_adjoint_f32',
           'cbx_draw_lines_adjoint_f64', 'cbx_label', 'cbx_label_threshold_f32', 'cbx_timing_enable',
           'cbx_timing_reset', 'cbx_timing_read', 'cbx_workspace_release', 'cbx_workspace_bytes',
           'cbx_label_force_path', 'cbx_plan_create_f32', 'cbx_plan_draw_f32', 'cbx_plan_adjoint_f32',
           'cbx_plan_check', 'cbx_plan_bytes', 'cbx_plan_destroy', 'cbx_label_noalloc', 'cbx_label_noalloc_capacity', 'cbx_region_moments_f32',
           'cbx_region_moments_f64', 'cbx_p_values_f32', 'cbx_p_values_f64', 'cbx_draw_lines3_f32',
           'cbx_draw_lines3_f64', 'cbx_accumulate_lines3_f32', 'cbx_accumulate_lines3_f64',
           'cbx_binary_dilation', 'cbx_exclusive_scan_i32', 'cbx_exclusive_scan_i64')
MOMENTS = {'total_mass': 0, 'mean': 1, 'center_of_mass': 2, 'moment_of_inertia': 3, 'covariance_matrix': 4,
           'line_fit': 5}


class DrawOpts(ctypes.Structure):
    """``cbx_draw_opts``: tile side, columns per walk checkpoint, deterministic
    flag, tile-kernel workgroup size, exact-value flag, items per phase,
    per-line profile ids (device pointer or None) and the event `out` waits
    for (hipEvent_t or None)."""
    _fields_ = [('tile', ctypes.c_int32), ('seg', ctypes.c_int32), ('deterministic', ctypes.c_int32),
                ('threads', ctypes.c_int32), ('exact_values', ctypes.c_int32), ('chunk', ctypes.c_int32),
                ('kernels', ctypes.c_void_p), ('out_ready', ctypes.c_void_p)]


class CapacityError(RuntimeError):
    """A plan's segment or tile-item capacity was exceeded (create a larger plan)."""


_lib = None
_lock = threading.Lock()


def _declare(L: ctypes.CDLL) -> None:
    P, I64, I32, F, D, C = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_double,
                            ctypes.c_char_p)
    OP = ctypes.POINTER(DrawOpts)
    L.cbx_last_error.restype = C
    L.cbx_last_error.argtypes = []
    L.cbx_version.restype = C
    L.cbx_version.argtypes = []
    L.cbx_kernel_id.restype = I32
    L.cbx_kernel_id.argtypes = [C]
    for name, R in (('f32', F), ('f64', D)):
        fn = getattr(L, 'cbx_draw_lines_' + name)
        fn.restype = I32
        fn.argtypes = [P, I64, I64, I64, P, I64, P, R, I32, I32, OP, P]
        fn = getattr(L, 'cbx_draw_pixels_' + name)
        fn.restype = I32
        fn.argtypes = [I64, I64, I64, P, I64, P, R, I32, I32, I32, P, P, P, P, P, I64,
                       ctypes.POINTER(ctypes.c_int64), P]
        fn = getattr(L, 'cbx_accumulate_lines_' + name)
        fn.restype = I32
        fn.argtypes = [P, I64, I64, I64, P, I64, P, P, I64, R, I32, I32, I32, OP, P]
        fn = getattr(L, 'cbx_draw_lines_adjoint_' + name)
        fn.restype = I32
        fn.argtypes = [P, I64, I64, I64, P, I64, P, R, I32, P, OP, P]
        fn = getattr(L, 'cbx_draw_lines3_' + name)
        fn.restype = I32
        fn.argtypes = [P, I64, I64, I64, I64, P, I64, P, R, I32, I32, P]
        fn = getattr(L, 'cbx_accumulate_lines3_' + name)
        fn.restype = I32
        fn.argtypes = [P, I64, I64, I64, I64, P, I64, P, P, I64, R, I32, I32, I32, P]
    L.cbx_label.restype = I32
    L.cbx_label.argtypes = [P, I32, I32, P, P, I64, I64, P, ctypes.POINTER(ctypes.c_int64), P]
    for name, R in (('f32', F), ('f64', D)):
        fn = getattr(L, 'cbx_region_moments_' + name)
        fn.restype = I32
        fn.argtypes = [P, P, I32, P, P, I64, I32, P, P]
        fn = getattr(L, 'cbx_p_values_' + name)
        fn.restype = I32
        fn.argtypes = [P, P, I32, P, P, I64, P, R, R, R, P, P, P]
    L.cbx_binary_dilation.restype = I32
    L.cbx_binary_dilation.argtypes = [P, P, I32, P, P, I64, I64, P, P]
    L.cbx_label_noalloc.restype = I32
    L.cbx_label_noalloc.argtypes = [P, I32, I32, P, P, I64, I64, P, ctypes.POINTER(ctypes.c_int64), P]
    L.cbx_label_noalloc_capacity.restype = I64
    L.cbx_label_noalloc_capacity.argtypes = []
    L.cbx_label_threshold_f32.restype = I32
    L.cbx_label_threshold_f32.argtypes = [P, F, I32, P, P, I64, I64, P, ctypes.POINTER(ctypes.c_int64), P]
    L.cbx_timing_enable.restype = None
    L.cbx_timing_enable.argtypes = [I32]
    L.cbx_timing_reset.restype = None
    L.cbx_timing_reset.argtypes = []
    L.cbx_timing_read.restype = I32
    L.cbx_timing_read.argtypes = [C, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    L.cbx_plan_create_f32.restype = I32
    L.cbx_plan_create_f32.argtypes = [ctypes.POINTER(ctypes.c_void_p), I64, I64, I64, P, I64, P, P, I32, OP, D, P]
    L.cbx_plan_draw_f32.restype = I32
    L.cbx_plan_draw_f32.argtypes = [P, P, P, I64, P, P, F, I32, I32, P]
    L.cbx_exclusive_scan_i32.restype = I32
    L.cbx_exclusive_scan_i32.argtypes = [P, P, I64, I32, P]
    L.cbx_exclusive_scan_i64.restype = I32
    L.cbx_exclusive_scan_i64.argtypes = [P, P, I64, I32, P]
    L.cbx_plan_adjoint_f32.restype = I32
    L.cbx_plan_adjoint_f32.argtypes = [P, P, P, I64, P, F, I32, P, P]
    L.cbx_plan_check.restype = I32
    L.cbx_plan_check.argtypes = [P, P]
    L.cbx_plan_bytes.restype = I64
    L.cbx_plan_bytes.argtypes = [P]
    L.cbx_plan_destroy.restype = None
    L.cbx_plan_destroy.argtypes = [P]
    L.cbx_label_force_path.restype = None
    L.cbx_label_force_path.argtypes = [I32]
    L.cbx_workspace_release.restype = I64
    L.cbx_workspace_release.argtypes = []
    L.cbx_workspace_bytes.restype = I64
    L.cbx_workspace_bytes.argtypes = []


def lib() -> ctypes.CDLL:
    """The loaded library; raises RuntimeError if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(f"cbclib_v2_amd: the HIP library {LIB_PATH} is not built; run "
                                       "`python -c 'import __graft_entry__ as g; g.build()'` or "
                                       "`make -C cbclib_v2_amd/csrc`")
                L = ctypes.CDLL(LIB_PATH)
                _declare(L)
                _lib = L
    return _lib


def check(status: int) -> None:
    """Raise the exception the reference raises for a failed call."""
    if status == CBX_OK:
        return
    msg = lib().cbx_last_error().decode(errors='replace')
    if status == CBX_EVALUE:
        raise ValueError(msg)
    if status == CBX_EINDEX:
        raise IndexError(msg)
    if status == CBX_ECAPACITY:
        raise CapacityError(msg)
    raise RuntimeError(msg)


def kernel_id(kernel: str) -> int:
    if kernel not in KERNELS:
        # kernels::get_type (include.hpp:495-503)
        raise ValueError(f"kernel is missing for {kernel}")
    return KERNELS[kernel]


def kernel_ids(kernels) -> 'object':
    """Per-line profiles as uint8 ids: names, ids, or an integer array."""
    import numpy as np
    if hasattr(kernels, 'dtype') and not isinstance(kernels, np.ndarray):
        return kernels                      # a device tensor of ids
    arr = np.asarray(kernels)
    if arr.dtype.kind in 'US':
        arr = np.array([kernel_id(str(k)) for k in arr.ravel()], np.uint8).reshape(arr.shape)
    if arr.size and (arr.min() < 0 or arr.max() > 4):
        raise ValueError("kernel is missing for a per-line profile id")
    return arr.astype(np.uint8)


def overlap_id(overlap: str) -> int:
    if overlap not in OVERLAPS:
        # bresenham.cpp:148
        raise ValueError(f"Invalid overlap keyword: {overlap}")
    return OVERLAPS[overlap]


def timing_enable(on: bool = True, tiles_only: bool = False) -> None:
    """Record HIP events around the library's main kernels (see cbx_timing_read);
    ``tiles_only``: around the tile kernels only (``draw_tiles``, ``adjoint_tiles``, ...)."""
    lib().cbx_timing_enable((2 if tiles_only else 1) if on else 0)


def timing_reset() -> None:
    lib().cbx_timing_reset()


def timing_read(name: str):
    """(total milliseconds, launches) of kernel ``name`` since the last reset."""
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    lib().cbx_timing_read(name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def workspace_release() -> int:
    """Free the library's cached intermediate device memory; returns the bytes freed."""
    return int(lib().cbx_workspace_release())


def workspace_bytes() -> int:
    """Bytes of intermediate device memory the library keeps cached for reuse."""
    return int(lib().cbx_workspace_bytes())


def label_force_path(path: int) -> None:
    """Label algorithm: 0 fastest applicable, 1 no run-length path, 2 element union-find only."""
    lib().cbx_label_force_path(int(path))
