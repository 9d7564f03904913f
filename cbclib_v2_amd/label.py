"""``label()`` on MI355X, a drop-in for cbclib_v2.label
(cbclib_v2/label.py:1-4 -> cbclib_v2/_src/functions.py:354-399, 481-507).

* :class:`Structure` mirrors the reference's neighbourhood class
  (label.hpp:291-313, binding label.cpp:303-393): the shifts of the
  (2r+1)^N box with 0 < L1 <= connectivity, the centre first, then row-major.
* :func:`label` returns, like the reference, a :class:`CPLabelResult`
  (int32 label map + index) for GPU arrays and a :class:`LabelResult`
  (regions) for NumPy arrays.  Ids are assigned in raster order of each
  component's smallest flat index and components smaller than ``npts`` are
  dropped (label.cpp:139-168; cuda_label.cu:294-357): the map is
  bit-identical to the reference's.
"""
from __future__ import annotations

import ctypes
from itertools import product
from typing import Any, Iterator, List, NamedTuple, Optional, Sequence, Union

import numpy as np
import torch

from . import _native, device
from ._arrays import NUMPY, TORCH, dtype_of, namespace, shape_of, stream_ptr, to_device

__all__ = ['Structure', 'Region', 'LabelResult', 'NPLabelResult', 'CPLabelResult', 'label', 'label_threshold',
           'binary_dilation', 'labels', 'index', 'center_of_mass', 'covariance_matrix', 'ellipse_fit', 'line_fit', 'p_values',
           'total_mass', 'mean', 'moment_of_inertia']


def _compute_index(axis: int, size: int, msg: str) -> int:
    if axis < -size or axis >= size:
        raise IndexError(msg)
    return axis % size


class Structure:
    """Pixel connectivity structure (label.hpp:135-339).

    Args:
        radii: radius of the neighbourhood box along each axis.
        connectivity: largest L1 distance of a neighbour from the centre.
    """

    def __init__(self, radii: Sequence[int], connectivity: int):
        self.radii = [int(r) for r in radii]
        self.connectivity = int(connectivity)
        self._shifts: List[List[int]] = [[0] * len(self.radii)]
        for point in product(*(range(2 * r + 1) for r in self.radii)):
            p = [c - r for c, r in zip(point, self.radii)]
            a = sum(abs(c) for c in p)
            if 0 < a <= self.connectivity:
                self._shifts.append(p)

    @property
    def rank(self) -> int:
        return len(self.radii)

    @property
    def shape(self) -> List[int]:
        return [2 * r + 1 for r in self.radii]

    def __iter__(self) -> Iterator[List[int]]:
        return iter([list(s) for s in self._shifts])

    def __len__(self) -> int:
        return len(self._shifts)

    def __repr__(self) -> str:
        return f"<Structure, connectivity = {self.connectivity}, rank = {self.rank}, size = {len(self)}>"

    @classmethod
    def from_reference(cls, structure: Any) -> 'Structure':
        """This package's Structure for a reference one (cbclib_v2.label.Structure).

        The reference binding exposes ``shape`` (2r + 1 per axis),
        ``connectivity`` and ``rank`` (the number of axes), not the radii
        (label.cpp:303-306), so the radii are recovered from ``shape``.
        """
        if isinstance(structure, cls):
            return structure
        return cls([d // 2 for d in structure.shape], structure.connectivity)

    def shifts(self) -> np.ndarray:
        """The non-zero shifts, (len(self) - 1, rank) int64."""
        return np.asarray(self._shifts[1:], dtype=np.int64).reshape(-1, self.rank)

    def squeeze(self) -> 'Structure':
        return Structure([d // 2 for d in self.shape if d > 1], self.connectivity)

    def expand_dims(self, axis: Union[int, Sequence[int]] = 0) -> 'Structure':
        radii = [d // 2 for d in self.shape]
        axes = [axis] if isinstance(axis, (int, np.integer)) else list(axis)
        n = len(radii) + len(axes)
        axes = sorted(_compute_index(int(a), n, "axis out of bounds for expand_dims") for a in axes)
        for a in axes:
            radii.insert(a, 0)
        return Structure(radii, self.connectivity)

    def to_array(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        if out is None:
            out = np.zeros(self.shape, dtype=bool)
            center = [d // 2 for d in self.shape]
        else:
            if out.ndim != self.rank:
                raise ValueError(f"output array dimension ({out.ndim}) does not match structure rank ({self.rank})")
            for n in range(out.ndim):
                if out.shape[n] < self.shape[n]:
                    raise ValueError(f"output array shape is smaller than structure shape at dimension {n}")
            center = [d // 2 for d in out.shape]
        for s in self._shifts:
            out[tuple(c + d for c, d in zip(s, center))] = True
        return out


class Region:
    """A set of flat indices (label.hpp:342-523), sorted ascending."""

    def __init__(self, indices: Optional[np.ndarray] = None):
        self._idx = np.zeros(0, np.int64) if indices is None else np.asarray(indices, np.int64)

    def __iter__(self) -> Iterator[int]:
        return iter(int(i) for i in self._idx)

    def __len__(self) -> int:
        return int(self._idx.size)

    def __repr__(self) -> str:
        return f"<Region, size = {len(self)}>"

    def first(self) -> int:
        return int(self._idx[0]) if self._idx.size else -1


class LabelResult:
    """Regions of a labelling (label.hpp:525-554, binding label.cpp:395-493)."""

    def __init__(self, shape: Sequence[int], label_map: Optional[np.ndarray] = None, n_labels: int = 0,
                 regions: Optional[List[Region]] = None):
        self._shape = [int(s) for s in shape]
        self._map = label_map
        self._n = int(n_labels) if regions is None else len(regions)
        self._regions = regions

    @property
    def shape(self) -> List[int]:
        return list(self._shape)

    @property
    def regions(self) -> List[Region]:
        if self._regions is None:
            flat = self._map.ravel()
            order = np.argsort(flat, kind='stable')
            sorted_ids = flat[order]
            bounds = np.searchsorted(sorted_ids, np.arange(1, self._n + 2))
            self._regions = [Region(order[bounds[k]:bounds[k + 1]]) for k in range(self._n)]
        return self._regions

    @regions.setter
    def regions(self, value: List[Region]) -> None:
        self._regions = list(value)
        self._n = len(self._regions)
        self._map = None

    def __len__(self) -> int:
        return self._n

    @classmethod
    def from_array(cls, labels: np.ndarray, index: Optional[np.ndarray] = None) -> 'LabelResult':
        labels = np.asarray(labels)
        flat = labels.ravel()
        if index is None:
            ids = [i for i in range(1, int(flat.max(initial=0)) + 1)]
            regions = [Region(np.flatnonzero(flat == i)) for i in ids]
            regions = [r for r in regions if len(r)]
        else:
            index = np.asarray(index)
            if np.any(index <= 0):
                raise ValueError("index array must contain positive label ids")
            regions = [Region(np.flatnonzero(flat == i)) for i in index]
        return cls(labels.shape, regions=regions)

    def to_array(self, index: Any, out: Optional[np.ndarray] = None) -> np.ndarray:
        index = np.asarray(index)
        if index.size != self._n:
            raise ValueError("Index array size does not match number of regions")
        if out is None:
            out = np.zeros(self._shape, dtype=np.int64)
        elif list(out.shape) != self._shape:
            raise ValueError("out array shape does not match region shape")
        if self._map is not None and self._regions is None:
            lut = np.zeros(self._n + 1, dtype=out.dtype)
            lut[1:] = index
            out[...] = lut[self._map]
            return out
        flat = out.reshape(-1)
        for r, v in zip(self.regions, index):
            flat[r._idx] = v
        return out


NPLabelResult = LabelResult


class CPLabelResult(NamedTuple):
    """GPU labelling result (functions.py:350-352): int32 map and ids."""
    labels: Any
    index: Any


def _shifts_for(structure: Structure, ndim: int) -> np.ndarray:
    structure = Structure.from_reference(structure)
    if structure.rank != ndim:
        # label.cpp:114-118
        raise ValueError(f"input array dimension ({ndim}) does not match structure rank ({structure.rank})")
    return np.ascontiguousarray(structure.shifts(), dtype=np.int64)


def _finish(t_out: torch.Tensor, n: int, like: str, shape):
    if like == NUMPY:
        return LabelResult(shape, t_out.cpu().numpy(), n)
    return CPLabelResult(labels=t_out, index=torch.arange(1, n + 1, dtype=torch.int64, device=t_out.device))


def label(inp: Any, structure: Structure, npts: int = 1):
    """Label connected components of equal-valued non-zero elements."""
    device.require_rocm('label')
    shape = shape_of(inp)
    shifts = _shifts_for(structure, len(shape))
    dt = dtype_of(inp)
    if dt == np.int32:
        t_in, code = to_device(inp, torch.int32), _native.DTYPES['int32']
    elif dt == np.int64:
        t_in, code = to_device(inp, torch.int64), _native.DTYPES['int64']
    else:
        t_in, code = to_device(inp, torch.bool), _native.DTYPES['bool']
    like = namespace(inp)
    t_out = torch.empty(shape, dtype=torch.int32, device=t_in.device)
    n = ctypes.c_int64(0)
    shp = np.asarray(shape, dtype=np.int64)
    _native.check(_native.lib().cbx_label(t_in.data_ptr(), code, len(shape), shp.ctypes.data, shifts.ctypes.data,
                                          shifts.shape[0], int(npts), t_out.data_ptr(), ctypes.byref(n),
                                          stream_ptr()))
    return _finish(t_out, n.value, like, shape)


def label_threshold(image: Any, vmin: float, structure: Structure, npts: int = 1):
    """``label(image > vmin, structure, npts)`` with the threshold fused into
    the labelling pass (the detection pipeline's pattern,
    data_processing.py:630-631).  ``image`` is cast to float32."""
    device.require_rocm('label_threshold')
    shape = shape_of(image)
    shifts = _shifts_for(structure, len(shape))
    t_in = to_device(image, torch.float32)
    like = namespace(image)
    t_out = torch.empty(shape, dtype=torch.int32, device=t_in.device)
    n = ctypes.c_int64(0)
    shp = np.asarray(shape, dtype=np.int64)
    _native.check(_native.lib().cbx_label_threshold_f32(t_in.data_ptr(), float(vmin), len(shape), shp.ctypes.data,
                                                        shifts.ctypes.data, shifts.shape[0], int(npts),
                                                        t_out.data_ptr(), ctypes.byref(n), stream_ptr()))
    return _finish(t_out, n.value, like, shape)


def binary_dilation(inp: Any, structure: Structure, iterations: int = 1, mask: Any = None):
    """Dilate a boolean array ``iterations`` times by ``structure``
    (functions.py:304-343 -> label.binary_dilation, label.cpp:8-107).

    Every true element of ``inp`` grows by the structure's shifts, inside the
    array; with ``mask`` an element joins only where the mask is true.
    Returns a new bool array (NumPy for NumPy input, a GPU tensor otherwise).
    ``iterations = 0`` returns a copy, as the reference's CPU module does.
    Non-boolean input is cast to bool.  A mask of another shape or a
    structure of another rank raises ValueError.
    """
    device.require_rocm('binary_dilation')
    shape = shape_of(inp)
    shifts = _shifts_for(structure, len(shape))
    iterations = int(iterations)
    if iterations < 0:
        raise ValueError(f"iterations must be non-negative, got {iterations}")
    t_in = to_device(inp, torch.bool)
    t_mask = None
    if mask is not None:
        if tuple(shape_of(mask)) != tuple(shape):
            # label.cpp:62-64
            raise ValueError("input and mask must have the same shape")
        t_mask = to_device(mask, torch.bool)
    t_out = torch.empty(shape, dtype=torch.bool, device=t_in.device)
    shp = np.asarray(shape, dtype=np.int64)
    _native.check(_native.lib().cbx_binary_dilation(t_in.data_ptr(), None if t_mask is None else t_mask.data_ptr(),
                                                    len(shape), shp.ctypes.data, shifts.ctypes.data,
                                                    shifts.shape[0], iterations, t_out.data_ptr(), stream_ptr()))
    return t_out.cpu().numpy() if namespace(inp) == NUMPY else t_out


def index(labels: Union[LabelResult, CPLabelResult]):
    """Array of region ids (functions.py:481-495)."""
    if isinstance(labels, LabelResult):
        return np.arange(1, len(labels) + 1, dtype=int)
    if isinstance(labels, CPLabelResult):
        return labels.index
    raise ValueError("Invalid labels type. Expected NPLabelResult or CPLabelResult.")


def labels(labels: Union[LabelResult, CPLabelResult]):
    """Dense label map (functions.py:497-507)."""
    if isinstance(labels, LabelResult):
        return labels.to_array(index(labels))
    if isinstance(labels, CPLabelResult):
        return labels.labels
    raise ValueError("Invalid labels type. Expected NPLabelResult or CPLabelResult.")


# ---------------------------------------------------------------------------
# region moments and p-values (functions.py:401-641; label.cpp:176-255,
# 534-583; cuda_label.cu:380-829)
# ---------------------------------------------------------------------------
def _label_map(lab: Union[LabelResult, CPLabelResult]):
    """(int32 label map, int64 ids) on the GPU, and the namespace of results."""
    if isinstance(lab, CPLabelResult):
        return to_device(lab.labels, torch.int32), to_device(lab.index, torch.int64).reshape(-1), namespace(lab.labels)
    if isinstance(lab, LabelResult):
        n = len(lab)
        m = lab.to_array(np.arange(1, n + 1), np.zeros(lab.shape, np.int32))
        return to_device(m, torch.int32), to_device(np.arange(1, n + 1, dtype=np.int64), torch.int64), NUMPY
    raise ValueError("Invalid labels type. Expected NPLabelResult or CPLabelResult.")


def _data_for(data: Any, t_map: torch.Tensor) -> torch.Tensor:
    shape = shape_of(data)
    if len(shape) != t_map.dim():
        # cuda_label.cu:495-499
        raise ValueError(f"labels array dimension ({t_map.dim()}) does not match data array dimension "
                         f"({len(shape)})")
    if tuple(shape) != tuple(t_map.shape):
        raise ValueError(f"data shape {tuple(shape)} does not match the label map shape {tuple(t_map.shape)}")
    dt = dtype_of(data)
    return to_device(data, torch.float32 if dt == np.float32 else torch.float64)


def _region_moments(what: str, lab: Union[LabelResult, CPLabelResult], data: Any):
    device.require_rocm(what)
    t_map, t_idx, like = _label_map(lab)
    t_data = _data_for(data, t_map)
    N = t_map.dim()
    n = t_idx.numel()
    M = {'total_mass': 1, 'mean': N, 'center_of_mass': N, 'moment_of_inertia': N * N,
         'covariance_matrix': N * N, 'line_fit': 2 * N}[what]
    out = torch.empty((n, M), dtype=t_data.dtype, device=t_data.device)
    f64 = t_data.dtype == torch.float64
    fn = _native.lib().cbx_region_moments_f64 if f64 else _native.lib().cbx_region_moments_f32
    shp = np.asarray(t_map.shape, dtype=np.int64)
    _native.check(fn(t_map.data_ptr(), t_data.data_ptr(), N, shp.ctypes.data, t_idx.data_ptr(), n,
                     _native.MOMENTS[what], out.data_ptr(), stream_ptr()))
    if what == 'total_mass':
        out = out.reshape(n)
    elif what in ('moment_of_inertia', 'covariance_matrix'):
        out = out.reshape(n, N, N)
    from ._arrays import result
    return result(out, like)


def total_mass(labels: Union[LabelResult, CPLabelResult], data: Any):
    """Sum of max(data, 0) over each region (label.cpp:534-541)."""
    return _region_moments('total_mass', labels, data)


def mean(labels: Union[LabelResult, CPLabelResult], data: Any):
    """Raw first moment sum max(data, 0) * p of each region, zyx order (label.cpp:543-549)."""
    return _region_moments('mean', labels, data)


def center_of_mass(labels: Union[LabelResult, CPLabelResult], data: Any):
    """Centre of mass of each region, (n, ndim) in the input's axis order
    (functions.py:401-439); the region's first pixel when its mass is zero."""
    return _region_moments('center_of_mass', labels, data)


def moment_of_inertia(labels: Union[LabelResult, CPLabelResult], data: Any):
    """Raw second moment matrix of each region, (n, ndim, ndim) (label.cpp:559-565)."""
    return _region_moments('moment_of_inertia', labels, data)


def covariance_matrix(labels: Union[LabelResult, CPLabelResult], data: Any):
    """Covariance matrix of each region, (n, ndim, ndim) in the input's axis
    order (functions.py:441-479); zero when the region's mass is zero."""
    return _region_moments('covariance_matrix', labels, data)


def _to_ellipse(matrix):
    """functions.py:534-545 on a numpy array or a torch tensor."""
    if isinstance(matrix, torch.Tensor):
        xp, atan, stack = torch, torch.atan, torch.stack
    else:
        xp, atan, stack = np, np.arctan, np.stack
    if matrix.shape[0] == 0:
        return xp.zeros((0, 3), dtype=matrix.dtype) if xp is np else torch.zeros((0, 3), dtype=matrix.dtype,
                                                                                   device=matrix.device)
    mu_xx, mu_xy, mu_yy = matrix[..., -1, -1], matrix[..., -1, -2], matrix[..., -2, -2]
    theta = 0.5 * atan(2 * mu_xy / (mu_xx - mu_yy))
    delta = xp.sqrt(4 * mu_xy ** 2 + (mu_xx - mu_yy) ** 2)
    a = xp.sqrt(2 * np.log(2) * (mu_xx + mu_yy + delta))
    b = xp.sqrt(2 * np.log(2) * (mu_xx + mu_yy - delta))
    return stack((a, b, theta), -1)


def ellipse_fit(labels: Union[LabelResult, CPLabelResult], data: Any):
    """FWHM major and minor axes and orientation (a, b, theta) of each region's
    covariance ellipse, (n, 3) (functions.py:547-580)."""
    return _to_ellipse(covariance_matrix(labels, data))


def line_fit(labels: Union[LabelResult, CPLabelResult], data: Any):
    """The major axis of each region's covariance ellipse as a segment,
    (n, 2 ndim) end points in xyz order (functions.py:582-620;
    CentralMomentsND::line, label.hpp:603-619)."""
    return _region_moments('line_fit', labels, data)


def p_values(labels: Union[LabelResult, CPLabelResult], lines: Any, data: Any, p0: float, vmin: float,
             xtol: float, counts: bool = False):
    """Log binomial tail of each region (functions.py:622-657, label.cpp:228-255):
    of the n pixels of region j within ``xtol`` of ``lines[j]`` (shape
    (n_regions, 2 ndim), xyz end points), k reach ``vmin``; returns
    log P(X >= k), X ~ Bin(n, p0).  ``counts=True`` also returns the (n, k)
    pairs (no reference counterpart)."""
    device.require_rocm('p_values')
    t_map, t_idx, like = _label_map(labels)
    t_data = _data_for(data, t_map)
    N = t_map.dim()
    lshape = shape_of(lines)
    if len(lshape) != 2 or lshape[1] != 2 * N:
        # label.cpp:246-249
        raise ValueError("lines array must have shape (n_lines, data.ndim() * 2)")
    n = t_idx.numel()
    if lshape[0] < n:
        raise ValueError(f"lines has {lshape[0]} rows for {n} regions")
    t_lines = to_device(lines, t_data.dtype)
    out = torch.empty(n, dtype=t_data.dtype, device=t_data.device)
    t_cnt = torch.empty((n, 2), dtype=torch.int64, device=t_data.device) if counts else None
    f64 = t_data.dtype == torch.float64
    fn = _native.lib().cbx_p_values_f64 if f64 else _native.lib().cbx_p_values_f32
    shp = np.asarray(t_map.shape, dtype=np.int64)
    _native.check(fn(t_map.data_ptr(), t_data.data_ptr(), N, shp.ctypes.data, t_idx.data_ptr(), n,
                     t_lines.data_ptr(), float(p0), float(vmin), float(xtol), out.data_ptr(),
                     None if t_cnt is None else t_cnt.data_ptr(), stream_ptr()))
    from ._arrays import result
    if counts:
        return result(out, like), result(t_cnt, like)
    return result(out, like)
