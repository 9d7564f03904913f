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
           'labels', 'index']


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
