"""cbx_label_noalloc (label with no device allocation: its intermediate
buffers live in a fixed device array) through the C ABI, against cbx_label
on the same inputs: label maps and counts bit-identical, on two calls in a
row with the second input smaller, for each label algorithm (run-length,
band and element union-find); an input that needs more than the fixed
scratch fails with ValueError and leaves the output untouched."""
import ctypes

import numpy as np
import pytest
import torch

from cbclib_v2_amd import _native
from cbclib_v2_amd import label as cl

pytestmark = pytest.mark.gpu


def call(name, inp: torch.Tensor, code: int, shifts: np.ndarray, npts: int):
    out = torch.full(tuple(inp.shape), -7, dtype=torch.int32, device=inp.device)
    shp = np.asarray(inp.shape, np.int64)
    sh = np.ascontiguousarray(shifts, np.int64).reshape(-1, inp.dim())
    n = ctypes.c_int64(-1)
    rc = getattr(_native.lib(), name)(inp.data_ptr(), code, inp.dim(), shp.ctypes.data, sh.ctypes.data,
                                      sh.shape[0], int(npts), out.data_ptr(), ctypes.byref(n),
                                      torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, out, n.value


def inputs(kind, shape, rng):
    if kind == 'bool':
        return torch.from_numpy(rng.random(shape) > 0.55).cuda(), _native.DTYPES['bool']
    a = rng.integers(0, 3, shape)
    if kind == 'int32':
        return torch.from_numpy(a.astype(np.int32)).cuda(), _native.DTYPES['int32']
    return torch.from_numpy(a.astype(np.int64)).cuda(), _native.DTYPES['int64']


# (input kind, structure, first shape, second (smaller) shape): run-length path
# (binary, horizontal neighbours), band path (equal-value 2D 3x3), element
# union-find (3D structure)
CASES = [('bool', ([0, 1, 1], 2), (3, 700, 900), (2, 300, 257)),
         ('int32', ([1, 1], 1), (900, 1100), (130, 77)),
         ('int64', ([1, 1, 1], 3), (24, 60, 70), (9, 13, 11))]


@pytest.mark.parametrize('case', CASES, ids=['runs', 'bands', 'elements'])
def test_noalloc_matches_label(cuda, case):
    kind, (radii, conn), first, second = case
    st = cl.Structure(radii, conn)
    rng = np.random.default_rng(sum(first))
    for shape, npts in ((first, 1), (second, 3)):
        inp, code = inputs(kind, shape, rng)
        print(f"noalloc {kind} {shape} npts={npts}", flush=True)
        rc0, ref, n0 = call('cbx_label', inp, code, st.shifts(), npts)
        _native.check(rc0)
        rc1, got, n1 = call('cbx_label_noalloc', inp, code, st.shifts(), npts)
        _native.check(rc1)
        assert n1 == n0
        assert torch.equal(got, ref)


def test_noalloc_rejects_large_inputs(cuda):
    cap = int(_native.lib().cbx_label_noalloc_capacity())
    assert cap == 64 << 20
    st = cl.Structure([0, 1, 1], 2)
    inp = torch.zeros((8, 2048, 2048), dtype=torch.bool, device='cuda')   # needs ~2.1 bytes per pixel
    rc, out, n = call('cbx_label_noalloc', inp, _native.DTYPES['bool'], st.shifts(), 1)
    assert rc == _native.CBX_EVALUE
    assert b'scratch' in _native.lib().cbx_last_error()
    assert bool((out == -7).all()), "a rejected call must not write its output"
    # and the entry point still works afterwards
    small = torch.ones((2, 64, 64), dtype=torch.bool, device='cuda')
    rc, out, n = call('cbx_label_noalloc', small, _native.DTYPES['bool'], st.shifts(), 1)
    assert rc == 0 and n == 2 and int(out.max()) == 2
