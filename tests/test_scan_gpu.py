"""The device-wide exclusive scan (one kernel, decoupled look-back) that the
draw, adjoint and label pipelines use, against NumPy's cumulative sums: sizes
inside one tile (2048 entries), across tiles, at the switch to 4096-entry
tiles above 512k entries, and past 256 tiles (one look-back poll reads 256
predecessors); int32 sums that wrap; int64; the form that also writes the
total; scans of falling and rising sizes on one stream (each clears the
look-back state the previous one left) and on a second stream."""
import numpy as np
import pytest
import torch

from cbclib_v2_amd import _native
from cbclib_v2_amd._arrays import stream_ptr

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 5, 2047, 2048, 2049, 300_001, 524_288, 524_289, 1_048_577, 3_000_017]


def expected(a, total, dtype):
    c = np.cumsum(a.astype(np.int64)) if len(a) else np.zeros(0, np.int64)
    out = np.concatenate([[0], c[:-1]]) if len(a) else np.zeros(0, np.int64)
    if total:
        out = np.concatenate([out, [c[-1] if len(a) else 0]])
    if dtype == np.int32:
        out = (out + 2 ** 31) % 2 ** 32 - 2 ** 31     # int32 sums wrap
    return out.astype(dtype)


def scan(a, total):
    L = _native.lib()
    fn = L.cbx_exclusive_scan_i32 if a.dtype == np.int32 else L.cbx_exclusive_scan_i64
    n = len(a)
    t_in = torch.from_numpy(a).cuda()
    t_out = torch.full((n + total,), -7, dtype=t_in.dtype, device='cuda')
    _native.check(fn(t_in.data_ptr() if n else None, t_out.data_ptr() if n + total else None, n, total,
                     stream_ptr()))
    return t_out.cpu().numpy()


@pytest.mark.parametrize('dtype', [np.int32, np.int64])
@pytest.mark.parametrize('total', [0, 1])
def test_exclusive_scan(cuda, dtype, total):
    rng = np.random.default_rng(7)
    for n in SIZES + SIZES[::-1]:
        a = rng.integers(-50, 1000, n).astype(dtype)
        print(f"scan {np.dtype(dtype).name} n={n} total={total}", flush=True)
        assert np.array_equal(scan(a, total), expected(a, total, dtype)), (n, dtype, total)


def test_exclusive_scan_int32_wraps(cuda):
    a = np.full(700_001, 2 ** 30, np.int32)
    assert np.array_equal(scan(a, 1), expected(a, 1, np.int32))


def test_exclusive_scan_second_stream(cuda):
    rng = np.random.default_rng(8)
    s2 = torch.cuda.Stream()
    for n in (3000, 1_500_000, 70_000):
        a = rng.integers(0, 100, n).astype(np.int32)
        with torch.cuda.stream(s2):
            got2 = scan(a, 1)
        got1 = scan(a[::-1].copy(), 1)
        assert np.array_equal(got2, expected(a, 1, np.int32))
        assert np.array_equal(got1, expected(a[::-1], 1, np.int32))


def test_exclusive_scan_errors(cuda):
    L = _native.lib()
    with pytest.raises(ValueError):
        _native.check(L.cbx_exclusive_scan_i32(None, None, -1, 0, stream_ptr()))
    with pytest.raises(ValueError):
        _native.check(L.cbx_exclusive_scan_i32(None, None, 10, 0, stream_ptr()))
