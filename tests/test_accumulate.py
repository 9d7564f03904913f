"""accumulate_lines: lines grouped into terms, terms into frames.

CPU tests pin the oracle's restatement (bresenham.cpp:151-252) against the
reference's compiled CPU module and against draw_lines identities that fix
each overlap mode; GPU tests compare cbx_accumulate_lines with the oracle
(f32 within rtol 1e-5 / atol 1e-6 * max_val: equal overlaps run the tile
path, whose float values are a few ulp from the reference's; f64 within
1e-12).
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.lines_gen import random_segments

KERNELS = ['biweight', 'gaussian', 'parabolic', 'rectangular', 'triangular']
MODES = [('sum', 'sum'), ('max', 'sum'), ('sum', 'max'), ('max', 'max')]


def acc_case(seed, dtype, n_lines=None, n_terms=None):
    rng = np.random.default_rng(seed)
    F, H, W = int(rng.integers(1, 4)), int(rng.integers(8, 60)), int(rng.integers(8, 60))
    n = int(rng.integers(1, 50)) if n_lines is None else n_lines
    T = int(rng.integers(1, 12)) if n_terms is None else n_terms
    lines = random_segments(rng, n, H, W, 0.0, 40.0, 0.2, 6.0, pad=8.0)
    if seed % 3 == 0:
        lines[:, :4] = np.round(lines[:, :4] * 2) / 2
    it = np.int64 if dtype == np.float64 else np.int32
    terms = rng.integers(0, T, n).astype(it)
    frames = rng.integers(0, F, T).astype(it)
    return (F, H, W), lines.astype(dtype), terms, frames


def tol(dtype):
    return dict(rtol=1e-5, atol=1e-6) if dtype == np.float32 else dict(rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------- CPU

@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('kernel', ['gaussian', 'rectangular', 'biweight'])
@pytest.mark.parametrize('modes', MODES)
@pytest.mark.parametrize('max_val', [0.8, -0.8])
def test_oracle_matches_compiled_reference(oracle, ref_bresenham, dtype, kernel, modes, max_val):
    """Negative max_val and negative frame values exercise the zero start of
    the term buffer (std::max(0, v) in max mode)."""
    for seed in range(6):
        shape, lines, terms, frames = acc_case(100 + seed, dtype)
        lo = 0.0 if max_val > 0 else -0.5
        base = np.random.default_rng(seed).uniform(lo, 0.5, shape).astype(dtype)
        ref = ref_bresenham.accumulate_lines(base.copy(), lines, terms, frames, max_val, kernel, modes[0], modes[1],
                                             1)
        got = oracle.accumulate_lines(base.copy(), lines, terms, frames, max_val, kernel, modes[0], modes[1])
        np.testing.assert_allclose(got, ref, **tol(dtype), err_msg=f"seed {seed}")


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('modes', MODES)
@pytest.mark.parametrize('max_val', [1.1, -0.7])
def test_oracle_matches_python_formulation(oracle, dtype, modes, max_val):
    """The C restatement (recorded-index term buffer) against the independent
    Python one (touched mask over the pixel list): same order, same bits."""
    for seed in range(5):
        shape, lines, terms, frames = acc_case(150 + seed, dtype)
        base = np.random.default_rng(seed).uniform(-0.5, 0.5, shape).astype(dtype)
        a = oracle.accumulate_lines(base.copy(), lines, terms, frames, max_val, 'triangular', *modes)
        b = oracle.accumulate_lines_py(base.copy(), lines, terms, frames, max_val, 'triangular', *modes)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"seed {seed}"


@pytest.mark.parametrize('kernel', KERNELS)
def test_oracle_equal_modes_are_draw_lines(oracle, kernel):
    """sum/sum is draw_lines(idxs = frames[terms]) up to summation order (the
    reference's own test_accumulate_lines); max/max equals it exactly."""
    for seed in range(6):
        shape, lines, terms, frames = acc_case(200 + seed, np.float64)
        for ov in ('sum', 'max'):
            ref = oracle.draw_lines(np.zeros(shape), lines, frames[terms], 1.3, kernel, ov)
            got = oracle.accumulate_lines(np.zeros(shape), lines, terms, frames, 1.3, kernel, ov, ov)
            if ov == 'max':
                assert np.array_equal(got, ref), f"seed {seed}"
            else:
                np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12, err_msg=f"seed {seed}")


@pytest.mark.parametrize('kernel', ['gaussian', 'triangular'])
def test_oracle_mixed_modes_identities(oracle, kernel):
    """One line per term: the in-term overlap is the identity, so max/sum is
    draw_lines sum and sum/max is draw_lines max.  A term holding two copies
    of every line: in 'max' draws each line once, in 'sum' twice."""
    for seed in range(6):
        shape, lines, _, _ = acc_case(300 + seed, np.float64)
        n = lines.shape[0]
        frames = np.random.default_rng(seed).integers(0, shape[0], n)
        one = np.arange(n)
        ref_sum = oracle.draw_lines(np.zeros(shape), lines, frames, 1.0, kernel, 'sum')
        ref_max = oracle.draw_lines(np.zeros(shape), lines, frames, 1.0, kernel, 'max')
        got = oracle.accumulate_lines(np.zeros(shape), lines, one, frames, 1.0, kernel, 'max', 'sum')
        np.testing.assert_allclose(got, ref_sum, rtol=1e-12, atol=1e-12)
        got = oracle.accumulate_lines(np.zeros(shape), lines, one, frames, 1.0, kernel, 'sum', 'max')
        assert np.array_equal(got, ref_max)

        two = np.concatenate([lines, lines])
        terms = np.concatenate([one, one])
        once = oracle.accumulate_lines(np.zeros(shape), lines, one, frames, 1.0, kernel, 'max', 'sum')
        got = oracle.accumulate_lines(np.zeros(shape), two, terms, frames, 1.0, kernel, 'max', 'sum')
        np.testing.assert_allclose(got, once, rtol=1e-12, atol=1e-12)
        got = oracle.accumulate_lines(np.zeros(shape), two, terms, frames, 1.0, kernel, 'sum', 'sum')
        np.testing.assert_allclose(got, 2 * once, rtol=1e-12, atol=1e-12)


def test_oracle_index_errors(oracle):
    shape, lines, terms, frames = acc_case(400, np.float64, n_lines=10, n_terms=4)
    with pytest.raises(IndexError):
        oracle.accumulate_lines(np.zeros(shape), lines, np.full(10, 4), frames)
    with pytest.raises(IndexError):
        oracle.accumulate_lines(np.zeros(shape), lines, terms, np.full(4, shape[0]))
    with pytest.raises(ValueError):
        oracle.accumulate_lines(np.zeros(shape), lines, terms[:-1], frames)


# --------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('kernel', KERNELS)
@pytest.mark.parametrize('modes', MODES)
@pytest.mark.parametrize('max_val', [0.8, -0.8])
def test_accumulate_matches_oracle(oracle, cuda, dtype, kernel, modes, max_val):
    from cbclib_v2_amd import ndimage
    for seed in range(5):
        shape, lines, terms, frames = acc_case(500 + seed, dtype)
        lo = 0.0 if max_val > 0 else -0.5
        base = np.random.default_rng(seed).uniform(lo, 0.5, shape).astype(dtype)
        print(f"accumulate seed={seed} shape={shape} n={lines.shape[0]} modes={modes} max_val={max_val}",
              flush=True)
        ref = oracle.accumulate_lines(base.copy(), lines, terms, frames, max_val, kernel, *modes)
        got = ndimage.accumulate_lines(base.copy(), lines, terms, frames, max_val, kernel, *modes)
        t = tol(dtype)
        np.testing.assert_allclose(got, ref, rtol=t['rtol'], atol=t['atol'] * abs(max_val), err_msg=f"seed {seed}")


def many_terms(seed, n, T, F, H, W):
    rng = np.random.default_rng(seed)
    lines = random_segments(rng, n, H, W, 2.0, 30.0, 0.5, 4.0).astype(np.float32)
    terms = rng.integers(0, T, n).astype(np.int32)
    frames = rng.integers(0, F, T).astype(np.int32)
    return lines, terms, frames


@pytest.mark.gpu
@pytest.mark.parametrize('modes', MODES)
@pytest.mark.parametrize('size', [(7, 2000, 300, 3, 128, 96), (8, 6000, 300, 4, 256, 256)])
def test_accumulate_many_terms(oracle, cuda, modes, size):
    """Many lines per term and many terms per frame (2000 lines in 300 terms
    over 3 frames of 128 x 96; 6000 lines in 300 terms over 4 frames of 256^2).
    Mixed overlaps reduce each term exactly in the reference's order and differ
    from the oracle only in the order terms reach a frame; sum/sum runs the
    fixed-point tile path (per term at most 2^-(clz(items)+1) * max_val from
    the float sum, items <= 4095 per round): all within rtol 1e-5 / atol
    1e-6 * max_val, with the support exact."""
    from cbclib_v2_amd import ndimage
    seed, n, T, F, H, W = size
    lines, terms, frames = many_terms(seed, n, T, F, H, W)
    ref = oracle.accumulate_lines(np.zeros((F, H, W), np.float32), lines, terms, frames, 1.0, 'gaussian', *modes)
    got = ndimage.accumulate_lines(np.zeros((F, H, W), np.float32), lines, terms, frames, 1.0, 'gaussian', *modes)
    assert np.array_equal(got != 0, ref != 0)
    nz = ref != 0
    print(f"{modes} {size}: max-abs {np.abs(got - ref).max():.3e}, max-rel "
          f"{(np.abs(got - ref)[nz] / np.abs(ref[nz])).max():.3e}", flush=True)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize('modes', [('max', 'sum'), ('sum', 'max')])
def test_accumulate_term_batches(oracle, cuda, modes):
    """More (term, tile) pairs than one plan holds (40 000 terms over 2 frames
    of 1024^2: 41 M pairs, three batches of 16 384 terms): the lines of each
    batch are drawn with the others skipped, and the result equals the
    oracle's as for one batch."""
    from cbclib_v2_amd import ndimage
    rng = np.random.default_rng(21)
    F, H, W, T = 2, 1024, 1024, 40000
    n = 60000
    lines = random_segments(rng, n, H, W, 2.0, 20.0, 0.5, 2.0).astype(np.float32)
    terms = rng.integers(0, T, n).astype(np.int32)
    frames = rng.integers(0, F, T).astype(np.int32)
    ref = oracle.accumulate_lines(np.zeros((F, H, W), np.float32), lines, terms, frames, 1.0, 'parabolic', *modes)
    got = ndimage.accumulate_lines(np.zeros((F, H, W), np.float32), lines, terms, frames, 1.0, 'parabolic', *modes)
    assert np.array_equal(got != 0, ref != 0)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_accumulate_errors_and_empty(cuda):
    import torch
    from cbclib_v2_amd import ndimage
    shape, lines, terms, frames = acc_case(600, np.float32, n_lines=12, n_terms=5)
    for modes in MODES:
        with pytest.raises(IndexError):
            ndimage.accumulate_lines(np.zeros(shape, np.float32), lines, np.full(12, 5, np.int32), frames,
                                     in_overlap=modes[0], out_overlap=modes[1])
        with pytest.raises(IndexError):
            ndimage.accumulate_lines(np.zeros(shape, np.float32), lines, terms,
                                     np.full(5, shape[0], np.int32), in_overlap=modes[0], out_overlap=modes[1])
        with pytest.raises(ValueError):
            ndimage.accumulate_lines(np.zeros(shape, np.float32), lines, terms[:-1], frames,
                                     in_overlap=modes[0], out_overlap=modes[1])
        out = ndimage.accumulate_lines(np.ones(shape, np.float32), np.zeros((0, 5), np.float32),
                                       np.zeros(0, np.int32), frames, in_overlap=modes[0], out_overlap=modes[1])
        assert np.array_equal(out, np.ones(shape, np.float32))
    with pytest.raises(ValueError):
        ndimage.accumulate_lines(np.zeros(shape, np.float32), lines, terms, frames, in_overlap='mean')
    # device tensors in, the same tensor out
    t = torch.zeros(shape, dtype=torch.float32, device='cuda')
    r = ndimage.accumulate_lines(t, torch.as_tensor(lines, device='cuda'), torch.as_tensor(terms, device='cuda'),
                                 torch.as_tensor(frames, device='cuda'), in_overlap='max')
    assert r is t and float(t.sum()) > 0
