"""binary_dilation (label.cpp:8-107).

CPU: the oracle (one breadth-first search per seed, the reference's
single-threaded procedure) bit for bit against the reference's outputs in
tests/golden/dilation_cases.npz, against the compiled reference on fresh
inputs where oracle/_ref is built, and against scipy.ndimage.binary_dilation
applied one masked step at a time (scipy's own multi-iteration masked mode
is a different operation, so the rounds are iterated here).
GPU: cbx_binary_dilation against the oracle, bit-exact, on every rank 1-7,
radii up to 70 along the packed axis (shifts across several 64-bit words),
rows that are not multiples of 64, empty arrays and iterations 0.
"""
from __future__ import annotations

import os

import numpy as np
import pytest
import scipy.ndimage as ndi

from cbclib_v2_amd.label import Structure

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    return np.load(os.path.join(HERE, 'golden', 'dilation_cases.npz'))


def golden_cases(g):
    for c in range(len([k for k in g.files if k.endswith('_meta')])):
        meta = g[f'c{c}_meta']
        mask = g[f'c{c}_mask'] if f'c{c}_mask' in g.files else None
        yield c, g[f'c{c}_inp'], mask, [int(r) for r in meta[2:]], int(meta[0]), int(meta[1]), g[f'c{c}_out']


def scipy_rounds(inp, st: Structure, iterations, mask):
    x = inp.copy()
    for _ in range(iterations):
        d = ndi.binary_dilation(x, structure=st.to_array())
        x = d if mask is None else np.where(mask, d, x)
    return x


def random_case(seed, ranks=(1, 2, 3, 4)):
    rng = np.random.default_rng(seed)
    nd = ranks[seed % len(ranks)]
    top = {1: 300, 2: 90, 3: 20, 4: 9, 5: 6, 6: 5, 7: 4}[nd]
    shape = tuple(int(s) for s in rng.integers(1, top, nd))
    radii = [int(r) for r in rng.integers(0, 3, nd)]
    conn = int(rng.integers(1, sum(radii) + 2))
    inp = rng.random(shape) < rng.uniform(0.01, 0.2)
    mask = rng.random(shape) < rng.uniform(0.4, 0.9) if seed % 3 else None
    return inp, mask, radii, conn, int(rng.integers(0, 5))


# --------------------------------------------------------------------- CPU

def test_oracle_golden_dilation(oracle):
    for c, inp, mask, radii, conn, it, out in golden_cases(golden()):
        got = oracle.binary_dilation(inp, Structure(radii, conn).shifts(), it, mask)
        assert np.array_equal(got, out), f"case {c}"


def test_oracle_dilation_matches_compiled_reference(oracle, ref_label):
    for seed in range(80):
        inp, mask, radii, conn, it = random_case(seed)
        ref = np.asarray(ref_label.binary_dilation(inp, ref_label.Structure(radii, conn), it, mask, 1))
        got = oracle.binary_dilation(inp, Structure(radii, conn).shifts(), it, mask)
        assert np.array_equal(got, ref), f"seed {seed}"


def test_oracle_dilation_matches_scipy(oracle):
    for seed in range(80):
        inp, mask, radii, conn, it = random_case(seed)
        st = Structure(radii, conn)
        got = oracle.binary_dilation(inp, st.shifts(), it, mask)
        assert np.array_equal(got, scipy_rounds(inp, st, it, mask)), f"seed {seed}"


def test_oracle_dilation_edges(oracle):
    """iterations 0 copies (non-zero input -> True); a mask of another shape
    is a ValueError; an all-false input stays all false."""
    inp = np.array([[0, 3], [0, 0]], np.int32)
    st = Structure([1, 1], 1)
    assert np.array_equal(oracle.binary_dilation(inp, st.shifts(), 0), inp != 0)
    with pytest.raises(ValueError):
        oracle.binary_dilation(inp, st.shifts(), 1, np.ones((3, 2), bool))
    assert not oracle.binary_dilation(np.zeros((4, 5), bool), st.shifts(), 3).any()


# --------------------------------------------------------------------- GPU

def _gpu(inp, structure, iterations, mask):
    import torch
    from cbclib_v2_amd import label as cl
    t = cl.binary_dilation(torch.from_numpy(inp).cuda(), structure, iterations,
                           None if mask is None else torch.from_numpy(mask).cuda())
    return t.cpu().numpy()


@pytest.mark.gpu
def test_gpu_golden_dilation(cuda):
    for c, inp, mask, radii, conn, it, out in golden_cases(golden()):
        assert np.array_equal(_gpu(inp, Structure(radii, conn), it, mask), out), f"case {c}"


@pytest.mark.gpu
def test_gpu_dilation_vs_oracle(cuda, oracle):
    for seed in range(140):
        inp, mask, radii, conn, it = random_case(seed, ranks=(1, 2, 3, 4, 5, 6, 7))
        st = Structure(radii, conn)
        ref = oracle.binary_dilation(inp, st.shifts(), it, mask)
        assert np.array_equal(_gpu(inp, st, it, mask), ref), f"seed {seed}"


@pytest.mark.gpu
def test_gpu_dilation_wide_shifts(cuda, oracle):
    """Radii 63, 64, 65 and 70 along the packed axis: x shifts that reach one
    and two words away; rows of 1, 63, 64, 65, 200 and 1000 elements."""
    rng = np.random.default_rng(4)
    for W in (1, 63, 64, 65, 200, 1000):
        for r in (63, 64, 65, 70):
            for conn in (1, r + 1):
                inp = rng.random((5, W)) < 0.01
                mask = rng.random((5, W)) < 0.8
                st = Structure([1, r], conn)
                for it, m in ((1, None), (2, mask)):
                    ref = oracle.binary_dilation(inp, st.shifts(), it, m)
                    assert np.array_equal(_gpu(inp, st, it, m), ref), (W, r, conn, it)


@pytest.mark.gpu
def test_gpu_dilation_large_and_api(cuda, oracle):
    """A 4 x 512 x 777 stack (one pass and 40 rounds, which converge early
    under a sparse mask), numpy in -> numpy out, int input, empty arrays,
    iterations 0, and the reference's ValueErrors."""
    import torch
    from cbclib_v2_amd import label as cl
    rng = np.random.default_rng(8)
    inp = rng.random((4, 512, 777)) < 0.002
    mask = rng.random((4, 512, 777)) < 0.55
    st = Structure([0, 1, 1], 2)
    for it, m in ((1, None), (1, mask), (40, mask), (3, None)):
        ref = oracle.binary_dilation(inp, st.shifts(), it, m)
        assert np.array_equal(_gpu(inp, st, it, m), ref), it
    got = cl.binary_dilation(inp.astype(np.int64), st, 2, mask)
    assert isinstance(got, np.ndarray) and got.dtype == np.bool_
    assert np.array_equal(got, oracle.binary_dilation(inp, st.shifts(), 2, mask))
    assert np.array_equal(cl.binary_dilation(inp, st, 0), inp)
    assert cl.binary_dilation(np.zeros((0, 5), bool), Structure([1, 1], 1)).shape == (0, 5)
    with pytest.raises(ValueError):
        cl.binary_dilation(inp, Structure([1, 1], 1))
    with pytest.raises(ValueError):
        cl.binary_dilation(inp, st, 1, mask[:, :10])
    with pytest.raises(ValueError):
        cl.binary_dilation(inp, st, -1)
    t = torch.from_numpy(inp).cuda()
    assert bool((cl.binary_dilation(t, st, 1) >= t).all())
