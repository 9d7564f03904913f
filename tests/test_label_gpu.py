"""label() on the GPU against the oracle (pinned to the compiled reference and
to scipy): label maps and counts bit-exact."""
import numpy as np
import pytest
import torch

from cbclib_v2_amd import label as cl

pytestmark = pytest.mark.gpu


def masks(seed):
    rng = np.random.default_rng(seed)
    nd = 2 + seed % 2
    shape = tuple(int(s) for s in rng.integers(1, 30 if nd == 3 else 90, nd))
    radii = [int(r) for r in rng.integers(0, 3, nd)]
    conn = int(rng.integers(1, sum(radii) + 2))
    kind = seed % 3
    if kind == 0:
        inp = rng.random(shape) > rng.uniform(0.3, 0.8)
    elif kind == 1:
        inp = rng.integers(0, 3, shape).astype(np.int32)
    else:
        inp = rng.integers(0, 3, shape).astype(np.int64)
    return inp, cl.Structure(radii, conn), int(rng.integers(1, 4))


def test_label_random(cuda, oracle):
    for seed in range(60):
        inp, st, npts = masks(seed)
        ref, n = oracle.label(inp, st.shifts(), npts)
        res = cl.label(torch.from_numpy(inp).cuda(), st, npts)
        assert int(res.index.numel()) == n, f"seed {seed}"
        assert np.array_equal(res.labels.cpu().numpy(), ref), f"seed {seed}"


def test_label_per_frame_stack(cuda, oracle):
    """Config-2 pattern: a stack of frames, per-frame 8-connectivity Structure([0,1,1],2)."""
    rng = np.random.default_rng(5)
    inp = rng.random((6, 120, 140)) > 0.62
    st = cl.Structure([0, 1, 1], 2)
    ref, n = oracle.label(inp, st.shifts(), 1)
    res = cl.label(torch.from_numpy(inp).cuda(), st)
    assert int(res.index.numel()) == n
    assert np.array_equal(res.labels.cpu().numpy(), ref)


def test_label_threshold_fused(cuda, oracle):
    rng = np.random.default_rng(6)
    img = rng.random((4, 100, 130)).astype(np.float32)
    st = cl.Structure([0, 1, 1], 2)
    ref, n = oracle.label(img > 0.5, st.shifts(), 3)
    res = cl.label_threshold(torch.from_numpy(img).cuda(), 0.5, st, 3)
    assert int(res.index.numel()) == n
    assert np.array_equal(res.labels.cpu().numpy(), ref)


def test_label_numpy_result(cuda, ref_label):
    """NumPy input returns regions like the reference CPU LabelResult."""
    rng = np.random.default_rng(7)
    mask = rng.random((50, 50)) > 0.6
    st = cl.Structure([2, 2], 2)
    res = cl.label(mask, st)
    ref = ref_label.label(mask, ref_label.Structure([2, 2], 2), 1, 1)
    assert len(res.regions) == len(ref.regions)
    for a, b in zip(res.regions, ref.regions):
        assert list(a) == list(b)
    assert np.array_equal(cl.labels(res), ref.to_array(np.arange(1, len(ref.regions) + 1)))


def test_label_errors(cuda):
    with pytest.raises(ValueError):
        cl.label(torch.zeros(4, 4, dtype=torch.bool, device='cuda'), cl.Structure([1, 1, 1], 1))


@pytest.mark.parametrize('shape', [(64, 2048), (3, 40, 1024), (12, 8192), (2, 5, 8193), (700, 13)])
@pytest.mark.parametrize('conn', [1, 2])
@pytest.mark.parametrize('kind', ['bool', 'int32', 'int64'])
def test_label_bands(cuda, oracle, shape, conn, kind):
    """Many bands per frame (band = whole rows of <= 8192 pixels), the border
    merge, equal-value inputs, the npts filter, and the general fallback for
    rows wider than a band."""
    rng = np.random.default_rng(sum(shape) + conn)
    if kind == 'bool':
        inp = rng.random(shape) > 0.55
    else:
        inp = rng.integers(0, 3, shape).astype(np.int32 if kind == 'int32' else np.int64)
    radii = [0] * (len(shape) - 2) + [1, 1]
    st = cl.Structure(radii, conn)
    for npts in (1, 4):
        ref, n = oracle.label(inp, st.shifts(), npts)
        res = cl.label(torch.from_numpy(inp).cuda(), st, npts)
        assert int(res.index.numel()) == n, (shape, conn, kind, npts)
        assert np.array_equal(res.labels.cpu().numpy(), ref), (shape, conn, kind, npts)
