"""draw_lines on the GPU against the oracle (the C restatement of the
reference CPU path, itself pinned against the compiled reference).

Bar: pixel indices bit-exact; images from the deterministic mode bit-exact;
images from the tile (fast) path within rtol 1e-5 / atol 1e-6 * max_val of
the oracle; max overlap exact.
"""
import numpy as np
import pytest
import torch

from cbclib_v2_amd import ndimage
from tests.lines_gen import config1, random_segments

pytestmark = pytest.mark.gpu

KERNELS = ['biweight', 'gaussian', 'parabolic', 'rectangular', 'triangular']
RTOL, ATOL = 1e-5, 1e-6


def case(seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    F = int(rng.integers(1, 4))
    H, W = (int(v) for v in rng.integers(20, 180, 2))
    n = int(rng.integers(1, 80))
    lines = random_segments(rng, n, H, W, 0.0, 150.0, 0.3, 7.0, pad=25.0)
    if seed % 4 == 0:
        lines[:, :4] = np.round(lines[:, :4] * 2) / 2
    if seed % 5 == 0:
        lines[rng.random(n) < 0.2, 4] = rng.choice([0.0, -1.0, 0.5, 1.0])
    idxs = rng.integers(0, F, n)
    it = np.int64 if dtype == np.float64 else np.int32
    return (F, H, W), lines.astype(dtype), idxs.astype(it)


def gpu_image(shape, lines, idxs, kernel, overlap, dtype, init=0.0, max_val=1.7, **kw):
    out = torch.full(shape, init, dtype=torch.float64 if dtype == np.float64 else torch.float32, device='cuda')
    ndimage.draw_lines(out, torch.from_numpy(lines).cuda(), None if idxs is None else torch.from_numpy(idxs).cuda(),
                       max_val, kernel, overlap, **kw)
    return out.cpu().numpy()


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('kernel', KERNELS)
def test_pixel_list_raw_bit_exact(cuda, oracle, dtype, kernel):
    for seed in range(6):
        shape, lines, idxs = case(seed, dtype)
        ref = oracle.pixel_list(shape, lines, idxs, 1.7, kernel, dtype=dtype)
        pl = ndimage.draw_pixels(shape, torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(), 1.7, kernel,
                                 merge=False, dtype=dtype)
        got = [t.cpu().numpy() for t in pl]
        assert len(got[0]) == len(ref[0]), f"seed {seed}: {len(got[0])} entries vs {len(ref[0])}"
        assert np.array_equal(got[4], ref[0]), f"seed {seed}: line ids differ"
        assert np.array_equal(got[0], ref[1]), f"seed {seed}: frames differ"
        assert np.array_equal(got[1], ref[2]), f"seed {seed}: rows differ"
        assert np.array_equal(got[2], ref[3]), f"seed {seed}: columns differ"
        if kernel == 'gaussian':   # exp of the device math library vs glibc: last-ulp differences allowed
            np.testing.assert_allclose(got[3], ref[4], rtol=2e-7 if dtype == np.float32 else 1e-15, atol=0)
        else:
            assert np.array_equal(got[3], ref[4]), f"seed {seed}: values differ"


@pytest.mark.parametrize('overlap', ['sum', 'max'])
def test_pixel_list_merged(cuda, oracle, overlap):
    for seed in range(6):
        shape, lines, idxs = case(seed)
        raw = oracle.pixel_list(shape, lines, idxs, 1.7, 'triangular')
        F, H, W = shape
        keys = (raw[1].astype(np.int64) * H + raw[2]) * W + raw[3]
        uk = np.unique(keys)
        ref_val = {}
        for k, v in zip(keys, raw[4]):
            if k in ref_val:
                ref_val[k] = np.float32(ref_val[k] + v) if overlap == 'sum' else max(ref_val[k], v)
            else:
                ref_val[k] = v
        pl = ndimage.draw_pixels(shape, torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(), 1.7,
                                 'triangular', overlap, merge=True)
        got = [t.cpu().numpy() for t in pl]
        gk = (got[0].astype(np.int64) * H + got[1]) * W + got[2]
        assert np.array_equal(gk, uk), f"seed {seed}: merged keys differ"
        assert np.array_equal(got[3], np.array([ref_val[k] for k in uk], np.float32)), f"seed {seed}: values"


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('overlap', ['sum', 'max'])
@pytest.mark.parametrize('kernel', KERNELS)
def test_image_deterministic_bit_exact(cuda, oracle, dtype, overlap, kernel):
    for seed in range(4):
        shape, lines, idxs = case(seed, dtype)
        init = -0.25 if overlap == 'max' else 0.5
        ref = np.full(shape, init, dtype)
        oracle.draw_lines(ref, lines, idxs, 1.7, kernel, overlap)
        got = gpu_image(shape, lines, idxs, kernel, overlap, dtype, init, deterministic=True)
        if kernel == 'gaussian':
            np.testing.assert_allclose(got, ref, rtol=3e-7 if dtype == np.float32 else 1e-15, atol=0)
        else:
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), f"seed {seed}"


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('overlap', ['sum', 'max'])
@pytest.mark.parametrize('kernel', KERNELS)
@pytest.mark.parametrize('tile,seg,threads,chunk', [(64, 16, 256, 64), (32, 8, 64, 64), (64, 8, 128, 64),
                                                     (32, 16, 256, 64), (64, 8, 256, 256), (32, 16, 256, 256)])
@pytest.mark.parametrize('exact', [False, True])
def test_image_tiles(cuda, oracle, dtype, overlap, kernel, tile, seg, threads, chunk, exact):
    """Tile path: same support as the reference always; values exact per term with
    exact_values (max overlap then bit-exact), within RTOL/ATOL otherwise."""
    for seed in range(3):
        shape, lines, idxs = case(100 + seed, dtype)
        init = -0.25 if overlap == 'max' else 0.0
        ref = np.full(shape, init, dtype)
        oracle.draw_lines(ref, lines, idxs, 1.7, kernel, overlap)
        got = gpu_image(shape, lines, idxs, kernel, overlap, dtype, init, tile=tile, seg=seg, threads=threads,
                        chunk=chunk, exact_values=exact)
        if overlap == 'max' and kernel != 'gaussian' and (exact or dtype == np.float64):
            assert np.array_equal(got, ref), f"seed {seed}"
        else:
            np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL * 1.7)
        # support: every pixel the reference changed is changed, and no other
        assert np.array_equal(got != init, ref != init)


def test_many_lines_parity(cuda, oracle):
    """300000 short lines over 4 frames of 256^2 (1172 workgroups of the
    segment walk), gaussian sum and rectangular max, against the oracle."""
    rng = np.random.default_rng(7)
    n, F, H, W = 300_000, 4, 256, 256
    p0 = rng.uniform(-4, 260, (n, 2))
    ang = rng.uniform(0, 2 * np.pi, n)
    ln = rng.uniform(0, 10, n)
    lines = np.empty((n, 5), np.float32)
    lines[:, :2] = p0
    lines[:, 2] = p0[:, 0] + ln * np.cos(ang)
    lines[:, 3] = p0[:, 1] + ln * np.sin(ang)
    lines[:, 4] = rng.uniform(0.3, 2.5, n)
    idxs = rng.integers(0, F, n).astype(np.int32)
    for kernel, overlap in (('gaussian', 'sum'), ('rectangular', 'max')):
        ref = oracle.draw_lines(np.zeros((F, H, W), np.float32), lines, idxs, 1.0, kernel, overlap)
        got = gpu_image((F, H, W), lines, idxs, kernel, overlap, np.float32, max_val=1.0)
        if overlap == 'max':
            assert np.array_equal(got, ref)
        else:
            np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL)
        assert np.array_equal(got != 0, ref != 0)


SLOT_SEQUENCE = r"""
import sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
from cbclib_v2_amd import ndimage
from oracle import oracle as O
from tests.lines_gen import random_segments
F, H, W = 2, 512, 512
for n, lmax, seed in ((30, 20.0, 11), (3000, 400.0, 12), (30, 20.0, 13), (3000, 400.0, 14)):
    rng = np.random.default_rng(seed)
    lines = random_segments(rng, n, H, W, 1.0, lmax, 0.5, 5.0, pad=10.0).astype(np.float32)
    idxs = rng.integers(0, F, n).astype(np.int32)
    ref = O.draw_lines(np.zeros((F, H, W), np.float32), lines, idxs, 1.0, 'gaussian', 'sum')
    out = torch.zeros((F, H, W), dtype=torch.float32, device='cuda')
    ndimage.draw_lines(out, torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(), 1.0, 'gaussian', 'sum')
    got = out.cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6, err_msg=f"{n} lines")
    assert np.array_equal(got != 0, ref != 0), f"{n} lines: support"
    print(n, "lines ok", flush=True)
"""


def test_slot_capacity_from_previous_call(cuda):
    """An eager draw walks at the largest slot total of the previous calls; a
    call that needs more runs out of slots, is cleared and walks again.  In a
    fresh process (no previous calls) draws of 30 short lines, 3000 long ones
    (more slots than the first call's total: a second walk), 30 and 3000
    again, into 2 frames of 512^2, each against the oracle."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, '-c', SLOT_SEQUENCE, root], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("lines ok") == 4


def test_config1_parity(cuda, oracle):
    """Config 1: one 512x512 frame, 1000 segments, triangular, sum; plus its pixel list."""
    lines = config1()
    ref = oracle.draw_lines(np.zeros((512, 512), np.float32), lines, None, 1.0, 'triangular', 'sum')
    det = gpu_image((512, 512), lines, None, 'triangular', 'sum', np.float32, max_val=1.0, deterministic=True)
    assert np.array_equal(det, ref)
    fast = gpu_image((512, 512), lines, None, 'triangular', 'sum', np.float32, max_val=1.0)
    np.testing.assert_allclose(fast, ref, rtol=RTOL, atol=ATOL)
    raw = oracle.pixel_list((512, 512), lines, None, 1.0, 'triangular')
    pl = ndimage.draw_pixels((512, 512), torch.from_numpy(lines).cuda(), None, 1.0, 'triangular', merge=False)
    assert np.array_equal(pl.y.cpu().numpy(), raw[2]) and np.array_equal(pl.x.cpu().numpy(), raw[3])
    assert len(raw[0]) > 1_000_000


def test_config1_golden(cuda):
    """Config 1 against the reference's own render (tests/golden/config1.npz,
    written by the compiled reference): deterministic images bit-identical,
    the raw pixel list identical by digest."""
    import hashlib
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'config1.npz'))
    lines = config1()
    assert hashlib.sha256(lines.tobytes()).digest() == g['lines_sha256'].tobytes(), "config1 generator changed"
    rect = gpu_image((512, 512), lines, None, 'rectangular', 'sum', np.float32, max_val=1.0, deterministic=True)
    assert np.array_equal(rect, g['rectangular'])
    tri = gpu_image((512, 512), lines, None, 'triangular', 'sum', np.float32, max_val=1.0, deterministic=True)
    assert np.array_equal(tri[200:264], g['triangular_rows_200_263'])
    assert hashlib.sha256(tri.tobytes()).digest() == g['triangular_sha256'].tobytes()
    pl = ndimage.draw_pixels((512, 512), torch.from_numpy(lines).cuda(), None, 1.0, 'triangular', merge=False)
    assert pl.line.numel() == int(g['pixel_list_count'][0])
    h = hashlib.sha256()
    for a in (pl.line, pl.y, pl.x):
        h.update(a.cpu().numpy().astype(np.int32).tobytes())
    assert h.digest() == g['pixel_list_sha256'].tobytes()


def test_numpy_roundtrip_in_place(cuda, oracle):
    shape, lines, idxs = case(7)
    out = np.zeros(shape, np.float32)
    res = ndimage.draw_lines(out, lines, idxs, 1.0, 'parabolic')
    assert res is out
    ref = oracle.draw_lines(np.zeros(shape, np.float32), lines, idxs, 1.0, 'parabolic')
    np.testing.assert_allclose(out, ref, rtol=RTOL, atol=ATOL)


def test_non_contiguous_out(cuda, oracle):
    shape, lines, idxs = case(8)
    big = torch.zeros(shape[0], shape[1], shape[2] * 2, device='cuda')
    view = big[:, :, ::2]
    res = ndimage.draw_lines(view, torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(), 1.0, 'biweight')
    assert res is view
    ref = oracle.draw_lines(np.zeros(shape, np.float32), lines, idxs, 1.0, 'biweight')
    np.testing.assert_allclose(view.cpu().numpy(), ref, rtol=RTOL, atol=ATOL)
    assert float(big[:, :, 1::2].abs().sum()) == 0.0


def test_empty_and_degenerate(cuda):
    out = torch.zeros(16, 22, device='cuda')
    ndimage.draw_lines(out, torch.zeros(0, 5, device='cuda'))
    assert float(out.sum()) == 0.0
    lines = torch.tensor([[5.0, 5.0, 5.0, 5.0, 2.0], [1.0, 2.0, 10.0, 12.0, 0.0], [1.0, 2.0, 10.0, 12.0, -1.0]],
                         device='cuda')
    for k in KERNELS:
        ndimage.draw_lines(out, lines, None, 1.0, k)
    assert float(out.sum()) == 0.0


def test_errors(cuda):
    out = torch.zeros(3, 16, 22, device='cuda')
    lines = torch.tensor([[1.0, 2.0, 10.0, 12.0, 2.0]] * 4, device='cuda')
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines, torch.zeros(3, dtype=torch.int32, device='cuda'))
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines)  # 3 frames, 4 lines, no idxs
    with pytest.raises(IndexError):
        ndimage.draw_lines(out, lines, torch.tensor([0, 1, 2, 3], dtype=torch.int32, device='cuda'))
    with pytest.raises(IndexError):
        ndimage.draw_lines(out, lines, torch.tensor([0, -1, 2, 1], dtype=torch.int32, device='cuda'))
    assert float(out.abs().sum()) == 0.0, "a failed call must not modify out"
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines, None, kernel='cosine')
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines, None, overlap='min')
    with pytest.raises(RuntimeError):
        ndimage.draw_lines(out, torch.zeros(4, 6, device='cuda'))


def test_repeatable(cuda):
    """Two runs of the deterministic mode are bit-identical; so are two tile-path runs with max overlap."""
    shape, lines, idxs = case(11)
    a = gpu_image(shape, lines, idxs, 'gaussian', 'sum', np.float32, deterministic=True)
    b = gpu_image(shape, lines, idxs, 'gaussian', 'sum', np.float32, deterministic=True)
    assert np.array_equal(a, b)
    a = gpu_image(shape, lines, idxs, 'gaussian', 'max', np.float32)
    b = gpu_image(shape, lines, idxs, 'gaussian', 'max', np.float32)
    assert np.array_equal(a, b)


def test_out_ready_event(cuda, oracle):
    """`out` filled on a second stream (slowly: several passes over 256 MiB
    first), the draw told to wait for it with out_ready: the image is the
    fill plus the render, as if the two ran one after the other."""
    shape, lines, idxs = case(11)
    F, H, W = shape
    ref = oracle.draw_lines(np.full(shape, 0.25, np.float32), lines, idxs, 1.7, 'gaussian', 'sum')
    main = torch.cuda.current_stream()
    side = torch.cuda.Stream()
    big = torch.empty(64 * 1024 * 1024, device='cuda')
    t_lines, t_idxs = torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda()
    for rep in range(3):
        out = torch.zeros(shape, dtype=torch.float32, device='cuda')
        ready = torch.cuda.Event()
        side.wait_stream(main)
        with torch.cuda.stream(side):
            for _ in range(4):
                big.fill_(float(rep))
            out.fill_(0.25)
            ready.record(side)
        ndimage.draw_lines(out, t_lines, t_idxs, 1.7, 'gaussian', 'sum', out_ready=ready)
        got = out.cpu().numpy()
        assert np.array_equal(got != 0.25, ref != 0.25), f"rep {rep}: support"
        np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL * 1.7, err_msg=f"rep {rep}")
    torch.cuda.synchronize()
