"""Pins of the oracle (the C restatement of the reference CPU path) on the CPU.

Each oracle function is checked against something independent of it:

* draw_lines (walk, pixel set, value bits, frame indexing, sum/max): bit for
  bit against the committed golden fixtures (written from the reference's
  compiled ``bresenham`` module by ``tests/golden/make_golden.py``) and, where
  ``oracle/_ref`` is built, against that module on fresh random inputs;
  independently of the reference code, against the FP64 closed form (pixel
  set = {error <= 1} away from the boundary, values within rounding) and the
  pixel-list invariants;
* label: golden fixtures, the compiled ``label`` module, scipy.ndimage.label
  (partition and min-index numbering) and a brute-force BFS;
* profiles: their closed forms (include.hpp:446-459);
* adjoint (no reference counterpart): FP64 central finite differences of the
  FP64 closed-form render;
* every pin above fails for a planted mistake (dropped term, wrong sign,
  transposed operands, wrong connectivity norm, wrong profile argument).
"""
import os
from itertools import product

import numpy as np
import pytest
import scipy.ndimage as ndi

from cbclib_v2_amd.label import Structure
from tests.lines_gen import random_segments

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
KERNELS = ['biweight', 'gaussian', 'parabolic', 'rectangular', 'triangular']


def golden(name):
    with np.load(os.path.join(GOLD, name), allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    n = len({k.split('_')[0] for k in data})
    return [{k.split('_', 1)[1]: v for k, v in data.items() if k.split('_')[0] == f'c{i}'} for i in range(n)]


DRAW = golden('draw_cases.npz')
LABEL = golden('label_cases.npz')


def run_draw(oracle, case):
    out = case['init'].copy()
    kernel, overlap = KERNELS[int(case['meta'][0])], ('sum', 'max')[int(case['meta'][1])]
    oracle.draw_lines(out, case['lines'], case['idxs'], 1.7, kernel, overlap)
    return out


def run_label(oracle, case):
    conn, npts, _ = (int(v) for v in case['meta'])
    st = Structure([int(r) for r in case['radii']], conn)
    return oracle.label(case['inp'], st.shifts(), npts)


# ---------------------------------------------------------------------------
# draw_lines
# ---------------------------------------------------------------------------
@pytest.mark.parametrize('i', range(len(DRAW)))
def test_draw_matches_golden(oracle, i):
    case = DRAW[i]
    got = run_draw(oracle, case)
    assert got.dtype == case['out'].dtype
    assert np.array_equal(got.view(np.uint8), case['out'].view(np.uint8)), f"case {i}"


def random_case(seed, dtype):
    rng = np.random.default_rng(seed)
    F, H, W = int(rng.integers(1, 3)), int(rng.integers(8, 70)), int(rng.integers(8, 70))
    n = int(rng.integers(1, 40))
    lines = random_segments(rng, n, H, W, 0.0, 60.0, 0.2, 7.0, pad=10.0)
    if seed % 3 == 0:
        lines[:, :4] = np.round(lines[:, :4] * 2) / 2
    it = np.int64 if dtype == np.float64 else np.int32
    return (F, H, W), lines.astype(dtype), rng.integers(0, F, n).astype(it)


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('kernel', KERNELS)
@pytest.mark.parametrize('overlap', ['sum', 'max'])
def test_draw_matches_compiled_reference(oracle, ref_bresenham, dtype, kernel, overlap):
    for seed in range(8):
        shape, lines, idxs = random_case(1000 + seed, dtype)
        ref = ref_bresenham.draw_lines(np.zeros(shape, dtype), lines, idxs, 0.8, kernel, overlap, 1)
        got = oracle.draw_lines(np.zeros(shape, dtype), lines, idxs, 0.8, kernel, overlap)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), f"seed {seed}"


@pytest.mark.parametrize('dtype,rtol', [(np.float64, 1e-9), (np.float32, 2e-5)])
@pytest.mark.parametrize('kernel', ['biweight', 'parabolic', 'triangular', 'gaussian'])
def test_draw_matches_closed_form(oracle, dtype, rtol, kernel):
    """Values = max_val K(dist / w) summed over lines (FP64 closed form), and
    the pixel set = {dist <= w}, except within 1e-4 of the profile's edge."""
    for seed in range(5):
        shape, lines, idxs = random_case(2000 + seed, dtype)
        lines[:, 4] = np.abs(lines[:, 4]) + 0.5
        got = oracle.draw_lines(np.zeros(shape, dtype), lines, idxs, 1.3, kernel, 'sum').astype(np.float64)
        ref = oracle.render_closed_f64(shape, lines.astype(np.float64), idxs, 1.3, kernel)
        # distance of every pixel to the nearest profile edge, over all lines
        F, H, W = shape
        yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
        edge = np.full(shape, np.inf)
        for (x0, y0, x1, y1, w), f in zip(lines.astype(np.float64), idxs):
            tx, ty = x1 - x0, y1 - y0
            L2 = tx * tx + ty * ty
            t = np.clip(((xx - x0) * tx + (yy - y0) * ty) / L2, 0.0, 1.0) if L2 > 0 else 0.0
            d = np.hypot(x0 + t * tx - xx, y0 + t * ty - yy) / w
            edge[f] = np.minimum(edge[f], np.abs(d * d - 1.0))
        safe = edge > 1e-4
        np.testing.assert_allclose(got[safe], ref[safe], rtol=rtol, atol=rtol * 1.3)


def test_pixel_list_invariants(oracle):
    """Raw list: a line emits a pixel at most once; per-pixel sums of the list
    are the image; every entry's error is <= 1 and its value is K(sqrt(err))."""
    for seed in range(6):
        shape, lines, idxs = random_case(3000 + seed, np.float64)
        ln, fr, y, x, val, err = oracle.pixel_list(shape, lines, idxs, 1.0, 'parabolic', dtype=np.float64)
        F, H, W = shape
        key = ((ln.astype(np.int64) * F + fr) * H + y) * W + x
        assert len(np.unique(key)) == len(key)
        assert np.all(err <= 1.0)
        np.testing.assert_allclose(val, 0.75 * (1.0 - err), rtol=1e-12, atol=1e-15)
        img = np.zeros(shape)
        np.add.at(img, (fr, y, x), val)
        ref = oracle.draw_lines(np.zeros(shape), lines, idxs.astype(np.int64), 1.0, 'parabolic', 'sum')
        np.testing.assert_allclose(img, ref, rtol=1e-12, atol=1e-12)


def test_idxs_rules(oracle):
    """bresenham.cpp:5-36: None with one frame or one line per frame; range check."""
    lines = np.array([[1, 2, 10, 12, 2.0]] * 3, np.float32)
    a = oracle.draw_lines(np.zeros((3, 16, 16), np.float32), lines, None)
    b = oracle.draw_lines(np.zeros((3, 16, 16), np.float32), lines, np.arange(3, dtype=np.int32))
    assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        oracle.draw_lines(np.zeros((2, 16, 16), np.float32), lines, None)
    with pytest.raises(IndexError):
        oracle.draw_lines(np.zeros((3, 16, 16), np.float32), lines, np.array([0, 3, 1], np.int32))


# ---------------------------------------------------------------------------
# profiles
# ---------------------------------------------------------------------------
def test_profiles_closed_form(oracle):
    closed = {
        'biweight': lambda x: 0.9375 * max(1 - x * x, 0.0) ** 2,
        'gaussian': lambda x: np.exp(-4.5 * x * x) / np.sqrt(2 * np.pi) if abs(x) <= 1 else 0.0,
        'parabolic': lambda x: 0.75 * max(1 - x * x, 0.0),
        'rectangular': lambda x: 1.0 if abs(x) <= 1 else 0.0,
        'triangular': lambda x: max(1 - abs(x), 0.0),
    }
    for k, f in closed.items():
        for x in np.linspace(-1.2, 1.2, 97):
            assert oracle.profile(k, x) == pytest.approx(f(x), rel=1e-13, abs=1e-15), (k, x)


# ---------------------------------------------------------------------------
# label
# ---------------------------------------------------------------------------
@pytest.mark.parametrize('i', range(len(LABEL)))
def test_label_matches_golden(oracle, i):
    case = LABEL[i]
    got, n = run_label(oracle, case)
    assert n == int(case['meta'][2])
    assert np.array_equal(got, case['out']), f"case {i}"


def test_label_matches_compiled_reference(oracle, ref_label):
    rng = np.random.default_rng(7)
    for shape, radii, conn in [((40, 33), [1, 1], 2), ((5, 17, 19), [0, 1, 1], 1), ((7, 8, 9), [1, 1, 1], 2),
                               ((31, 29), [2, 1], 3)]:
        for dens in (0.3, 0.6):
            inp = rng.random(shape) < dens
            for npts in (1, 4):
                res = ref_label.label(inp, ref_label.Structure(radii, conn), npts, 1)
                n = len(res.regions)
                ref = res.to_array(np.arange(1, n + 1)).astype(np.int32) if n else np.zeros(shape, np.int32)
                got, m = oracle.label(inp, Structure(radii, conn).shifts(), npts)
                assert m == n and np.array_equal(got, ref), (shape, radii, conn, npts)


def scipy_label(inp, st: Structure, npts):
    """scipy's partition renumbered in raster order of each component's
    smallest index, components smaller than npts dropped."""
    arr = st.to_array()  # scipy takes 3^N footprints: pad the radius-0 axes
    arr = np.pad(arr, [((3 - d) // 2, (3 - d) // 2) for d in arr.shape])
    lab, n = ndi.label(inp, structure=arr)
    if n == 0:
        return lab.astype(np.int32), 0
    first = ndi.minimum(np.arange(inp.size).reshape(inp.shape), lab, index=np.arange(1, n + 1))
    size = ndi.sum(np.ones(inp.shape), lab, index=np.arange(1, n + 1))
    new = np.zeros(n + 1, np.int32)
    k = 0
    for c in np.argsort(first):
        if size[c] >= npts:
            k += 1
            new[c + 1] = k
    return new[lab], k


def test_label_matches_scipy(oracle):
    rng = np.random.default_rng(11)
    for shape, radii, conn in [((50, 60), [1, 1], 1), ((50, 60), [1, 1], 2), ((6, 30, 31), [0, 1, 1], 2),
                               ((9, 10, 11), [1, 1, 1], 1), ((9, 10, 11), [1, 1, 1], 3)]:
        st = Structure(radii, conn)
        for dens in (0.2, 0.5, 0.8):
            inp = rng.random(shape) < dens
            for npts in (1, 5):
                ref, n = scipy_label(inp, st, npts)
                got, m = oracle.label(inp, st.shifts(), npts)
                assert m == n and np.array_equal(got, ref), (shape, radii, conn, dens, npts)


def bfs_label(inp, shifts):
    out = np.zeros(inp.shape, np.int32)
    n = 0
    for start in range(inp.size):
        idx = np.unravel_index(start, inp.shape)
        if not inp[idx] or out[idx]:
            continue
        n += 1
        out[idx] = n
        todo = [idx]
        while todo:
            p = todo.pop()
            for s in shifts:
                q = tuple(int(a) + int(b) for a, b in zip(p, s))
                if all(0 <= c < d for c, d in zip(q, inp.shape)) and inp[q] == inp[p] and not out[q]:
                    out[q] = n
                    todo.append(q)
    return out, n


def test_label_matches_bfs_on_equal_values(oracle):
    """Equal-valued components of int inputs (label.cpp:141)."""
    rng = np.random.default_rng(3)
    for shape, radii, conn in [((12, 13), [1, 1], 2), ((4, 6, 7), [1, 1, 1], 1), ((30,), [2], 2)]:
        st = Structure(radii, conn)
        inp = rng.integers(0, 3, shape).astype(np.int32)
        ref, n = bfs_label(inp, st.shifts())
        got, m = oracle.label(inp, st.shifts(), 1)
        assert m == n and np.array_equal(got, ref)


def test_structure_closed_form(ref_label=None):
    """Structure(radii, c): the box (2r+1)^N points with 0 < L1 <= c, plus the centre."""
    for radii, conn in [([1, 1], 1), ([1, 1], 2), ([0, 1, 1], 2), ([2, 3], 3), ([1, 1, 1], 3)]:
        st = Structure(radii, conn)
        want = {p for p in product(*(range(-r, r + 1) for r in radii)) if 0 < sum(map(abs, p)) <= conn}
        assert {tuple(s) for s in st.shifts()} == want
        arr = st.to_array()
        assert arr.shape == tuple(2 * r + 1 for r in radii) and arr.sum() == len(want) + 1


def test_structure_matches_compiled_reference(ref_label):
    for radii, conn in [([1, 1], 1), ([1, 1], 2), ([0, 1, 1], 2), ([2, 3], 3)]:
        assert np.array_equal(Structure(radii, conn).to_array(), ref_label.Structure(radii, conn).to_array())


# ---------------------------------------------------------------------------
# adjoint
# ---------------------------------------------------------------------------
def test_adjoint_matches_finite_differences(oracle):
    """d sum(g * render) / d(x0, y0, x1, y1, w): the oracle's analytic FP64
    gradient over the walk's pixel set vs central differences (h = 1e-3) of the
    FP64 closed-form render, biweight (smooth at the footprint edge)."""
    rng = np.random.default_rng(5)
    shape = (1, 48, 52)
    # end points inside the frame, lengths >= 8 (a zero-length line draws
    # nothing in the walk but a disk in the closed form)
    pts = rng.uniform(8, 40, (40, 4))
    pts = pts[np.hypot(pts[:, 2] - pts[:, 0], pts[:, 3] - pts[:, 1]) > 8][:6]
    lines = np.concatenate([pts, rng.uniform(2.0, 5.0, (6, 1))], 1)
    idxs = np.zeros(6, np.int64)
    g = rng.normal(size=shape)
    ana = oracle.adjoint_f64(g, lines, idxs, 1.0, 'biweight', walk='f64')
    h = 1e-3
    fd = np.zeros_like(lines)
    for i in range(lines.shape[0]):
        for j in range(5):
            lp, lm = lines.copy(), lines.copy()
            lp[i, j] += h
            lm[i, j] -= h
            sp = np.sum(g * oracle.render_closed_f64(shape, lp, idxs, 1.0, 'biweight'))
            sm = np.sum(g * oracle.render_closed_f64(shape, lm, idxs, 1.0, 'biweight'))
            fd[i, j] = (sp - sm) / (2 * h)
    np.testing.assert_allclose(ana, fd, rtol=2e-4, atol=2e-4 * np.abs(fd).max())


# ---------------------------------------------------------------------------
# the pins catch plausible mistakes
# ---------------------------------------------------------------------------
@pytest.mark.parametrize('mistake', [1, 2, 3, 5])
def test_draw_pins_catch_planted_mistakes(oracle, mistake):
    """1 dropped tangent term, 2 wrong sign of a normal component,
    3 transposed operands, 5 profile of err instead of sqrt(err)."""
    caught = 0
    with oracle.planted(mistake):
        for case in DRAW:
            got = run_draw(oracle, case)
            caught += not np.array_equal(got.view(np.uint8), case['out'].view(np.uint8))
    assert caught > 0
    assert all(np.array_equal(run_draw(oracle, c).view(np.uint8), c['out'].view(np.uint8)) for c in DRAW[:4])


def test_label_pins_catch_planted_mistake(oracle):
    """4: shifts of L1 norm > 1 ignored (8-connectivity labelled as 4)."""
    caught = 0
    with oracle.planted(4):
        for case in LABEL:
            got, n = run_label(oracle, case)
            caught += (n != int(case['meta'][2])) or not np.array_equal(got, case['out'])
        rng = np.random.default_rng(0)
        inp = rng.random((30, 30)) < 0.5
        st = Structure([1, 1], 2)
        caught += not np.array_equal(oracle.label(inp, st.shifts(), 1)[0], scipy_label(inp, st, 1)[0])
    assert caught > 0


def test_closed_form_windows_and_loss_fd(oracle):
    """The per-line windows of the FD helper sum to render_closed_f64, and
    its central differences of a quadratic loss agree with the oracle's FP64
    adjoint of the same closed form (double walk pixel set)."""
    rng = np.random.default_rng(41)
    H, W = 60, 70
    lines = random_segments(rng, 12, H, W, 5.0, 40.0, 2.0, 5.0)
    idxs = np.zeros(12, np.int64)
    img = oracle.render_closed_f64((1, H, W), lines, idxs, 1.0, 'biweight')[0]
    acc = np.zeros((H, W))
    for ln in lines:
        box = oracle._line_box(ln, H, W)
        ya, yb, xa, xb = box
        acc[ya:yb + 1, xa:xb + 1] += oracle.segment_window_f64(ln, box, 1.0, 'biweight')
    np.testing.assert_allclose(acc, img, rtol=1e-12, atol=1e-12)
    tgt = oracle.render_closed_f64((1, H, W), lines + np.r_[rng.normal(0, 0.5, 4), 0.0], idxs, 1.0, 'biweight')[0]
    fd = oracle.loss_fd_f64({0: img}, {0: tgt}, lines, idxs, range(12), 1e-4, 1.0, 'biweight')
    an = oracle.adjoint_f64(2.0 * (img - tgt)[None], lines, idxs, 1.0, 'biweight', walk='f64')
    nrm = np.linalg.norm(an, axis=1, keepdims=True)
    assert (np.abs(fd - an) / nrm).max() < 1e-3
