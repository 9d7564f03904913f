#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the reference's own CPU
code: ``bresenham.draw_lines`` (cbclib_v2/_src/src/bresenham.cpp:142-149,
329-330) and ``label.label`` + ``LabelResult.to_array`` (label.cpp:109-175,
459-493, 530-532), compiled from the reference sources by ``oracle/Makefile``
into ``oracle/_ref``.  The reference ships no golden vectors of its own, so
these fixtures are what pins the oracle where the compiled reference is not
available (the oracle's CPU tests compare against them bit for bit).

Inputs are small and seeded; outputs are stored as computed (single thread,
so sums are in ascending line order).  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py [draw label config1 ...]

(no names: every fixture file).  ``moments_cases.npz`` holds the reference's
region moments and p-values of labelled random masks; ``draw3_cases.npz`` its
draw_lines and accumulate_lines on 3D lines (rows of 7); ``label_hd_cases.npz``
its label on ranks 1 and 4-7; ``dilation_cases.npz`` its binary_dilation.  ``config1.npz`` holds the reference's render
of the 512x512 / 1000-segment workload (seed 0): the rectangular image whole,
the triangular image as a SHA-256 of its bytes plus rows 200-263, the SHA-256
of its raw pixel list (line, y, x as int32), and a LabelResult case.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

KERNELS = ['biweight', 'gaussian', 'parabolic', 'rectangular', 'triangular']


def draw_case(rng, dtype):
    """Two 40x56 frames; segments of every kind the walk distinguishes:
    random, half-integer end points (ties of the rounding), axis-aligned,
    leaving the frame, zero-length, non-positive and sub-pixel widths."""
    F, H, W = 2, 40, 56
    n = 36
    x0 = rng.uniform(-8, W + 8, n)
    y0 = rng.uniform(-8, H + 8, n)
    th = rng.uniform(0, 2 * np.pi, n)
    L = rng.uniform(0, 45, n)
    lines = np.stack([x0, y0, x0 + L * np.cos(th), y0 + L * np.sin(th), rng.uniform(0.3, 6.0, n)], 1)
    lines[:8, :4] = np.round(lines[:8, :4] * 2) / 2          # half-integer end points
    lines[8:11, 3] = lines[8:11, 1]                            # horizontal
    lines[11:14, 2] = lines[11:14, 0]                          # vertical
    lines[14, 2:4] = lines[14, 0:2]                            # zero length
    lines[15, 4] = 0.0
    lines[16, 4] = -1.0
    idxs = rng.integers(0, F, n)
    it = np.int64 if dtype == np.float64 else np.int32
    return (F, H, W), lines.astype(dtype), idxs.astype(it)


def sha256(*arrays) -> np.ndarray:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8).copy()


def config1_fixtures(bres, lab):
    """The 512x512 / 1000-segment workload (tests/lines_gen.py:config1) through
    the reference, and a regions case for the NumPy LabelResult path."""
    from tests.lines_gen import config1
    lines = config1()
    fx = {'lines_sha256': sha256(lines)}
    rect = np.zeros((512, 512), np.float32)
    bres.draw_lines(rect, lines, None, 1.0, 'rectangular', 'sum', 1)
    fx['rectangular'] = rect
    tri = np.zeros((512, 512), np.float32)
    bres.draw_lines(tri, lines, None, 1.0, 'triangular', 'sum', 1)
    fx['triangular_sha256'] = sha256(tri)
    fx['triangular_rows_200_263'] = tri[200:264].copy()
    # raw pixel list: the reference keeps it internal (ImageBuffer), so its
    # digest comes from the oracle, whose emission order is pinned by the
    # images above and the draw fixtures
    line, _, y, x, _, _ = O.pixel_list((512, 512), lines, None, 1.0, 'triangular')
    fx['pixel_list_count'] = np.array([line.size], np.int64)
    fx['pixel_list_sha256'] = sha256(line.astype(np.int32), y.astype(np.int32), x.astype(np.int32))
    rng = np.random.default_rng(7)
    mask = rng.random((50, 50)) > 0.6
    res = lab.label(mask, lab.Structure([2, 2], 2), 1, 1)
    n = len(res.regions)
    fx['regions_mask'] = mask
    fx['regions_map'] = res.to_array(np.arange(1, n + 1)).astype(np.int32)
    fx['regions_first'] = np.array([list(r)[0] for r in res.regions], np.int64)
    fx['regions_size'] = np.array([len(list(r)) for r in res.regions], np.int64)
    np.savez_compressed(os.path.join(HERE, 'config1.npz'), **fx)
    print("wrote config1.npz")


MOMENT_FUNCS = ('total_mass', 'mean', 'center_of_mass', 'moment_of_inertia', 'covariance_matrix', 'line_fit')


def moments_fixtures(lab):
    """The reference's per-region functions and p_values (label.cpp:176-255,
    534-583) on labelled random masks: 2D f32/f64, a per-frame 3D stack and a
    4D case; each case stores the label map, the data, the outputs and the
    p_values arguments."""
    rng = np.random.default_rng(99)
    fx = {}
    cases = [((40, 60), [1, 1], 2, np.float32), ((40, 60), [1, 1], 1, np.float64),
             ((3, 30, 35), [0, 1, 1], 2, np.float64), ((3, 30, 35), [0, 1, 1], 2, np.float32),
             ((3, 4, 6, 7), [1, 1, 1, 1], 1, np.float64)]
    for c, (shape, radii, conn, dt) in enumerate(cases):
        mask = rng.random(shape) > 0.5
        data = (rng.random(shape) * 4.0 - 1.0).astype(dt)
        res = lab.label(mask, lab.Structure(radii, conn), 1, 1)
        n = len(res.regions)
        fx[f'c{c}_labels'] = res.to_array(np.arange(1, n + 1)).astype(np.int32)
        fx[f'c{c}_data'] = data
        for name in MOMENT_FUNCS:
            fx[f'c{c}_{name}'] = np.asarray(getattr(lab, name)(res, data))
        lines = np.asarray(lab.line_fit(res, data)).astype(dt)
        fx[f'c{c}_lines'] = lines
        fx[f'c{c}_pargs'] = np.array([0.3, 1.0, 1.5], np.float64)      # p0, vmin, xtol
        # Python floats, as functions.py passes them: pybind11 then picks the overload of the
        # data's type (NumPy float32 scalars would fail its no-convert pass and select double)
        fx[f'c{c}_p_values'] = np.asarray(lab.p_values(res, lines, data, 0.3, 1.0, 1.5))
    np.savez_compressed(os.path.join(HERE, 'moments_cases.npz'), **fx)
    print(f"wrote moments_cases.npz ({len(cases)} cases)")


def draw3_case(rng, shape, n, dtype):
    """3D lines in volumes ``shape`` (..., D, H, W): random, half-integer end
    points, axis-aligned, zero length, non-positive and sub-pixel widths,
    lines leaving the volume."""
    D, H, W = shape[-3:]
    p0 = rng.uniform(-4, [W + 4, H + 4, D + 4], (n, 3))
    p1 = p0 + rng.normal(0, 7, (n, 3))
    w = rng.uniform(0.3, 4.0, (n, 1))
    lines = np.concatenate([p0, p1, w], 1)
    lines[:3, :6] = np.round(lines[:3, :6] * 2) / 2            # half-integer end points
    lines[3, 3:5] = lines[3, 0:2]                               # along z
    lines[4, 4:6] = lines[4, 1:3]                               # along x
    lines[5, 3:6] = lines[5, 0:3]                               # zero length
    lines[6, 6] = 0.0
    lines[7, 6] = -1.0
    lines[8, 6] = 0.25
    return lines.astype(dtype)


def draw3_fixtures(bres):
    """The reference's bresenham.draw_lines and accumulate_lines on rows of 7
    (draw_lines_2d_3d / accumulate_lines_2d_3d, bresenham.cpp:126-149,
    284-309, single thread) on the shapes of its own tests
    (tests/draw_line_test.py: (4, 3, 16, 22) and (2, 20, 16, 22)).
    Accumulate cases hold at most 16 lines, where the reference's std::sort of
    the lines by term keeps their order (insertion sort), so its sums are in
    ascending line order."""
    rng = np.random.default_rng(7)
    fx = {}
    i = 0
    for dtype in (np.float32, np.float64):
        it = np.int64 if dtype == np.float64 else np.int32
        for k, kernel in enumerate(KERNELS):
            for overlap in ('sum', 'max'):
                shape = (4, 3, 16, 22) if (k + i) % 2 else (2, 20, 16, 22)
                lines = draw3_case(rng, shape, 24, dtype)
                idxs = rng.integers(0, int(np.prod(shape[:-3])), len(lines)).astype(it)
                init = np.zeros(shape, dtype)
                init.flat[rng.integers(0, init.size, 64)] = rng.uniform(-1, 1, 64)
                out = init.copy()
                bres.draw_lines(out, lines, idxs, 1.7, kernel, overlap, 1)
                fx.update({f'd{i}_lines': lines, f'd{i}_idxs': idxs, f'd{i}_init': init, f'd{i}_out': out,
                           f'd{i}_meta': np.array([k, 0 if overlap == 'sum' else 1], np.int64)})
                i += 1
    j = 0
    for dtype in (np.float32, np.float64):
        it = np.int64 if dtype == np.float64 else np.int32
        for m, (iov, oov) in enumerate((('sum', 'sum'), ('max', 'sum'), ('sum', 'max'), ('max', 'max'))):
            for max_val in (1.3, -0.6):
                shape = (3, 5, 12, 14)
                lines = draw3_case(rng, shape, 16, dtype)
                terms = rng.integers(0, 5, len(lines)).astype(it)
                frames = rng.integers(0, shape[0], 5).astype(it)
                init = np.round(rng.uniform(-0.4, 0.6, shape), 1).astype(dtype)
                out = init.copy()
                kernel = KERNELS[(m + j) % 5]
                bres.accumulate_lines(out, lines, terms, frames, max_val, kernel, iov, oov, 1)
                fx.update({f'a{j}_lines': lines, f'a{j}_terms': terms, f'a{j}_frames': frames, f'a{j}_init': init,
                           f'a{j}_out': out, f'a{j}_meta': np.array([KERNELS.index(kernel), m], np.int64),
                           f'a{j}_max_val': np.array(max_val, np.float64)})
                j += 1
    np.savez_compressed(os.path.join(HERE, 'draw3_cases.npz'), **fx)
    print(f"wrote draw3_cases.npz ({i} draw and {j} accumulate cases)")


HD_CASES = [((64,), [2], 1), ((5, 4, 6, 7), [1, 1, 1, 1], 1), ((5, 4, 6, 7), [1, 0, 1, 1], 3),
            ((3, 4, 3, 5, 6), [1, 1, 1, 1, 1], 2), ((3, 3, 4, 3, 5, 4), [1, 1, 1, 1, 1, 1], 6),
            ((2, 3, 3, 2, 4, 3, 4), [1, 1, 1, 1, 1, 1, 1], 1), ((2, 3, 3, 2, 4, 3, 4), [1, 1, 0, 1, 1, 1, 1], 4),
            ((2, 2, 3, 2, 3, 3, 4), [1, 1, 1, 1, 1, 1, 1], 7)]


def label_hd_fixtures(lab):
    """The reference's label (label.cpp:109-175, 459-493) on ranks 1 and 4-7,
    the ranks its Structure and label accept beyond the 2D/3D cases."""
    rng = np.random.default_rng(31)
    fx = {}
    for c, (shape, radii, conn) in enumerate(HD_CASES):
        for kind in ('bool', 'int64'):
            if kind == 'bool':
                inp = rng.random(shape) < 0.4
            else:
                inp = (rng.random(shape) < 0.5) * rng.integers(1, 3, shape)
            for npts in (1, 3):
                res = lab.label(inp, lab.Structure(radii, conn), npts, 1)
                n = len(res.regions)
                out = res.to_array(np.arange(1, n + 1)).astype(np.int32) if n else np.zeros(shape, np.int32)
                key = f'c{c}_{kind}_{npts}'
                fx[key + '_inp'] = inp
                fx[key + '_out'] = out
                fx[key + '_meta'] = np.array([conn, npts, n] + list(radii), np.int64)
    np.savez_compressed(os.path.join(HERE, 'label_hd_cases.npz'), **fx)
    print(f"wrote label_hd_cases.npz ({len(fx) // 3} cases)")


def dilation_fixtures(lab):
    """The reference's binary_dilation (label.cpp:8-107, num_threads = 1) with
    and without a mask, iterations 0-3, on ranks 1-4."""
    rng = np.random.default_rng(47)
    fx = {}
    cases = [((70,), [2], 1), ((33, 41), [1, 1], 1), ((33, 41), [1, 1], 2), ((30, 37), [2, 1], 2),
             ((5, 19, 23), [0, 1, 1], 2), ((9, 10, 11), [1, 1, 1], 3), ((6, 5, 7, 8), [1, 0, 1, 1], 2),
             ((20, 130), [0, 1], 1)]
    c = 0
    for shape, radii, conn in cases:
        for it in (0, 1, 2, 3):
            for masked in (False, True):
                inp = rng.random(shape) < 0.08
                mask = rng.random(shape) < 0.75 if masked else None
                out = np.asarray(lab.binary_dilation(inp, lab.Structure(radii, conn), it, mask, 1))
                fx[f'c{c}_inp'] = inp
                if masked:
                    fx[f'c{c}_mask'] = mask
                fx[f'c{c}_out'] = out
                fx[f'c{c}_meta'] = np.array([conn, it] + list(radii), np.int64)
                c += 1
    np.savez_compressed(os.path.join(HERE, 'dilation_cases.npz'), **fx)
    print(f"wrote dilation_cases.npz ({c} cases)")


def main():
    bres = O.ref_module('bresenham')
    lab = O.ref_module('label')
    if bres is None or lab is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference sources exist")
    names = sys.argv[1:] or ['draw', 'label', 'config1', 'moments', 'draw3', 'label_hd', 'dilation']
    if 'dilation' in names:
        dilation_fixtures(lab)
    if 'label_hd' in names:
        label_hd_fixtures(lab)
    if 'draw3' in names:
        draw3_fixtures(bres)
    if 'config1' in names:
        config1_fixtures(bres, lab)
    if 'moments' in names:
        moments_fixtures(lab)
    if 'draw' not in names and 'label' not in names:
        return
    rng = np.random.default_rng(20240601)
    draw = {}
    i = 0
    for dtype in (np.float32, np.float64):
        for kernel in KERNELS:
            for overlap in ('sum', 'max'):
                shape, lines, idxs = draw_case(rng, dtype)
                init = np.where(rng.random(shape) < 0.05, rng.uniform(-1, 1, shape), 0.0).astype(dtype)
                out = init.copy()
                bres.draw_lines(out, lines, idxs, 1.7, kernel, overlap, 1)
                draw[f'c{i}_lines'] = lines
                draw[f'c{i}_idxs'] = idxs
                draw[f'c{i}_init'] = init
                draw[f'c{i}_out'] = out
                draw[f'c{i}_meta'] = np.array([KERNELS.index(kernel), 0 if overlap == 'sum' else 1], np.int64)
                i += 1
    np.savez_compressed(os.path.join(HERE, 'draw_cases.npz'), **draw)

    label = {}
    j = 0
    for shape, radii, conn, dens in [((30, 41), [1, 1], 1, 0.45), ((30, 41), [1, 1], 2, 0.45),
                                     ((4, 20, 23), [0, 1, 1], 2, 0.5), ((6, 9, 11), [1, 1, 1], 3, 0.35),
                                     ((25, 25), [2, 2], 2, 0.3), ((64,), [1], 1, 0.5)]:
        for kind in ('bool', 'int32'):
            if kind == 'bool':
                inp = rng.random(shape) < dens
            else:
                inp = (rng.random(shape) < dens) * rng.integers(1, 3, shape).astype(np.int32)
            for npts in (1, 3):
                st = lab.Structure(radii, conn)
                res = lab.label(inp, st, npts, 1)
                n = len(res.regions)
                out = res.to_array(np.arange(1, n + 1)).astype(np.int32) if n else np.zeros(shape, np.int32)
                label[f'c{j}_inp'] = inp
                label[f'c{j}_radii'] = np.array(radii, np.int64)
                label[f'c{j}_meta'] = np.array([conn, npts, n], np.int64)
                label[f'c{j}_out'] = out
                j += 1
    np.savez_compressed(os.path.join(HERE, 'label_cases.npz'), **label)
    print(f"wrote {i} draw cases and {j} label cases")


if __name__ == '__main__':
    main()
