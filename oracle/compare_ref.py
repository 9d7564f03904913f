#!/usr/bin/env python3
"""Times the CPU baseline (the oracle's OpenMP restatement of the reference's
CPU draw_lines + label, ``oracle.draw_lines_omp`` / ``oracle.label_omp``)
against the reference's own compiled CPU modules (``oracle/_ref``, built by
``oracle/Makefile`` where the reference sources exist) on the benchmark's CPU
sample: frames 0-3 of the 64-frame 2048x2048 workload (2000 segments per
frame, seed 1), gaussian sum, then label(out > 0.5, Structure([0, 1, 1], 2)).

The compiled reference never travels to the GPU host, so this ratio can only
be measured here; the benchmark reports it beside the port's own timing.
Writes ``profiles/cpu_port_vs_reference.txt``.  Run from the repository root:

    make -C oracle && python oracle/compare_ref.py
"""
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:] = [ROOT] + [p for p in sys.path if os.path.abspath(p or '.') != HERE]

from oracle import oracle as O  # noqa: E402
from tests.lines_gen import config2  # noqa: E402

SAMPLE_FRAMES = 4


def cpu_model() -> str:
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or 'unknown'


def median_time(fn, runs=5):
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def main():
    ref = O.ref_module('bresenham')
    lab = O.ref_module('label')
    if ref is None or lab is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle` where the reference sources exist")
    from cbclib_v2_amd.label import Structure
    lines, idxs = config2(frames=64)
    sel = idxs < SAMPLE_FRAMES
    ln, ix = lines[sel], idxs[sel]
    shape = (SAMPLE_FRAMES, 2048, 2048)
    st = Structure([0, 1, 1], 2)
    rst = lab.Structure([0, 1, 1], 2)
    mask = O.draw_lines(np.zeros(shape, np.float32), ln, ix, 1.0, 'gaussian', 'sum') > 0.5
    rows = []
    for threads in (1, os.cpu_count() or 1):
        t_port_d = median_time(lambda: O.draw_lines_omp(np.zeros(shape, np.float32), ln, ix, 1.0, 'gaussian', 'sum',
                                                        threads))
        t_ref_d = median_time(lambda: ref.draw_lines(np.zeros(shape, np.float32), ln, ix, 1.0, 'gaussian', 'sum',
                                                     threads))
        t_port_l = median_time(lambda: O.label_omp(mask, st.shifts(), 1, threads))
        t_ref_l = median_time(lambda: lab.label(mask, rst, 1, threads))
        rows.append((threads, t_port_d, t_ref_d, t_port_l, t_ref_l))
    out = os.path.join(ROOT, 'profiles', 'cpu_port_vs_reference.txt')
    with open(out, 'w') as f:
        f.write("# CPU baseline (oracle OpenMP restatement, 'port') vs the compiled reference modules\n")
        f.write(f"# host: {cpu_model()}, {os.cpu_count()} logical CPUs (the build container, not the GPU host)\n")
        f.write(f"# sample: frames 0-{SAMPLE_FRAMES - 1} of the 64 x 2048^2 workload ({len(ln)} segments), gaussian "
                "sum; label(out > 0.5, Structure([0,1,1],2)); median of 5 runs after 1 warm-up\n")
        f.write("# command: make -C oracle && python oracle/compare_ref.py\n")
        f.write("# threads  port_draw_s  ref_draw_s  port_label_s  ref_label_s  ref/port(draw+label)\n")
        for th, pd, rd, pl, rl in rows:
            f.write(f"{th:8d}  {pd:11.4f}  {rd:10.4f}  {pl:12.4f}  {rl:11.4f}  {(rd + rl) / (pd + pl):8.3f}\n")
    print(open(out).read())


if __name__ == '__main__':
    main()
