#!/usr/bin/env python3
"""Benchmark of the streak-rasterization hot path on one MI355X.

Default workload (plain forward + label, 64 frames of 2048 x 2048 float32):
128 000 segments, 2 000 per frame, lengths U[20, 400] px, random angles,
widths U[2, 6] px, gaussian profile, sum overlap; the render is thresholded
at 0.5 and labelled with per-frame 8-connectivity (Structure([0, 1, 1], 2)).
One step = zero the frames + draw_lines + label(out > 0.5), inputs resident
in HBM.  value = segments per second of the whole step (Mlines/s).

Prints ONE JSON line.  Extra fields: forward-only rate, label time, the
roofline of the dominant kernel (the tile kernel, timed with HIP events on
its launch stream) and the reference CPU path timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/traffic.txt: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by
    the same counters on a kernel of known traffic), or None."""
    path = os.path.join(ROOT, 'profiles', 'traffic.txt')
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 4 and parts[0] == kernel:
            try:
                return float(parts[1]) + float(parts[2])
            except ValueError:
                return None
    return None


def fail_multi(n: int) -> None:
    print(json.dumps({"error": f"--gpus {n}: this benchmark measures one GPU only; the path shards by frame "
                      "(see the sharding paragraph of the design notes) and no multi-GPU run is made"}))
    sys.exit(2)


def host_cpu():
    """Model, sockets, physical cores and logical CPUs of this host (from /proc/cpuinfo)."""
    model, phys = 'unknown', set()
    cur = {}
    try:
        for line in open('/proc/cpuinfo'):
            if ':' not in line:
                if 'physical id' in cur and 'core id' in cur:
                    phys.add((cur['physical id'], cur['core id']))
                cur = {}
                continue
            k, v = (x.strip() for x in line.split(':', 1))
            cur[k] = v
            if k == 'model name':
                model = v
        if 'physical id' in cur and 'core id' in cur:
            phys.add((cur['physical id'], cur['core id']))
    except OSError:
        pass
    sockets = len({p for p, _ in phys}) or 1
    return {"model": model, "sockets": sockets, "physical_cores": len(phys) or (os.cpu_count() or 1),
            "logical_cpus": os.cpu_count() or 1}


def port_vs_reference():
    """Ratio of the compiled reference's CPU time to the port's on the same
    sample, measured in the build container (profiles/cpu_port_vs_reference.txt)."""
    path = os.path.join(ROOT, 'profiles', 'cpu_port_vs_reference.txt')
    if not os.path.exists(path):
        return None
    rows = {}
    for line in open(path):
        parts = line.split()
        if line.startswith('#') or len(parts) != 6:
            continue
        rows[int(parts[0])] = float(parts[5])
    return {"reference_time_over_port_time": rows, "where": "build container (8-core Xeon), not this host",
            "source": "profiles/cpu_port_vs_reference.txt"} if rows else None


def cpu_baseline(lines, idxs, frames, H, W, runs=5):
    """CPU baseline: the oracle's OpenMP restatement of the reference's CPU
    draw_lines + label (bresenham.cpp:62-94, label.cpp:109-175: per-thread
    pixel buffers and a critical merge; per-thread chunks grown into regions)
    on frames 0-3 of the workload, timed at this host's core share and at one
    thread: median of `runs` runs after one warm-up, wall clock around the
    native calls (the mask `out > 0.5` included)."""
    import numpy as np
    from oracle import oracle as O
    from cbclib_v2_amd.label import Structure
    cpu = host_cpu()
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    cores = max(1, min(16, share, cpu["physical_cores"]))
    sample_frames = 4
    sel = idxs < sample_frames
    ln, ix = lines[sel], idxs[sel]
    shifts = Structure([0, 1, 1], 2).shifts()
    out = np.zeros((sample_frames, H, W), np.float32)

    def one(threads):
        out[...] = 0
        t0 = time.perf_counter()
        O.draw_lines_omp(out, ln, ix, 1.0, 'gaussian', 'sum', threads)
        t1 = time.perf_counter()
        O.label_omp(out > 0.5, shifts, 1, threads)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    res = {}
    for threads in (cores, 1):
        one(threads)
        ts = [one(threads) for _ in range(runs)]
        tot = sorted(a + b for a, b in ts)[len(ts) // 2]
        draw = sorted(a for a, _ in ts)[len(ts) // 2]
        res[threads] = (len(ln) / tot / 1e6, len(ln) / draw / 1e6)
    return {"value": round(res[cores][0], 6), "unit": "Mlines/s", "cores": cores, "kind": "port",
            "single_thread": round(res[1][0], 6),
            "draw_only": {"cores": round(res[cores][1], 6), "single_thread": round(res[1][1], 6)},
            "cpu_model": cpu["model"], "sockets": cpu["sockets"], "physical_cores": cpu["physical_cores"],
            "logical_cpus": cpu["logical_cpus"],
            "port_vs_reference": port_vs_reference(),
            "sample": f"frames 0-{sample_frames - 1} of the {frames} ({len(ln)} segments): draw_lines (gaussian, sum) "
                      "+ label(out > 0.5, Structure([0,1,1],2)) through the oracle's OpenMP restatement of the "
                      f"reference CPU path; median of {runs} runs after 1 warm-up, at {cores} threads and at 1"}


def sweep(out, lines, idxs, F, H, W, N):
    """Tile-kernel time over launch shapes (tile, seg, threads) and value modes."""
    import torch
    from cbclib_v2_amd import _native, ndimage
    rows = []
    shapes = [(t, s, n, c) for t in (32, 64) for s in (8, 16) for n, c in ((64, 64), (128, 64), (256, 64), (256, 256))]
    for exact in (False, True):
        for tile, seg, threads, chunk in shapes:
            if True:
                if True:
                    kw = dict(tile=tile, seg=seg, threads=threads, chunk=chunk, exact_values=exact)
                    for _ in range(2):
                        ndimage.draw_lines(out, lines, idxs, 1.0, 'gaussian', 'sum', **kw)
                    torch.cuda.synchronize()
                    _native.timing_reset()
                    _native.timing_enable(True)
                    for _ in range(5):
                        ndimage.draw_lines(out, lines, idxs, 1.0, 'gaussian', 'sum', **kw)
                    torch.cuda.synchronize()
                    _native.timing_enable(False)
                    ms, n = _native.timing_read('draw_tiles')
                    avg = ms / max(n, 1)
                    gbs = (8.0 * F * H * W + 24.0 * N) / (avg * 1e-3) / 1e9
                    rows.append(dict(kw, ms=round(avg, 4), GBs=round(gbs, 1)))
                    print(json.dumps(rows[-1]), flush=True)


def timed(fn, steps, warmup):
    """Average milliseconds of fn() over `steps` calls after `warmup`, with the
    device synchronised on both sides of the timed region."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def workload_c3(args):
    """Batched uneven forward: 256 frames of 1024^2, per-frame counts
    log-uniform in [100, 20000] (seed 2), C2 segment distributions, each line
    triangular or gaussian.  One autotuned plan draw replayed from a HIP graph
    vs the per-frame loop of plain draw_lines calls (two per frame: the
    reference takes one profile per call)."""
    import numpy as np
    import torch
    from cbclib_v2_amd import _native, batch, ndimage
    from tests.lines_gen import config3
    dev = torch.device('cuda', 0)
    lines_np, idxs_np, prof, counts = config3()
    F, H, W = 256, 1024, 1024
    N = lines_np.shape[0]
    kern_np = np.where(prof == 0, 4, 1).astype(np.uint8)      # triangular / gaussian
    lines = torch.from_numpy(lines_np).to(dev)
    idxs = torch.from_numpy(idxs_np).to(dev)
    kern = torch.from_numpy(kern_np).to(dev)
    out = torch.zeros((F, H, W), dtype=torch.float32, device=dev)

    best = batch.tune((F, H, W), lines, idxs, kern, repeats=3, use_cache=False)
    opts = {k: best[k] for k in ('tile', 'seg', 'threads', 'chunk')}
    plan = batch.Plan((F, H, W), lines, idxs, kern, headroom=1.1, **opts)
    # tile-kernel time on eager plan draws (the graph replays record no timers)
    _native.timing_reset()
    _native.timing_enable(True)
    eager_ms = timed(lambda: (out.zero_(), plan.draw(out, lines, idxs, kern)), 5, 2)
    _native.timing_enable(False)
    tiles_ms, tiles_n = _native.timing_read('draw_tiles')
    plan.check()
    graph = batch.GraphDraw(plan, out, lines, idxs, kern)
    graph_ms = timed(graph.replay, args.steps, args.warmup)
    plan.check()
    ref = out.clone()

    # per-frame loop: draw_lines per frame and profile
    per = []
    off = np.concatenate([[0], np.cumsum(counts)])
    for f in range(F):
        sl = slice(off[f], off[f + 1])
        for kid, name in ((4, 'triangular'), (1, 'gaussian')):
            sel = np.nonzero(kern_np[sl] == kid)[0] + off[f]
            if len(sel):
                per.append((f, name, lines[torch.from_numpy(sel).to(dev)]))

    def loop():
        out.zero_()
        for f, name, ln in per:
            ndimage.draw_lines(out[f], ln, None, 1.0, name, 'sum')

    loop_ms = timed(loop, max(1, args.steps // 5), 1)
    same = float((out - ref).abs().max()) <= 1e-5 * max(1.0, float(ref.abs().max()))
    tile_avg = tiles_ms / max(tiles_n, 1)
    fwd_bytes = 8.0 * F * H * W + 25.0 * N
    return {"metric": "batched draw_lines Mlines/s (C3)", "value": round(N / (graph_ms * 1e-3) / 1e6, 4),
            "unit": "Mlines/s", "ms_per_step": round(graph_ms, 4),
            "frames_per_s": round(F / (graph_ms * 1e-3), 1),
            "per_frame_loop": {"ms_per_step": round(loop_ms, 3), "Mlines_per_s": round(N / (loop_ms * 1e-3) / 1e6, 4),
                               "calls": len(per)},
            "speedup_vs_per_frame_loop": round(loop_ms / graph_ms, 2),
            "eager_plan_ms": round(eager_ms, 4), "images_agree": same,
            "autotuned": opts, "tune_table": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                                              for r in best['table']],
            "roofline": {"bound": "hbm", "achieved": round(fwd_bytes / (tile_avg * 1e-3) / 1e9, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(fwd_bytes / (tile_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "kernel": "k_tiles", "avg_launch_ms": round(tile_avg, 4), "bytes_per_launch": fwd_bytes},
            "config": {"workload": f"batched uneven forward: {F} frames of {H}x{W} f32, {N} segments (per-frame counts "
                                   "log-uniform in [100, 20000], seed 2), per-line triangular/gaussian, sum; "
                                   "step = zero + one plan draw replayed from a HIP graph",
                       "segments": N}}


def workload_c4(args):
    """Forward + adjoint: 32 frames of 1024^2, 5000 segments each (C2
    distributions, seed 3), biweight; loss = sum((render - target)^2) with the
    target rendered from end points perturbed by N(0, 0.5)."""
    import numpy as np
    import torch
    from cbclib_v2_amd import _native, autograd, batch
    from tests.lines_gen import random_segments
    dev = torch.device('cuda', 0)
    F, H, W, per = 32, 1024, 1024, 5000
    rng = np.random.default_rng(3)
    truth = random_segments(rng, F * per, H, W, 20.0, 400.0, 2.0, 6.0)
    pert = truth.copy()
    pert[:, :4] += rng.normal(0.0, 0.5, (F * per, 4))
    idxs = torch.from_numpy(np.repeat(np.arange(F), per).astype(np.int32)).to(dev)
    lines = torch.tensor(truth, dtype=torch.float32, device=dev, requires_grad=True)
    target = autograd.render(torch.tensor(pert, dtype=torch.float32, device=dev), (F, H, W), idxs,
                             kernel='biweight').detach()

    # the step through a plan (sized once on the lines, headroom for their
    # drift): no host synchronisation, and the backward reuses the forward's
    # segments and tile lists; the eager step (draw_lines + adjoint, each
    # walking the lines) is timed beside it
    plan = batch.Plan((F, H, W), lines.detach(), idxs, kernel='biweight', headroom=1.25)

    def step():
        lines.grad = None
        img = autograd.render_plan(plan, lines, idxs, kernel='biweight')
        loss = torch.nn.functional.mse_loss(img, target, reduction="sum")   # sum((img - target)^2)
        loss.backward()

    def step_eager():
        lines.grad = None
        img = autograd.render(lines, (F, H, W), idxs, kernel='biweight')
        loss = torch.nn.functional.mse_loss(img, target, reduction="sum")   # sum((img - target)^2)
        loss.backward()

    eager_ms = timed(step_eager, args.steps, args.warmup)
    g_eager = lines.grad.clone()
    _native.timing_reset()
    _native.timing_enable(True)
    ms = timed(step, args.steps, args.warmup)
    _native.timing_enable(False)
    plan.check()
    same_grad = bool(torch.equal(lines.grad, g_eager))
    plan.close()
    adj_ms, adj_n = _native.timing_read('adjoint_tiles')
    N = F * per
    step_bytes = 12.0 * F * H * W + 68.0 * N
    adj_avg = adj_ms / max(adj_n, 1)
    adj_bytes = 4.0 * F * H * W + 44.0 * N
    adj_gbs = adj_bytes / (adj_avg * 1e-3) / 1e9 if adj_n else 0.0
    return {"metric": "fwd+adjoint frames/s (C4)", "value": round(F / (ms * 1e-3), 2), "unit": "frames/s",
            "adjoint_roofline": {"bound": "hbm", "achieved": round(adj_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(adj_gbs / HBM_PEAK_GBS, 4), "kernel": "k_adjoint_tiles",
                                 "bytes_per_launch": adj_bytes, "avg_launch_ms": round(adj_avg, 4)},
            "ms_per_step": round(ms, 4), "Mlines_per_s": round(N / (ms * 1e-3) / 1e6, 4),
            "config": {"workload": f"forward + adjoint: {F} frames of {H}x{W} f32, {N} segments ({per}/frame, "
                                   "C2 distributions, seed 3), biweight, loss sum((render - target)^2)"},
            "step_GBs": round(step_bytes / (ms * 1e-3) / 1e9, 2), "step_bytes": step_bytes,
            "adjoint_tiles_ms": round(adj_ms / max(adj_n, 1), 4), "path": "batch.Plan + autograd.render_plan",
            "eager": {"ms_per_step": round(eager_ms, 4), "frames_per_s": round(F / (eager_ms * 1e-3), 2),
                      "path": "autograd.render (draw_lines + draw_lines_adjoint)"},
            "grad_equals_eager": same_grad}


def workload_c5(args):
    """Refinement: 16 frames of 2048^2, 10000 segments each, biweight; start =
    truth + N(0, 1) end points, widths x U[0.8, 1.2]; target = Poisson(50 x
    render) / 50; Adam (lr 0.05) for `--adam-steps` steps."""
    import numpy as np
    import torch
    from cbclib_v2_amd import autograd, batch
    from tests.lines_gen import random_segments
    dev = torch.device('cuda', 0)
    F, H, W, per = 16, 2048, 2048, 10000
    rng = np.random.default_rng(4)
    truth = random_segments(rng, F * per, H, W, 20.0, 400.0, 2.0, 6.0)
    init = truth.copy()
    init[:, :4] += rng.normal(0.0, 1.0, (F * per, 4))
    init[:, 4] *= rng.uniform(0.8, 1.2, F * per)
    idxs = torch.from_numpy(np.repeat(np.arange(F), per).astype(np.int32)).to(dev)
    t_truth = torch.tensor(truth, dtype=torch.float32, device=dev)
    clean = autograd.render(t_truth, (F, H, W), idxs, kernel='biweight').detach()
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    target = torch.poisson(50.0 * clean, generator=gen) / 50.0
    lines = torch.nn.Parameter(torch.tensor(init, dtype=torch.float32, device=dev))
    opt = torch.optim.Adam([lines], lr=0.05, betas=(0.9, 0.999), eps=1e-8)
    # one plan for the whole refinement (headroom for the lines' drift; a
    # draw beyond it would be raised by the check after the loop)
    plan = batch.Plan((F, H, W), lines.detach(), idxs, kernel='biweight', headroom=1.5)

    def step(ln, o):
        o.zero_grad(set_to_none=True)
        img = autograd.render_plan(plan, ln, idxs, kernel='biweight')
        loss = torch.nn.functional.mse_loss(img, target, reduction="sum")   # sum((img - target)^2)
        loss.backward()
        o.step()

    # warm-up steps on a copy of the start (its own optimizer), so that the
    # timed refinement starts from `init` with the kernels and allocations warm
    warm = torch.nn.Parameter(lines.detach().clone())
    warm_opt = torch.optim.Adam([warm], lr=0.05, betas=(0.9, 0.999), eps=1e-8)
    for _ in range(max(args.warmup, 3)):
        step(warm, warm_opt)
    del warm, warm_opt

    def rmse(ln):
        a = ln[:, :4].detach().cpu().numpy().astype(np.float64)
        d1 = ((a - truth[:, :4]) ** 2).sum(1)
        d2 = ((a[:, [2, 3, 0, 1]] - truth[:, :4]) ** 2).sum(1)
        return float(np.sqrt(np.minimum(d1, d2).mean() / 2.0))

    r0 = rmse(lines)
    n = args.adam_steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        step(lines, opt)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    plan.check()
    plan.close()
    N = F * per
    return {"metric": "Adam refinement steps/s (C5)", "value": round(n / sec, 3), "unit": "steps/s",
            "path": "batch.Plan + autograd.render_plan",
            "ms_per_step": round(sec * 1e3 / n, 4), "adam_steps": n,
            "endpoint_rmse_px": {"start": round(r0, 4), "end": round(rmse(lines), 4)},
            "config": {"workload": f"refinement: {F} frames of {H}x{W} f32, {N} segments ({per}/frame, seed 4), "
                                   "biweight, target Poisson(50 render)/50, Adam lr 0.05"},
            "step_GBs": round((12.0 * F * H * W + 68.0 * N + 140.0 * N) / (sec / n) / 1e9, 2)}


def workload_acc(args):
    """accumulate_lines as the rotomap caller uses it (in-term max, out-term
    sum, gaussian): 16 frames of 512^2, 4000 terms of 16 segments (lengths
    U[5, 60], widths U[1, 3], seed 5), each term in a random frame.  Also
    times the equal-overlap (sum/sum) path on the same input; the output is
    compared with the oracle on the whole workload."""
    import numpy as np
    import torch
    from cbclib_v2_amd import ndimage
    from tests.lines_gen import random_segments
    dev = torch.device('cuda', 0)
    F, H, W, T, per = 16, 512, 512, 4000, 16
    rng = np.random.default_rng(5)
    lines_np = random_segments(rng, T * per, H, W, 5.0, 60.0, 1.0, 3.0).astype(np.float32)
    terms_np = np.repeat(np.arange(T), per).astype(np.int32)
    frames_np = rng.integers(0, F, T).astype(np.int32)
    lines = torch.from_numpy(lines_np).to(dev)
    terms = torch.from_numpy(terms_np).to(dev)
    frames = torch.from_numpy(frames_np).to(dev)
    out = torch.zeros((F, H, W), dtype=torch.float32, device=dev)
    N = T * per

    def step(i_ov, o_ov):
        out.zero_()
        ndimage.accumulate_lines(out, lines, terms, frames, 1.0, 'gaussian', i_ov, o_ov)

    ms = timed(lambda: step('max', 'sum'), args.steps, args.warmup)
    ms_eq = timed(lambda: step('sum', 'sum'), args.steps, args.warmup)
    res = {"metric": "accumulate_lines segments/s (in max, out sum)", "value": round(N / (ms * 1e-3) / 1e6, 4),
           "unit": "Mlines/s", "ms_per_step": round(ms, 4), "steps": args.steps, "warmup": args.warmup,
           "sum_sum_ms_per_step": round(ms_eq, 4),
           "config": {"workload": f"accumulate_lines: {F} frames of {H}x{W} f32, {T} terms x {per} segments "
                                  "(lengths U[5, 60], widths U[1, 3], seed 5), gaussian, in_overlap max, "
                                  "out_overlap sum"}}
    # the output at this scale against the oracle (single-threaded restatement
    # of bresenham.accumulate_lines), which is also the CPU baseline
    from oracle import oracle as O
    step('max', 'sum')
    got = out.cpu().numpy()
    o = np.zeros((F, H, W), np.float32)
    t0 = time.perf_counter()
    O.accumulate_lines(o, lines_np, terms_np, frames_np, 1.0, 'gaussian', 'max', 'sum')
    dt = time.perf_counter() - t0
    nz = o != 0
    res["parity"] = {"support_equal": bool(np.array_equal(got != 0, nz)),
                     "max_abs": float(np.abs(got - o).max()),
                     "max_rel": float((np.abs(got - o)[nz] / np.abs(o[nz])).max()) if nz.any() else 0.0,
                     "within_rtol_1e-5_atol_1e-6": bool(np.allclose(got, o, rtol=1e-5, atol=1e-6))}
    if not args.no_cpu:
        res["cpu_baseline"] = {"value": round(N / dt / 1e6, 6), "unit": "Mlines/s", "cores": 1, "kind": "port",
                               "sample": f"the whole workload once ({N} segments, {dt:.2f} s), the oracle's "
                                         "single-threaded restatement of bresenham.accumulate_lines"}
    return res


def _kernel_ms(name, fn, steps, warmup):
    """Average launch time (HIP events on the launch stream) of the library
    timer `name` over `steps` calls of fn, and the wall-clock ms per call."""
    from cbclib_v2_amd import _native
    for _ in range(warmup):
        fn()
    _native.timing_reset()
    _native.timing_enable(True)
    ms = timed(fn, steps, 0)
    _native.timing_enable(False)
    k_ms, k_n = _native.timing_read(name)
    return ms, (k_ms / k_n if k_n else None)


def _hbm(bytes_, ms, kernel):
    gbs = bytes_ / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel": kernel, "bytes_per_launch": bytes_,
            "avg_launch_ms": round(ms, 4)}


def workload_rows(args):
    """The rows beside the hot path, each on a realistic input, with its
    dominant kernel's roofline (HIP events) and the oracle, timed on a bounded
    sample, as CPU baseline:

    * region moments (center_of_mass, covariance_matrix, line_fit) and
      p_values of the C2 label output (64 frames of 2048^2: the labels of
      render > 0.5, the render as data);
    * binary_dilation of the C2 mask (render > 0.5, Structure([0, 1, 1], 2),
      1 iteration, and 4 iterations under a mask);
    * draw_lines on 3D lines: 16 volumes of 64 x 128 x 128 f32, 50 000
      segments (lengths U[5, 40], widths U[1, 3], seed 9), gaussian, sum and
      max."""
    import numpy as np
    import torch
    from cbclib_v2_amd import label as cl, ndimage
    from oracle import oracle as O
    from tests.lines_gen import config2, segments3
    dev = torch.device('cuda', 0)
    steps, warm = max(3, args.steps // 2), 2
    res = {}

    # -- region moments / p-values on the C2 label output
    F, H, W = 64, 2048, 2048
    lines_np, idxs_np = config2(frames=F, per_frame=2000, H=H, W=W)
    img = torch.zeros((F, H, W), dtype=torch.float32, device=dev)
    ndimage.draw_lines(img, torch.from_numpy(lines_np).to(dev), torch.from_numpy(idxs_np).to(dev), 1.0,
                       'gaussian', 'sum')
    st = cl.Structure([0, 1, 1], 2)
    lab = cl.label_threshold(img, 0.5, st)
    n_reg = int(lab.index.numel())
    P = F * H * W
    mom = {}
    for what in ('center_of_mass', 'covariance_matrix', 'line_fit'):
        fn = getattr(cl, what)
        ms, kms = _kernel_ms('region_moments', lambda: fn(lab, img), steps, warm)
        mom[what] = {"ms_per_call": round(ms, 4), "roofline": _hbm(8.0 * P, kms, "k_region_accumulate")}
    fit = cl.line_fit(lab, img)
    ms, kms = _kernel_ms('region_counts', lambda: cl.p_values(lab, fit, img, 0.3, 0.6, 1.5), steps, warm)
    mom["p_values"] = {"ms_per_call": round(ms, 4), "roofline": _hbm(8.0 * P, kms, "k_region_count")}
    # parity and CPU baseline on frame 0 (the oracle's single-threaded restatement)
    f0_lab = lab.labels[0].cpu().numpy()
    f0_img = img[0].cpu().numpy()
    ids = np.unique(f0_lab[f0_lab > 0]).astype(np.int64)
    t0 = time.perf_counter()
    ref = O.region_moments('line_fit', f0_lab, ids, f0_img)
    cpu_dt = time.perf_counter() - t0
    got = cl.line_fit(cl.CPLabelResult(labels=lab.labels[0].contiguous(), index=torch.from_numpy(ids).to(dev)),
                      img[0].contiguous()).cpu().numpy()
    scale = max(1.0, float(np.abs(ref).max()) if ref.size else 1.0)
    res["moments"] = {"n_regions": n_reg, **mom,
                      "parity_frame0": {"regions": int(ids.size),
                                        "line_fit_max_abs": float(np.nanmax(np.abs(got - ref))) if ids.size else 0.0,
                                        "within_1e-4_of_scale": bool(np.allclose(got, ref, atol=1e-4 * scale,
                                                                                 rtol=1e-4, equal_nan=True))},
                      "cpu_baseline": {"value": round(H * W / cpu_dt / 1e6, 3), "unit": "Mpx/s", "cores": 1,
                                       "kind": "port", "sample": "line_fit of frame 0 (2048^2 labels + data)"},
                      "gpu_Mpx_per_s_line_fit": round(P / (mom['line_fit']['ms_per_call'] * 1e-3) / 1e6, 1),
                      "config": {"workload": f"region moments of label(render > 0.5) on the C2 render: "
                                             f"{F} frames of {H}x{W} f32, {n_reg} regions"}}

    # -- binary_dilation of the C2 mask
    mask = img > 0.5
    sel = torch.rand((F, H, W), device=dev) < 0.7
    dil = {}
    for name, it, m in (("1_iteration", 1, None), ("4_iterations_masked", 4, sel)):
        ms, kms = _kernel_ms('dilation', lambda: cl.binary_dilation(mask, st, it, m), steps, warm)
        bytes_ = (3.0 if m is not None else 2.0) * P
        dil[name] = {"ms_per_call": round(ms, 4), "roofline": _hbm(bytes_, kms, "binary_dilation (pack, rounds, unpack)"),
                     "Gpx_per_s": round(P / (ms * 1e-3) / 1e9, 2)}
    m0 = mask[0].cpu().numpy()
    t0 = time.perf_counter()
    ref = O.binary_dilation(m0[None], st.shifts(), 1)[0]
    cpu_dt = time.perf_counter() - t0
    got = cl.binary_dilation(mask[0:1], st, 1)[0].cpu().numpy()
    res["dilation"] = {**dil, "parity_frame0_bit_exact": bool(np.array_equal(got, ref)),
                       "mask_density": round(float(mask.float().mean()), 5),
                       "cpu_baseline": {"value": round(H * W / cpu_dt / 1e9, 4), "unit": "Gpx/s", "cores": 1,
                                        "kind": "port", "sample": "frame 0, 1 iteration"},
                       "config": {"workload": f"binary_dilation(render > 0.5, Structure([0,1,1],2)) on the C2 "
                                              f"render: {F} frames of {H}x{W} bool"}}
    del img, mask, sel, lab
    torch.cuda.empty_cache()

    # -- 3D lines
    shape = (16, 64, 128, 128)
    rng = np.random.default_rng(9)
    n3 = 50000
    l3 = segments3(rng, n3, shape, 40.0, 1.0, 3.0, 2.0)
    l3[:, 3:6] = l3[:, 0:3] + (l3[:, 3:6] - l3[:, 0:3]) * (5.0 + 35.0 * rng.random((n3, 1))) / np.maximum(
        np.linalg.norm(l3[:, 3:6] - l3[:, 0:3], axis=1, keepdims=True), 1e-9)
    l3 = l3.astype(np.float32)
    i3 = rng.integers(0, shape[0], n3).astype(np.int32)
    t_l3, t_i3 = torch.from_numpy(l3).to(dev), torch.from_numpy(i3).to(dev)
    vol = torch.zeros(shape, dtype=torch.float32, device=dev)
    V = float(np.prod(shape))
    d3 = {}
    for ov in ('sum', 'max'):
        def call():
            vol.zero_()
            ndimage.draw_lines(vol, t_l3, t_i3, 1.0, 'gaussian', ov)
        ms, kms = _kernel_ms('draw3', call, steps, warm)
        d3[ov] = {"ms_per_call": round(ms, 4), "Mlines_per_s": round(n3 / (ms * 1e-3) / 1e6, 3),
                  "roofline": _hbm(8.0 * V + 28.0 * n3, kms, "k_draw3")}
    sel3 = i3 == 0
    o3 = np.zeros((1,) + shape[1:], np.float32)
    t0 = time.perf_counter()
    O.draw_lines3(o3, l3[sel3], np.zeros(int(sel3.sum()), np.int32), 1.0, 'gaussian', 'max')
    cpu_dt = time.perf_counter() - t0
    vol.zero_()
    ndimage.draw_lines(vol, t_l3, t_i3, 1.0, 'gaussian', 'max')
    res["draw3"] = {**d3, "parity_volume0_max_bit_exact": bool(np.array_equal(vol[0].cpu().numpy(), o3[0])),
                    "voxels_drawn_fraction": round(float((vol != 0).float().mean()), 4),
                    "cpu_baseline": {"value": round(int(sel3.sum()) / cpu_dt / 1e6, 4), "unit": "Mlines/s",
                                     "cores": 1, "kind": "port",
                                     "sample": f"volume 0 ({int(sel3.sum())} segments), max"},
                    "config": {"workload": f"3D draw_lines: {shape[0]} volumes of {shape[1]}x{shape[2]}x{shape[3]} "
                                           f"f32, {n3} segments (lengths U[5, 40], widths U[1, 3], seed 9), "
                                           "gaussian"}}
    del vol
    torch.cuda.empty_cache()
    return {"metric": "rows beside the hot path (moments, dilation, 3D lines)", "rows": res,
            "steps": steps, "warmup": warm}


def pmc_summary(src: str) -> str:
    """Per-kernel mean per dispatch of a rocprofv3 counter CSV
    (--pmc ... --output-format csv), as the text kept under profiles/."""
    import collections
    import csv
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.Counter()
    meta = {}
    for r in csv.DictReader(open(src)):
        k = r['Kernel_Name']
        agg[k][r['Counter_Name']] += float(r['Counter_Value'])
        cnt[(k, r['Counter_Name'])] += 1
        meta[k] = (r['VGPR_Count'], r['Scratch_Size'], r['LDS_Block_Size'], r['Workgroup_Size'], r['Grid_Size'])
    lines = []
    for k, d in sorted(agg.items(), key=lambda kv: -kv[1].get('SQ_BUSY_CYCLES', 0)):
        n = max(cnt[(k, c)] for c in d)
        lines.append(f"{k}\n  dispatches {n}  vgpr,scratch,lds,wg,grid {meta[k]}")
        for c, v in sorted(d.items()):
            lines.append(f"  {c:26s} {v / cnt[(k, c)]:18.1f}")
        if d.get('SQ_ACTIVE_INST_LDS'):
            lines.append(f"  {'conflict/active LDS':26s} {d['SQ_LDS_BANK_CONFLICT'] / d['SQ_ACTIVE_INST_LDS']:18.3f}")
    return "\n".join(lines)


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--frames', type=int, default=64)
    ap.add_argument('--tile', type=int, default=0)
    ap.add_argument('--seg', type=int, default=0)
    ap.add_argument('--threads', type=int, default=0)
    ap.add_argument('--sweep', action='store_true', help='time the tile kernel over launch shapes and exit')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--workload', choices=['c2', 'c3', 'c4', 'c5', 'acc', 'rows'], default='c2',
                    help='c2 (default, the headline line); c3 batched uneven forward; c4 forward + adjoint; '
                         'c5 Adam refinement; acc accumulate_lines (in max, out sum); rows region moments, '
                         'binary_dilation and 3D lines')
    ap.add_argument('--adam-steps', type=int, default=200)
    ap.add_argument('--headline-only', action='store_true',
                    help='the default workload only (no C3 / C4 summaries in the line)')
    ap.add_argument('--calibrate', action='store_true',
                    help='also run an elementwise kernel of known traffic (256 MiB read + 256 MiB written), the '
                         'reference for scaling FETCH_SIZE / WRITE_SIZE counters')
    ap.add_argument('--deterministic', action='store_true',
                    help='draw in the deterministic (reference-order) mode; for profiling k_tiles_ordered')
    ap.add_argument('--pmc-summary', metavar='CSV',
                    help='print the per-kernel means of a rocprofv3 counter CSV and exit (no GPU use)')
    args = ap.parse_args()
    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary))
        return
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        fail_multi(max(args.gpus, world))

    if args.workload in ('c3', 'c4', 'c5', 'acc', 'rows'):
        res = {'c3': workload_c3, 'c4': workload_c4, 'c5': workload_c5, 'acc': workload_acc,
               'rows': workload_rows}[args.workload](args)
        res.update({"n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic"})
        print(json.dumps(res))
        return

    import numpy as np
    import torch
    from cbclib_v2_amd import _native, label as cl, ndimage
    from tests.lines_gen import config2

    dev = torch.device('cuda', 0)
    F, H, W, per = args.frames, 2048, 2048, 2000
    lines_np, idxs_np = config2(frames=F, per_frame=per, H=H, W=W)
    N = lines_np.shape[0]
    lines = torch.from_numpy(lines_np).to(dev)
    idxs = torch.from_numpy(idxs_np).to(dev)
    out = torch.empty((F, H, W), dtype=torch.float32, device=dev)
    st = cl.Structure([0, 1, 1], 2)
    kw = {}
    if args.tile:
        kw['tile'] = args.tile
    if args.seg:
        kw['seg'] = args.seg
    if args.threads:
        kw['threads'] = args.threads
    if args.deterministic:
        kw['deterministic'] = True
    if args.sweep:
        sweep(out, lines, idxs, F, H, W, N)
        return

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    acc = {'zero': 0.0, 'fwd': 0.0, 'label': 0.0}
    n_labels = 0

    def step(measure: bool):
        nonlocal n_labels
        if measure:
            ev[0].record()
        out.zero_()
        if measure:
            ev[1].record()
        ndimage.draw_lines(out, lines, idxs, 1.0, 'gaussian', 'sum', **kw)
        if measure:
            ev[2].record()
        res = cl.label_threshold(out, 0.5, st)
        n_labels = int(res.index.numel())
        if measure:
            ev[3].record()
            torch.cuda.synchronize()
            acc['zero'] += ev[0].elapsed_time(ev[1])
            acc['fwd'] += ev[1].elapsed_time(ev[2])
            acc['label'] += ev[2].elapsed_time(ev[3])

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    # the timed steps record HIP events around the tile kernel only (each
    # event pair costs the step a few microseconds)
    _native.timing_reset()
    _native.timing_enable(True, tiles_only=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _native.timing_enable(False)
    tiles_ms, tiles_n = _native.timing_read('draw_tiles')
    # per-part breakdown (and the walk's and label's own timers) on separate,
    # untimed-for-value steps
    _native.timing_reset()
    _native.timing_enable(True)
    for _ in range(2):
        step(True)
    torch.cuda.synchronize()
    _native.timing_enable(False)
    label_ms, label_n = _native.timing_read('label_total')
    walk_ms, walk_n = _native.timing_read('segments_walk')

    ms_step = (t1 - t0) * 1e3 / args.steps
    tile_avg = tiles_ms / max(tiles_n, 1)
    fwd_bytes = 8.0 * F * H * W + 24.0 * N
    achieved = fwd_bytes / (tile_avg * 1e-3) / 1e9 if tile_avg > 0 else 0.0
    mask_density = float((out > 0.5).float().mean())
    result = {
        "metric": "draw_lines Mlines/s and fwd+bwd frames/s on 1 MI355X; % HBM peak; vs CPU",
        "value": round(N / (ms_step * 1e-3) / 1e6, 4),
        "unit": "Mlines/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain forward + label: {F} frames of {H}x{W} f32, {N} segments ({per}/frame, "
                               "L U[20,400], w U[2,6], random angles, seed 1), gaussian, sum; label(out > 0.5, "
                               "Structure([0,1,1],2)); step = zero + draw_lines + label",
                   "frames": F, "segments": N, "tile": args.tile or 64, "seg": args.seg or 8},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic('k_tiles'),
                     "kernel": "k_tiles (draw_lines tile accumulation)",
                     "bytes_per_launch": fwd_bytes, "avg_launch_ms": round(tile_avg, 4), "launches": tiles_n},
        "fwd_Mlines_per_s": round(N / (acc['fwd'] / 2 * 1e-3) / 1e6, 4),
        "fwd_label_frames_per_s": round(F / (ms_step * 1e-3), 2),
        "breakdown_ms": {k: round(v / 2, 4) for k, v in acc.items()},
        "segments_walk_ms": round(walk_ms / max(walk_n, 1), 4),
        "label_total_ms": round(label_ms / max(label_n, 1), 4),
        "label": {"n_labels": n_labels, "mask_density": round(mask_density, 6),
                  "algorithmic_GBs": round(8.0 * F * H * W / (label_ms / max(label_n, 1) * 1e-3) / 1e9, 2)},
    }
    if args.calibrate:
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(lines_np, idxs_np, F, H, W)
    else:
        result["cpu_baseline"] = None
    if not args.headline_only:
        # the other configurations, measured in the same run (summaries)
        del out, lines, idxs
        torch.cuda.empty_cache()
        c3 = workload_c3(args)
        result["c3"] = {k: c3[k] for k in ("ms_per_step", "value", "frames_per_s", "speedup_vs_per_frame_loop",
                                           "per_frame_loop", "images_agree", "autotuned", "roofline", "config")}
        result["c3"]["Mlines_per_s"] = result["c3"].pop("value")
        torch.cuda.empty_cache()
        c4 = workload_c4(args)
        result["c4"] = {"fwd_adjoint_frames_per_s": c4["value"],
                        **{k: c4[k] for k in ("ms_per_step", "Mlines_per_s", "step_GBs", "step_bytes",
                                              "adjoint_tiles_ms", "adjoint_roofline", "config", "path", "eager",
                                              "grad_equals_eager")}}
    print(json.dumps(result))


if __name__ == '__main__':
    main()
