"""The benchmark workloads themselves, on the GPU, against the oracle.

* C2 (64 frames of 2048^2, 128 000 segments, gaussian sum, then
  label(out > 0.5, Structure([0, 1, 1], 2))): every frame of the tile-path
  render (support bit-exact; values within rtol 1e-5 with atol 1e-6 *
  max_val, max-abs / max-rel printed, also against the FP64 render), the
  deterministic render bit-exact, and every frame's labels bit-exact.
* C3 (256 frames of 1024^2, log-uniform per-frame counts, per-line
  triangular / gaussian): a plan draw replayed from a HIP graph against the
  oracle on all frames, and graph replay == eager plan draw == the per-frame
  loop of plain draw_lines calls.
* C4 (32 frames of 1024^2, 5000 segments each, biweight, loss
  sum((render - target)^2)): the adjoint against the oracle's FP64 gradient
  on frames 0-1, and against FP64 central finite differences (h = 1e-3) of
  the closed-form loss on 200 segments of those frames.
* C5 (16 frames of 2048^2, 10 000 segments each): 5 Adam steps at full size;
  the first step's gradient on frame 0 against the oracle; the loss falls.

The oracle runs single-threaded per frame (each frame keeps the reference's
order), frames spread over a thread pool of the host's CPU share.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from cbclib_v2_amd import autograd, batch, ndimage
from cbclib_v2_amd import label as cl
from tests.lines_gen import config2, config3, random_segments

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))
RTOL, ATOL = 1e-5, 1e-6


def say(msg):
    print(msg, flush=True)


def frame_groups(idxs, n_frames):
    order = np.argsort(idxs, kind='stable')
    bounds = np.searchsorted(idxs[order], np.arange(n_frames + 1))
    return [order[bounds[f]:bounds[f + 1]] for f in range(n_frames)]


def oracle_frames(oracle, shape, lines, idxs, max_val, kernel_of, overlap='sum', frames=None):
    """The oracle's render of the given frames (all by default), one
    single-threaded oracle call per frame and profile, frames in parallel.
    kernel_of: a profile name, or an array of per-line profile names."""
    F, H, W = shape
    groups = frame_groups(idxs, F)
    frames = range(F) if frames is None else frames
    out = {f: np.zeros((H, W), lines.dtype) for f in frames}

    def one(f):
        sel = groups[f]
        if isinstance(kernel_of, str):
            oracle.draw_lines(out[f], lines[sel], None, max_val, kernel_of, overlap)
        else:
            for k in ('triangular', 'gaussian', 'biweight', 'parabolic', 'rectangular'):
                s = sel[kernel_of[sel] == k]
                if s.size:
                    oracle.draw_lines(out[f], lines[s], None, max_val, k, overlap)
        return f

    with ThreadPoolExecutor(WORKERS) as ex:
        list(ex.map(one, frames))
    return out


def compare_image(got, ref, max_val, what):
    """Support bit-exact, values within RTOL / ATOL * max_val; returns (max_abs, max_rel)."""
    assert np.array_equal(got != 0, ref != 0), f"{what}: support differs"
    d = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    nz = ref != 0
    rel = d[nz] / np.abs(ref[nz].astype(np.float64)) if nz.any() else np.zeros(1)
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=ATOL * max_val, err_msg=what)
    return float(d.max()), float(rel.max())


# ----------------------------------------------------------------------- C2

def test_config2_full(cuda, oracle):
    F, H, W = 64, 2048, 2048
    lines, idxs = config2()
    say(f"C2: {lines.shape[0]} segments over {F} frames of {H}x{W}")
    t_lines, t_idxs = torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda()
    out = torch.zeros((F, H, W), dtype=torch.float32, device='cuda')
    ndimage.draw_lines(out, t_lines, t_idxs, 1.0, 'gaussian', 'sum')
    torch.cuda.synchronize()
    say("C2: tile-path render done; oracle render")
    ref = oracle_frames(oracle, (F, H, W), lines, idxs, 1.0, 'gaussian')
    worst_abs = worst_rel = 0.0
    for f in range(F):
        a, r = compare_image(out[f].cpu().numpy(), ref[f], 1.0, f"C2 frame {f}")
        worst_abs, worst_rel = max(worst_abs, a), max(worst_rel, r)
    say(f"C2 tile path vs oracle (f32, reference order): max-abs {worst_abs:.3e}, max-rel {worst_rel:.3e}")

    # the same frames against the FP64 render (the double walk): report only
    groups = frame_groups(idxs, F)
    d_abs, n_diff, n_px = 0.0, 0, 0
    for f in (0, 1):
        r64 = oracle.draw_lines(np.zeros((H, W)), lines[groups[f]].astype(np.float64), None, 1.0, 'gaussian', 'sum')
        g = out[f].cpu().numpy().astype(np.float64)
        both = (g != 0) & (r64 != 0)
        d_abs = max(d_abs, float(np.abs(g - r64)[both].max()))
        n_diff += int(((g != 0) != (r64 != 0)).sum())
        n_px += int(both.sum())
    say(f"C2 tile path vs FP64 oracle (frames 0-1): max-abs {d_abs:.3e} over {n_px} common pixels, "
        f"{n_diff} pixels in only one support (boundary decisions of the f32 vs f64 walk)")

    # label of the GPU render, every frame against the oracle (per-frame ids plus the running offset)
    st = cl.Structure([0, 1, 1], 2)
    res = cl.label_threshold(out, 0.5, st)
    torch.cuda.synchronize()
    say(f"C2: label done, {int(res.index.numel())} labels; oracle per frame")
    imgs = [out[f].cpu().numpy() for f in range(F)]

    def lab(f):
        return oracle.label(imgs[f] > 0.5, st.squeeze().shifts(), 1)

    with ThreadPoolExecutor(WORKERS) as ex:
        per = list(ex.map(lab, range(F)))
    off = 0
    for f in range(F):
        m, n = per[f]
        got = res.labels[f].cpu().numpy()
        exp = np.where(m > 0, m + off, 0)
        assert np.array_equal(got, exp), f"C2 frame {f}: label map differs"
        off += n
    assert int(res.index.numel()) == off
    say(f"C2: labels bit-exact on all {F} frames ({off} labels)")

    # deterministic mode: the single-threaded reference's order, every frame
    out.zero_()
    ndimage.draw_lines(out, t_lines, t_idxs, 1.0, 'gaussian', 'sum', deterministic=True)
    torch.cuda.synchronize()
    worst = 0.0
    for f in range(F):
        g = out[f].cpu().numpy()
        assert np.array_equal(g != 0, ref[f] != 0), f"C2 deterministic frame {f}: support"
        # device exp vs glibc exp: last-ulp differences of the gaussian values only
        np.testing.assert_allclose(g, ref[f], rtol=3e-7, atol=0, err_msg=f"C2 deterministic frame {f}")
        nz = ref[f] != 0
        worst = max(worst, float((np.abs(g - ref[f])[nz] / ref[f][nz]).max()))
    say(f"C2 deterministic vs oracle: max-rel {worst:.3e} (0 = bit-identical)")


# ----------------------------------------------------------------------- C3

def test_config3_full(cuda, oracle):
    F, H, W = 256, 1024, 1024
    lines, idxs, prof, counts = config3()
    N = lines.shape[0]
    names = np.where(prof == 0, 'triangular', 'gaussian')
    kern = torch.from_numpy(np.where(prof == 0, 4, 1).astype(np.uint8)).cuda()
    t_lines, t_idxs = torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda()
    say(f"C3: {N} segments over {F} frames of {H}x{W}, {counts.min()}..{counts.max()} per frame")
    plan = batch.Plan((F, H, W), t_lines, t_idxs, kern, headroom=1.1, tile=64, seg=16, threads=256, chunk=64)
    out_eager = torch.zeros((F, H, W), dtype=torch.float32, device='cuda')
    plan.draw(out_eager, t_lines, t_idxs, kern)
    plan.check()
    out_graph = torch.zeros_like(out_eager)
    graph = batch.GraphDraw(plan, out_graph, t_lines, t_idxs, kern)     # zero + plan draw per replay
    out_loop = torch.zeros_like(out_eager)
    groups = frame_groups(idxs, F)
    per_frame = []
    for f in range(F):
        for k, name in ((0, 'triangular'), (1, 'gaussian')):
            sel = groups[f][prof[groups[f]] == k]
            if sel.size:
                per_frame.append((f, name, t_lines[torch.from_numpy(sel).cuda()]))
    identical = []
    for rep in range(3):
        out_graph.fill_(7.0)
        graph.replay()
        torch.cuda.synchronize()
        plan.check()
        identical.append(bool(torch.equal(out_graph, out_eager)))
        torch.testing.assert_close(out_graph, out_eager, rtol=RTOL, atol=ATOL)
        # eager work of the program between replays: a third of the per-frame loop
        for f, name, ln in per_frame[rep::3]:
            ndimage.draw_lines(out_loop[f], ln, None, 1.0, name, 'sum')
        torch.cuda.synchronize()
    say(f"C3: graph replays bit-identical to the eager plan draw: {identical}")
    torch.testing.assert_close(out_loop, out_eager, rtol=RTOL, atol=ATOL)
    say("C3: per-frame loop (plain draw_lines per frame and profile) agrees with the plan draw")

    ref = oracle_frames(oracle, (F, H, W), lines, idxs, 1.0, names)
    worst_abs = worst_rel = 0.0
    for f in range(F):
        a, r = compare_image(out_eager[f].cpu().numpy(), ref[f], 1.0, f"C3 frame {f}")
        worst_abs, worst_rel = max(worst_abs, a), max(worst_rel, r)
    say(f"C3 plan draw vs oracle, all {F} frames: max-abs {worst_abs:.3e}, max-rel {worst_rel:.3e}")
    heavy, light = np.argsort(counts)[-4:], np.argsort(counts)[:4]
    for f in np.concatenate([heavy, light]):
        compare_image(out_loop[f].cpu().numpy(), ref[f], 1.0, f"C3 per-frame loop frame {f}")
    say(f"C3: per-frame loop vs oracle on heaviest {list(heavy)} and lightest {list(light)} frames")


# ----------------------------------------------------------------------- C4

def config4():
    F, H, W, per = 32, 1024, 1024, 5000
    rng = np.random.default_rng(3)
    truth = random_segments(rng, F * per, H, W, 20.0, 400.0, 2.0, 6.0)
    pert = truth.copy()
    pert[:, :4] += rng.normal(0.0, 0.5, (F * per, 4))
    idxs = np.repeat(np.arange(F), per).astype(np.int32)
    return (F, H, W), truth.astype(np.float32), pert.astype(np.float32), idxs


def test_config4_adjoint(cuda, oracle):
    shape, truth, pert, idxs = config4()
    F, H, W = shape
    t_idxs = torch.from_numpy(idxs).cuda()
    lines = torch.from_numpy(truth).cuda().requires_grad_(True)
    target = autograd.render(torch.from_numpy(pert).cuda(), shape, t_idxs, kernel='biweight').detach()
    img = autograd.render(lines, shape, t_idxs, kernel='biweight')
    loss = ((img - target) ** 2).sum()
    loss.backward()
    grad = lines.grad.cpu().numpy()
    dimg = (2.0 * (img - target)).detach()
    say(f"C4: loss {float(loss.detach()):.6e}; oracle FP64 adjoint on frames 0-1")
    sel = np.flatnonzero(idxs < 2)
    g2 = dimg[:2].cpu().numpy().astype(np.float64)
    ref = oracle.adjoint_f64(g2, truth[sel], idxs[sel], 1.0, 'biweight', walk='f32')
    got = grad[sel]
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref)
    big = np.abs(ref) > 1e-3 * scale
    say(f"C4 adjoint vs oracle: max-abs {err.max():.3e} (component scale {scale.max():.3e}), "
        f"max-rel {float((err[big] / np.abs(ref[big])).max()):.3e} over components above 1e-3 of their scale")
    # f32 output of double sums: rtol 1e-5; atol 1e-7 of each component's scale for cancelled sums
    bad = err > RTOL * np.abs(ref) + 1e-7 * scale
    assert not bad.any(), f"C4 adjoint: {int(bad.sum())} components outside rtol 1e-5 / atol 1e-7 * scale"

    # FP64 central finite differences of the closed-form loss on 200 segments of frames 0-1
    rng = np.random.default_rng(33)
    pick = rng.choice(sel, 200, replace=False)
    groups = frame_groups(idxs, F)
    img64, tgt64 = {}, {}
    for f in (0, 1):
        img64[f] = oracle.render_closed_f64((H, W), truth[groups[f]].astype(np.float64), None, 1.0, 'biweight')
        tgt64[f] = oracle.render_closed_f64((H, W), pert[groups[f]].astype(np.float64), None, 1.0, 'biweight')
    fd = oracle.loss_fd_f64(img64, tgt64, truth.astype(np.float64), idxs, pick, 1e-3, 1.0, 'biweight')
    nrm = np.linalg.norm(fd, axis=1)
    rows = np.searchsorted(sel, pick)

    def dev(an):
        return np.abs(an - fd).max(axis=1) / nrm

    # (a) the adjoint itself: the GPU adjoint of the closed-form FP64 residual 2 (I - T).  Central
    # differences at h = 1e-3 agree with the FP64 adjoint to ~2.4e-5 of each segment's gradient
    # norm on this input (the closed form's second derivatives jump at the profile edge and the caps)
    g64 = np.stack([2.0 * (img64[0] - tgt64[0]), 2.0 * (img64[1] - tgt64[1])])
    an_c = autograd.adjoint(torch.from_numpy(g64.astype(np.float32)).cuda(), torch.from_numpy(truth[sel]).cuda(),
                            torch.from_numpy(idxs[sel]).cuda(), 1.0, 'biweight').cpu().numpy()[rows]
    # (b) end to end: the gradient of the loss of the f32 renders (the reference's float walk), which
    # differ from the closed form by float rounding of the walk's errors: ~4e-4 of the norm
    an_e = grad[pick].astype(np.float64)
    da, de = dev(an_c), dev(an_e)
    say(f"C4 vs FP64 central differences (h = 1e-3, 200 segments of frames 0-1), max |an - fd| / |fd_segment|: "
        f"adjoint of the closed-form residual {da.max():.3e} (median {np.median(da):.3e}); "
        f"end-to-end gradient of the f32 renders {de.max():.3e} (median {np.median(de):.3e})")
    assert da.max() < 1e-4
    assert de.max() < 1e-3


# ----------------------------------------------------------------------- C5

def test_config5_adam(cuda, oracle):
    F, H, W, per = 16, 2048, 2048, 10000
    rng = np.random.default_rng(4)
    truth = random_segments(rng, F * per, H, W, 20.0, 400.0, 2.0, 6.0)
    init = truth.copy()
    init[:, :4] += rng.normal(0.0, 1.0, (F * per, 4))
    init[:, 4] *= rng.uniform(0.8, 1.2, F * per)
    idxs = np.repeat(np.arange(F), per).astype(np.int32)
    t_idxs = torch.from_numpy(idxs).cuda()
    clean = autograd.render(torch.from_numpy(truth.astype(np.float32)).cuda(), (F, H, W), t_idxs,
                            kernel='biweight').detach()
    gen = torch.Generator(device='cuda')
    gen.manual_seed(4)
    target = torch.poisson(50.0 * clean, generator=gen) / 50.0
    lines = torch.nn.Parameter(torch.from_numpy(init.astype(np.float32)).cuda())
    opt = torch.optim.Adam([lines], lr=0.05, betas=(0.9, 0.999), eps=1e-8)
    losses = []
    for step in range(5):
        opt.zero_grad(set_to_none=True)
        img = autograd.render(lines, (F, H, W), t_idxs, kernel='biweight')
        loss = ((img - target) ** 2).sum()
        loss.backward()
        if step == 0:
            sel = np.flatnonzero(idxs == 0)
            g0 = (2.0 * (img - target))[0].detach().cpu().numpy().astype(np.float64)
            l0 = lines.detach().cpu().numpy()[sel]
            ref = oracle.adjoint_f64(g0[None], l0, None, 1.0, 'biweight', walk='f32')
            got = lines.grad.cpu().numpy()[sel]
            scale = np.abs(ref).max(axis=0)
            bad = np.abs(got - ref) > RTOL * np.abs(ref) + 1e-7 * scale
            assert not bad.any(), f"C5 gradient: {int(bad.sum())} components outside rtol 1e-5 / atol 1e-7 * scale"
            say(f"C5 step 0: frame-0 gradient vs oracle max-abs {np.abs(got - ref).max():.3e}")
        opt.step()
        losses.append(float(loss))
        say(f"C5 step {step}: loss {losses[-1]:.6e}")
    assert all(b < a for a, b in zip(losses, losses[1:])), losses
