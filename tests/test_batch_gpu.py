"""Per-line profiles, plans, graph replay and the autotuner on the GPU.

Float sums of the tile path accumulate in fixed point, which does not depend
on the order of the terms, so a plan's draw is bit-identical to draw_lines
with the same launch shape; against the oracle the bar is the tile path's
(rtol 1e-5, atol 1e-6 * max_val; support exact).
"""
import numpy as np
import pytest
import torch

from cbclib_v2_amd import _native, autograd, batch, ndimage
from tests.lines_gen import random_segments

pytestmark = pytest.mark.gpu

KERNELS = ['biweight', 'gaussian', 'parabolic', 'rectangular', 'triangular']


def frames_case(seed, F=3, H=96, W=80, n=120):
    rng = np.random.default_rng(seed)
    lines = random_segments(rng, n, H, W, 0.0, 70.0, 0.5, 6.0, pad=10.0).astype(np.float32)
    idxs = rng.integers(0, F, n).astype(np.int32)
    kern = rng.integers(0, 5, n).astype(np.uint8)
    return (F, H, W), lines, idxs, kern


def test_per_line_profiles_match_oracle(cuda, oracle):
    for seed in range(4):
        shape, lines, idxs, kern = frames_case(seed)
        ref = np.zeros(shape, np.float32)
        for k in range(5):
            sel = kern == k
            if sel.any():
                oracle.draw_lines(ref, lines[sel], idxs[sel], 1.3, KERNELS[k], 'sum')
        out = torch.zeros(shape, device='cuda')
        ndimage.draw_lines(out, torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(), 1.3,
                           kernels=torch.from_numpy(kern).cuda())
        got = out.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1.3e-6)
        assert np.array_equal(got != 0, ref != 0)
        names = [KERNELS[k] for k in kern]
        out2 = ndimage.draw_lines(np.zeros(shape, np.float32), lines, idxs, 1.3, kernels=names, overlap='max')
        ref2 = np.zeros(shape, np.float32)
        for k in range(5):
            sel = kern == k
            if sel.any():
                oracle.draw_lines(ref2, lines[sel], idxs[sel], 1.3, KERNELS[k], 'max')
        np.testing.assert_allclose(out2, ref2, rtol=1e-5, atol=1.3e-6)


def test_per_line_profile_errors(cuda):
    out = torch.zeros(2, 16, 16, device='cuda')
    lines = torch.tensor([[1.0, 2.0, 10.0, 12.0, 2.0]] * 2, device='cuda')
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines, torch.tensor([0, 1], dtype=torch.int32, device='cuda'), kernels=[0, 7])
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines, torch.tensor([0, 1], dtype=torch.int32, device='cuda'),
                           kernels=torch.tensor([0, 9], dtype=torch.uint8, device='cuda'))
    with pytest.raises(ValueError):
        ndimage.draw_lines(out, lines, None, deterministic=True, kernels=['gaussian', 'biweight'])
    assert float(out.abs().sum()) == 0.0


@pytest.mark.parametrize('tile,seg', [(64, 8), (32, 16)])
def test_plan_matches_draw_lines(cuda, tile, seg):
    for seed in range(3):
        shape, lines, idxs, kern = frames_case(10 + seed)
        t_lines, t_idxs, t_kern = (torch.from_numpy(a).cuda() for a in (lines, idxs, kern))
        ref = torch.zeros(shape, device='cuda')
        ndimage.draw_lines(ref, t_lines, t_idxs, 0.9, kernels=t_kern, tile=tile, seg=seg)
        plan = batch.Plan(shape, t_lines, t_idxs, t_kern, tile=tile, seg=seg)
        out = torch.zeros(shape, device='cuda')
        plan.draw(out, t_lines, t_idxs, t_kern, max_val=0.9)
        plan.check()
        assert torch.equal(out, ref)
        # a perturbed line set through the same plan (headroom)
        pert = t_lines.clone()
        pert[:, :4] += torch.randn_like(pert[:, :4]) * 0.7
        ref.zero_()
        out.zero_()
        ndimage.draw_lines(ref, pert, t_idxs, 0.9, kernels=t_kern, tile=tile, seg=seg)
        plan.draw(out, pert, t_idxs, t_kern, max_val=0.9)
        plan.check()
        assert torch.equal(out, ref)
        # single profile, max overlap
        ref.zero_()
        out.zero_()
        ndimage.draw_lines(ref, t_lines, t_idxs, 0.9, 'gaussian', 'max', tile=tile, seg=seg)
        plan.draw(out, t_lines, t_idxs, None, 0.9, 'gaussian', 'max')
        plan.check()
        assert torch.equal(out, ref)
        plan.close()


def test_plan_errors_leave_out_untouched(cuda):
    shape = (2, 96, 80)
    rng = np.random.default_rng(20)
    n = 200
    small = np.concatenate([rng.uniform(5, 75, (n, 2)), rng.uniform(5, 75, (n, 2)), np.full((n, 1), 0.5)], 1)
    small[:, 2:4] = small[:, :2] + 0.8                      # tiny lines: ~1 slot, ~1 item each
    t_small = torch.tensor(small, dtype=torch.float32, device='cuda')
    t_idxs = torch.from_numpy(rng.integers(0, 2, n).astype(np.int32)).cuda()
    plan = batch.Plan(shape, t_small, t_idxs, headroom=1.0, kernel='parabolic')
    out = torch.full(shape, 0.5, device='cuda')
    # more lines than the plan was made for
    with pytest.raises(ValueError):
        plan.draw(out, torch.cat([t_small, t_small]), torch.cat([t_idxs, t_idxs]))
    # long, wide lines: far more segments and tile items than the sizing set
    big = t_small.clone()
    big[:, 0] = 0.0
    big[:, 2] = 79.0
    big[:, 4] = 20.0
    plan.draw(out, big, t_idxs)
    # the adjoint of the failed draw: a zero gradient, the error at the check
    g_big = plan.adjoint(torch.ones(shape, device='cuda'), big, t_idxs)
    with pytest.raises(_native.CapacityError):
        plan.check()
    assert float((out - 0.5).abs().max()) == 0.0
    assert float(g_big.abs().max()) == 0.0
    # an index outside the frames
    bad = t_idxs.clone()
    bad[3] = shape[0]
    plan.draw(out, t_small, bad)
    with pytest.raises(IndexError):
        plan.check()
    assert float((out - 0.5).abs().max()) == 0.0
    # and a good draw after the errors
    plan.draw(out, t_small, t_idxs)
    plan.check()
    assert float((out - 0.5).abs().max()) > 0.0
    plan.close()


@pytest.mark.parametrize('tile,seg', [(64, 8), (32, 16)])
def test_plan_adjoint_matches_eager(cuda, tile, seg):
    """The plan's adjoint runs on its last draw's segments and tile lists; a
    line's partial sums are added in the same slot order as in the eager
    adjoint, so the gradients are identical."""
    for seed, kernel in enumerate(['biweight', 'gaussian', 'parabolic', 'triangular', 'rectangular']):
        shape, lines, idxs, _ = frames_case(50 + seed)
        t_lines, t_idxs = torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda()
        g_img = torch.from_numpy(np.random.default_rng(seed).normal(size=shape).astype(np.float32)).cuda()
        plan = batch.Plan(shape, t_lines, t_idxs, kernel=kernel, tile=tile, seg=seg, headroom=1.5)
        out = torch.zeros(shape, device='cuda')
        for rep in range(2):
            if rep:   # a perturbed set through the same plan
                t_lines = t_lines.clone()
                t_lines[:, :4] += 0.6 * torch.randn_like(t_lines[:, :4])
            ref = autograd.adjoint(g_img, t_lines, t_idxs, 1.2, kernel, tile, seg)
            out.zero_()
            plan.draw(out, t_lines, t_idxs, max_val=1.2)
            got = plan.adjoint(g_img, t_lines, t_idxs, 1.2)
            plan.check()
            assert torch.equal(got, ref), f"{kernel} rep {rep}: max |diff| {float((got - ref).abs().max())}"
        plan.close()


def test_render_plan_matches_render(cuda):
    shape, lines, idxs, _ = frames_case(60, F=2, H=128, W=112, n=150)
    t_idxs = torch.from_numpy(idxs).cuda()
    target = torch.rand(shape, device='cuda')
    plan = batch.Plan(shape, torch.from_numpy(lines).cuda(), t_idxs, kernel='biweight', headroom=1.5)
    res = []
    for use_plan in (False, True):
        x = torch.from_numpy(lines).cuda().requires_grad_(True)
        img = autograd.render_plan(plan, x, t_idxs) if use_plan else autograd.render(x, shape, t_idxs)
        ((img - target) ** 2).sum().backward()
        res.append((img.detach(), x.grad))
    plan.check()
    assert torch.equal(res[0][0], res[1][0])
    assert torch.equal(res[0][1], res[1][1])
    # a draw between the forward and its backward; a different line count
    x = torch.from_numpy(lines).cuda().requires_grad_(True)
    img = autograd.render_plan(plan, x, t_idxs)
    plan.draw(torch.zeros(shape, device='cuda'), x.detach(), t_idxs)
    with pytest.raises(RuntimeError):
        img.sum().backward()
    with pytest.raises(ValueError):
        plan.adjoint(torch.zeros(shape, device='cuda'), x.detach()[:10], t_idxs[:10])
    plan.close()


def test_graph_replays_match_eager(cuda):
    """Replays with eager work and synchronisations between them; the graph's
    inputs are updated in place; every replay equals the eager draw."""
    shape, lines, idxs, kern = frames_case(30, F=4, H=128, W=96, n=300)
    t_lines, t_idxs, t_kern = (torch.from_numpy(a).cuda() for a in (lines, idxs, kern))
    plan = batch.Plan(shape, t_lines, t_idxs, t_kern, headroom=1.5)
    out = torch.zeros(shape, device='cuda')
    g = batch.GraphDraw(plan, out, t_lines, t_idxs, t_kern, max_val=1.1)
    rng = np.random.default_rng(0)
    for rep in range(4):
        if rep:
            t_lines[:, :4] += torch.from_numpy(rng.normal(0, 0.3, (300, 4)).astype(np.float32)).cuda()
        junk = torch.randn(1000, device='cuda').sum()    # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        plan.check()
        ref = torch.zeros(shape, device='cuda')
        ndimage.draw_lines(ref, t_lines, t_idxs, 1.1, kernels=t_kern)
        assert torch.equal(out, ref), f"replay {rep}"
        del junk
    plan.close()


def test_tune(cuda):
    shape, lines, idxs, kern = frames_case(40, F=2, H=128, W=128, n=200)
    best = batch.tune(shape, torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(),
                      torch.from_numpy(kern).cuda(), repeats=2, use_cache=False)
    assert (best['tile'], best['seg'], best['threads'], best['chunk']) in batch.CANDIDATES
    assert len(best['table']) == len(batch.CANDIDATES) and best['ms'] > 0
