"""Adjoint of draw_lines on the GPU against the oracle's FP64 gradient over
the same pixel set (rtol 1e-5, atol 1e-7 of the gradient scale), and against
FP64 central finite differences (h = 1e-3) of the closed-form render: that
render and the walk's f32 render differ on boundary pixels, hence rtol 2e-2
there (the C4 test checks the FD of the closed-form residual at 1e-5 scale)."""
import numpy as np
import pytest
import torch

from cbclib_v2_amd import autograd
from tests.lines_gen import random_segments

pytestmark = pytest.mark.gpu


def setup(seed, n=40, F=2, H=90, W=110):
    rng = np.random.default_rng(seed)
    lines = random_segments(rng, n, H, W, 5.0, 60.0, 2.0, 5.0).astype(np.float32)
    idxs = rng.integers(0, F, n).astype(np.int32)
    return lines, idxs, (F, H, W), rng


@pytest.mark.parametrize('kernel', ['biweight', 'parabolic', 'gaussian', 'triangular'])
def test_adjoint_vs_oracle(cuda, oracle, kernel):
    lines, idxs, shape, rng = setup(3)
    g = rng.standard_normal(shape).astype(np.float32)
    got = autograd.adjoint(torch.from_numpy(g).cuda(), torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(),
                           1.3, kernel).cpu().numpy()
    ref = oracle.adjoint_f64(g.astype(np.float64), lines, idxs, 1.3, kernel, walk='f32')
    scale = np.abs(ref).max()
    big = np.abs(ref) > 1e-3 * scale
    print(f"{kernel}: max-rel {float((np.abs(got - ref)[big] / np.abs(ref[big])).max()):.3e} over components "
          f"above 1e-3 of the scale", flush=True)
    # the bar of the north star (1e-5 relative), with atol 1e-7 of the scale for cancelled sums
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-7 * scale)


def test_adjoint_rectangular_is_zero(cuda):
    lines, idxs, shape, rng = setup(4)
    g = rng.standard_normal(shape).astype(np.float32)
    got = autograd.adjoint(torch.from_numpy(g).cuda(), torch.from_numpy(lines).cuda(), torch.from_numpy(idxs).cuda(),
                           1.0, 'rectangular')
    assert float(got.abs().sum()) == 0.0


def test_autograd_loss_vs_fd(cuda, oracle):
    """loss = sum (render(lines) - target)^2, biweight; FP64 central differences of the closed form."""
    lines, idxs, shape, rng = setup(5, n=20)
    target = oracle.render_closed_f64(shape, lines.astype(np.float64) + np.r_[rng.normal(0, 0.5, 4), 0.0], idxs,
                                      1.0, 'biweight')
    t = torch.from_numpy(lines).cuda().requires_grad_(True)
    img = autograd.render(t, shape, torch.from_numpy(idxs).cuda(), 1.0, 'biweight')
    loss = (img - torch.from_numpy(target.astype(np.float32)).cuda()).square().sum()
    loss.backward()
    got = t.grad.cpu().numpy()

    def f(l64):
        return np.sum((oracle.render_closed_f64(shape, l64, idxs, 1.0, 'biweight') - target) ** 2)

    base = lines.astype(np.float64)
    h = 1e-3
    fd = np.zeros_like(base)
    for i in range(base.shape[0]):
        for j in range(5):
            p = base.copy(); p[i, j] += h
            m = base.copy(); m[i, j] -= h
            fd[i, j] = (f(p) - f(m)) / (2 * h)
    scale = np.abs(fd).max()
    np.testing.assert_allclose(got, fd, rtol=2e-2, atol=2e-3 * scale)
