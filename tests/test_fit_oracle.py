"""CPU checks of the texture-fitting oracle (film L2 loss, projected Adam) and of the argument
checks of their C-ABI entry points (which return before any device work).

The loss and Adam have no reference counterpart. The oracle's binary32 restatements are pinned
here by torch's independent implementations (torch.optim.Adam, autograd of the loss), a few
ulps apart because their operation order differs, and by closed forms."""
import ctypes as C

import numpy as np
import pytest
import torch


def test_adam_matches_torch_optim_adam(oracle_module):
    rng = np.random.default_rng(1)
    p0 = rng.standard_normal(4099).astype(np.float32)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=0.01, betas=(0.9, 0.999), eps=1e-8, foreach=False)
    for step in range(1, 8):
        g = (rng.standard_normal(p0.shape) * 10.0 ** rng.uniform(-3, 1)).astype(np.float32)
        p, m, v = oracle_module.adam_step(p, g, m, v, 0.01, step)
        tp.grad = torch.tensor(g)
        opt.step()
        want = tp.detach().numpy()
        # a few ulps of the values (|p| ~ 1) or 5e-6 of one lr-sized step
        assert np.all(np.abs(p - want) <= 1e-6 * np.abs(want) + 5e-8), (step, np.abs(p - want).max())
        state = opt.state[tp]
        # torch forms m by lerp: a few ulps of the largest term (m cancels where g changes sign)
        want_m, want_v = state["exp_avg"].numpy(), state["exp_avg_sq"].numpy()
        assert np.allclose(m, want_m, rtol=1e-6, atol=1e-6 * np.abs(want_m).max())
        assert np.allclose(v, want_v, rtol=1e-6, atol=0.0)


def test_adam_first_step_closed_form(oracle_module):
    """With m = v = 0, the bias-corrected first step is p - lr * g / (|g| + eps): a step of lr
    against the gradient's sign."""
    rng = np.random.default_rng(2)
    p0 = rng.uniform(0.2, 0.8, 1000).astype(np.float32)
    g = (rng.choice([-1.0, 1.0], 1000) * rng.uniform(1e-3, 1e3, 1000)).astype(np.float32)
    p, m, v = oracle_module.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 0.05, 1)
    assert np.allclose(p, p0 - 0.05 * np.sign(g), rtol=0, atol=2e-7)
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, 0.001 * g.astype(np.float64) ** 2, rtol=1e-6)


def test_adam_projection(oracle_module):
    p0 = np.array([0.0, 0.01, 0.99, 1.0, 0.5], dtype=np.float32)
    g = np.array([1.0, 1.0, -1.0, -1.0, 0.0], dtype=np.float32)
    p, _, _ = oracle_module.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 0.05, 1, lower=0.0, upper=1.0)
    assert p.tolist() == [0.0, 0.0, 1.0, 1.0, 0.5]


def test_film_loss_matches_autograd(oracle_module):
    rng = np.random.default_rng(3)
    pixels = 777
    counts = rng.integers(0, 9, pixels).astype(np.uint32)
    film = (rng.uniform(0, 2, pixels * 3) * np.repeat(counts, 3)).astype(np.float32)
    target = rng.uniform(0, 1, pixels * 3).astype(np.float32)
    scale = 1.0 / (3 * pixels)
    loss, gradient = oracle_module.film_l2_loss(film, counts, target, scale)
    f = torch.tensor(film.astype(np.float64), requires_grad=True)
    c = torch.tensor(np.repeat(np.maximum(counts, 1), 3).astype(np.float64))
    want = float(np.float32(scale)) * ((f / c - torch.tensor(target.astype(np.float64))) ** 2).sum()
    want.backward()
    assert abs(loss - float(want)) <= 1e-6 * float(want)
    assert np.allclose(gradient, f.grad.numpy(), rtol=1e-5, atol=1e-12)
    # pixels without samples: mean 0, gradient (-2 scale target) / 1
    empty = np.repeat(counts == 0, 3)
    assert np.allclose(gradient[empty], -2 * np.float32(scale) * target[empty], rtol=1e-6)


def test_planted_adam_mistake_is_caught(oracle_module):
    """A restatement with the bias correction of v dropped differs from torch.optim.Adam: the
    comparison above is sensitive to that class of mistake."""
    p0 = np.full(8, 0.5, dtype=np.float32)
    g = np.linspace(-1, 1, 8).astype(np.float32)
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=0.01)
    tp.grad = torch.tensor(g)
    opt.step()
    m = 0.1 * g
    wrong = p0 - (0.01 / 0.1) * m / (np.sqrt(0.001 * g * g) + 1e-8)  # v not bias-corrected
    assert not np.allclose(wrong, tp.detach().numpy(), rtol=0, atol=1e-5)


def test_fit_entry_points_reject_bad_arguments(hip, abi):
    lib = hip.library()
    param = abi.AdamParam()
    param.count = 0
    opts = abi.AdamOptions(0.9, 0.999, 1e-8, 0, 0)
    assert lib.blackframe_hip_adam_step(C.byref(param), 1, C.byref(opts), None) == 1  # step 0
    opts.step = 1
    opts.beta1 = 1.0
    assert lib.blackframe_hip_adam_step(C.byref(param), 1, C.byref(opts), None) == 1
    opts.beta1 = 0.9
    param.reserved = 1
    assert lib.blackframe_hip_adam_step(C.byref(param), 1, C.byref(opts), None) == 1
    param.reserved, param.count = 0, 4
    assert lib.blackframe_hip_adam_step(C.byref(param), 1, C.byref(opts), None) == 1  # NULL buffers
    job = abi.FilmLossJob()
    job.pixel_count = 4
    assert lib.blackframe_hip_film_l2_loss(C.byref(job), 1, 1.0, C.c_void_p(8), None) == 1  # NULL buffers
    assert lib.blackframe_hip_film_l2_loss(C.byref(job), 1, float("nan"), C.c_void_p(8), None) == 1
    assert lib.blackframe_hip_adam_step(None, 0, None, None) == 0  # nothing to do
