"""Pins the CPU oracle of the warm-start network (oracle/mlp_oracle.py) before it is
used to check the GPU kernels: BF16 rounding against torch's independent conversion,
the float64 backward pass against central finite differences of the loss, and the Adam
update against its closed form at the first step."""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        # exact ties between two BF16 neighbours: both parities
                        (1.0 + (2 * np.arange(64) + 1) * 2.0 ** -8).astype(np.float32),
                        np.array([0.0, -0.0, 1e-30, -3.0e38, 65504.0], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(MO.bf16(x), ref)


def test_float64_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    widths = [3, 32, 32, 32]
    P = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(len(widths) - 1))
    p = rng.standard_normal(P) * 0.3
    x = rng.standard_normal((6, 3))
    t = rng.standard_normal((6, 32))
    loss, g = MO.loss_grad(p, widths, x, t, bf16_ops=False)
    h = 1e-6
    idx = rng.choice(P, 60, replace=False)
    for i in idx:
        e = np.zeros(P)
        e[i] = h
        lp, _ = MO.loss_grad(p + e, widths, x, t, bf16_ops=False)
        lm, _ = MO.loss_grad(p - e, widths, x, t, bf16_ops=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_bf16_mode_stays_close_to_float64():
    """BF16 operands perturb the gradient by about their 2^-9 relative rounding."""
    rng = np.random.default_rng(9)
    widths = [9, 64, 64, 32]
    P = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(len(widths) - 1))
    p = rng.standard_normal(P) * 0.1
    x = rng.uniform(-1, 1, (64, 9))
    t = rng.uniform(-1, 1, (64, 32))
    _, g64 = MO.loss_grad(p, widths, x, t, bf16_ops=False)
    _, g16 = MO.loss_grad(p, widths, x, t, bf16_ops=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 2e-2


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(1)
    p, g = rng.standard_normal(100), rng.standard_normal(100)
    p1, m, v = MO.adam_step(p, g, np.zeros(100), np.zeros(100), 1, 1e-3)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
