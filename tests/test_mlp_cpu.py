"""Pins the warm-start MLP's CPU oracle (tests/mlp_ref.py): its analytic
gradient against central finite differences of its own loss, bfloat16
rounding against the definition, GELU's derivative against differences."""
import numpy as np

import mlp_ref as M


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.3e38],
                 dtype=np.float32)
    r = M.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even
    assert r[2] == np.float32(1.0 + 2 ** -7)      # exact
    assert r[3] == np.float32(1.0 + 2 ** -6)      # tie -> even (up)
    assert r[4] == -2.5
    assert np.all((M.bf16_round(x).view(np.uint32) & 0xFFFF) == 0)


def test_gelu_grad_matches_differences():
    x = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(x), fd, atol=1e-8)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64)
    p[640:704] = rng.normal(0, 0.1, 64)      # nonzero biases exercise every term
    p[-20:] = rng.normal(0, 0.5, 20)
    B = 16
    s = np.abs(rng.normal(0, 20, (10, B))) + rng.normal(0, 1, (10, B))
    xref = np.where(rng.uniform(0, 1, (10, B)) < 0.6, rng.exponential(10, (10, B)), 0.0)
    loss, g = M.loss_and_grad(p, s, xref, quantize=False)
    idx = rng.choice(M.NPARAM, 120, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, s, xref, quantize=False)
        lm, _ = M.loss_and_grad(p - e, s, xref, quantize=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_adam_matches_closed_form_first_step():
    p = np.ones(5)
    g = np.array([1.0, -2.0, 0.5, 0.0, 3.0])
    q, m, v = M.adam(p, g, np.zeros(5), np.zeros(5), 1, 0.01)
    # first bias-corrected step is lr * sign(g) (for |g| >> eps)
    assert np.allclose(q[[0, 1, 2, 4]], p[[0, 1, 2, 4]] - 0.01 * np.sign(g[[0, 1, 2, 4]]),
                       atol=1e-7)
    assert q[3] == 1.0
