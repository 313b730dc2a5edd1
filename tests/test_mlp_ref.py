"""Pins for the FP64 numpy statement of the learned closure (tests/mlp_ref.py):
central finite differences of the loss against the analytic gradient, the
SiLU derivative, and the first Adam step in closed form."""
import numpy as np

from tests import mlp_ref as M


def test_param_count():
    assert M.NPARAM == 64 * 10 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def test_dsilu_matches_finite_difference():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    np.testing.assert_allclose(M.dsilu(z), fd, rtol=1e-8, atol=1e-10)


def test_gradient_matches_finite_difference():
    rng = np.random.default_rng(3)
    theta = M.init_params(1) + 0.05 * rng.standard_normal(M.NPARAM)
    X = rng.standard_normal((37, M.NIN))
    t = rng.standard_normal(37) * 0.1
    g = M.grad(theta, X, t)
    idx = rng.choice(M.NPARAM, 200, replace=False)
    idx = np.r_[idx, M.NPARAM - 1, 0, 640, 704]  # include b4, W1[0], b1[0], W2[0]
    h = 1e-6
    for q in idx:
        e = np.zeros(M.NPARAM)
        e[q] = h
        fd = (M.loss(theta + e, X, t) - M.loss(theta - e, X, t)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])) + 1e-9, (q, fd, g[q])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    theta = rng.standard_normal(M.NPARAM)
    g = rng.standard_normal(M.NPARAM)
    m = np.zeros(M.NPARAM)
    v = np.zeros(M.NPARAM)
    th1, m1, v1 = M.adam(theta, m, v, g, 1, lr=1e-3)
    # bias-corrected first moments equal g and g^2, so the step is lr*g/(|g|+eps)
    np.testing.assert_allclose(th1, theta - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(v1, 0.001 * g * g)
