"""Learned closure training kernels on the device against the FP64 numpy
statement (tests/mlp_ref.py): forward values, the gradient of the batch loss,
the batch drawn by the device's counter hash, one Adam update, and a short
training run that lowers the loss on a smooth regression target.

Tolerances: FP32 device arithmetic against FP64: forward 1e-5 relative,
gradient 1e-4 relative in the L2 norm, parameters after one Adam step 1e-6
absolute (the step is lr * sign-like ratio, lr = 1e-3).
"""
import sys

import numpy as np
import pytest

from tests import mlp_ref as M
from tests.blgpu_import import blgpu

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def batch_rows(n, B, seed, step):
    """Rows the device gathers for training step `step` (0-based)."""
    return np.array([mix64(seed ^ mix64((step * 0x100000001B3 + b) & MASK)) % n for b in range(B)])


def data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, M.NIN))
    t = 0.05 * np.tanh(X[:, 0] * X[:, 1] + 0.5 * X[:, 9]) + 0.02 * X[:, 4] ** 2
    return X, t


@pytest.fixture
def net():
    c = blgpu.Closure(capacity=4096, batch_max=1024)
    theta = M.init_params(7).astype(np.float32)
    c.params = theta
    yield c, theta.astype(np.float64)
    c.close()


def test_forward_matches_reference(net):
    c, theta = net
    X, _ = data(777)
    y = c.forward(X)
    yr, _ = M.forward(theta, X.astype(np.float32).astype(np.float64))
    e = np.abs(y - yr).max() / np.abs(yr).max()
    say("forward rel err", e)
    assert e < 1e-5


@pytest.mark.parametrize("n", [1, 100, 1024])
def test_gradient_matches_reference(net, n):
    c, theta = net
    X, t = data(n, seed=n)
    g, loss = c.gradient(X, t)
    Xr = X.astype(np.float32).astype(np.float64)
    tr = t.astype(np.float32).astype(np.float64)
    gr = M.grad(theta, Xr, tr)
    e = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say("n", n, "grad rel err", e, "loss", loss, M.loss(theta, Xr, tr))
    assert e < 1e-4
    assert loss == pytest.approx(M.loss(theta, Xr, tr), rel=1e-5)


def test_one_adam_step_matches_reference(net):
    c, theta = net
    n, B, seed = 3000, 512, 12345
    X, t = data(n, seed=3)
    rows = np.zeros((n, blgpu.MLP_ROW), dtype=np.float32)
    rows[:, :10] = X
    rows[:, 10] = t
    rows[:, 11] = 1.0
    c.add_rows(rows)
    assert c.count == n
    np.testing.assert_array_equal(c.rows(0, 5), rows[:5])
    hist = c.train(1, B, lr=1e-3, seed=seed)
    idx = batch_rows(n, B, seed, 0)
    Xb = rows[idx, :10].astype(np.float64)
    tb = rows[idx, 10].astype(np.float64)
    assert hist[0] == pytest.approx(M.loss(theta, Xb, tb), rel=1e-5)
    g = M.grad(theta, Xb, tb)
    th1, _, _ = M.adam(theta, np.zeros_like(theta), np.zeros_like(theta), g, 1, lr=1e-3)
    got = c.params.astype(np.float64)
    # components with a tiny gradient may take a different sign-like step
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    err = np.abs(got - th1)[big].max()
    say("adam step max err", err, "moved", np.abs(got - theta).max())
    assert err < 1e-6


def test_training_lowers_the_loss(net):
    c, _ = net
    n = 4096
    X, t = data(n, seed=5)
    rows = np.zeros((n, blgpu.MLP_ROW), dtype=np.float32)
    rows[:, :10] = X
    rows[:, 10] = t
    rows[:, 11] = 1.0
    c.add_rows(rows)
    hist = c.train(400, 1024, lr=2e-3, seed=1)
    first, last = hist[:20].mean(), hist[-20:].mean()
    say("loss first", first, "last", last)
    assert np.all(np.isfinite(hist))
    assert last < 0.2 * first
