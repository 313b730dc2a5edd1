"""FP64 numpy statement of the learned subgrid closure (test helper).

The network maps 10 inputs (the velocity-gradient tensor g_ij / |g| of the
filtered field and the wall distance y / delta99) to the dimensionless eddy
viscosity c = nu_t / (Delta^2 |S|):

    h1 = silu(W1 x + b1)   (64)
    h2 = silu(W2 h1 + b2)  (64)
    h3 = silu(W3 h2 + b3)  (64)
    c  = W4 h3 + b4        (1)

Loss: mean squared error over the batch.  Adam with bias correction.  This
module is the checker of the device kernels (forward values, gradients, one
Adam update) and is itself pinned by central finite differences.

Parameter vector layout (shared with the device, FP32 there):
W1[64][10], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4[1].
"""
import numpy as np

NIN, NH = 10, 64
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)),
          ("b3", (NH,)), ("W4", (NH,)), ("b4", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def init_params(seed=0):
    """He-style uniform init, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[1] if len(shp) == 2 else shp[0]
            lim = np.sqrt(6.0 / fan_in) * 0.5
            parts.append(rng.uniform(-lim, lim, size=int(np.prod(shp))))
        else:
            parts.append(np.zeros(int(np.prod(shp))))
    return np.concatenate(parts)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, X):
    p = unpack(theta)
    z1 = X @ p["W1"].T + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = silu(z2)
    z3 = h2 @ p["W3"].T + p["b3"]
    h3 = silu(z3)
    y = h3 @ p["W4"] + p["b4"][0]
    return y, (z1, h1, z2, h2, z3, h3)


def loss(theta, X, t):
    y, _ = forward(theta, X)
    return np.mean((y - t) ** 2)


def grad(theta, X, t):
    """Gradient of the mean squared error, in the parameter layout."""
    p = unpack(theta)
    y, (z1, h1, z2, h2, z3, h3) = forward(theta, X)
    B = X.shape[0]
    dy = 2.0 * (y - t) / B
    g = {}
    g["W4"] = h3.T @ dy
    g["b4"] = np.array([dy.sum()])
    d3 = np.outer(dy, p["W4"]) * dsilu(z3)
    g["W3"] = d3.T @ h2
    g["b3"] = d3.sum(0)
    d2 = (d3 @ p["W3"]) * dsilu(z2)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * dsilu(z1)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in SHAPES])


def adam(theta, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new theta, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
