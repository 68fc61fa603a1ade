"""CPU ORACLE for the learned-sizing MLP — test infrastructure only.

float64 numpy forward/backward of the 32-64-64-2 ReLU network with the
parameter layout of include/minilake_gpu.h (W1[64][32], b1, W2[64][64], b2,
W3[2][64], b3) and loss = mean over samples and outputs of (y - t)^2. Its
gradients are themselves pinned by central finite differences in the tests.
"""

import numpy as np

IN, H1, H2, OUT = 32, 64, 64, 2
SHAPES = [("W1", (H1, IN)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)), ("W3", (OUT, H2)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SHAPES])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / IN), (H1, IN)), "b1": np.zeros(H1),
         "W2": rng.normal(0, np.sqrt(2 / H1), (H2, H1)), "b2": np.zeros(H2),
         "W3": rng.normal(0, np.sqrt(1 / H2), (OUT, H2)), "b3": np.zeros(OUT)}
    return pack(d).astype(np.float32)


def forward(p, x):
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (x, a1, h1, a2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def loss_grad(p, x, t):
    w = unpack(p)
    y, (x, a1, h1, a2, h2) = forward(p, x)
    n = x.shape[0]
    e = y - np.asarray(t, np.float64)
    dy = 2.0 * e / (n * OUT)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ w["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ x
    g["b1"] = dh1.sum(0)
    return float(np.mean(e ** 2)), pack(g)
