"""Pins the numpy MLP oracle (oracle/mlp_ref.py) by central finite differences,
so the GPU gradients are compared against gradients that are themselves checked
against the definition of the derivative."""

import numpy as np

from oracle import mlp_ref as M


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1).astype(np.float64)
    x = rng.normal(size=(16, 32))
    t = rng.normal(size=(16, 2))
    _, g = M.loss_grad(p, x, t)
    eps = 1e-6
    for i in rng.choice(M.NPARAM, 60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss(pp, x, t) - M.loss(pm, x, t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd)


def test_oracle_layout_matches_header():
    assert M.NPARAM == 6402
    d = M.unpack(np.arange(M.NPARAM, dtype=np.float64))
    assert d["W1"][1, 0] == 32 and d["b1"][0] == 2048 and d["W3"].shape == (2, 64) and d["b3"][1] == 6401
