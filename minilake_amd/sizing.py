"""Learned hash-table sizing.

A 32-64-64-2 FP32 ReLU MLP, trained and evaluated by the library's own HIP
kernels (ml_mlp_*), predicts from cheap input features
  * log2 of the number of distinct GROUP BY keys, and
  * log2 of the number of rows a join produces,
so the aggregate's group table can be allocated once at its final size instead
of growing by grow-and-rehash (MiniLake's SwissTable::reserve,
crates/minilake-hashtable/src/swiss.rs:219-232), and a join's output buffer
without the exact counting pass. It replaces MiniLake's fixed estimate
(aggregate = input / 10, crates/minilake-sql/src/physical_planner.rs:251-295)
for sizing only: a prediction that is too small makes the exact path rerun the
batch with plain resize-on-overflow sizing, so results never depend on it.

Features (ml_sizing_features): row-group min/max statistics (122,880-row row
groups, the shape of crates/minilake-storage/src/stats.rs:15-22) and a
4,096-row sample with its KMV sketch and frequency profile. Join-side
features 27-31 are filled here from the build table.

Training pairs come from the plain engine itself: synthetic key columns are
aggregated and joined on the GPU with exact sizing, and the true counts are
the labels.
"""

from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional, Tuple

import numpy as np

from . import abi
from . import workloads as W

ROW_GROUP_ROWS = 122_880
SAMPLE = 4096
_HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(_HERE, "estimator_weights.npy")


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialized parameters in the header's layout (W1, b1, W2, b2, W3, b3)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, math.sqrt(2 / 32), (64, 32)), np.zeros(64), rng.normal(0, math.sqrt(2 / 64), (64, 64)),
             np.zeros(64), rng.normal(0, math.sqrt(1 / 64), (2, 64)), np.zeros(2)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


class Mlp:
    """The device-resident network (ml_mlp) of one GpuEngine."""

    def __init__(self, engine, params: Optional[np.ndarray] = None):
        from .gpu import lib
        self.engine = engine
        p = np.ascontiguousarray(init_params() if params is None else params, dtype=np.float32)
        assert p.size == abi.MLP_NUM_PARAMS
        h = C.c_void_p()
        engine._check(lib().ml_mlp_create(engine.ctx, p.ctypes.data, C.byref(h)))
        self.handle = h.value

    def params(self) -> np.ndarray:
        from .gpu import lib
        out = np.zeros(abi.MLP_NUM_PARAMS, dtype=np.float32)
        self.engine._check(lib().ml_mlp_get_params(self.handle, out.ctypes.data))
        return out

    def _dev(self, a):
        torch = self.engine.torch
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(f"cuda:{self.engine.device}")

    def forward(self, x: np.ndarray) -> np.ndarray:
        from .gpu import lib
        xd = self._dev(x)
        y = self.engine.torch.zeros((len(x), 2), dtype=self.engine.torch.float32, device=xd.device)
        self.engine._check(lib().ml_mlp_forward(self.handle, xd.data_ptr(), len(x), y.data_ptr()))
        return y.cpu().numpy()

    def loss_grad(self, x: np.ndarray, t: np.ndarray) -> Tuple[float, np.ndarray]:
        from .gpu import lib
        xd, td = self._dev(x), self._dev(t)
        loss = C.c_float()
        g = np.zeros(abi.MLP_NUM_PARAMS, dtype=np.float32)
        self.engine._check(lib().ml_mlp_loss_grad(self.handle, xd.data_ptr(), td.data_ptr(), len(x), C.byref(loss),
                                                  g.ctypes.data))
        return loss.value, g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 200, batch: int = 256, lr: float = 2e-3,
              seed: int = 0, log=None) -> float:
        """Adam over shuffled minibatches; the data stays in HBM."""
        from .gpu import lib
        torch = self.engine.torch
        xd, td = self._dev(x), self._dev(t)
        n = len(x)
        rng = np.random.default_rng(seed)
        loss = C.c_float()
        last = float("nan")
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).to(xd.device)
            xs, ts = xd[perm].contiguous(), td[perm].contiguous()
            step_lr = lr * (0.5 * (1 + math.cos(math.pi * ep / epochs)))
            for b in range(0, n, batch):
                m = min(batch, n - b)
                want = (b + batch >= n)
                self.engine._check(lib().ml_mlp_train_step(self.handle, xs[b].data_ptr(), ts[b].data_ptr(), m,
                                                           step_lr, C.byref(loss) if want else None))
            last = loss.value
            if log and (ep % 20 == 0 or ep == epochs - 1):
                log(f"epoch {ep}: batch loss {last:.4f}")
        return last

    def __del__(self):
        from . import gpu
        if getattr(self, "handle", None) and gpu._lib is not None:
            gpu._lib.ml_mlp_free(self.handle)
            self.handle = None


def column_features(engine, keys_dev, n: int, seed: int = 0) -> np.ndarray:
    """ml_sizing_features over a device int64 key column."""
    from .gpu import lib
    f = np.zeros(32, dtype=np.float32)
    engine._check(lib().ml_sizing_features(engine.ctx, keys_dev.data_ptr(), n, ROW_GROUP_ROWS, SAMPLE, seed,
                                           f.ctypes.data))
    return f


def join_features(f: np.ndarray, build_rows: int, build_distinct: int, sample_hits: int, sample_matches: int):
    """Fill features 27-31 from the build side: rows, distinct keys, share of the
    probe sample that finds a key, matches per sampled probe row."""
    f = f.copy()
    f[27] = math.log2(max(1, build_rows))
    f[28] = math.log2(max(1, build_distinct))
    f[29] = build_rows / max(1, build_distinct)
    f[30] = sample_hits / SAMPLE
    f[31] = math.log2(1 + sample_matches * (2 ** float(f[0])) / SAMPLE)
    return f


# ----------------------------------------------------------- training tables
def synthetic_table(seed: int, rows: int, domain: int, s: float, build_mult: int, build_share: float):
    """A key column of `rows` Zipf(s) draws over `domain` keys (scrambled so key
    order carries no information) and a build table holding a `build_share`
    fraction of the domain's keys, each `build_mult` times."""
    ranks = W.zipf(seed, 901, rows, s, domain)
    perm_seed = seed * 7919 + 13
    keys = W.permutation_of(perm_seed, 902, domain, ranks - 1)
    nb = max(1, int(domain * build_share))
    bkeys = W.permutation(perm_seed, 903, domain, count=nb)
    bkeys = np.repeat(bkeys, build_mult)
    return keys.astype(np.int64), bkeys.astype(np.int64)
