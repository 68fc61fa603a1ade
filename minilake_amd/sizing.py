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
        self._free = lib().ml_mlp_free  # bound now: module globals may be gone when __del__ runs at exit

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
        self.engine.synchronize()
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

    def close(self):
        if getattr(self, "handle", None):
            self._free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def column_features(engine, keys_dev, n: int, seed: int = 0) -> np.ndarray:
    """ml_sizing_features over a device int64 key column."""
    from .gpu import lib
    f = np.zeros(32, dtype=np.float32)
    engine._check(lib().ml_sizing_features(engine.ctx, keys_dev.data_ptr(), n, ROW_GROUP_ROWS, SAMPLE, seed,
                                           f.ctypes.data))
    return f


# ----------------------------------------------------------- training tables
def _splitmix_np(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def sample_rows(n: int, seed: int, s: int = SAMPLE) -> np.ndarray:
    """Row ids of a `s`-row sample of a table of n rows (seeded splitmix64)."""
    with np.errstate(over="ignore"):
        r = _splitmix_np(np.arange(s, dtype=np.uint64) + np.uint64(seed))
    return (r % np.uint64(max(n, 1))).astype(np.int64)


def features(engine, keys_dev, n: int, build_table, build_rows: int, build_distinct: int, seed: int = 0):
    """All 32 features for GROUP BY key and probe(key) against `build_table`."""
    from .gpu import DeviceTable
    from .abi import col, probe_op
    f = column_features(engine, keys_dev, n, seed)
    idx = engine.torch.from_numpy(sample_rows(n, seed + 1)).to(keys_dev.device)
    samp = DeviceTable.from_tensors([keys_dev[idx].contiguous()], [abi.INT64])
    matches = engine.count(samp, [probe_op(build_table, [col(0)])]) if build_table is not None else 0
    return join_features(f, build_rows, build_distinct, matches, matches)


def join_features(f: np.ndarray, build_rows: int, build_distinct: int, sample_hits: int, sample_matches: int):
    """Features 27-31 from the build side: rows, distinct keys, rows per key,
    matches per sampled probe row, and the scaled-up match estimate."""
    f = f.copy()
    f[27] = math.log2(max(1, build_rows))
    f[28] = math.log2(max(1, build_distinct))
    f[29] = build_rows / max(1, build_distinct)
    f[30] = sample_matches / SAMPLE
    f[31] = math.log2(1 + sample_matches * (2 ** float(f[0])) / SAMPLE)
    return f


class KeyGen:
    """Synthetic key columns generated in HBM (device twin of workloads.zipf_keys)."""

    def __init__(self, engine, max_rows: int):
        self.engine = engine
        torch = engine.torch
        self.keys = torch.empty(max(1, max_rows), dtype=torch.int64, device=f"cuda:{engine.device}")
        self.bkeys = torch.empty(max(1, max_rows), dtype=torch.int64, device=f"cuda:{engine.device}")
        self.stream = engine.stream.cuda_stream

    def probe(self, n, seed, s, domain):
        W.gpu_zipf_keys(self.keys.data_ptr(), n, seed, 11, s, domain, 12, self.stream)
        return self.keys[:n]

    def build(self, nkeys, mult, seed, domain, hot: bool = True):
        """hot: the build side holds the most frequent probe keys; else random keys of the domain."""
        W.gpu_build_keys(self.bkeys.data_ptr(), nkeys, mult, seed, 12 if hot else 13, domain, self.stream)
        return self.bkeys[:nkeys * mult]


def labelled_table(engine, gen: KeyGen, seed: int, rows: int, domain: int, s: float, share: float, mult: int,
                   hot: bool = True):
    """One training pair: features and the exact (log2 distinct, log2 1+join rows)
    computed by the plain engine."""
    from .gpu import DeviceTable
    from .abi import INT64, col, probe_op
    keys = gen.probe(rows, seed, s, domain)
    nb = max(1, min(int(domain * share), gen.bkeys.numel() // mult))
    bkeys = gen.build(nb, mult, seed, domain, hot)
    probe_t = DeviceTable.from_tensors([keys], [INT64])
    build_t = engine.join_build(DeviceTable.from_tensors([bkeys], [INT64]), [], [col(0)], [INT64])
    x = features(engine, keys, rows, build_t, nb * mult, nb, seed=seed)
    distinct = engine.count_distinct(probe_t, [], [col(0)], [INT64], sizing=abi.SIZING_HINT,
                                     capacity_hint=min(rows, domain))
    join_rows = engine.count(probe_t, [probe_op(build_t, [col(0)])])
    return x, np.array([math.log2(max(1, distinct)), math.log2(1 + join_rows)], dtype=np.float32)


def training_set(engine, ntables: int, seed: int = 1, min_rows: float = 1e4, max_rows: float = 1e8,
                 max_domain: float = 1e9, log=None):
    rng = np.random.default_rng(seed)
    gen = KeyGen(engine, int(max_rows))
    X = np.zeros((ntables, 32), np.float32)
    Y = np.zeros((ntables, 2), np.float32)
    for t in range(ntables):
        rows = int(10 ** rng.uniform(math.log10(min_rows), math.log10(max_rows)))
        domain = max(1, int(10 ** rng.uniform(0, math.log10(max_domain))))
        s = float(rng.uniform(0, 1.5))
        share = float(10 ** rng.uniform(-3, 0))
        mult = int(rng.integers(1, 4))
        hot = bool(rng.integers(0, 2))
        X[t], Y[t] = labelled_table(engine, gen, seed * 1_000_003 + t, rows, domain, s, share, mult, hot)
        if log and (t % 500 == 0 or t == ntables - 1):
            log(f"training table {t}/{ntables}")
    return X, Y


def load_params(path: str = WEIGHTS_PATH) -> Optional[np.ndarray]:
    if not os.path.exists(path):
        return None
    return np.load(path, allow_pickle=False).astype(np.float32)


def capacity_from_log2(pred: float, margin: float = 1.15) -> int:
    return int(math.ceil((2.0 ** float(pred)) * margin))


# ------------------------------------------------------------ evaluation
def zipf_workload(engine, rows: int = 60_000_000, domain: int = 100_000_000, s: float = 1.1,
                  build_keys: int = 2_000_000, mult: int = 2, seed: int = 42):
    """The learned-sizing evaluation: a `rows`-row fact table of Zipf(s) int64 keys
    (scrambled over `domain`), a measure w = key mod 1009, and a build table of
    `build_keys` random keys, each `mult` times, with payload b = key mod 7."""
    from .gpu import DeviceTable
    from .abi import INT64
    if build_keys > domain:
        raise ValueError(f"{build_keys} distinct build keys do not fit a key domain of {domain}")
    torch = engine.torch
    gen = KeyGen(engine, max(rows, build_keys * mult))
    keys = gen.probe(rows, seed, s, domain).clone()
    bkeys = gen.build(build_keys, mult, seed, domain, hot=False).clone()
    w = torch.remainder(keys, 1009)
    b = torch.remainder(bkeys, 7)
    fact = DeviceTable.from_tensors([keys, w], [INT64, INT64])
    build = DeviceTable.from_tensors([bkeys, b], [INT64, INT64])
    return fact, build


def run_plain(engine, fact, build, stats: Optional[dict] = None):
    """GROUP BY key (COUNT, SUM) with resize-on-overflow sizing, and the join
    collected after an exact counting pass."""
    from .abi import AGG_SUM, INT64, Agg, col, count_star, probe_op
    agg = engine.hashagg(fact, [], [col(0)], [INT64], [count_star(), Agg(AGG_SUM, col(1), INT64)],
                         sizing=abi.SIZING_PLAIN, stats=stats)
    jt = engine.join_build(build, [], [col(0)], [INT64])
    joined = engine.collect(fact, [probe_op(jt, [col(0)])])
    return agg, joined


# The group table plain sizing reserves before its first device batch (2^21
# rows: reserve(len + batch), the same 4M-slot capacity).
PLAIN_FIRST_RESERVE = 1 << 21


def run_plain_partitioned(engine, fact, build, stats: Optional[dict] = None):
    """run_plain's queries with the aggregate on the learned path's algorithm
    (the partitioned LDS aggregation) but no estimate: it starts from the
    capacity plain sizing reserves for its first device batch and doubles
    whenever a partition's table fills up (the rows partitioned again), so
    learned vs this compares sizing alone. The join as in run_plain."""
    from .abi import AGG_SUM, INT64, Agg, col, count_star, probe_op
    agg = engine.hashagg(fact, [], [col(0)], [INT64], [count_star(), Agg(AGG_SUM, col(1), INT64)],
                         sizing=abi.SIZING_LEARNED, capacity_hint=PLAIN_FIRST_RESERVE, stats=stats)
    jt = engine.join_build(build, [], [col(0)], [INT64])
    joined = engine.collect(fact, [probe_op(jt, [col(0)])])
    return agg, joined


def predict(engine, mlp: "Mlp", fact, build_rows: int, build_distinct: int, jt, seed: int = 0):
    x = features(engine, fact.tensors[0], fact.num_rows, jt, build_rows, build_distinct, seed=seed)
    return mlp.forward(x[None, :])[0]


def run_learned(engine, mlp: "Mlp", fact, build, build_distinct: int, margin: float = 1.15,
                stats: Optional[dict] = None):
    """The same two queries with both outputs pre-sized from the network's
    prediction (features + inference included); too-small predictions fall
    back to the exact path inside the library."""
    from .abi import AGG_SUM, INT64, Agg, col, count_star, probe_op
    jt = engine.join_build(build, [], [col(0)], [INT64])
    y = predict(engine, mlp, fact, build.num_rows, build_distinct, jt)
    groups = capacity_from_log2(y[0], margin)
    out_rows = max(1, capacity_from_log2(y[1], margin) - 1)
    agg = engine.hashagg(fact, [], [col(0)], [INT64], [count_star(), Agg(AGG_SUM, col(1), INT64)],
                         sizing=abi.SIZING_LEARNED, capacity_hint=groups, stats=stats)
    joined, fell_back = engine.collect_sized(fact, [probe_op(jt, [col(0)])], out_rows)
    return agg, joined, y, fell_back


def run_sized(engine, fact, build, jt, groups: int, join_rows: int, stats: Optional[dict] = None):
    """The evaluation queries with the group table and the join output sized
    from given estimates (plus margin), through the same path as run_learned:
    a too-small estimate falls back to the exact path inside the library."""
    from .abi import AGG_SUM, INT64, Agg, col, count_star, probe_op
    agg = engine.hashagg(fact, [], [col(0)], [INT64], [count_star(), Agg(AGG_SUM, col(1), INT64)],
                         sizing=abi.SIZING_LEARNED, capacity_hint=max(1, int(groups)), stats=stats)
    joined, fell_back = engine.collect_sized(fact, [probe_op(jt, [col(0)])], max(1, int(join_rows)))
    return agg, joined, fell_back


# Estimators compared on the same sizing path: (name, needs the features pass,
# f(features, fact rows, build rows) -> (log2 groups, log2 (1 + join rows))).
def _heuristic(x, rows, build_rows):
    """MiniLake's planner estimate (estimate_rows, physical_planner.rs:283-292):
    an aggregate keeps 1/10 of its input, a join yields its larger input."""
    return math.log2(max(1.0, rows / 10.0)), math.log2(1.0 + max(rows, build_rows))


def _sample_sketch(x, rows, build_rows):
    """The sample's own estimates, no model: GEE (sqrt(N/n) f1 + sum of f_j for
    j >= 2, feature 15) for distinct keys, the sampled match rate scaled to the
    table (feature 31) for the join."""
    return float(x[15]), float(x[31])


def _kmv_sample(x, rows, build_rows):
    """The raw KMV sketch of the sampled keys (k = 256, feature 23): distinct
    keys of the sample, not scaled to the table; the join as _sample_sketch."""
    return float(x[23]), float(x[31])


ESTIMATORS = [("reference_heuristic", False, _heuristic), ("sample_gee", True, _sample_sketch),
              ("kmv_sample", True, _kmv_sample)]


def ablation(engine, mlp: "Mlp", fact, build, build_distinct: int, reps: int = 3, margin: float = 1.15) -> dict:
    """Time each estimator's sizing of the same two queries on the same path
    (features and inference inside the timed region where they are used), the
    MLP and exact sizing (the true counts) among them; every leg's results
    must equal the plain run's."""
    import time
    from .abi import INT64, col
    torch = engine.torch
    ps = {}
    agg_p, join_p = run_plain(engine, fact, build, stats=ps)
    true_groups, true_join = agg_p.num_rows, join_p.num_rows
    legs = list(ESTIMATORS) + [("mlp", True, None), ("exact", False, None)]
    out = {"true_groups": true_groups, "true_join_rows": true_join, "legs": {}}
    for name, needs_features, f in legs:
        def one():
            jt = engine.join_build(build, [], [col(0)], [INT64])
            if name == "mlp":
                y = predict(engine, mlp, fact, build.num_rows, build_distinct, jt)
                g, j = float(y[0]), float(y[1])
            elif name == "exact":
                g, j = math.log2(max(1, true_groups)), math.log2(1 + true_join)
            else:
                x = features(engine, fact.tensors[0], fact.num_rows, jt, build.num_rows, build_distinct) \
                    if needs_features else None
                g, j = f(x, fact.num_rows, build.num_rows)
            groups = capacity_from_log2(g, margin)
            out_rows = max(1, capacity_from_log2(j, margin) - 1)
            st = {}
            agg, joined, fb = run_sized(engine, fact, build, jt, groups, out_rows, stats=st)
            return agg, joined, fb, groups, out_rows, st
        one()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            agg, joined, fb, groups, out_rows, st = one()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        out["legs"][name] = {"ms": sorted(ts)[len(ts) // 2], "groups_capacity": groups, "join_capacity": out_rows,
                             "groups_ratio": groups / max(1, true_groups), "join_ratio": out_rows / max(1, true_join),
                             "agg_resizes": st.get("resizes"), "join_fell_back": bool(fb),
                             "identical": same_rows(agg, agg_p) and same_rows(joined, join_p)}
    return out


def same_rows(a, b) -> bool:
    """Bit-identical results as multisets of rows (device fingerprints)."""
    if a.num_rows != b.num_rows or a.types != b.types:
        return False
    return bool(a.engine.torch.equal(a.fingerprint(), b.fingerprint()))


def compare_plain_learned(engine, mlp: "Mlp", fact, build, build_distinct: int, reps: int = 3, log=None) -> dict:
    """Time the plain and learned runs of the evaluation queries (after one
    warm-up each) and check that their results are bit-identical."""
    import time
    torch = engine.torch

    def timed(f):
        f()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = f()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return sorted(ts)[len(ts) // 2], out

    ps, pps, ls = {}, {}, {}
    plain_ms, (agg_p, join_p) = timed(lambda: run_plain(engine, fact, build, stats=ps))
    part_ms, (agg_pp, join_pp) = timed(lambda: run_plain_partitioned(engine, fact, build, stats=pps))
    learned_ms, (agg_l, join_l, y, fell_back) = timed(lambda: run_learned(engine, mlp, fact, build, build_distinct,
                                                                          stats=ls))
    out = {"plain_ms": plain_ms, "learned_ms": learned_ms, "speedup": plain_ms / learned_ms,
           # the same partitioned algorithm without an estimate: sizing alone
           "plain_partitioned_ms": part_ms, "speedup_vs_plain_partitioned": part_ms / learned_ms,
           "plain_partitioned_resizes": pps.get("resizes"),
           "groups": agg_p.num_rows, "predicted_groups": float(2.0 ** y[0]),
           "join_rows": join_p.num_rows, "predicted_join_rows": float(2.0 ** y[1] - 1),
           "plain_resizes": ps.get("resizes"), "learned_resizes": ls.get("resizes"),
           "join_fell_back": bool(fell_back),
           "identical": (same_rows(agg_p, agg_l) and same_rows(join_p, join_l) and same_rows(agg_p, agg_pp)
                         and same_rows(join_p, join_pp))}
    if log:
        log(f"learned sizing: plain {plain_ms:.2f} ms, plain partitioned {part_ms:.2f} ms "
            f"({out['plain_partitioned_resizes']} doublings), learned {learned_ms:.2f} ms "
            f"({out['speedup']:.2f}x / {out['speedup_vs_plain_partitioned']:.2f}x), identical={out['identical']}")
    return out
