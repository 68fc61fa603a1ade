"""GPU checks of the learned-sizing network.

* MLP forward and gradients from the library's HIP kernels against the float64
  numpy oracle (itself pinned by finite differences in test_mlp_oracle.py);
  tolerance: FP32 arithmetic, 2e-5 of the largest gradient.
* ml_sizing_features' device statistics against numpy on the same rows.
"""

import ctypes as C

import numpy as np
import pytest

from minilake_amd import abi, sizing
from minilake_amd.abi import AGG_COUNT_STAR, AGG_SUM, FLOAT64, INT64, Agg, col
from minilake_amd.table import HostColumn, HostTable
from oracle import mlp_ref
from tests.compare import rows_of

pytestmark = pytest.mark.gpu


def test_mlp_forward_and_gradients_match_oracle(gpu_engine):
    rng = np.random.default_rng(0)
    p = sizing.init_params(3)
    m = sizing.Mlp(gpu_engine, p)
    for n in (1, 7, 300):
        x = rng.normal(0, 1, (n, 32)).astype(np.float32)
        t = rng.normal(0, 1, (n, 2)).astype(np.float32)
        y_ref, _ = mlp_ref.forward(p.astype(np.float64), x)
        y = m.forward(x)
        assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
        loss, g = m.loss_grad(x, t)
        loss_ref, g_ref = mlp_ref.loss_grad(p.astype(np.float64), x, t)
        assert abs(loss - loss_ref) <= 1e-4 * max(1.0, loss_ref)
        assert np.max(np.abs(g - g_ref)) <= 2e-5 * np.max(np.abs(g_ref)), n
    # deterministic: the same batch gives bit-identical gradients
    _, g2 = m.loss_grad(x, t)
    assert np.array_equal(g, g2)


def test_mlp_training_reduces_loss(gpu_engine):
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (2048, 32)).astype(np.float32)
    w = rng.normal(0, 1, (32, 2))
    t = (np.tanh(x @ w / 4) * 3).astype(np.float32)
    m = sizing.Mlp(gpu_engine)
    before, _ = m.loss_grad(x, t)
    after = m.train(x, t, epochs=30, batch=128, lr=3e-3)
    final, _ = m.loss_grad(x, t)
    assert final < 0.2 * before
    # the trained parameters reproduce the oracle's forward
    y_ref, _ = mlp_ref.forward(m.params().astype(np.float64), x[:50])
    assert np.max(np.abs(m.forward(x[:50]) - y_ref)) <= 1e-3


def _splitmix(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def test_sizing_features_statistics(gpu_engine):
    rng = np.random.default_rng(2)
    n = 500_000
    keys = rng.integers(-(1 << 40), 1 << 40, n)
    keys[::5] = rng.integers(0, 1000, len(keys[::5]))
    dev = gpu_engine.torch.from_numpy(keys).to("cuda")
    f = sizing.column_features(gpu_engine, dev, n, seed=9)
    rg = sizing.ROW_GROUP_ROWS
    ng = (n + rg - 1) // rg
    assert f[0] == pytest.approx(np.log2(n), rel=1e-6)
    assert f[1] == pytest.approx(np.log2(ng), rel=1e-6)
    assert f[2] == pytest.approx(np.log2(float(keys.max()) - float(keys.min()) + 1), rel=1e-6)
    idx = [(_splitmix(9 + k) * n) >> 64 for k in range(sizing.SAMPLE)]
    sample = keys[idx]
    assert f[9] == pytest.approx(np.log2(len(np.unique(sample))), rel=1e-6)
    mins = [keys[g * rg:(g + 1) * rg].min() for g in range(ng)]
    maxs = [keys[g * rg:(g + 1) * rg].max() for g in range(ng)]
    ml = np.mean([np.log2(float(b) - float(a) + 1) for a, b in zip(mins, maxs)])
    assert f[3] == pytest.approx(ml, rel=1e-5)


def test_learned_sizing_bit_identical_and_faster(gpu_engine):
    """The trained estimator pre-sizes the group table and the join output:
    results are bit-identical to plain resize-on-overflow sizing, the group
    table never grows, and the run is faster."""
    from minilake_amd import sizing
    params = sizing.load_params()
    assert params is not None, "trained weights (minilake_amd/estimator_weights.npy) are missing"
    mlp = sizing.Mlp(gpu_engine, params)
    # 16M Zipf(1.1) keys over 1e8 values: ~2.2M groups, so plain sizing grows its
    # table (2^21-row device batches reserve len + batch) while learned does not
    fact, build = sizing.zipf_workload(gpu_engine, rows=16_000_000, domain=100_000_000, build_keys=1_000_000)
    r = sizing.compare_plain_learned(gpu_engine, mlp, fact, build, build_distinct=1_000_000, reps=3)
    print(r)
    assert r["identical"]
    assert r["plain_resizes"] >= 1
    assert r["learned_resizes"] == 0
    # the partitioned aggregation started from plain sizing's first reservation
    # (4M slots) holds ~2.2M groups without doubling
    assert r["plain_partitioned_resizes"] == 0
    assert r["learned_ms"] < r["plain_ms"]


def test_learned_sizing_full_size_matches_oracle(gpu_engine):
    """The learned-sizing evaluation at its configured size: 60M Zipf(1.1) keys
    over 1e8 (about 6.2M groups) aggregated and joined with 2M keys x 2 rows.
    Learned and plain sizing give bit-identical results; the aggregate (keys and
    integer states exact) and the join's row count equal the oracle's."""
    import os
    from minilake_amd.abi import probe_op
    from oracle.oracle import OracleEngine
    params = sizing.load_params()
    assert params is not None
    mlp = sizing.Mlp(gpu_engine, params)
    fact, build = sizing.zipf_workload(gpu_engine, rows=60_000_000)
    agg_p, join_p = sizing.run_plain(gpu_engine, fact, build)
    agg_l, join_l, _, _ = sizing.run_learned(gpu_engine, mlp, fact, build, build_distinct=2_000_000)
    assert sizing.same_rows(agg_p, agg_l) and sizing.same_rows(join_p, join_l)
    host_fact = HostTable([HostColumn(INT64, fact.tensors[0].cpu().numpy()),
                           HostColumn(INT64, fact.tensors[1].cpu().numpy())])
    host_build = HostTable([HostColumn(INT64, build.tensors[0].cpu().numpy()),
                            HostColumn(INT64, build.tensors[1].cpu().numpy())])
    oe = OracleEngine(threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    want = oe.hashagg(host_fact, [], [col(0)], [INT64], [Agg(AGG_COUNT_STAR, None, INT64),
                                                         Agg(AGG_SUM, col(1), INT64)]).to_host()
    got = agg_l.to_host()
    assert got.num_rows == want.num_rows > 5_000_000
    og, ow = np.argsort(got.columns[0].values), np.argsort(want.columns[0].values)
    for c in range(3):
        np.testing.assert_array_equal(got.columns[c].values[og], want.columns[c].values[ow])
    jo = oe.join_build(host_build, [], [col(0)], [INT64])
    assert join_l.num_rows == oe.count(host_fact, [probe_op(jo, [col(0)])])
