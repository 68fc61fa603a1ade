"""Batched fragments (ml_fragments_*) against the CPU oracle's join +
aggregate (join.rs:283-300 emits every chain match; aggregate/hash.rs).

* Small fragments with duplicate build keys, NULL build and probe keys,
  INT64_MIN and INT64_MAX keys, one-group and empty-probe fragments, in every
  run mode (one persistent launch sequence, a replayed hipGraph, launches per
  fragment) and under several tunings (probe width, LDS array / cache /
  global atomics).
* A build group id outside [0, num_groups) is skipped on the device and the
  run returns ML_ERR_PLAN naming the fragment; the other fragments' results
  stay correct.
* The full batched-fragment workload: 256 fragments, probe rows
  floor(2^U[10,24]), build rows = probe / U[4,40], groups floor(2^U[0,20]),
  Zipf s in {0, 0.8, 1.2}, after the autotuner, every mode against the oracle.
Sums within 1e-12 relative (FP64 atomics add in another order than the
oracle's sequential sums), counts exact.
"""

import numpy as np
import pytest

from minilake_amd import abi
from minilake_amd import fragments as FR
from minilake_amd import workloads as W
from minilake_amd.abi import AGG_COUNT_STAR, AGG_SUM, FLOAT64, INT64, Agg, col
from minilake_amd.table import HostColumn, HostTable
from tests.compare import rows_of

pytestmark = pytest.mark.gpu

ML_ERR_PLAN = 2
I64_MIN, I64_MAX = -(1 << 63), (1 << 63) - 1


def oracle_fragment(oracle_engine, bk, bg, pk, pv, G, bvalid=None, pvalid=None):
    build = HostTable([HostColumn(INT64, bk, None, bvalid), HostColumn(INT64, bg)])
    probe = HostTable([HostColumn(INT64, pk, None, pvalid), HostColumn(FLOAT64, pv)])
    jt = oracle_engine.join_build(build, [], [col(0)], [INT64])
    # layout after the probe: probe (key, value) ++ build (key, group)
    r = oracle_engine.hashagg(probe, [abi.probe_op(jt, [col(0)])], [col(3)], [INT64],
                              [Agg(AGG_SUM, col(1), FLOAT64), Agg(AGG_COUNT_STAR)]).to_host()
    sums, counts = np.zeros(G), np.zeros(G, dtype=np.int64)
    g = r.columns[0].values
    sums[g] = r.columns[1].values
    counts[g] = r.columns[2].values
    return sums, counts


def assert_fragment_equal(got, want, i):
    (gs, gc), (ws, wc) = got, want
    np.testing.assert_array_equal(gc, wc, err_msg=f"fragment {i} counts")
    np.testing.assert_allclose(gs, ws, rtol=1e-12, atol=1e-12 * max(1.0, float(np.abs(ws).max(initial=0))),
                               err_msg=f"fragment {i} sums")


def small_fragments(seed=7, n=18):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        prow = 0 if i == 5 else int(2 ** rng.uniform(4, 17))
        brow = max(1, prow // int(rng.integers(2, 41)))
        G = 1 if i == 3 else int(2 ** rng.uniform(0, 16))
        if i % 4 == 0:  # duplicate build keys: a small key domain
            bk = rng.integers(0, max(2, brow // 3), brow).astype(np.int64)
        else:
            bk = rng.permutation(1 << 22)[:brow].astype(np.int64) - (1 << 21)
        bk[:2] = [I64_MIN, I64_MAX][: min(2, brow)] if brow >= 1 else bk[:2]
        bg = rng.integers(0, G, brow).astype(np.int64)
        pk = bk[rng.integers(0, brow, prow)] if prow else np.zeros(0, np.int64)
        pk = pk.copy()
        pk[::5] = rng.integers(-(1 << 30), 1 << 30, len(pk[::5]))  # many probe keys miss
        pk[1::11] = I64_MIN
        pv = rng.random(prow)
        bvalid = (rng.random(brow) > 0.1) if i % 3 == 1 else None
        pvalid = (rng.random(prow) > 0.1) if i % 3 == 2 else None
        out.append((bk, bg, pk, pv, G, bvalid, pvalid))
    return out


TUNINGS = [None, (256, 1024, 2, 0, 1, 0), (512, 4096, 4, 2048, 8, 1), (1024, 2048, 1, 512, 2, 0)]


@pytest.mark.parametrize("tuning", range(len(TUNINGS)))
def test_fragments_duplicates_nulls_extremes_match_oracle(gpu_engine, oracle_engine, tuning):
    frags = small_fragments()
    want = [oracle_fragment(oracle_engine, *f) for f in frags]
    dev = [FR.device_fragment(gpu_engine, *f) for f in frags]
    t = TUNINGS[tuning]
    fs = FR.FragmentSet(gpu_engine, dev, abi.MlFragmentTuning(*t) if t else None)
    try:
        for mode in FR.MODES:
            for rep in range(2 if mode == "graph" else 1):  # a replay sees the reset of the first
                fs.run(mode)
                for i, (g, w) in enumerate(zip(fs.results(), want)):
                    assert_fragment_equal(g, w, i)
    finally:
        fs.close()


def test_fragments_bad_group_id_is_refused(gpu_engine, oracle_engine):
    frags = small_fragments(seed=8, n=6)
    bk, bg, pk, pv, G, _, pv_ = frags[4]
    bv = None  # every build key valid, so both bad rows are seen
    bg = bg.copy()
    bg[len(bg) // 2] = G  # one past the domain
    bg[0] = -1
    frags[4] = (bk, bg, pk, pv, G, bv, pv_)
    dev = [FR.device_fragment(gpu_engine, *f) for f in frags]
    fs = FR.FragmentSet(gpu_engine, dev)
    try:
        for mode in FR.MODES:
            with pytest.raises(Exception) as ei:
                fs.run(mode)
            assert getattr(ei.value, "code", ML_ERR_PLAN) == ML_ERR_PLAN
            assert "fragment 4" in str(ei.value)
            res = fs.results()
            for i in (0, 1, 2, 3, 5):  # the other fragments are unaffected
                assert_fragment_equal(res[i], oracle_fragment(oracle_engine, *frags[i]), i)
    finally:
        fs.close()


def test_batched_fragment_workload_matches_oracle(gpu_engine, oracle_engine):
    shapes = W.fragment_shapes(seed=42, n=256)
    dev = FR.generate(gpu_engine, shapes, seed=42)
    print(f"probe rows {sum(s.probe_rows for s in shapes):,}, build rows {sum(s.build_rows for s in shapes):,}")
    host = [(f.build_keys.cpu().numpy(), f.build_groups.cpu().numpy(), f.probe_keys.cpu().numpy(),
             f.probe_values.cpu().numpy(), f.groups) for f in dev]
    for (bk, bg, pk, pv, G), sh in zip(host[:3], shapes[:3]):  # the generator's contract
        assert len(np.unique(bk)) == len(bk) and bg.min() >= 0 and bg.max() < G
        assert np.isin(pk, bk).all() and (pv >= 0).all() and (pv < 1).all()
    from oracle.oracle import OracleEngine
    big = OracleEngine(threads=16)
    want = [oracle_fragment(big, *h) for h in host]
    fs = FR.FragmentSet(gpu_engine, dev)
    try:
        best = fs.autotune()
        print("tuning", {k: getattr(best, k) for k, _ in best._fields_})
        for mode in ("persistent", "graph", "graph", "per_fragment"):
            fs.run(mode)
            for i, (g, w) in enumerate(zip(fs.results(), want)):
                assert_fragment_equal(g, w, i)
    finally:
        fs.close()
