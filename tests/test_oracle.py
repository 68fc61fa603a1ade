"""Pins the CPU oracle before it is trusted as the GPU's checker.

Each check is independent of the oracle's own code: closed-form golden vectors
(big-integer arithmetic, tests/golden/), the reference's in-source known answers
(file:line cited), brute force in plain Python/numpy on small inputs, and
invariants (equivalence with a dict, thread-count independence).
"""

import json
import math
import os
import struct

import numpy as np
import pytest

from minilake_amd import abi, queries as Q, workloads as W
from minilake_amd.abi import (AGG_AVG, AGG_COUNT_STAR, AGG_MIN, AGG_SUM, DATE, FLOAT64, INT32, INT64, UTF8, Agg, col,
                              date_lit, lit)
from minilake_amd.table import HostColumn, HostTable
from oracle import oracle as O
from tests.compare import assert_tables_equal, concat_tables, rows_of

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hash_vectors.json")))
M64 = (1 << 64) - 1


# --------------------------------------------------------------------- hash
def test_hash_constants_and_golden_vectors():
    for a, b, want in GOLD["fold_mul"]:
        assert O.fold_mul(a, b) == want
    for x, want in GOLD["hash_u64"]:
        assert O.hash_u64(x) == want
    for h, v, want in GOLD["combine"]:
        assert O.combine(h, v) == want
    for s, want, n in GOLD["hash_bytes"]:
        assert O.hash_bytes(s.encode()) == want


def test_hash_distinct_inputs_distinct_hashes():
    # hash.rs:60-68
    hs = {O.hash_u64(i) for i in range(10_000)}
    assert len(hs) == 10_000
    assert O.hash_bytes(b"abc") != O.hash_bytes(b"abd")
    assert O.hash_bytes(b"") != O.hash_bytes(b"\0")


def _col_table(vals, t, d=None):
    return HostTable([HostColumn(t, np.array(vals), d)])


@pytest.mark.parametrize("kind,t", [("column_i32", INT32), ("column_i64", INT64), ("column_f64", FLOAT64),
                                    ("column_bool", abi.BOOLEAN)])
def test_hash_columns_golden(oracle_engine, kind, t):
    vals = [v for v, _ in GOLD[kind]]
    want = np.array([h for _, h in GOLD[kind]], dtype=np.uint64)
    got = oracle_engine.hash_columns([col(0)], _col_table(vals, t))
    np.testing.assert_array_equal(got, want)


def test_hash_columns_strings_and_nulls(oracle_engine):
    strs = [s for s, _ in GOLD["column_str"]]
    t = HostTable([HostColumn(UTF8, np.arange(len(strs), dtype=np.uint32), strs)])
    np.testing.assert_array_equal(oracle_engine.hash_columns([col(0)], t),
                                  np.array([h for _, h in GOLD["column_str"]], dtype=np.uint64))
    # a NULL key (x / 0) hashes as NULL_HASH (kernels/hash.rs:40-44)
    t2 = HostTable([HostColumn(FLOAT64, np.array([1.0, 2.0]))])
    got = oracle_engine.hash_columns([col(0) / lit(0.0)], t2)
    assert got.tolist() == [GOLD["null_single"]] * 2


def test_hash_columns_three_keys(oracle_engine):
    a = np.array([r[0] for r in GOLD["three_keys"]], dtype=np.int64)
    b = np.array([r[1] for r in GOLD["three_keys"]], dtype=np.int32)
    c = np.array([r[2] for r in GOLD["three_keys"]], dtype=np.int32)
    t = HostTable([HostColumn(INT64, a), HostColumn(DATE, b), HostColumn(INT32, c)])
    np.testing.assert_array_equal(oracle_engine.hash_columns([col(0), col(1), col(2)], t),
                                  np.array([r[3] for r in GOLD["three_keys"]], dtype=np.uint64))


# ---------------------------------------------------------------- SwissTable
def test_swar_match_known_answer():
    # swiss.rs:308-316
    group = int.from_bytes(bytes([5, 0x80, 7, 5, 0, 1, 2, 3]), "little")
    m = O.match_tag(group, 5)
    hits = [i for i in range(8) if m & (0x80 << (i * 8))]
    assert 0 in hits and 3 in hits
    assert O.match_empty(group) == 0x80 << 8


def test_swar_match_brute_force():
    # every true match is reported; a false positive only sits above a true match
    rng = np.random.default_rng(0)
    for _ in range(20000):
        bs = rng.integers(0, 128, 8)
        bs[rng.integers(0, 8, 2)] = rng.integers(0, 128)
        tag = int(rng.integers(0, 128))
        m = O.match_tag(int.from_bytes(bytes(bs.tolist()), "little"), tag)
        true = {i for i in range(8) if bs[i] == tag}
        rep = {i for i in range(8) if m & (0x80 << (8 * i))}
        assert true <= rep
        for i in rep - true:
            assert any(j < i for j in true)


def _first_seen(keys):
    ref, out = {}, []
    for k in keys:
        out.append(ref.setdefault(int(k), len(ref)))
    return out, len(ref)


@pytest.mark.parametrize("weak", [False, True])
@pytest.mark.parametrize("seed", range(5))
def test_swiss_matches_dict_first_seen(weak, seed):
    # crates/minilake-hashtable/tests/proptest_tables.rs:14-40,63-79
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 3000))
    keys = rng.integers(0, 200 if weak else 500, n).astype(np.uint64)
    ids, distinct, cap, _ = O.swiss_group_ids(keys, batch=int(rng.integers(1, 300)), weak=weak)
    want, wn = _first_seen(keys)
    assert ids.tolist() == want and distinct == wn
    assert distinct * 8 <= cap * 7  # load factor <= 7/8


def test_swiss_grow_keeps_entries():
    # swiss.rs:318-335 (10,000 keys from capacity 0: grows by doubling)
    keys = np.arange(10_000, dtype=np.uint64)
    ids, n, cap, resizes = O.swiss_group_ids(keys, batch=1)
    assert n == 10_000 and ids.tolist() == list(range(10_000))
    # reserve(1) per insert: grow to with_capacity(max(len+1, 2*len)) whenever
    # len+1 passes 7/8 (swiss.rs:86-96,219-232): 16 -> 64 -> 256 -> ... -> 16384
    c, r = 16, 0
    for length in range(10_000):
        if (length + 1) * 8 > c * 7:
            c = 1 << (max(length + 1, 2 * length) * 8 // 7).bit_length()
            r += 1
    assert (cap, resizes) == (c, r) == (16384, 5)


def test_swiss_full_collisions():
    # swiss.rs:338-350 with the collision hash x % 4
    keys = (np.arange(100, dtype=np.uint64) * 4)
    ids, n, _, _ = O.swiss_group_ids(keys, batch=7, weak=True)
    assert n == 100 and ids.tolist() == list(range(100))


# ----------------------------------------------------------------- kernels
def test_compare_scalar_and_select(oracle_engine):
    # crates/minilake-exec/src/kernels/cmp.rs:124-133
    t = _col_table([5, 1, 7, 3], INT64)
    assert oracle_engine.filter(col(0).lt(lit(4)), t).tolist() == [1, 3]
    assert oracle_engine.filter(col(0).lt(lit(4)), t, np.array([0, 3], np.uint32)).tolist() == [3]


def test_empty_and_all(oracle_engine):
    # cmp.rs:135-140
    e = HostTable([HostColumn(INT64, np.zeros(0, np.int64))])
    assert oracle_engine.filter(col(0).lt(lit(4)), e).tolist() == []
    t = _col_table([1, 2, 3], INT64)
    assert oracle_engine.filter(col(0).gt(lit(0)), t).tolist() == [0, 1, 2]
    assert oracle_engine.filter(col(0).gt(lit(9)), t).tolist() == []


def _batch():
    return HostTable([HostColumn(INT64, np.array([1, 5, 10, 20])), HostColumn(FLOAT64, np.array([0.5, 1.5, 2.5, 3.5]))])


def test_filter_and_arith(oracle_engine):
    # crates/minilake-exec/src/expr/eval.rs:548-572
    b = _batch()
    pred = col(0).ge(lit(5)) & col(1).lt(lit(3.0))
    assert oracle_engine.filter(pred, b).tolist() == [1, 2]
    c = oracle_engine.project([col(0) * col(1)], b)
    assert c.columns[0].type == FLOAT64 and c.columns[0].values[3] == 70.0


def test_divide_by_zero_is_null(oracle_engine):
    # eval.rs:574-584
    c = oracle_engine.project([col(1) / lit(0)], _batch())
    assert c.columns[0].validity is not None and not c.columns[0].validity.any()


def test_arith_ops(oracle_engine):
    # crates/minilake-exec/src/kernels/arith.rs:136-153
    t = HostTable([HostColumn(INT64, np.array([1, 2, 3])), HostColumn(INT64, np.array([10, 20, 30]))])
    r = oracle_engine.project([col(0) + col(1), lit(100) - col(1), col(1) % lit(0)], t)
    assert r.columns[0].values.tolist() == [11, 22, 33]
    assert r.columns[1].values.tolist() == [90, 80, 70]
    assert r.columns[2].values.tolist() == [0, 0, 0] and not r.columns[2].validity.any()
    m = oracle_engine.project([col(0) * (lit(1.0) - col(1))],
                              HostTable([HostColumn(FLOAT64, np.array([10.0])), HostColumn(FLOAT64, np.array([0.1]))]))
    assert m.columns[0].values.tolist() == [9.0]


def test_kleene_logic(oracle_engine):
    # crates/minilake-exec/src/kernels/boolean.rs:1-13 truth table; NULL from 1/0
    t = HostTable([HostColumn(INT64, np.array([0, 1, 2] * 3)), HostColumn(INT64, np.array([0] * 3 + [1] * 3 + [2] * 3))])
    def tri(c):  # 0 -> FALSE, 1 -> TRUE, 2 -> NULL
        return (lit(1.0) / (lit(2) - c)).gt(lit(0.75))  # c=0: 0.5>0.75 F; c=1: 1>0.75 T; c=2: NULL
    r = oracle_engine.project([tri(col(0)) & tri(col(1)), tri(col(0)) | tri(col(1))], t)
    val = lambda c, i: None if (c.validity is not None and not c.validity[i]) else bool(c.values[i])
    F, T, N = False, True, None
    AND = {(F, F): F, (F, T): F, (F, N): F, (T, T): T, (T, N): N, (N, N): N}
    OR = {(F, F): F, (F, T): T, (F, N): N, (T, T): T, (T, N): T, (N, N): N}
    enc = {0: F, 1: T, 2: N}
    for i in range(9):
        a, b = enc[i % 3], enc[i // 3]
        key = (a, b) if (a, b) in AND else (b, a)
        assert val(r.columns[0], i) == AND[key]
        assert val(r.columns[1], i) == OR[key]


def test_wrapping_and_saturating_casts(oracle_engine):
    t = HostTable([HostColumn(INT32, np.array([2147483647, -2147483648], np.int32)),
                   HostColumn(FLOAT64, np.array([1e300, float("nan")]))])
    r = oracle_engine.project([col(0) + lit(1, INT32), col(1).cast(INT64), col(1).cast(INT32)], t)
    assert r.columns[0].values.tolist() == [-2147483648, -2147483647]  # wrapping_add
    assert r.columns[1].values.tolist() == [2**63 - 1, 0]  # Rust `as`: saturate, NaN -> 0
    assert r.columns[2].values.tolist() == [2**31 - 1, 0]


# ------------------------------------------------------------- aggregation
def test_sum_avg_min_with_groups(oracle_engine):
    # accumulator.rs:600-623: SUM per group {1.0, 6.0}; merged AVG group 1 = 3.0
    t = HostTable([HostColumn(FLOAT64, np.array([1.0, 2.0, 3.0, 4.0])), HostColumn(INT64, np.array([0, 1, 0, 1]))])
    # row 2 made NULL with x / (x - 3)*0 + x : use the NULL from division by zero
    v = col(0) / (lit(1.0) - (col(0).eq(lit(3.0))).cast(FLOAT64))
    r = oracle_engine.hashagg(t, [], [col(1)], [INT64], [Agg(AGG_SUM, v, FLOAT64), Agg(AGG_AVG, v, FLOAT64)]).to_host()
    got = {k: (s, a) for k, s, a in rows_of(r)}
    assert got[0][0] == 1.0 and got[1][0] == 6.0 and got[1][1] == 3.0


@pytest.mark.parametrize("seed", range(6))
def test_hash_aggregate_matches_reference(seed):
    # crates/minilake-exec/tests/aggregate_proptest.rs:45-82: BTreeMap reference,
    # with several worker threads so the partial-state merge path runs
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 5000))
    k = rng.integers(0, 50, n)
    v = rng.integers(-1000, 1000, n)
    t = HostTable([HostColumn(INT64, k), HostColumn(INT64, v)])
    eng = O.OracleEngine(threads=int(rng.integers(1, 5)), batch_size=int(rng.integers(1, 200)), morsel_rows=int(rng.integers(1, 700)))
    r = eng.hashagg(t, [], [col(0)], [INT64], [Agg(AGG_SUM, col(1), INT64), Agg(AGG_COUNT_STAR), Agg(AGG_MIN, col(1), INT64)]).to_host()
    ref = {}
    for kk, vv in zip(k.tolist(), v.tolist()):
        s, c, m = ref.get(kk, (0, 0, 10**9))
        ref[kk] = (s + vv, c + 1, min(m, vv))
    got = {row[0]: tuple(row[1:]) for row in rows_of(r)}
    assert got == ref


def test_thread_count_does_not_change_integer_results():
    # every morsel is consumed exactly once (scheduler.rs:152-172)
    li = W.filter_groupby(seed=3, n=50_000)
    a = Q.run_q1(O.OracleEngine(threads=1), li).to_host()
    b = Q.run_q1(O.OracleEngine(threads=8, morsel_rows=1000), li).to_host()
    assert_tables_equal(a, b, rtol=1e-12)


# ---------------------------------------------------- end-to-end closed forms
# crates/minilake-sql/tests/end_to_end.rs:19-68: 10,000 lines over 1,000 orders.
N_LINE, N_ORD = 10_000, 1_000
DAY0 = 9131  # 1995-01-01


def e2e_tables():
    i = np.arange(N_LINE)
    qty_valid = (i % 13) != 0
    lines = HostTable([
        HostColumn(INT64, (i % N_ORD).astype(np.int64)),
        HostColumn(FLOAT64, (i % 50).astype(np.float64), validity=qty_valid),
        HostColumn(FLOAT64, (i % 97) + 0.5),
        HostColumn(UTF8, (i % 3).astype(np.uint32), ["A", "N", "R"]),
        HostColumn(DATE, (DAY0 + i % 30).astype(np.int32)),
    ])
    j = np.arange(N_ORD)
    orders = HostTable([
        HostColumn(INT64, j.astype(np.int64)),
        HostColumn(DATE, (DAY0 + j // 100).astype(np.int32)),
        HostColumn(UTF8, (j % 3).astype(np.uint32), ["1-URGENT", "2-HIGH", "3-MEDIUM"]),
    ])
    return lines, orders


def test_group_by_order_by_matches_reference():
    # end_to_end.rs:233-267
    lines, _ = e2e_tables()
    cutoff = DAY0 + 9
    ref = {}
    for k in range(N_LINE):
        ship = DAY0 + k % 30
        if ship >= cutoff:
            continue
        f = "ANR"[k % 3]
        e = ref.setdefault(f, [0.0, 0, 0.0, 0, 10**9])
        e[0] += (k % 97) + 0.5
        e[1] += 1
        if k % 13:
            e[2] += k % 50
            e[3] += 1
        e[4] = min(e[4], ship)
    for threads in (1, 4):
        eng = O.OracleEngine(threads=threads, morsel_rows=2500)
        r = eng.hashagg(lines, [abi.filter_op(col(4).lt(date_lit(cutoff)))], [col(3)], [UTF8],
                        [Agg(AGG_SUM, col(2), FLOAT64), Agg(AGG_COUNT_STAR), Agg(AGG_AVG, col(1), FLOAT64),
                         Agg(AGG_MIN, col(4), DATE)]).to_host()
        got = {row[0]: row[1:] for row in rows_of(r)}
        assert sorted(got) == sorted(ref)
        for f, (s, c, qs, qc, md) in ref.items():
            gs, gc, ga, gm = got[f]
            assert abs(gs - s) <= 1e-9 * max(abs(s), 1) and gc == c and gm == md
            assert abs(ga - qs / qc) <= 1e-9 * max(abs(qs / qc), 1)


def test_join_top_n_matches_reference():
    # end_to_end.rs:269-298: revenue per order priority, top 2
    lines, orders = e2e_tables()
    cutoff = DAY0 + 4
    rev = {}
    for k in range(N_LINE):
        o = k % N_ORD
        if DAY0 + o // 100 >= cutoff:
            p = ["1-URGENT", "2-HIGH", "3-MEDIUM"][o % 3]
            rev[p] = rev.get(p, 0.0) + (k % 97) + 0.5
    expected = sorted(rev.items(), key=lambda kv: -kv[1])[:2]
    for threads in (1, 3):
        eng = O.OracleEngine(threads=threads, morsel_rows=2500)
        jt = eng.join_build(orders, [abi.filter_op(col(1).ge(date_lit(cutoff)))], [col(0)], [INT64])
        agg = eng.hashagg(lines, [abi.probe_op(jt, [col(0)], build_is_left=True)], [col(2)], [UTF8],
                          [Agg(AGG_SUM, col(5), FLOAT64)])
        top = rows_of(eng.topn(agg, [], [(col(1), False, False)], 2).to_host())
        assert [r[0] for r in top] == [p for p, _ in expected]
        for r, (_, v) in zip(top, expected):
            assert abs(r[1] - v) <= 1e-9 * abs(v)


# ------------------------------------------------ query shapes vs brute force
def test_filter_groupby_query_brute_force():
    li = W.filter_groupby(seed=5, n=60_000)
    r = Q.run_q1(O.OracleEngine(threads=4, morsel_rows=5000), li).to_host()
    c = [x.values for x in li.columns]
    m = c[6] <= W.FILTER_GROUPBY_CUTOFF
    got = {(row[0], row[1]): row[2:] for row in rows_of(r)}
    assert len(got) == 6
    for a, fa in enumerate("ANR"):
        for b, fb in enumerate("FO"):
            s = m & (c[4] == a) & (c[5] == b)
            qty, p, d, t = c[0][s], c[1][s], c[2][s], c[3][s]
            dp = p * (1.0 - d)
            ch = dp * (1.0 + t)
            want = [int(qty.sum()), math.fsum(p), math.fsum(dp), math.fsum(ch), qty.mean(), math.fsum(p) / len(p),
                    math.fsum(d) / len(d), int(s.sum())]
            g = got[(fa, fb)]
            assert g[0] == want[0] and g[7] == want[7]
            for x, y in zip(g[1:7], want[1:7]):
                assert abs(x - y) <= 1e-12 * abs(y)


def test_join_topn_query_brute_force():
    d = W.join_topn(seed=2, scale=0.02)
    top, (_, _, agg) = Q.run_q3(O.OracleEngine(threads=4, morsel_rows=5000), d.customer, d.orders, d.lineitem)
    ck, cs = d.customer.columns[0].values, d.customer.columns[1].values
    ok, oc, od, op = [x.values for x in d.orders.columns]
    lk, lp, ld, ls = [x.values for x in d.lineitem.columns]
    building = set(ck[cs == 1].tolist())
    orders = {int(k): (int(dd), int(pp)) for k, c_, dd, pp in zip(ok, oc, od, op) if c_ in building and dd < W.D_1995_03_15}
    ref = {}
    for k, p, dd, s in zip(lk.tolist(), lp.tolist(), ld.tolist(), ls.tolist()):
        if s > W.D_1995_03_15 and k in orders:
            key = (k, *orders[k])
            ref.setdefault(key, []).append(p * (1.0 - dd))
    got = {tuple(r[:3]): r[3] for r in rows_of(agg.to_host())}
    assert set(got) == set(ref)
    for kk, v in ref.items():
        assert abs(got[kk] - math.fsum(v)) <= 1e-12 * math.fsum(v)
    best = sorted(ref.items(), key=lambda kv: (-math.fsum(kv[1]), kv[0][1]))[:10]
    tr = rows_of(top.to_host())
    assert [(r[1], round(r[3], 6)) for r in tr] == [(k[1], round(math.fsum(v), 6)) for k, v in best]


def test_star_join_query_brute_force():
    sj = W.star_join(seed=4, scale=0.02)
    top, _ = Q.run_q5(O.OracleEngine(threads=4, morsel_rows=5000), sj)
    cn = sj.customer.columns[1].values
    ok, oc, od = [x.values for x in sj.orders.columns]
    sn = sj.supplier.columns[1].values
    lo, lsup, lp, ld = [x.values for x in sj.lineitem.columns]
    in_year = (od >= W.D_1994_01_01) & (od < W.D_1995_01_01)
    onat = np.full(len(ok) + 1, -1)
    onat[ok[in_year]] = cn[oc[in_year] - 1]
    ln = onat[lo]
    sup_n = sn[lsup - 1]
    asia = np.array([W.NATION_REGION[x] == 2 for x in range(25)])
    keep = (ln >= 0) & (ln == sup_n) & asia[np.maximum(ln, 0)]
    ref = {}
    for n, p, dd in zip(ln[keep].tolist(), lp[keep].tolist(), ld[keep].tolist()):
        ref.setdefault(W.NATIONS[n], []).append(p * (1.0 - dd))
    want = sorted(((k, math.fsum(v)) for k, v in ref.items()), key=lambda kv: -kv[1])
    got = rows_of(top.to_host())
    assert [r[0] for r in got] == [k for k, _ in want]
    for r, (_, v) in zip(got, want):
        assert abs(r[1] - v) <= 1e-12 * v


def test_workload_generators_are_deterministic():
    a = W.uniform_int(42, 7, 1000, 0, 9)
    b = W.uniform_int(42, 7, 1000, 0, 9)
    assert (a == b).all() and a.min() == 0 and a.max() == 9
    assert (W.uniform_int(42, 7, 10, 0, 9, offset=5) == a[5:15]).all()
    p = W.permutation(1, 2, 1000)
    assert sorted(p.tolist()) == list(range(1000))
    z = W.zipf(1, 3, 100_000, 1.1, 1000)
    assert z.min() >= 1 and z.max() <= 1000
    counts = np.bincount(z)[1:4]
    # Zipf(1.1): P(1)/P(2) = 2^1.1 ~ 2.14
    assert 1.9 < counts[0] / counts[1] < 2.4


# ------------------------------------------------ UngroupedAggregate, LIMIT
def test_ungrouped_aggregate_closed_forms():
    """UngroupedAggregateSink (ungrouped.rs:36-139): one row always; over no
    rows COUNT = 0 and SUM/AVG/MIN NULL (ungrouped.rs:82-91); over 0..99 the
    closed forms (sum 4950, mean of 0.5*i = 24.75), with several threads and
    tiny morsels so the per-thread merge runs."""
    t = HostTable([HostColumn(INT64, np.arange(100, dtype=np.int64)), HostColumn(FLOAT64, np.arange(100) * 0.5)])
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(0), INT64), Agg(AGG_AVG, col(1), FLOAT64),
            Agg(AGG_MIN, col(1), FLOAT64)]
    for threads, morsel in ((1, 1 << 20), (4, 7)):
        eng = O.OracleEngine(threads=threads, morsel_rows=morsel)
        empty = rows_of(eng.hashagg(t, [abi.filter_op(col(0).lt(lit(0)))], [], [], aggs).to_host())
        assert empty == [(0, None, None, None)]
        full = rows_of(eng.hashagg(t, [], [], [], aggs).to_host())
        assert full == [(100, 4950, 24.75, 0.0)]


def test_limit_offset_slices_source_order():
    """LimitSink (limit.rs:26-117) with one thread: rows [offset, offset + limit)
    of the filtered input in order; offsets past the end give no rows."""
    t = HostTable([HostColumn(INT64, np.arange(100, dtype=np.int64))])
    eng = O.OracleEngine(threads=1)
    ops = [abi.filter_op(col(0).gt(lit(10)))]
    assert eng.limit(t, ops, 5, 3).to_host().columns[0].values.tolist() == [14, 15, 16, 17, 18]
    assert eng.limit(t, ops, 1000, 80).to_host().columns[0].values.tolist() == list(range(91, 100))
    assert eng.limit(t, ops, 5, 500).to_host().num_rows == 0
    assert eng.limit(t, ops, 0, 0).to_host().num_rows == 0


def _modes_table(n=3000, seed=21):
    rng = np.random.default_rng(seed)
    return HostTable([HostColumn(INT32, rng.integers(-3, 4, n).astype(np.int32)),
                      HostColumn(FLOAT64, rng.uniform(1, 100, n)),
                      HostColumn(INT64, rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)),
                      HostColumn(DATE, rng.integers(8000, 9000, n).astype(np.int32))])


_MODE_AGGS = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(2), INT64), Agg(AGG_SUM, col(1) / col(0), FLOAT64),
              Agg(AGG_AVG, col(1) / col(0), FLOAT64), Agg(AGG_MIN, col(3), DATE), Agg(abi.AGG_MAX, col(2), INT64),
              Agg(abi.AGG_COUNT, col(1) / col(0), FLOAT64), Agg(abi.AGG_MAX, col(1) / col(0), FLOAT64)]


@pytest.mark.parametrize("grouped", [True, False])
def test_partial_then_final_equals_single(grouped):
    """AggMode (aggregate/mod.rs:50-57; split.rs:34-125): Partial states of
    three slices, concatenated and merged by a Final aggregate, equal the
    Single result (SUM/COUNT/MIN/MAX/AVG are decomposable). Group 0 has only
    NULL arguments (x / 0), so its SUM/AVG/MIN states are NULL or (0.0, 0)."""
    t = _modes_table()
    eng = O.OracleEngine(threads=3, morsel_rows=257)
    keys, kt = ([col(0)], [INT32]) if grouped else ([], [])
    single = eng.hashagg(t, [], keys, kt, _MODE_AGGS).to_host()
    cuts = [0, 1000, 1001, t.num_rows]
    parts = [eng.hashagg(t.take(np.arange(a, b)), [], keys, kt, _MODE_AGGS, mode=abi.AGG_MODE_PARTIAL).to_host()
             for a, b in zip(cuts, cuts[1:])]
    states = concat_tables(parts)
    # group keys, then 1+1+1+2+1+1+1+1 state columns
    assert len(states.columns) == len(kt) + 9
    final = eng.hashagg(states, [], keys, kt, _MODE_AGGS, mode=abi.AGG_MODE_FINAL).to_host()
    assert_tables_equal(final, single, rtol=1e-12)


def test_partial_states_closed_forms():
    """Accumulator::state (accumulator.rs:382-404): COUNT a count; SUM and MIN
    the value with NULL when no argument was valid; AVG (sum, count) with no
    NULLs; integer MIN/MAX states are Int64. Ungrouped over no rows: the fresh
    states (COUNT 0, AVG (0.0, 0), the others NULL)."""
    t = HostTable([HostColumn(INT32, np.array([1, 1, 2, 2], np.int32)),
                   HostColumn(FLOAT64, np.array([1.0, 3.0, 5.0, 7.0])),
                   HostColumn(INT32, np.array([1, 1, 0, 0], np.int32))])
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(1) / col(2), FLOAT64), Agg(AGG_AVG, col(1) / col(2), FLOAT64),
            Agg(AGG_MIN, col(0), INT32)]
    eng = O.OracleEngine(threads=1)
    got = eng.hashagg(t, [], [col(0)], [INT32], aggs, mode=abi.AGG_MODE_PARTIAL).to_host()
    assert [c.type for c in got.columns] == [INT32, INT64, FLOAT64, FLOAT64, INT64, INT64]
    assert sorted(rows_of(got)) == [(1, 2, 4.0, 4.0, 2, 1), (2, 2, None, 0.0, 0, 2)]
    empty = eng.hashagg(t, [abi.filter_op(col(0).gt(lit(5)))], [], [], aggs, mode=abi.AGG_MODE_PARTIAL).to_host()
    assert rows_of(empty) == [(0, None, 0.0, 0, None)]
    fin = eng.hashagg(empty, [], [], [], aggs, mode=abi.AGG_MODE_FINAL).to_host()
    assert rows_of(fin) == [(0, None, None, None)]
    assert fin.columns[3].type == INT32


def test_mlb1_known_bytes_and_round_trip():
    """MLB1 (ipc.rs:1-13): the bytes of a small batch written out by hand from
    the format (magic, u32 columns, u64 rows; per column type tag, validity
    flag, LSB-first words, payload; strings as u32 offsets, u64 length and the
    bytes), with the selection applied first and a validity section only where
    a kept row is NULL. Reading them back gives the rows; a stream shorter than
    the magic is a clean end; a cut batch is an I/O error; two batches in a row
    are read one after the other."""
    B = abi.BOOLEAN
    t = HostTable([HostColumn(INT32, np.array([5, -1, 7, 9]), validity=np.array([True, False, True, True])),
                   HostColumn(UTF8, np.array([1, 0, 1, 2], np.uint32), ["a", "", "hello"]),
                   HostColumn(FLOAT64, np.array([0.5, 1.5, -2.0, 3.0])),
                   HostColumn(B, np.array([1, 0, 0, 1], np.uint8)),
                   HostColumn(DATE, np.array([10, 20, 30, 40])), HostColumn(INT64, np.array([1, 2, 3, -4]))])
    eng = O.OracleEngine(threads=1)
    sel = np.array([0, 1, 3], np.uint32)
    want = b"MLB1" + struct.pack("<IQ", 6, 3)
    want += bytes([1, 1]) + struct.pack("<Q", 0b101) + struct.pack("<3i", 5, -1, 9)
    want += bytes([5, 0]) + struct.pack("<4I", 0, 0, 1, 6) + struct.pack("<Q", 6) + b"ahello"
    want += bytes([3, 0]) + struct.pack("<3d", 0.5, 1.5, 3.0)
    want += bytes([0, 0]) + bytes([1, 0, 1])
    want += bytes([4, 0]) + struct.pack("<3i", 10, 20, 40)
    want += bytes([2, 0]) + struct.pack("<3q", 1, 2, -4)
    got = eng.mlb1_write(t, sel)
    assert got == want
    back, used = eng.mlb1_read(got)
    assert used == len(got)
    assert rows_of(back.to_host()) == rows_of(t.take(sel))
    # the NULL row not kept: no validity section
    assert eng.mlb1_write(t, np.array([0, 2], np.uint32))[16:18] == bytes([1, 0])
    assert eng.mlb1_read(b"ML") == (None, 0)
    with pytest.raises(O.OracleError) as e:
        eng.mlb1_read(got[:-1])
    assert e.value.code == abi.ML_ERR_IO
    two = got + eng.mlb1_write(t)
    first, used = eng.mlb1_read(two)
    second, used2 = eng.mlb1_read(two[used:])
    assert first.to_host().num_rows == 3 and rows_of(second.to_host()) == rows_of(t) and used + used2 == len(two)


def _nan_key_table(n=3000, seed=31):
    rng = np.random.default_rng(seed)
    pool = np.array([1.5, -0.0, 0.0, np.nan, 2.5, -7.25, np.inf, -np.inf])
    k = pool[rng.integers(0, len(pool), n)]
    v = rng.integers(-100, 100, n)
    g = rng.integers(0, 3, n)
    return HostTable([HostColumn(FLOAT64, k), HostColumn(INT64, v), HostColumn(INT64, g)]), k, v, g


def test_group_keys_nan_never_equal_negative_zero_equal(oracle_engine):
    # values_equal (kernels/hash.rs:84-101): f64 keys compare with ==, so every
    # NaN row is a group of its own and -0.0 joins +0.0's group
    t, k, v, g = _nan_key_table()
    r = oracle_engine.hashagg(t, [], [col(0), col(2)], [FLOAT64, INT64],
                              [Agg(AGG_SUM, col(1), INT64), Agg(AGG_COUNT_STAR)]).to_host()
    rows = rows_of(r)
    nan_rows = sorted((gg, s, c) for kk, gg, s, c in rows if math.isnan(kk))
    assert nan_rows == sorted((int(gg), int(vv), 1) for kk, vv, gg in zip(k, v, g) if math.isnan(kk))
    ref = {}
    for kk, vv, gg in zip(k.tolist(), v.tolist(), g.tolist()):
        if math.isnan(kk):
            continue
        key = (0.0 if kk == 0 else kk, gg)
        s, c = ref.get(key, (0, 0))
        ref[key] = (s + vv, c + 1)
    got = {(0.0 if kk == 0 else kk, gg): (s, c) for kk, gg, s, c in rows if not math.isnan(kk)}
    assert got == ref


def test_join_keys_nan_never_match(oracle_engine):
    # join.rs:283-300 finds candidates with values_equal: a NaN key matches
    # nothing (not even NaN); -0.0 matches +0.0
    bk = np.array([np.nan, -0.0, 1.0, np.nan, 2.0, 0.0])
    pk = np.array([np.nan, 0.0, -0.0, 1.0, 3.0, np.nan, 2.0])
    build = HostTable([HostColumn(FLOAT64, bk)])
    probe = HostTable([HostColumn(FLOAT64, pk)])
    jt = oracle_engine.join_build(build, [], [col(0)], [FLOAT64])
    p, b = oracle_engine.join_probe(jt, [col(0)], probe)
    pairs = sorted(zip(p.tolist(), b.tolist()))
    want = sorted((i, j) for i, x in enumerate(pk.tolist()) for j, y in enumerate(bk.tolist()) if x == y)
    assert pairs == want and len(want) == 6


# ------------------------------------------- compare_rows (sort.rs:43-82)
SORT_DICT = ["pear", "apple", "fig", "", "Zebra", "apple2", "date"]  # not in byte order


def _sort_table(n=400, seed=61):
    rng = np.random.default_rng(seed)
    pool = np.array([np.nan, -np.nan, -np.inf, np.inf, -0.0, 0.0, 1.5, -2.5, 1e-300, -1e-300])
    f = pool[rng.integers(0, len(pool), n)]
    f_valid = rng.random(n) > 0.15
    s = rng.integers(0, len(SORT_DICT), n).astype(np.uint32)
    s_valid = rng.random(n) > 0.15
    i = rng.integers(-2, 3, n).astype(np.int32)
    return HostTable([HostColumn(FLOAT64, f, None, f_valid), HostColumn(UTF8, s, SORT_DICT, s_valid),
                      HostColumn(INT32, i), HostColumn(INT64, np.arange(n, dtype=np.int64))])


def _f64_rank(x):
    # f64::total_cmp: -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN
    if math.isnan(x):
        return (2 if math.copysign(1.0, x) > 0 else -2, 0.0, 0.0)
    return (0, x, math.copysign(1.0, x))


def _brute_sort(rows, keys):
    """Python's stable sort applied key by key from the last to the first,
    with each key's NULL placement and direction."""
    out = list(rows)
    for c, asc, nulls_first in reversed(keys):
        def k(r, c=c):
            v = r[c]
            if v is None:
                return (0,)
            if isinstance(v, float):
                return (1,) + _f64_rank(v)
            if isinstance(v, str):
                return (1, v.encode())
            return (1, v)
        nonnull = sorted([r for r in out if r[c] is not None], key=k, reverse=not asc)
        if not asc:  # a reversed stable sort must keep ties in input order
            nonnull = []
            groups = {}
            for r in out:
                if r[c] is not None:
                    groups.setdefault(k(r), []).append(r)
            for kk in sorted(groups, reverse=True):
                nonnull.extend(groups[kk])
        nulls = [r for r in out if r[c] is None]
        out = nulls + nonnull if nulls_first else nonnull + nulls
    return out


@pytest.mark.parametrize("asc0,nf0,asc1,nf1", [(a0, n0, a1, n1) for a0 in (True, False) for n0 in (True, False)
                                               for a1 in (True, False) for n1 in (True, False)])
def test_compare_rows_matches_brute_force(asc0, nf0, asc1, nf1):
    # ORDER BY f64 (total_cmp, NaN/-0/inf), then a dictionary string (byte
    # order, not code order): the full sort is stable (sort_by keeps ties in
    # input order; one thread, so input order is source order)
    t = _sort_table()
    eng = O.OracleEngine(threads=1)
    keys = [(col(0), asc0, nf0), (col(1), asc1, nf1)]
    got = rows_of(eng.topn(t, [], keys, -1).to_host())
    want = _brute_sort(rows_of(t), [(0, asc0, nf0), (1, asc1, nf1)])
    same = lambda a, b: (a is None and b is None) or (isinstance(a, float) and isinstance(b, float) and math.isnan(a)
                                                       and math.isnan(b)) or a == b
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[3] == w[3], (g, w)  # the same row at every position
        assert all(same(x, y) for x, y in zip(g, w))
    # with a limit: the first rows of the same order (ties at the cut may differ)
    top = rows_of(eng.topn(t, [], keys, 25).to_host())
    kf = lambda r: tuple(None if v is None else (_f64_rank(v) if isinstance(v, float) else v) for v in r[:2])
    assert [kf(r) for r in top] == [kf(r) for r in want[:25]]


STR_ENTRIES = ["pear", "Apple", "", "apple", "éclair", "z", "zz", "Zebra", "pea"]  # not in byte order


def _str_agg_table(n=4000, seed=41):
    """key i64 in [0, 60) (key 59: only NULL strings), a dictionary string with
    10% NULLs, and a small dictionary key (direct-table grouping)."""
    rng = np.random.default_rng(seed)
    key = rng.integers(0, 60, n).astype(np.int64)
    codes = rng.integers(0, len(STR_ENTRIES), n).astype(np.uint32)
    valid = (rng.random(n) > 0.1) & (key != 59)
    small = rng.integers(0, 3, n).astype(np.uint32)
    return HostTable([HostColumn(INT64, key), HostColumn(UTF8, codes, STR_ENTRIES, valid),
                      HostColumn(UTF8, small, ["x", "y", "w"])])


def _str_minmax_brute(t, key_col):
    keys, strs = t.columns[key_col].decoded(), t.columns[1].decoded()
    out = {}
    for k, s in zip(keys, strs):
        lo, hi = out.get(k, (None, None))
        if s is not None:  # NULL arguments are skipped; bytes compare unsigned (Vec<u8> Ord)
            b = s.encode("utf-8")
            lo = s if lo is None or b < lo.encode("utf-8") else lo
            hi = s if hi is None or b > hi.encode("utf-8") else hi
        out[k] = (lo, hi)
    return out


STR_AGGS = [Agg(AGG_MIN, col(1), UTF8), Agg(abi.AGG_MAX, col(1), UTF8), Agg(AGG_COUNT_STAR)]


@pytest.mark.parametrize("key_col", [0, 2])
def test_string_min_max_matches_brute_force(key_col):
    """State::MinMaxStr (accumulator.rs:94, 352-368, 470-484, 546-553): MIN/MAX
    of strings by unsigned byte order ('é' sorts after 'z', upper case before
    lower case, '' first), NULL arguments skipped, a group with none is NULL.
    Threads and small morsels make the partial-state merge run; Partial states
    of three slices merged by Final equal Single."""
    t = _str_agg_table()
    eng = O.OracleEngine(threads=3, morsel_rows=301)
    kt = [INT64] if key_col == 0 else [UTF8]
    got = eng.hashagg(t, [], [col(key_col)], kt, STR_AGGS).to_host()
    assert got.columns[1].type == UTF8 and got.columns[2].type == UTF8
    want = _str_minmax_brute(t, key_col)
    rows = {r[0]: (r[1], r[2]) for r in rows_of(got)}
    assert rows == want
    if key_col == 0:
        assert rows[59] == (None, None)
    cuts = [0, 1500, 1501, t.num_rows]
    parts = [eng.hashagg(t.take(np.arange(a, b)), [], [col(key_col)], kt, STR_AGGS, mode=abi.AGG_MODE_PARTIAL).to_host()
             for a, b in zip(cuts, cuts[1:])]
    final = eng.hashagg(concat_tables(parts), [], [col(key_col)], kt, STR_AGGS, mode=abi.AGG_MODE_FINAL).to_host()
    assert_tables_equal(final, got, rtol=0)
    ungrouped = eng.hashagg(t, [], [], [], STR_AGGS).to_host()
    allv = [s for s in t.columns[1].decoded() if s is not None]
    assert rows_of(ungrouped) == [(min(allv, key=lambda s: s.encode()), max(allv, key=lambda s: s.encode()), t.num_rows)]
