"""GPU parity of the single operators against the CPU oracle: filter
(ascending selection, bit-exact), projection (per-row values bit-exact,
NULLs from x/0 and x%0), hash_columns (bit-identical to the reference hash),
join probe with duplicate build keys, hash aggregate edge cases, TopN."""

import numpy as np
import pytest

from minilake_amd import abi
from minilake_amd.abi import (AGG_AVG, AGG_COUNT, AGG_COUNT_STAR, AGG_MAX, AGG_MIN, AGG_SUM, BOOLEAN, DATE,
                              FLOAT64, INT32, INT64, UTF8, Agg, col, date_lit, lit)
from minilake_amd.table import HostColumn, HostTable
from tests.compare import assert_tables_equal, concat_tables, rows_of

pytestmark = pytest.mark.gpu


def mixed_table(n=5000, seed=3):
    rng = np.random.default_rng(seed)
    i32 = rng.integers(-1000, 1000, n).astype(np.int32)
    i64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    f64 = rng.normal(0, 100, n)
    f64[::17] = 0.0
    f64[5::31] = -0.0
    b = (rng.integers(0, 2, n)).astype(np.uint8)
    d = rng.integers(8000, 9000, n).astype(np.int32)
    s = rng.integers(0, 4, n).astype(np.uint32)
    small = rng.integers(-3, 4, n).astype(np.int32)  # includes 0 divisors
    return HostTable([
        HostColumn(INT32, i32), HostColumn(INT64, i64), HostColumn(FLOAT64, f64), HostColumn(BOOLEAN, b),
        HostColumn(DATE, d), HostColumn(UTF8, s, ["apple", "banana", "cherry", "date"]), HostColumn(INT32, small),
    ])


PREDICATES = [
    col(0).lt(lit(0)),
    col(1).ge(lit(0)) & col(2).lt(lit(50.0)),
    col(4).le(date_lit(8500)) | col(3).eq(lit(True)),
    col(5).eq(lit("banana")),
    col(5).gt(lit("b")),
    ~(col(6).eq(lit(0))),
    (col(0) % col(6)).eq(lit(1)),          # NULL where col6 == 0
    (col(2) / col(6)).gt(lit(1.0)),         # NULL where col6 == 0
    col(0).cast(FLOAT64).gt(col(2)),
    lit(True) & col(3).eq(lit(False)),
]


@pytest.mark.parametrize("k", range(len(PREDICATES)))
def test_filter_matches_oracle(gpu_engine, oracle_engine, k):
    t = mixed_table()
    pred = PREDICATES[k]
    want = oracle_engine.filter(pred, t)
    got = gpu_engine.filter(pred, t)
    np.testing.assert_array_equal(got, want)
    sel = np.arange(0, t.num_rows, 3, dtype=np.uint32)
    np.testing.assert_array_equal(gpu_engine.filter(pred, t, sel), oracle_engine.filter(pred, t, sel))


EXPRS = [
    col(2) * (lit(1.0) - col(2)),
    (col(2) * (lit(1.0) - col(2))) * (lit(1.0) + col(2)),
    col(0) + col(6),
    col(0) * lit(2147483) ,                # wrapping i32 multiply
    col(1) * lit(12345678901),             # wrapping i64 multiply
    col(0) / col(6),                       # f64, NULL on zero
    col(0) % col(6),                       # NULL on zero
    col(1) % col(6),
    col(2) % lit(3.0),
    col(4) + lit(30),                      # date arithmetic stays a date
    col(4) - col(4),
    -col(2),
    -col(0),
    col(2).cast(INT64),
    col(2).cast(INT32),
    col(1).cast(INT32),
    col(3).cast(FLOAT64),
    col(0) + col(1),
    col(0) + lit(1.5),
]


@pytest.mark.parametrize("k", range(len(EXPRS)))
def test_project_matches_oracle_bit_exact(gpu_engine, oracle_engine, k):
    t = mixed_table()
    e = EXPRS[k]
    want = oracle_engine.project([e], t)
    got = gpu_engine.project([e], t)
    assert got.types == want.types
    g, w = got.columns[0], want.columns[0]
    gv = np.ones(t.num_rows, bool) if g.validity is None else g.validity
    wv = np.ones(t.num_rows, bool) if w.validity is None else w.validity
    np.testing.assert_array_equal(gv, wv)
    # values must be bit-identical wherever valid
    np.testing.assert_array_equal(g.values.view(np.uint8).reshape(t.num_rows, -1)[wv],
                                  w.values.view(np.uint8).reshape(t.num_rows, -1)[wv])


HASH_KEYS = [[col(0)], [col(1)], [col(2)], [col(3)], [col(4)], [col(5)], [col(1), col(0), col(5)],
             [col(0) / col(6)], [col(2), col(4)]]


@pytest.mark.parametrize("k", range(len(HASH_KEYS)))
def test_hash_columns_bit_identical(gpu_engine, oracle_engine, k):
    t = mixed_table()
    keys = HASH_KEYS[k]
    np.testing.assert_array_equal(gpu_engine.hash_columns(keys, t), oracle_engine.hash_columns(keys, t))
    sel = np.arange(1, t.num_rows, 7, dtype=np.uint32)
    np.testing.assert_array_equal(gpu_engine.hash_columns(keys, t, sel), oracle_engine.hash_columns(keys, t, sel))


def test_join_probe_with_duplicate_build_keys(gpu_engine, oracle_engine):
    rng = np.random.default_rng(5)
    bkeys = rng.integers(0, 300, 1000)
    build = HostTable([HostColumn(INT64, bkeys), HostColumn(INT32, np.arange(1000, dtype=np.int32))])
    probe = HostTable([HostColumn(INT64, rng.integers(-10, 320, 4000))])
    jo = oracle_engine.join_build(build, [], [col(0)], [INT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT64])
    assert not jg.unique() and jg.num_rows() == 1000
    po, bo = oracle_engine.join_probe(jo, [col(0)], probe)
    pg, bg = gpu_engine.join_probe(jg, [col(0)], probe)
    # build row ids are table-internal: compare (probe row, original build row id)
    ids_g = jg.column(1)
    assert sorted(zip(pg.tolist(), ids_g[bg].tolist())) == sorted(zip(po.tolist(), bo.tolist()))
    # every pair joins equal keys
    assert (jg.column(0)[bg] == probe.columns[0].values[pg]).all()


def test_join_build_many_matches_oracle(gpu_engine):
    """Builds finalized together (one device read for all layout decisions):
    a large dense table, a small one (one-workgroup build), one with duplicate
    keys (hashed layout) and a two-key table; every table probes like the
    oracle's, pairs compared as (probe row, build payload). The oracle runs one
    thread so its build rows keep source order (with several threads the
    morsels of a multi-morsel build are concatenated in completion order)."""
    from oracle.oracle import OracleEngine
    oracle_engine = OracleEngine(threads=1)
    rng = np.random.default_rng(8)
    big = HostTable([HostColumn(INT64, rng.permutation(200_000).astype(np.int64)), HostColumn(FLOAT64, rng.random(200_000))])
    small = HostTable([HostColumn(INT32, np.arange(50, dtype=np.int32)), HostColumn(FLOAT64, rng.random(50))])
    dup = HostTable([HostColumn(INT64, rng.integers(0, 5000, 20_000)), HostColumn(FLOAT64, rng.random(20_000))])
    two = HostTable([HostColumn(INT64, rng.permutation(100_000).astype(np.int64)),
                     HostColumn(FLOAT64, rng.random(100_000)), HostColumn(INT32, rng.integers(0, 25, 100_000).astype(np.int32))])
    specs = [(big, [], [col(0)], [INT64]), (small, [], [col(0)], [INT32]), (dup, [], [col(0)], [INT64]),
             (two, [], [col(0), col(2)], [INT64, INT32])]
    tables = gpu_engine.join_build_many(specs)
    probe = HostTable([HostColumn(INT64, rng.integers(-10, 210_000, 30_000)),
                       HostColumn(INT32, rng.integers(0, 25, 30_000).astype(np.int32))])
    for (src, ops, keys, kt), jg in zip(specs, tables):
        jo = oracle_engine.join_build(src, ops, keys, kt)
        pk = [col(1)] if kt[0] == INT32 else ([col(0)] if len(kt) == 1 else [col(0), col(1)])
        po, bo = oracle_engine.join_probe(jo, pk, probe)
        pg, bg = gpu_engine.join_probe(jg, pk, probe)
        want = sorted(zip(po.tolist(), src.columns[1].values[bo].tolist()))
        assert sorted(zip(pg.tolist(), jg.column(1)[bg].tolist())) == want and len(want) > 0
    assert not tables[2].unique() and tables[0].unique()


@pytest.mark.parametrize("spread", [3, 6])
def test_offset_indexed_dense_table(gpu_engine, spread):
    """A dense table with gaps in its key range (keys spread x3 and x6, some
    NULL keys and NULL payload values) keeps its rows at their key offset; the
    probe pairs and the column views still number the build rows one way,
    which must agree with each other and with the oracle's pairs (compared as
    (probe row, build payload)); one- and two-column keys."""
    from oracle.oracle import OracleEngine
    oracle_engine = OracleEngine(threads=1)
    rng = np.random.default_rng(21 + spread)
    nb = 90_000
    k64 = rng.permutation(nb).astype(np.int64) * spread - 77
    kv = rng.random(nb) >= 0.01  # NULL build keys
    ids = np.arange(nb, dtype=np.int64)
    pay = HostColumn(FLOAT64, rng.random(nb))
    pay.validity = rng.random(nb) >= 0.05
    k32 = rng.integers(0, 9, nb).astype(np.int32)
    kcol = HostColumn(INT64, k64)
    kcol.validity = kv
    build = HostTable([kcol, HostColumn(INT64, ids), pay, HostColumn(INT32, k32)])
    npr = 200_000
    pk64 = rng.integers(-200, nb * spread + 200, npr).astype(np.int64)
    pk32 = np.where(rng.random(npr) < 0.7, k32[rng.integers(0, nb, npr)], rng.integers(0, 9, npr)).astype(np.int32)
    probe = HostTable([HostColumn(INT64, pk64), HostColumn(INT32, pk32)])
    for keys, kt, pkeys in (([col(0)], [INT64], [col(0)]), ([col(0), col(3)], [INT64, INT32], [col(0), col(1)])):
        jo = oracle_engine.join_build(build, [], keys, kt)
        jg = gpu_engine.join_build(build, [], keys, kt)
        assert jg.unique() and jg.num_rows() == nb
        po, bo = oracle_engine.join_probe(jo, pkeys, probe)
        pg, bg = gpu_engine.join_probe(jg, pkeys, probe)
        want = sorted(zip(po.tolist(), ids[bo].tolist()))
        assert len(want) > 1000
        assert sorted(zip(pg.tolist(), jg.column(1)[bg].tolist())) == want
        # the views hold every build row once
        assert sorted(jg.column(1).tolist()) == ids.tolist()
        got_k = jg.column(0)
        by_id = np.empty(nb, dtype=np.int64)
        by_id[jg.column(1)] = got_k
        assert (by_id[kv] == k64[kv]).all()
        # a pipeline reading the build payload through the probe
        got = gpu_engine.collect(probe, [abi.probe_op(jg, pkeys)]).to_host()
        exp = oracle_engine.collect(probe, [abi.probe_op(jo, pkeys)]).to_host()
        assert_tables_equal(got, exp, rtol=0.0)


def test_expanding_probe_into_aggregate(gpu_engine, oracle_engine):
    rng = np.random.default_rng(9)
    build = HostTable([HostColumn(INT64, rng.integers(0, 50, 400)), HostColumn(FLOAT64, rng.random(400))])
    probe = HostTable([HostColumn(INT64, rng.integers(0, 60, 3000)), HostColumn(INT32, rng.integers(0, 5, 3000).astype(np.int32))])
    aggs = [Agg(AGG_SUM, col(3), FLOAT64), count_star_agg(), Agg(AGG_MAX, col(3), FLOAT64)]
    jo = oracle_engine.join_build(build, [], [col(0)], [INT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT64])
    want = oracle_engine.hashagg(probe, [abi.probe_op(jo, [col(0)])], [col(1)], [INT32], aggs).to_host()
    got = gpu_engine.hashagg(probe, [abi.probe_op(jg, [col(0)])], [col(1)], [INT32], aggs).to_host()
    assert_tables_equal(got, want, rtol=1e-12)


def count_star_agg():
    return Agg(AGG_COUNT_STAR)


def test_hashagg_all_functions_and_null_keys(gpu_engine, oracle_engine):
    t = mixed_table(20000)
    aggs = [Agg(AGG_SUM, col(0), INT32), Agg(AGG_SUM, col(2), FLOAT64), Agg(AGG_AVG, col(0), INT32),
            Agg(AGG_AVG, col(2), FLOAT64), Agg(AGG_MIN, col(4), DATE), Agg(AGG_MAX, col(1), INT64),
            Agg(AGG_MIN, col(2), FLOAT64), Agg(AGG_COUNT, col(0) / col(6), FLOAT64), count_star_agg(),
            Agg(AGG_SUM, col(0) % col(6), INT32)]
    # group by an expression that is NULL when col6 == 0, and a dictionary column
    group = [col(6) % col(6).cast(INT32) + col(0) % lit(7), col(5)]
    want = oracle_engine.hashagg(t, [], group, [INT32, UTF8], aggs).to_host()
    got = gpu_engine.hashagg(t, [], group, [INT32, UTF8], aggs).to_host()
    assert_tables_equal(got, want, rtol=1e-12)


def test_hashagg_high_cardinality_plain_resizes(gpu_engine, oracle_engine):
    rng = np.random.default_rng(1)
    n = 400_000
    t = HostTable([HostColumn(INT64, rng.integers(0, 150_000, n)), HostColumn(FLOAT64, rng.random(n))])
    aggs = [Agg(AGG_SUM, col(1), FLOAT64), count_star_agg()]
    want = oracle_engine.hashagg(t, [], [col(0)], [INT64], aggs).to_host()
    stats = {}
    from minilake_amd.gpu import GpuEngine
    small = GpuEngine(batch_rows=50_000)
    try:
        got = small.hashagg(t, [], [col(0)], [INT64], aggs, stats=stats).to_host()
    finally:
        small.close()
    assert stats["resizes"] >= 2
    assert_tables_equal(got, want, rtol=1e-12)


def test_hashagg_learned_underestimate_falls_back(gpu_engine, oracle_engine):
    rng = np.random.default_rng(2)
    n = 200_000
    t = HostTable([HostColumn(INT64, rng.integers(0, 50_000, n)), HostColumn(INT64, rng.integers(0, 100, n))])
    aggs = [Agg(AGG_SUM, col(1), INT64), count_star_agg()]
    want = oracle_engine.hashagg(t, [], [col(0)], [INT64], aggs).to_host()
    got = gpu_engine.hashagg(t, [], [col(0)], [INT64], aggs, sizing=abi.SIZING_LEARNED, capacity_hint=100).to_host()
    assert_tables_equal(got, want)
    with pytest.raises(Exception):
        gpu_engine.hashagg(t, [], [col(0)], [INT64], aggs, sizing=abi.SIZING_HINT, capacity_hint=100)


def test_topn_keys_and_ties(gpu_engine, oracle_engine):
    rng = np.random.default_rng(4)
    n = 50_000
    t = HostTable([HostColumn(FLOAT64, np.round(rng.normal(0, 10, n), 1)), HostColumn(DATE, rng.integers(0, 50, n).astype(np.int32)),
                   HostColumn(INT64, np.arange(n, dtype=np.int64))])
    keys = [(col(0), False, False), (col(1), True, False)]
    for limit in (1, 10, 100):
        want = rows_of(oracle_engine.topn(t, [], keys, limit).to_host())
        got = rows_of(gpu_engine.topn(t, [], keys, limit).to_host())
        assert [(r[0], r[1]) for r in got] == [(r[0], r[1]) for r in want]


@pytest.mark.parametrize("n", [1, 5, 37, 1025, 2100])
def test_topn_short_tiles(gpu_engine, oracle_engine, n):
    """Tiles holding fewer rows than the tile size, and limits above the row count."""
    rng = np.random.default_rng(n)
    t = HostTable([HostColumn(FLOAT64, rng.permutation(n).astype(np.float64)),
                   HostColumn(INT64, np.arange(n, dtype=np.int64))])
    for limit in (1, 3, 25, 100):
        want = rows_of(oracle_engine.topn(t, [], [(col(0), False, False)], limit).to_host())
        got = rows_of(gpu_engine.topn(t, [], [(col(0), False, False)], limit).to_host())
        assert got == want


def test_collect_through_probe(gpu_engine, oracle_engine):
    rng = np.random.default_rng(8)
    build = HostTable([HostColumn(INT64, np.arange(1000, dtype=np.int64)), HostColumn(UTF8, rng.integers(0, 3, 1000).astype(np.uint32), ["x", "y", "z"])])
    probe = HostTable([HostColumn(INT64, rng.integers(0, 2000, 5000)), HostColumn(FLOAT64, rng.random(5000))])
    jo = oracle_engine.join_build(build, [], [col(0)], [INT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT64])
    ops_o = [abi.filter_op(col(1).lt(lit(0.7))), abi.probe_op(jo, [col(0)], build_is_left=True)]
    ops_g = [abi.filter_op(col(1).lt(lit(0.7))), abi.probe_op(jg, [col(0)], build_is_left=True)]
    want = oracle_engine.collect(probe, ops_o).to_host()
    got = gpu_engine.collect(probe, ops_g).to_host()
    assert_tables_equal(got, want, rtol=0.0)


@pytest.mark.parametrize("keys", ["i64", "i32", "i64+i32"])
def test_streaming_first_probe(gpu_engine, oracle_engine, keys):
    """A probe as the first pipeline step runs several rows per lane with the
    Bloom filter (>= 64K build rows) and linear-probe collisions; results must
    equal the oracle's for one- and two-column keys of both widths."""
    rng = np.random.default_rng(11)
    nb, npr = 120_000, 300_000
    bk64 = rng.permutation(400_000)[:nb].astype(np.int64) * 7 - 1_000_000
    bk32 = rng.integers(0, 5, nb).astype(np.int32)
    build = HostTable([HostColumn(INT64, bk64), HostColumn(INT32, bk32), HostColumn(FLOAT64, rng.random(nb))])
    pick = rng.integers(0, nb, npr)
    miss = rng.random(npr) < 0.6
    p64 = np.where(miss, rng.integers(-(1 << 40), 1 << 40, npr), bk64[pick]).astype(np.int64)
    p32 = np.where(rng.random(npr) < 0.2, rng.integers(0, 5, npr), bk32[pick]).astype(np.int32)
    probe = HostTable([HostColumn(INT64, p64), HostColumn(INT32, p32), HostColumn(INT64, np.arange(npr, dtype=np.int64))])
    if keys == "i64":
        bkeys, btypes, pkeys = [col(0)], [INT64], [col(0)]
    elif keys == "i32":
        bkeys, btypes, pkeys = [col(1)], [INT32], [col(1)]
        build = HostTable([HostColumn(INT64, bk64), HostColumn(INT32, np.arange(nb, dtype=np.int32) * 3),
                           HostColumn(FLOAT64, rng.random(nb))])
        probe = HostTable([HostColumn(INT64, p64), HostColumn(INT32, rng.integers(-1000, 3 * nb + 1000, npr).astype(np.int32)),
                           HostColumn(INT64, np.arange(npr, dtype=np.int64))])
    else:
        # (i64, i32) pairs are unique in the build side: the i64 part already is
        bkeys, btypes, pkeys = [col(0), col(1)], [INT64, INT32], [col(0), col(1)]
    jo = oracle_engine.join_build(build, [], bkeys, btypes)
    jg = gpu_engine.join_build(build, [], bkeys, btypes)
    ops_o = [abi.probe_op(jo, pkeys), abi.filter_op(col(2).lt(lit(0.9)) | col(5).gt(lit(0.5)))]
    ops_g = [abi.probe_op(jg, pkeys), abi.filter_op(col(2).lt(lit(0.9)) | col(5).gt(lit(0.5)))]
    want = oracle_engine.collect(probe, ops_o).to_host()
    got = gpu_engine.collect(probe, ops_g).to_host()
    assert want.num_rows > 1000
    assert_tables_equal(got, want, rtol=0.0)
    assert gpu_engine.count(probe, ops_g) == want.num_rows


def test_join_build_heavy_duplicates_falls_back_unpartitioned(gpu_engine, oracle_engine):
    """Build keys with thousands of copies overflow one LDS partition; the
    table is then rebuilt unpartitioned and probes still see every copy."""
    nb = 30_000
    bk = (np.arange(nb) % 3).astype(np.int64) * 1_000_003
    build = HostTable([HostColumn(INT64, bk), HostColumn(INT64, np.arange(nb, dtype=np.int64))])
    probe = HostTable([HostColumn(INT64, np.array([0, 1_000_003, 7, 2_000_006, 0], dtype=np.int64))])
    jo = oracle_engine.join_build(build, [], [col(0)], [INT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT64])
    assert not jg.unique()
    want = oracle_engine.count(probe, [abi.probe_op(jo, [col(0)])])
    assert want == 4 * 10_000
    assert gpu_engine.count(probe, [abi.probe_op(jg, [col(0)])]) == want
    agg_o = oracle_engine.hashagg(probe, [abi.probe_op(jo, [col(0)])], [col(0)], [INT64],
                                  [Agg(AGG_SUM, col(2), INT64), Agg(AGG_COUNT_STAR, None, INT64)])
    agg_g = gpu_engine.hashagg(probe, [abi.probe_op(jg, [col(0)])], [col(0)], [INT64],
                               [Agg(AGG_SUM, col(2), INT64), Agg(AGG_COUNT_STAR, None, INT64)])
    assert_tables_equal(agg_g.to_host(), agg_o.to_host(), rtol=0.0)


@pytest.mark.parametrize("nb", [1, 100, 70_000, 400_000])
def test_join_build_partitioned_sizes(gpu_engine, oracle_engine, nb):
    """Partitioned builds from one partition up to hundreds, with and without
    the Bloom filter (on from 64K rows), unique and with a few duplicates."""
    rng = np.random.default_rng(nb)
    bk = rng.integers(-(1 << 50), 1 << 50, nb).astype(np.int64)
    if nb > 10:
        bk[: nb // 50] = bk[nb // 50: 2 * (nb // 50)]  # 2% duplicated keys
    build = HostTable([HostColumn(INT64, bk), HostColumn(INT32, rng.integers(0, 9, nb).astype(np.int32))])
    npr = 50_000
    pk = np.where(rng.random(npr) < 0.5, bk[rng.integers(0, nb, npr)], rng.integers(-(1 << 50), 1 << 50, npr))
    probe = HostTable([HostColumn(INT64, pk.astype(np.int64))])
    jo = oracle_engine.join_build(build, [], [col(0)], [INT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT64])
    assert jg.unique() == (nb <= 10)
    ops_o, ops_g = [abi.probe_op(jo, [col(0)])], [abi.probe_op(jg, [col(0)])]
    assert gpu_engine.count(probe, ops_g) == oracle_engine.count(probe, ops_o)
    agg_o = oracle_engine.hashagg(probe, ops_o, [col(2)], [INT32], [Agg(AGG_COUNT_STAR, None, INT64)])
    agg_g = gpu_engine.hashagg(probe, ops_g, [col(2)], [INT32], [Agg(AGG_COUNT_STAR, None, INT64)])
    assert_tables_equal(agg_g.to_host(), agg_o.to_host(), rtol=0.0)


@pytest.mark.parametrize("case", ["i64_neg", "i32", "dup_falls_back", "null_keys", "two_keys", "two_keys_wide",
                                  "two_keys_dup_first", "two_keys_small"])
def test_join_dense_layout(gpu_engine, oracle_engine, case):
    """Unique integer keys over a small range use the direct-addressed layout;
    probes below, inside and above the range, NULL build keys, and a repeated
    key (which must fall back to the hashed layout) all match the oracle. Two
    key columns: dense first key with the second compared per row (behind a
    Bloom filter from 64K rows), a first key over a wide range (hashed), and a
    repeated first key with distinct full keys (falls back to hashed). Builds
    of up to 65,536 rows take the one-workgroup dense build."""
    rng = np.random.default_rng(21)
    nb = 20_000 if case == "two_keys_small" else 70_000 if case.startswith("two_keys") else 50_000
    keys = rng.permutation(200_000)[:nb].astype(np.int64) - 100_000
    if case == "two_keys_wide":
        keys = keys * 1_000_003
    if case in ("dup_falls_back", "two_keys_dup_first"):
        keys[7] = keys[8]
    ktype = INT32 if case == "i32" else INT64
    zero = np.zeros(nb, dtype=np.int64)
    one = np.ones(nb, dtype=np.int64)
    if case == "null_keys":
        one[::5] = 0  # key + (0 % 0) is NULL on every fifth row
    build = HostTable([HostColumn(ktype, keys.astype(np.int32) if ktype == INT32 else keys),
                       HostColumn(INT64, np.arange(nb, dtype=np.int64)), HostColumn(INT64, zero),
                       HostColumn(INT64, one)])
    npr = 200_000
    pk = rng.integers(-150_000, 150_000, npr)
    probe = HostTable([HostColumn(ktype, pk.astype(np.int32) if ktype == INT32 else pk.astype(np.int64))])
    bkey = [col(0) + (col(2) % col(3))] if case == "null_keys" else [col(0)]
    btypes, pkeys = [ktype], [col(0)]
    if case.startswith("two_keys"):
        # (first key, small int tag); the tag differs on the repeated first key
        tag = (keys % 3).astype(np.int32)
        if case == "two_keys_dup_first":
            tag[7] = (tag[8] + 1) % 3
        build = HostTable(build.columns + [HostColumn(INT32, tag)])
        bkey, btypes, pkeys = [col(0), col(4)], [INT64, INT32], [col(0), col(1)]
        if case == "two_keys_wide":
            pk = np.where(rng.random(npr) < 0.5, keys[rng.integers(0, nb, npr)], pk)
        probe = HostTable([HostColumn(INT64, pk.astype(np.int64)),
                           HostColumn(INT32, (pk % 3 + (rng.random(npr) < 0.3)).astype(np.int32))])
    jo = oracle_engine.join_build(build, [], bkey, btypes)
    jg = gpu_engine.join_build(build, [], bkey, btypes)
    assert jg.unique() == (case != "dup_falls_back")
    ops_o, ops_g = [abi.probe_op(jo, pkeys)], [abi.probe_op(jg, pkeys)]
    want = oracle_engine.collect(probe, ops_o).to_host()
    got = gpu_engine.collect(probe, ops_g).to_host()
    assert want.num_rows > (5_000 if case == "two_keys_small" else 10_000)
    assert_tables_equal(got, want, rtol=0.0)


FLOW_PREDICATES = PREDICATES + [
    col(4).ge(date_lit(8200)) & col(4).lt(date_lit(8600)),                      # range: two atoms
    (col(0).lt(lit(-500)) | col(0).gt(lit(500))) & ~col(5).eq(lit("apple")),    # 3 atoms, OR, NOT
    col(1).ne(lit(0)) & col(2).ge(lit(-10.0)) & col(2).le(lit(10.0)) & col(6).eq(lit(2)),  # 4 atoms
    col(2).gt(lit(float("nan"))) | col(2).ne(lit(float("nan"))),                # NaN compares
    col(0).lt(lit(0)) & lit(True),                                              # a constant operand
]


@pytest.mark.parametrize("k", range(len(FLOW_PREDICATES)))
def test_pipeline_filter_collect_matches_oracle(gpu_engine, oracle_engine, k):
    """Filters as the first pipeline step (column-atom fast path or the
    interpreter), collected through the wave-queued kernel in source order."""
    t = mixed_table(n=70_000, seed=9)
    ops = [abi.filter_op(FLOW_PREDICATES[k])]
    want = oracle_engine.collect(t, ops).to_host()
    got = gpu_engine.collect(t, ops).to_host()
    assert_tables_equal(got, want, rtol=0.0)
    assert gpu_engine.count(t, ops) == want.num_rows


def _agg_table(n, groups, seed, skew=False):
    rng = np.random.default_rng(seed)
    if skew:  # one key holds a third of the rows: its partition is split into chunks and merged
        g = np.where(rng.random(n) < 0.33, 7, rng.integers(0, groups, n))
    else:
        g = rng.integers(0, groups, n)
    k2 = (g % 3).astype(np.int32)
    v = rng.random(n) * 100.0  # positive (prices): FP64 sums in any order agree to ~n ulp
    x = rng.integers(-5, 5, n).astype(np.int64)
    return HostTable([HostColumn(INT64, (g * 2654435761) % (1 << 40)), HostColumn(INT32, k2), HostColumn(FLOAT64, v),
                      HostColumn(INT64, x)])


@pytest.mark.parametrize("case", ["uniform", "skew", "nullable"])
def test_hashagg_partitioned_matches_oracle(gpu_engine, oracle_engine, case):
    """A group-count hint of >= 65,536 groups aggregates in LDS partitions:
    every aggregate function, NULL keys / arguments (x / 0), and a hot key
    whose partition is split into chunks, against the oracle."""
    n, groups = 600_000, 150_000
    t = _agg_table(n, groups, 5, skew=(case == "skew"))
    keys = [col(0), col(1)]
    ktypes = [INT64, INT32]
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(2), FLOAT64), Agg(AGG_MIN, col(3), INT64),
            Agg(AGG_MAX, col(2), FLOAT64), Agg(AGG_AVG, col(3), INT64), Agg(AGG_COUNT, col(2), FLOAT64)]
    if case == "nullable":
        keys = [col(0), col(1) % col(3).cast(INT32)]               # NULL where x == 0
        # NULL where x == 0; positive terms, so any summation order agrees to ~n ulp
        aggs = aggs + [Agg(AGG_SUM, col(2) / (col(3) * col(3)).cast(FLOAT64), FLOAT64)]
    want = oracle_engine.hashagg(t, [], keys, ktypes, aggs).to_host()
    st = {}
    got = gpu_engine.hashagg(t, [], keys, ktypes, aggs, sizing=abi.SIZING_HINT, capacity_hint=int(want.num_rows * 1.1),
                             stats=st)
    assert st["resizes"] == 0
    assert_tables_equal(got.to_host(), want, rtol=1e-12)


def test_hashagg_partitioned_underestimate(gpu_engine, oracle_engine):
    """A hint far below the real group count overflows the partitions: HINT
    reports ML_ERR_CAPACITY; LEARNED doubles the partitioned table (the rows
    partitioned again) until every partition fits, and is exact."""
    t = _agg_table(800_000, 600_000, 6)  # ~440k groups, far above a 256k-slot table
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(3), INT64)]
    want = oracle_engine.hashagg(t, [], [col(0)], [INT64], aggs).to_host()
    with pytest.raises(Exception):
        gpu_engine.hashagg(t, [], [col(0)], [INT64], aggs, sizing=abi.SIZING_HINT, capacity_hint=70_000)
    st = {}
    got = gpu_engine.hashagg(t, [], [col(0)], [INT64], aggs, sizing=abi.SIZING_LEARNED, capacity_hint=70_000,
                             stats=st)
    assert_tables_equal(got.to_host(), want, rtol=0.0)
    # 256k slots -> 512k -> (at most) 1M for ~440k groups
    assert 1 <= st["resizes"] <= 2 and st["capacity"] >= 512 * 1024


def test_ungrouped_aggregate_matches_oracle(gpu_engine, oracle_engine):
    """No GROUP BY (UngroupedAggregate, ungrouped.rs:36-139): one row, also over
    no input rows (COUNT 0, the rest NULL); NULL arguments from x / 0."""
    t = mixed_table(30_000, seed=12)
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(0), INT32), Agg(AGG_SUM, col(2), FLOAT64), Agg(AGG_AVG, col(0), INT32),
            Agg(AGG_MIN, col(4), DATE), Agg(AGG_MAX, col(1), INT64), Agg(AGG_COUNT, col(0) / col(6), FLOAT64),
            Agg(AGG_SUM, col(2) / col(6), FLOAT64)]
    for ops in ([], [abi.filter_op(col(0).lt(lit(-100)))], [abi.filter_op(col(0).lt(lit(-5000)))]):
        want = oracle_engine.hashagg(t, ops, [], [], aggs).to_host()
        got = gpu_engine.hashagg(t, ops, [], [], aggs).to_host()
        assert got.num_rows == want.num_rows == 1
        assert_tables_equal(got, want, rtol=1e-12)


def _with_nulls(t, cols, seed, frac=0.2):
    rng = np.random.default_rng(seed)
    for i in cols:
        t.columns[i].validity = rng.random(t.num_rows) >= frac
    return t


def test_nullable_source_columns(gpu_engine, oracle_engine):
    """Input columns with validity bitmaps (Column, column.rs:126-200): filters,
    projections, group keys, aggregate arguments and join keys see the NULLs
    as the oracle does (NULL keys never join, hash.rs group NULL keys together),
    and a device result holding NULLs keeps them as the source of the next
    pipeline."""
    t = _with_nulls(mixed_table(20_000, seed=41), (0, 1, 2, 5), seed=41)
    for pred in (col(0).lt(lit(0)) | col(2).gt(lit(10.0)), col(5).eq(lit("banana")), ~(col(1).ge(lit(0)))):
        np.testing.assert_array_equal(gpu_engine.filter(pred, t), oracle_engine.filter(pred, t))
    for e in (col(0) + col(6), col(2) * lit(2.0), col(1) % col(6)):
        want, got = oracle_engine.project([e], t), gpu_engine.project([e], t)
        assert rows_of(got) == rows_of(want)
    aggs = [Agg(AGG_SUM, col(0), INT32), Agg(AGG_COUNT, col(2), FLOAT64), Agg(AGG_AVG, col(0), INT32),
            Agg(AGG_MAX, col(2), FLOAT64), Agg(AGG_MIN, col(1), INT64), Agg(AGG_COUNT_STAR)]
    for keys, kt in (([col(5), col(6)], [UTF8, INT32]), ([col(0)], [INT32]), ([], [])):
        want = oracle_engine.hashagg(t, [], keys, kt, aggs).to_host()
        got = gpu_engine.hashagg(t, [], keys, kt, aggs).to_host()
        assert_tables_equal(got, want, rtol=1e-12)
    # a collected device result with NULLs as the next source
    ops = [abi.filter_op(col(6).ne(lit(2)))]
    mid_o = oracle_engine.collect(t, ops).to_host()
    mid_g = gpu_engine.collect(t, ops)
    assert rows_of(mid_g.to_host()) == rows_of(mid_o)
    want = oracle_engine.hashagg(mid_o, [], [col(6)], [INT32], aggs).to_host()
    got = gpu_engine.hashagg(mid_g, [], [col(6)], [INT32], aggs).to_host()
    assert_tables_equal(got, want, rtol=1e-12)
    # NULL join keys on both sides, duplicate build keys
    build = HostTable([t.columns[0], HostColumn(INT32, np.arange(t.num_rows, dtype=np.int32))]).head(3000)
    probe = HostTable([t.columns[0]]).take(np.arange(3000, 9000))
    jo = oracle_engine.join_build(build, [], [col(0)], [INT32])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT32])
    po, bo = oracle_engine.join_probe(jo, [col(0)], probe)
    pg, bg = gpu_engine.join_probe(jg, [col(0)], probe)
    ids_g = jg.column(1)
    assert sorted(zip(pg.tolist(), ids_g[bg].tolist())) == sorted(zip(po.tolist(), bo.tolist()))
    assert len(po) > 0


def test_mlb1_bytes_match_oracle(gpu_engine, oracle_engine):
    """MLB1 batch bytes (ipc.rs:44-170): the device writer's bytes equal the
    oracle's for every column type with NULLs and a dictionary column, with
    and without a selection (one keeping no NULL drops that validity section),
    and for a device result as the source (an empty selection: a batch of no
    rows); reading them back on the device gives the rows; a cut batch is an
    I/O error, an empty stream a clean end."""
    from minilake_amd.gpu import GpuError
    t = _with_nulls(mixed_table(50_000, seed=51), (0, 2, 3, 5), seed=51)
    no_null_rows = np.nonzero(t.columns[0].validity)[0].astype(np.uint32)
    for sel in (None, np.arange(0, t.num_rows, 3, dtype=np.uint32), no_null_rows, np.array([], np.uint32)):
        want = oracle_engine.mlb1_write(t, sel)
        got = gpu_engine.mlb1_write(t, sel)
        assert got == want
        back, used = gpu_engine.mlb1_read(got)
        assert used == len(got)
        assert rows_of(back.to_host()) == rows_of(t if sel is None else t.take(sel))
    ops = [abi.filter_op(col(6).ne(lit(2)))]
    mid = gpu_engine.collect(t, ops)
    assert gpu_engine.mlb1_write(mid) == oracle_engine.mlb1_write(oracle_engine.collect(t, ops).to_host())
    two = got + gpu_engine.mlb1_write(mid)
    first, used = gpu_engine.mlb1_read(two)
    second, used2 = gpu_engine.mlb1_read(two[used:])
    assert used + used2 == len(two) and rows_of(second.to_host()) == rows_of(mid.to_host())
    with pytest.raises(GpuError) as e:
        gpu_engine.mlb1_read(got[:-1])
    assert e.value.code == abi.ML_ERR_IO
    assert gpu_engine.mlb1_read(b"") == (None, 0)


@pytest.mark.parametrize("mode", [abi.AGG_MODE_SINGLE, abi.AGG_MODE_PARTIAL, abi.AGG_MODE_FINAL])
def test_memory_budget_and_spill(oracle_engine, mode):
    """MemoryPool + spill (memory.rs:22-157, aggregate/hash.rs:189-282, 369-425):
    a GROUP BY of ~150K groups (an i64 and a dictionary string key) under a
    budget far below its table: without spill ResourcesExhausted; with spill
    the groups go to host memory in 16 hash partitions, are merged one
    partition at a time, and the result equals the oracle's (Partial mode:
    the same states); the budget is never exceeded and is all given back."""
    from minilake_amd.gpu import GpuEngine, GpuError
    rng = np.random.default_rng(77)
    n = 600_000
    t = HostTable([HostColumn(INT64, rng.integers(0, 40_000, n)),
                   HostColumn(UTF8, rng.integers(0, 4, n).astype(np.uint32), ["pear", "apple", "fig", "date"]),
                   HostColumn(FLOAT64, rng.uniform(1, 100, n)),
                   HostColumn(INT32, rng.integers(-3, 4, n).astype(np.int32)),
                   HostColumn(DATE, rng.integers(8000, 9000, n).astype(np.int32))])
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(2), FLOAT64), Agg(AGG_AVG, col(3), INT32),
            Agg(AGG_MIN, col(4), DATE), Agg(AGG_MAX, col(2) / col(3), FLOAT64), Agg(AGG_SUM, col(3), INT32)]
    keys, kt = [col(0), col(1)], [INT64, UTF8]
    if mode == abi.AGG_MODE_FINAL:  # the input is the partial states of three slices
        cuts = [0, 200_000, 400_000, n]
        t = concat_tables([oracle_engine.hashagg(t.take(np.arange(a, b)), [], keys, kt, aggs,
                                                 mode=abi.AGG_MODE_PARTIAL).to_host() for a, b in zip(cuts, cuts[1:])])
    want = oracle_engine.hashagg(t, [], keys, kt, aggs, mode=mode).to_host()
    assert want.num_rows > 100_000
    eng = GpuEngine(batch_rows=50_000)
    try:
        limit = 6 << 20
        assert_tables_equal(eng.hashagg(t, [], keys, kt, aggs, mode=mode).to_host(), want, rtol=1e-12)
        assert eng.memory_usage()["peak"] > limit  # unlimited: the whole table at once
        eng.set_memory_limit(limit, spill=False)
        assert eng.memory_usage() == {"used": 0, "peak": 0, "limit": limit}  # the peak restarts
        with pytest.raises(GpuError) as e:
            eng.hashagg(t, [], keys, kt, aggs, mode=mode)
        assert e.value.code == abi.ML_ERR_RESOURCES_EXHAUSTED
        assert eng.memory_usage()["used"] == 0
        eng.set_memory_limit(limit, spill=True)
        got = eng.hashagg(t, [], keys, kt, aggs, mode=mode).to_host()
        usage = eng.memory_usage()
        assert 0 < usage["peak"] <= limit and usage["used"] == 0
        assert_tables_equal(got, want, rtol=1e-12)
    finally:
        eng.close()


def test_join_table_reserved_against_budget():
    """JoinBuildSink::finalize reserves the built table (join.rs:120-153): over
    the budget ResourcesExhausted; otherwise held until the table is freed."""
    from minilake_amd.gpu import GpuEngine, GpuError
    rng = np.random.default_rng(5)
    t = HostTable([HostColumn(INT64, rng.permutation(100_000).astype(np.int64)), HostColumn(FLOAT64, rng.random(100_000))])
    eng = GpuEngine()
    try:
        eng.set_memory_limit(1 << 20)
        with pytest.raises(GpuError) as e:
            eng.join_build(t, [], [col(0)], [INT64])
        assert e.value.code == abi.ML_ERR_RESOURCES_EXHAUSTED
        assert eng.memory_usage()["used"] == 0
        eng.set_memory_limit(64 << 20)
        j = eng.join_build(t, [], [col(0)], [INT64])
        assert 1 << 20 < eng.memory_usage()["used"] <= 64 << 20
        j.close()
        assert eng.memory_usage()["used"] == 0
    finally:
        eng.close()


def _modes_table(n, seed):
    rng = np.random.default_rng(seed)
    return HostTable([HostColumn(INT32, rng.integers(-3, 4, n).astype(np.int32)),
                      HostColumn(FLOAT64, rng.uniform(1, 100, n)),
                      HostColumn(INT64, rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)),
                      HostColumn(DATE, rng.integers(8000, 9000, n).astype(np.int32)),
                      HostColumn(UTF8, rng.integers(0, 4, n).astype(np.uint32), ["pear", "apple", "fig", "date"]),
                      HostColumn(INT32, rng.integers(0, 1000, n).astype(np.int32))])


_NULLY = col(1) / (col(0) * col(0))  # positive, NULL where col 0 is 0
_MODE_AGGS = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(5), INT32), Agg(AGG_SUM, _NULLY, FLOAT64), Agg(AGG_AVG, col(5), INT32),
              Agg(AGG_AVG, _NULLY, FLOAT64), Agg(AGG_MIN, col(3), DATE), Agg(AGG_MAX, col(2), INT64),
              Agg(AGG_MIN, _NULLY, FLOAT64), Agg(AGG_COUNT, _NULLY, FLOAT64), Agg(AGG_MAX, col(0), INT32)]


@pytest.mark.parametrize("grouped", ["none", "small", "wide"])
def test_partial_final_aggregate_matches_oracle(gpu_engine, oracle_engine, grouped):
    """AggMode Partial / Final (aggregate/mod.rs:50-57, split.rs:34-125): the
    GPU's partial states of three slices equal the oracle's (Accumulator::state,
    accumulator.rs:382-404); a Final aggregate over their concatenation equals
    the oracle's Final over the same rows and the Single result over the whole
    input. Keys: none; an int and a dictionary string (a direct table; the
    group with key 0 has only NULL arguments); a wide i64 (a hashed table)."""
    t = _modes_table(40_000, seed=31)
    keys, kt = {"none": ([], []), "small": ([col(0), col(4)], [INT32, UTF8]),
                "wide": ([col(2) % lit(20_000), col(0)], [INT64, INT32])}[grouped]
    cuts = [0, 9_000, 9_001, t.num_rows]
    parts = []
    for a, b in zip(cuts, cuts[1:]):
        piece = t.take(np.arange(a, b))
        want = oracle_engine.hashagg(piece, [], keys, kt, _MODE_AGGS, mode=abi.AGG_MODE_PARTIAL).to_host()
        got = gpu_engine.hashagg(piece, [], keys, kt, _MODE_AGGS, mode=abi.AGG_MODE_PARTIAL).to_host()
        assert_tables_equal(got, want, rtol=1e-12)
        parts.append(want)
    states = concat_tables(parts)
    want = oracle_engine.hashagg(states, [], keys, kt, _MODE_AGGS, mode=abi.AGG_MODE_FINAL).to_host()
    got = gpu_engine.hashagg(states, [], keys, kt, _MODE_AGGS, mode=abi.AGG_MODE_FINAL).to_host()
    assert_tables_equal(got, want, rtol=1e-12)
    assert_tables_equal(got, oracle_engine.hashagg(t, [], keys, kt, _MODE_AGGS).to_host(), rtol=1e-12)
    # no input rows: Partial emits the fresh states, Final the empty finish
    none = [abi.filter_op(col(0).gt(lit(5)))]
    for mode, src in ((abi.AGG_MODE_PARTIAL, t), (abi.AGG_MODE_FINAL, states)):
        ops = none if mode == abi.AGG_MODE_PARTIAL else [abi.filter_op(col(len(kt)).lt(lit(0)))]
        want = oracle_engine.hashagg(src, ops, keys, kt, _MODE_AGGS, mode=mode).to_host()
        got = gpu_engine.hashagg(src, ops, keys, kt, _MODE_AGGS, mode=mode).to_host()
        assert got.num_rows == want.num_rows == (1 if grouped == "none" else 0)
        assert_tables_equal(got, want, rtol=0)


@pytest.mark.parametrize("limit,offset", [(10, 0), (7, 123), (100_000, 4_000), (5, 10_000_000), (0, 0)])
def test_limit_matches_oracle(gpu_engine, limit, offset):
    """LIMIT / OFFSET (limit.rs:26-117): the GPU keeps rows [offset, offset +
    limit) in source order, which is what the oracle returns with one thread."""
    from oracle.oracle import OracleEngine
    t = mixed_table(20_000, seed=13)
    ops = [abi.filter_op(col(0).lt(lit(300)) & ~col(6).eq(lit(0)))]
    want = OracleEngine(threads=1).limit(t, ops, limit, offset).to_host()
    got = gpu_engine.limit(t, ops, limit, offset).to_host()
    assert got.num_rows == want.num_rows
    assert rows_of(got) == rows_of(want)


@pytest.mark.parametrize("n", [1, 700, 1024, 1025, 5_000, 100_000])
def test_full_sort_matches_oracle(gpu_engine, n):
    """ORDER BY without LIMIT (sort_batches, sort.rs:85-132) and with a limit
    above the TopN range: a stable sort by f64 DESC, an f64 with NULLs (x / 0)
    first, a date, then a dictionary string; ties keep source order, as the
    one-threaded oracle's concatenation does."""
    from oracle.oracle import OracleEngine
    rng = np.random.default_rng(n)
    t = HostTable([HostColumn(FLOAT64, np.round(rng.normal(0, 10, n), 1)),
                   HostColumn(DATE, rng.integers(0, 40, n).astype(np.int32)),
                   HostColumn(INT32, rng.integers(-2, 3, n).astype(np.int32)),
                   HostColumn(UTF8, rng.integers(0, 4, n).astype(np.uint32), ["pear", "apple", "fig", "date"]),
                   HostColumn(INT64, np.arange(n, dtype=np.int64))])
    keys = [(col(0), False, False), (col(0) / col(2), True, True), (col(1), True, False), (col(3), True, False)]

    def key_of(r):
        return (r[0], r[0] / r[2] if r[2] != 0 else None, r[1], r[3])

    oracle = OracleEngine(threads=1)
    for limit in (-1, 1000):
        want = rows_of(oracle.topn(t, [], keys, limit).to_host())
        got = rows_of(gpu_engine.topn(t, [], keys, limit).to_host())
        assert len(got) == len(want) == (n if limit < 0 else min(limit, n))
        if limit < 0 or limit >= n:
            assert got == want  # stable_sort: ties keep source order
            continue
        # limit < n: the cut is select_nth_unstable_by (sort.rs:118-124), so the
        # order inside a run of equal keys, and which rows of the last run
        # survive the cut, are unspecified. The key sequence is exact; inside a
        # run the GPU keeps source order (the row index is its last sort word).
        assert [key_of(r) for r in got] == [key_of(r) for r in want]
        i = 0
        while i < len(got):
            j = i
            while j < len(got) and key_of(got[j]) == key_of(got[i]):
                j += 1
            assert [r[4] for r in got[i:j]] == sorted(r[4] for r in got[i:j])
            if j < len(got):
                assert sorted(got[i:j]) == sorted(want[i:j])
            i = j


def _split_nan_rows(t):
    rows = rows_of(t)
    import math
    return ([r for r in rows if not (isinstance(r[0], float) and math.isnan(r[0]))],
            [r for r in rows if isinstance(r[0], float) and math.isnan(r[0])])


@pytest.mark.parametrize("sizing", ["plain", "partitioned"])
def test_hashagg_nan_and_negative_zero_group_keys(gpu_engine, oracle_engine, sizing):
    """values_equal (kernels/hash.rs:84-101): each NaN key row is a group of its
    own, -0.0 groups with +0.0; the rest of the result equals the oracle's."""
    import math
    rng = np.random.default_rng(44)
    n = 200_000
    pool = np.array([1.5, -0.0, 0.0, np.nan, 2.5, -7.25, np.inf, -np.inf])
    k = pool[rng.integers(0, len(pool), n)]
    k[rng.integers(0, n, 5000)] = rng.normal(0, 1e6, 5000)  # many distinct keys too
    v = rng.integers(-100, 100, n)
    g = rng.integers(0, 3, n)
    t = HostTable([HostColumn(FLOAT64, k), HostColumn(INT64, v), HostColumn(INT64, g)])
    aggs = [Agg(AGG_SUM, col(1), INT64), count_star_agg()]
    want = oracle_engine.hashagg(t, [], [col(0), col(2)], [FLOAT64, INT64], aggs).to_host()
    kw = {} if sizing == "plain" else dict(sizing=abi.SIZING_HINT, capacity_hint=200_000)
    got = gpu_engine.hashagg(t, [], [col(0), col(2)], [FLOAT64, INT64], aggs, **kw).to_host()
    assert got.num_rows == want.num_rows
    gn, gnan = _split_nan_rows(got)
    wn, wnan = _split_nan_rows(want)
    assert len(gnan) == int(np.isnan(k).sum()) == len(wnan)
    assert sorted(r[1:] for r in gnan) == sorted(r[1:] for r in wnan)
    # every NaN key is emitted as the canonical quiet NaN
    kc = got.columns[0].values.view(np.uint64)
    assert set(kc[np.isnan(got.columns[0].values)].tolist()) == {0x7FF8000000000000}
    norm = lambda rows: sorted((0.0 if r[0] == 0 else r[0],) + tuple(r[1:]) for r in rows)
    assert norm(gn) == norm(wn)
    assert not any(math.copysign(1.0, r[0]) < 0 and r[0] == 0 for r in gn)  # one zero group per g: +0.0


def test_join_nan_keys_never_match(gpu_engine, oracle_engine):
    """A NaN join key matches nothing, -0.0 matches +0.0 (join.rs:283-300 with
    values_equal), in the probe operator and inside a fused pipeline."""
    rng = np.random.default_rng(45)
    nb, npr = 3000, 20000
    pool = np.array([np.nan, -0.0, 0.0, 1.0, 2.0, 3.5])
    bk = np.concatenate([pool, rng.normal(0, 1e3, nb - len(pool))])
    bk[rng.integers(0, nb, 100)] = np.nan
    pk = np.where(rng.random(npr) < 0.5, bk[rng.integers(0, nb, npr)], pool[rng.integers(0, len(pool), npr)])
    build = HostTable([HostColumn(FLOAT64, bk), HostColumn(INT32, np.arange(nb, dtype=np.int32))])
    probe = HostTable([HostColumn(FLOAT64, pk), HostColumn(INT64, rng.integers(0, 10, npr))])
    jo = oracle_engine.join_build(build, [], [col(0)], [FLOAT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [FLOAT64])
    po, bo = oracle_engine.join_probe(jo, [col(0)], probe)
    pg, bg = gpu_engine.join_probe(jg, [col(0)], probe)
    ids_g = jg.column(1)
    assert sorted(zip(pg.tolist(), ids_g[bg].tolist())) == sorted(zip(po.tolist(), bo.tolist()))
    assert not np.isnan(pk[pg]).any()
    aggs = [Agg(AGG_SUM, col(1), INT64), count_star_agg()]
    want = oracle_engine.hashagg(probe, [abi.probe_op(jo, [col(0)])], [col(3)], [INT32], aggs).to_host()
    got = gpu_engine.hashagg(probe, [abi.probe_op(jg, [col(0)])], [col(3)], [INT32], aggs).to_host()
    assert_tables_equal(got, want)


@pytest.mark.parametrize("asc0,nf0,asc1,nf1", [(True, False, True, False), (False, True, True, True),
                                               (True, True, False, False), (False, False, False, True)])
def test_sort_total_order_nan_zero_strings_nulls(gpu_engine, asc0, nf0, asc1, nf1):
    """The full sort and TopN against a brute-force sort: f64 in total_cmp
    order (NaN, -0.0, infinities), a dictionary string in byte order, NULLS
    FIRST/LAST, ties kept in source order (sort.rs:43-132)."""
    import math
    from tests.test_oracle import _brute_sort, _f64_rank, _sort_table
    t = _sort_table(n=3000, seed=62)
    keys = [(col(0), asc0, nf0), (col(1), asc1, nf1)]
    want = _brute_sort(rows_of(t), [(0, asc0, nf0), (1, asc1, nf1)])
    got = rows_of(gpu_engine.topn(t, [], keys, -1).to_host())
    assert [r[3] for r in got] == [r[3] for r in want]
    kf = lambda r: tuple(None if v is None else (_f64_rank(v) if isinstance(v, float) else v) for v in r[:2])
    for limit in (10, 200):
        top = rows_of(gpu_engine.topn(t, [], keys, limit).to_host())
        assert [kf(r) for r in top] == [kf(r) for r in want[:limit]]


def test_collect_sized_expanding_probe_matches_oracle(gpu_engine, oracle_engine):
    """ml_pipeline_collect_sized over a probe with duplicate build keys: a hint
    below the true row count falls back to the exact (counted) collect, a hint
    at or above it is used as is; the rows equal the oracle's collect either way
    (learned sizing of a join output, join.rs:283-300 emits every match)."""
    from minilake_amd.abi import probe_op
    rng = np.random.default_rng(12)
    bkeys = rng.integers(0, 2000, 6000)  # about 3 rows per key
    build = HostTable([HostColumn(INT64, bkeys), HostColumn(INT32, np.arange(6000, dtype=np.int32))])
    probe = HostTable([HostColumn(INT64, rng.integers(-50, 2100, 50_000)),
                       HostColumn(FLOAT64, rng.normal(0, 1, 50_000))])
    jo = oracle_engine.join_build(build, [], [col(0)], [INT64])
    jg = gpu_engine.join_build(build, [], [col(0)], [INT64])
    ops_o = [abi.probe_op(jo, [col(0)])]
    ops_g = [abi.probe_op(jg, [col(0)])]
    want = oracle_engine.collect(probe, ops_o).to_host()
    exact = gpu_engine.collect(probe, ops_g).to_host()
    assert_tables_equal(exact, want, rtol=0)
    n = want.num_rows
    assert n > probe.num_rows  # the probe expands
    for hint, fallback in ((n // 3, True), (n - 1, True), (n, False), (2 * n, False)):
        got, fell_back = gpu_engine.collect_sized(probe, ops_g, hint)
        assert fell_back == fallback, hint
        assert_tables_equal(got.to_host(), want, rtol=0)


@pytest.mark.parametrize("case", ["hashed", "direct", "partitioned", "ungrouped"])
def test_string_min_max_matches_oracle(gpu_engine, oracle_engine, case):
    """MIN/MAX of a dictionary string (State::MinMaxStr, accumulator.rs:94,
    352-368): the device folds each entry's byte-order rank and maps the
    result back to an entry of the same dictionary. A hashed group table, a
    direct one (dictionary key), the partitioned LDS path (pre-sized, >65,536
    groups) and no keys; NULL arguments and an all-NULL group; Partial states
    of three slices merged by Final."""
    from tests.test_oracle import STR_AGGS, _str_agg_table
    t = _str_agg_table(n=400_000 if case == "partitioned" else 40_000, seed=43)
    if case == "partitioned":  # ~200k distinct keys
        t.columns[0].values[:] = np.random.default_rng(5).integers(0, 200_000, t.num_rows)
    keys, kt = {"hashed": ([col(0)], [INT64]), "direct": ([col(2)], [UTF8]), "partitioned": ([col(0)], [INT64]),
                "ungrouped": ([], [])}[case]
    kw = dict(sizing=abi.SIZING_HINT, capacity_hint=250_000) if case == "partitioned" else {}
    want = oracle_engine.hashagg(t, [], keys, kt, STR_AGGS).to_host()
    got = gpu_engine.hashagg(t, [], keys, kt, STR_AGGS, **kw).to_host()
    assert [c.type for c in got.columns] == [c.type for c in want.columns]
    assert_tables_equal(got, want, rtol=0)
    if case == "partitioned":
        return
    cuts = [0, 9_000, 9_001, t.num_rows]
    parts = []
    for a, b in zip(cuts, cuts[1:]):
        piece = t.take(np.arange(a, b))
        w = oracle_engine.hashagg(piece, [], keys, kt, STR_AGGS, mode=abi.AGG_MODE_PARTIAL).to_host()
        g = gpu_engine.hashagg(piece, [], keys, kt, STR_AGGS, mode=abi.AGG_MODE_PARTIAL).to_host()
        assert_tables_equal(g, w, rtol=0)
        parts.append(w)
    states = concat_tables(parts)
    got = gpu_engine.hashagg(states, [], keys, kt, STR_AGGS, mode=abi.AGG_MODE_FINAL).to_host()
    assert_tables_equal(got, want, rtol=0)
