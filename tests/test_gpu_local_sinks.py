"""Per-thread sink states through the C ABI, driven the way MiniLake's
run_parallel drives LocalSinks (scheduler.rs:102-144, pipeline.rs:70-79):
three host threads, each with its own context, pull morsels (row ranges, as
selections) from one shared counter, sink them into their local state without
locking and combine once; the global sink is finalized after every combine.
Results are compared with the CPU oracle over the whole table: hash aggregate
(grouped, ungrouped, Partial mode, and a global that also takes batches
itself), join build (probed afterwards), TopN, and LIMIT with is_finished."""

import ctypes as C
import threading

import numpy as np
import pytest

from minilake_amd import abi
from minilake_amd.abi import AGG_AVG, AGG_COUNT, AGG_COUNT_STAR, AGG_MAX, AGG_MIN, AGG_SUM, FLOAT64, INT32, INT64, Agg, col, lit
from minilake_amd.gpu import GpuEngine, GpuError, lib
from minilake_amd.table import HostColumn, HostTable
from tests.compare import assert_tables_equal, rows_of

pytestmark = pytest.mark.gpu

THREADS = 3
MORSEL = 20_000


def table(n=150_000, seed=71):
    rng = np.random.default_rng(seed)
    return HostTable([HostColumn(INT64, rng.integers(0, 5_000, n)), HostColumn(FLOAT64, rng.normal(10, 3, n)),
                      HostColumn(INT32, rng.integers(-50, 50, n).astype(np.int32)),
                      HostColumn(INT64, np.arange(n, dtype=np.int64))])


def run_parallel(main, t, body):
    """body(eng, batch_for_morsel) per worker thread; morsels [i*MORSEL, ...) from a shared counter."""
    n = t.num_rows
    nmorsels = (n + MORSEL - 1) // MORSEL
    lock = threading.Lock()
    nxt = [0]
    errors = []
    dev = main.table(t)

    def worker():
        eng = GpuEngine()
        try:
            def morsels():
                while True:
                    with lock:
                        i = nxt[0]
                        nxt[0] += 1
                    if i >= nmorsels:
                        return
                    sel = np.arange(i * MORSEL, min(n, (i + 1) * MORSEL), dtype=np.uint32)
                    keep = abi.Keep()
                    _, b = eng._src(dev, keep, sel)
                    yield b, keep
            body(eng, morsels)
        except Exception as e:  # noqa: BLE001
            errors.append(e)
        finally:
            eng.synchronize()
            eng.close()

    ths = [threading.Thread(target=worker) for _ in range(THREADS)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    if errors:
        raise errors[0]


AGGS = [Agg(AGG_SUM, col(1), FLOAT64), Agg(AGG_AVG, col(2), INT32), Agg(AGG_MIN, col(1), FLOAT64),
        Agg(AGG_MAX, col(2), INT32), Agg(AGG_COUNT, col(1) / col(2).cast(FLOAT64), FLOAT64), Agg(AGG_COUNT_STAR)]
OPS = [abi.filter_op(col(2).ne(lit(7)))]


@pytest.mark.parametrize("case", ["grouped", "ungrouped", "partial", "global_sinks_too"])
def test_hashagg_locals_combine(gpu_engine, oracle_engine, case):
    L = lib()
    t = table()
    keys, kt = ([], []) if case == "ungrouped" else ([col(0)], [INT64])
    mode = abi.AGG_MODE_PARTIAL if case == "partial" else abi.AGG_MODE_SINGLE
    keep = abi.Keep()
    spec = abi.agg_spec(keys, kt, AGGS, keep, abi.SIZING_PLAIN, 0, mode)
    g = C.c_void_p()
    gpu_engine._check(L.ml_hashagg_create(gpu_engine.ctx, C.byref(spec), C.byref(g)))
    try:
        def body(eng, morsels):
            loc = C.c_void_p()
            eng._check(L.ml_hashagg_create_local(g, eng.ctx, C.byref(loc)))
            for b, k in morsels():
                eng._check(L.ml_pipeline_hashagg(eng.ctx, C.byref(b), eng._ops(OPS, k), len(OPS), loc))
            eng._check(L.ml_hashagg_combine(g, loc))
        if case == "global_sinks_too":  # the global takes the first half itself; the threads the rest
            half = t.num_rows // 2
            first, rest = t.take(np.arange(half)), t.take(np.arange(half, t.num_rows))
            keep2 = abi.Keep()
            _, b = gpu_engine._src(first, keep2)
            gpu_engine._check(L.ml_pipeline_hashagg(gpu_engine.ctx, C.byref(b), gpu_engine._ops(OPS, keep2), len(OPS), g))
            run_parallel(gpu_engine, rest, body)
        else:
            run_parallel(gpu_engine, t, body)
        out = C.c_void_p()
        gpu_engine._check(L.ml_hashagg_finalize(g, C.byref(out)))
        from minilake_amd.gpu import GpuResult
        got = GpuResult(gpu_engine, out.value).to_host()
    finally:
        L.ml_hashagg_free(g)
    want = oracle_engine.hashagg(t, OPS, keys, kt, AGGS, mode=mode).to_host()
    assert_tables_equal(got, want, rtol=1e-12)


def test_join_build_locals_combine(gpu_engine, oracle_engine):
    L = lib()
    build = table(60_000, seed=72)
    keep = abi.Keep()
    spec = abi.join_build_spec([col(0)], [INT64], keep)
    g = C.c_void_p()
    gpu_engine._check(L.ml_join_build_create(gpu_engine.ctx, C.byref(spec), C.byref(g)))

    def body(eng, morsels):
        loc = C.c_void_p()
        eng._check(L.ml_join_build_create_local(g, eng.ctx, C.byref(loc)))
        for b, k in morsels():
            eng._check(L.ml_pipeline_join_build(eng.ctx, C.byref(b), eng._ops(OPS, k), len(OPS), loc))
        eng._check(L.ml_join_build_combine(g, loc))
    run_parallel(gpu_engine, build, body)
    tbl = C.c_void_p()
    gpu_engine._check(L.ml_join_build_finalize(g, C.byref(tbl)))
    from minilake_amd.gpu import GpuJoin
    jg = GpuJoin(tbl.value)
    probe = table(40_000, seed=73)
    from oracle.oracle import OracleEngine
    jo = OracleEngine(threads=1).join_build(build, OPS, [col(0)], [INT64])  # build rows in source order
    assert jg.num_rows() == int((build.columns[2].values != 7).sum())
    po, bo = OracleEngine(threads=1).join_probe(jo, [col(0)], probe)
    pg, bg = gpu_engine.join_probe(jg, [col(0)], probe)
    # build rows are table-internal on both sides: compare (probe row, build row id) pairs;
    # the oracle's build rows are the filtered rows in source order
    got = sorted(zip(pg.tolist(), jg.column(3)[bg].tolist()))
    oracle_ids = build.columns[3].values[build.columns[2].values != 7]
    want = sorted(zip(po.tolist(), oracle_ids[bo].tolist()))
    assert got == want and len(got) > 0


def test_topn_locals_combine(gpu_engine, oracle_engine):
    L = lib()
    t = table(120_000, seed=74)
    keys = [(col(1), False, False), (col(3), True, False)]
    for limit in (25, -1):
        keep = abi.Keep()
        spec = abi.sort_spec(keys, limit, keep)
        g = C.c_void_p()
        gpu_engine._check(L.ml_topn_create(gpu_engine.ctx, C.byref(spec), C.byref(g)))
        try:
            def body(eng, morsels):
                loc = C.c_void_p()
                eng._check(L.ml_topn_create_local(g, eng.ctx, C.byref(loc)))
                for b, k in morsels():
                    eng._check(L.ml_pipeline_topn(eng.ctx, C.byref(b), eng._ops(OPS, k), len(OPS), loc))
                eng._check(L.ml_topn_combine(g, loc))
            run_parallel(gpu_engine, t, body)
            out = C.c_void_p()
            gpu_engine._check(L.ml_topn_finalize(g, C.byref(out)))
            from minilake_amd.gpu import GpuResult
            got = rows_of(GpuResult(gpu_engine, out.value).to_host())
        finally:
            L.ml_topn_free(g)
        want = rows_of(oracle_engine.topn(t, OPS, keys, limit).to_host())
        assert got == want  # the row id breaks every tie


def test_limit_sink_shared_and_finished(gpu_engine):
    L = lib()
    t = table(150_000, seed=75)
    keep0 = abi.Keep()
    for limit, offset in ((1_000, 500), (10**9, 0)):
        s = C.c_void_p()
        gpu_engine._check(L.ml_limit_create(gpu_engine.ctx, limit, offset, C.byref(s)))
        seen = []
        try:
            def body(eng, morsels):
                for b, k in morsels():
                    if L.ml_limit_is_finished(s):  # Pipeline::run_worker stops early (pipeline.rs:140-147)
                        seen.append(1)
                        continue
                    eng._check(L.ml_limit_sink(s, eng.ctx, C.byref(b)))
            run_parallel(gpu_engine, t, body)
            out = C.c_void_p()
            gpu_engine._check(L.ml_limit_finalize(s, C.byref(out)))
            from minilake_amd.gpu import GpuResult
            got = GpuResult(gpu_engine, out.value).to_host()
            finished = L.ml_limit_is_finished(s)
        finally:
            L.ml_limit_free(s)
        ids = got.columns[3].values
        if limit == 1_000:
            assert finished == 1 and got.num_rows == 1_000 and len(seen) > 0
        else:
            assert finished == 0 and got.num_rows == t.num_rows
        assert len(np.unique(ids)) == len(ids) and ids.min() >= 0 and ids.max() < t.num_rows
        # each kept row is the source row its id names
        np.testing.assert_array_equal(got.columns[1].values, t.columns[1].values[ids])
