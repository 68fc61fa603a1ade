"""Throughput of the operators around the star-join path on one MI355X:
full sort, Partial -> Final aggregation, MLB1 encoding, the device Parquet
scan and a high-cardinality GROUP BY with and without a memory budget (the
batched fragments are measured by bench.py at their full workload). Inputs are generated on the device (the Parquet file in host memory),
each operator runs `--iters` times after one warm-up, and one JSON line per
operator gives rows/s and the algorithmic bytes per second against the
8 TB/s HBM peak.

    python minilake_amd/bench_ops.py [--rows N] [--iters K]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _ROOT not in sys.path:  # run as a script
    sys.path.insert(0, _ROOT)

HBM_PEAK_GBS = 8000.0


def _timed(torch, fn, iters):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / iters, out


def _line(op, rows, sec, algo_bytes, **extra):
    gbs = algo_bytes / sec / 1e9
    rec = {"op": op, "rows": rows, "ms": round(sec * 1e3, 3), "rows_per_s": round(rows / sec),
           "algorithmic_bytes": int(algo_bytes), "achieved_gbs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
    rec.update(extra)
    print(json.dumps(rec), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=16_000_000)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--parquet-rows", type=int, default=4_000_000)
    a = ap.parse_args()
    import torch

    from minilake_amd import abi
    from minilake_amd.abi import AGG_AVG, AGG_COUNT_STAR, AGG_SUM, FLOAT64, INT32, Agg, col
    from minilake_amd.gpu import DeviceTable, GpuEngine, ParquetTable

    eng = GpuEngine()
    n = a.rows
    g = torch.Generator(device="cuda").manual_seed(7)
    f64 = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 1000
    i32 = torch.randint(0, 1000, (n,), generator=g, device="cuda", dtype=torch.int32)
    i64 = torch.randint(0, 1 << 40, (n,), generator=g, device="cuda", dtype=torch.int64)
    code = torch.randint(0, 7, (n,), generator=g, device="cuda", dtype=torch.int32)
    words = ["AIR", "MAIL", "RAIL", "SHIP", "TRUCK", "REG AIR", "FOB"]
    t = DeviceTable.from_tensors([f64, i32, i64, code], [FLOAT64, INT32, abi.INT64, abi.UTF8],
                                 dicts=[None, None, None, words], num_rows=n)
    torch.cuda.synchronize()

    # full sort: ORDER BY f64 DESC, i32 (every row out, every column gathered)
    keys = [(col(0), False, False), (col(1), True, False)]
    sec, _ = _timed(torch, lambda: eng.topn(t, [], keys, -1), a.iters)
    _line("full_sort", n, sec, n * (12 + 24), note="bytes: the two key columns read once, all 24 B/row written")

    # Partial then Final: GROUP BY i32 (1,000 groups), SUM / COUNT / AVG
    aggs = [Agg(AGG_SUM, col(0), FLOAT64), Agg(AGG_COUNT_STAR), Agg(AGG_AVG, col(0), FLOAT64)]

    def partial_final():
        p = eng.hashagg(t, [], [col(1)], [INT32], aggs, mode=abi.AGG_MODE_PARTIAL)
        return eng.hashagg(p, [], [col(0)], [INT32], aggs, mode=abi.AGG_MODE_FINAL)
    sec, _ = _timed(torch, partial_final, a.iters)
    _line("partial_final_agg", n, sec, n * 12, note="bytes: the key and argument columns read once")

    # MLB1 encoding of the whole batch into device memory
    from minilake_amd import gpu as G
    import ctypes as C
    keep = abi.Keep()
    b = t.ml_batch(keep)
    size = C.c_int64()
    L = G.lib()
    eng._check(L.ml_mlb1_write(eng.ctx, C.byref(b), None, 0, C.byref(size)))
    buf = torch.empty(size.value, dtype=torch.uint8, device="cuda")

    def mlb1():
        eng._check(L.ml_mlb1_write(eng.ctx, C.byref(b), buf.data_ptr(), buf.numel(), C.byref(size)))
    sec, _ = _timed(torch, mlb1, a.iters)
    _line("mlb1_write", n, sec, n * 24 + size.value, mlb1_bytes=size.value,
          note="bytes: the columns read once plus the MLB1 bytes written")

    # device Parquet scan of a file in host memory (UNCOMPRESSED, dictionary pages)
    import io

    import pyarrow as pa
    import pyarrow.parquet as pq
    m = a.parquet_rows
    rng = np.random.default_rng(3)
    tab = pa.table({"k": pa.array(rng.integers(0, 1 << 40, m)), "v": pa.array(rng.random(m)),
                    "d": pa.array(rng.integers(8000, 9000, m).astype(np.int32)).cast(pa.date32()),
                    "s": pa.array([words[x] for x in rng.integers(0, 7, m)])})
    bio = io.BytesIO()
    pq.write_table(tab, bio, compression="NONE", row_group_size=1 << 20)
    data = bio.getvalue()
    pt = ParquetTable(data, eng)

    def scan():
        return [pt.read_row_group(rg) for rg in range(pt.num_row_groups)]
    sec, _ = _timed(torch, scan, a.iters)
    _line("parquet_scan", m, sec, len(data) + m * 24, file_bytes=len(data),
          note="host file -> device columns; includes the pageable host-to-device copy of every chunk")
    del pt, data, bio, tab

    # GROUP BY a high-cardinality i64 key (n/8 groups), unlimited and under a
    # memory budget of about a quarter of the group table (spill to host in 16
    # partitions, merged one partition at a time); 256K-row device batches, so
    # that one batch fits the budget
    hk = torch.randint(0, n // 8, (n,), generator=g, device="cuda", dtype=torch.int64)
    ht = DeviceTable.from_tensors([hk, f64], [abi.INT64, FLOAT64], num_rows=n)
    hagg = [Agg(AGG_SUM, col(1), FLOAT64), Agg(AGG_COUNT_STAR)]
    torch.cuda.synchronize()
    spill_eng = GpuEngine(batch_rows=1 << 18)
    try:
        sec_u, res = _timed(torch, lambda: spill_eng.hashagg(ht, [], [col(0)], [abi.INT64], hagg), a.iters)
        groups = res.num_rows
        del res
        limit = max(32 << 20, groups * 64 // 4)
        spill_eng.set_memory_limit(limit, spill=True)
        sec_s, res_s = _timed(torch, lambda: spill_eng.hashagg(ht, [], [col(0)], [abi.INT64], hagg), a.iters)
        assert res_s.num_rows == groups
        del res_s  # results are freed before their context
        peak = spill_eng.memory_usage()["peak"]
        assert peak <= limit, (peak, limit)
    finally:
        spill_eng.close()
    _line("hashagg_unlimited", n, sec_u, n * 16, groups=groups, note="bytes: the key and argument columns read once")
    _line("hashagg_spill", n, sec_s, n * 16, groups=groups, memory_limit=limit, peak_reserved=peak,
          slowdown=round(sec_s / sec_u, 2), note="same GROUP BY under a budget of ~1/4 of its table, spilling to host")
    del ht, hk



if __name__ == "__main__":
    main()
