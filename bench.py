"""Benchmark: the 60M-row six-table star join + aggregate (TPC-H Q5 shape at 10x).

One step = the whole query on one MI355X with every input table resident in
HBM: five join builds (customer, orders filtered by a one-year range and
probed against customer, supplier with a two-column key, nation, region
filtered to one name), the 60M-row lineitem pipeline probing all four tables
into the nation-name aggregate, and the final ORDER BY revenue.

Prints ONE JSON line:
  value      input rows/s over the whole query (all six tables' rows)
  roofline   the dominant kernel (the lineitem pipeline): algorithmic bytes
             (lineitem columns read once at their Column widths) / its average
             launch time from HIP events on its stream, against 8 TB/s
  cpu_baseline  the C++ restatement of MiniLake's CPU executor on the same
             query at 1/10 scale (6M lineitem rows), all host threads
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "flow_kernel_star_join_q5.md")


def profiled_traffic():
    """HBM bytes per launch of the lineitem kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=float, default=10.0, help="TPC-H-like scale; 10 = 60M lineitem rows")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bloom", type=int, default=-1)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain sizing comparison")
    ap.add_argument("--learned-rows", type=int, default=60_000_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the star-join benchmark runs on one GPU; sharding over several GPUs "
                                   "(replicated build sides, row-sharded fact table, a gather of the partial "
                                   "aggregates) is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    import ctypes as C
    import numpy as np
    import torch

    from minilake_amd import abi, queries as Q, workloads as W
    from minilake_amd.gpu import GpuEngine, lib

    t0 = time.time()
    sj = W.star_join(seed=args.seed, scale=args.scale)
    log(f"generated {sj.input_rows():,} rows in {time.time() - t0:.1f}s")
    eng = GpuEngine(device=0, use_bloom=args.bloom)
    dev = W.StarJoin(*[eng.table(t) for t in sj.tables()])
    torch.cuda.synchronize()
    input_rows = sj.input_rows()
    fact_rows = sj.lineitem.num_rows
    fact_bytes = sj.lineitem.nbytes()  # l_orderkey i64, l_suppkey i64, price f64, disc f64 = 32 B/row

    L = lib()
    L.ml_ctx_enable_stats.argtypes = [C.c_void_p, C.c_int32]
    L.ml_ctx_enable_stats.restype = C.c_int32

    class Stat(C.Structure):
        _fields_ = [("sink", C.c_int32), ("_pad", C.c_int32), ("rows", C.c_int64), ("ms", C.c_double)]

    L.ml_ctx_take_stats.argtypes = [C.c_void_p, C.POINTER(Stat), C.c_int32]
    L.ml_ctx_take_stats.restype = C.c_int32

    for _ in range(args.warmup):
        top, _ = Q.run_q5(eng, dev)
    torch.cuda.synchronize()
    L.ml_ctx_enable_stats(eng.ctx, fact_rows)  # HIP events around the lineitem launches only
    stats_buf = (Stat * 4096)()
    L.ml_ctx_take_stats(eng.ctx, stats_buf, 4096)  # clear
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(eng.stream)
    t_wall = time.time()
    for _ in range(args.steps):
        top, _ = Q.run_q5(eng, dev)
    stop.record(eng.stream)
    torch.cuda.synchronize()
    wall = time.time() - t_wall
    ms_total = start.elapsed_time(stop)
    ms_per_step = ms_total / args.steps
    n = L.ml_ctx_take_stats(eng.ctx, stats_buf, 4096)
    L.ml_ctx_enable_stats(eng.ctx, 0)
    fact_ms = [stats_buf[i].ms for i in range(n) if stats_buf[i].rows == fact_rows]
    fact_ms_avg = sum(fact_ms) / max(1, len(fact_ms))
    achieved = fact_bytes / (fact_ms_avg * 1e-3) / 1e9 if fact_ms else None
    value = input_rows / (ms_per_step * 1e-3)
    result = top.to_host()
    log(f"{args.steps} steps: {ms_per_step:.3f} ms/step (wall {wall / args.steps * 1e3:.3f} ms), "
        f"lineitem kernel {fact_ms_avg:.3f} ms over {len(fact_ms)} launches; groups={result.num_rows}")

    # measured stream-copy bandwidth on this GPU (read + write bytes / time)
    copy_bytes = 1 << 31
    src = torch.empty(copy_bytes // 8, dtype=torch.int64, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record(eng.stream)
    for _ in range(5):
        dst.copy_(src)
    c1.record(eng.stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * copy_bytes * 5 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst
    log(f"device copy {copy_gbs:.0f} GB/s")

    # learned vs plain hash-table sizing on the 60M-row Zipf(1.1) aggregate + join
    learned = None
    if not args.no_learned:
        from minilake_amd import sizing
        params = sizing.load_params()
        if params is None:
            learned = {"error": "no trained estimator weights (python -m minilake_amd.train_estimator)"}
        else:
            del dev
            torch.cuda.empty_cache()
            mlp = sizing.Mlp(eng, params)
            fact, build = sizing.zipf_workload(eng, rows=args.learned_rows)
            learned = sizing.compare_plain_learned(eng, mlp, fact, build, build_distinct=2_000_000, log=log)
            learned["workload"] = (f"{args.learned_rows:,} int64 keys Zipf(1.1) over 1e8, GROUP BY key COUNT(*), "
                                   f"SUM(key mod 1009); join with 2,000,000 keys x 2 rows")
            del fact, build, mlp

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import OracleEngine
        threads = max(1, min(16, os.cpu_count() or 1))

        def timed(oe, data, runs):
            Q.run_q5(oe, data)  # warm-up, as the reference's bench does (plan once, 1 warm-up, timed runs)
            ts = []
            for _ in range(runs):
                t = time.time()
                Q.run_q5(oe, data)
                ts.append(time.time() - t)
            return sorted(ts)[len(ts) // 2]

        med = timed(OracleEngine(threads=threads), sj, 5)
        small = W.star_join(seed=args.seed, scale=args.cpu_scale)
        med1 = timed(OracleEngine(threads=1), small, 3)
        cpu = {"value": input_rows / med, "unit": "input rows/s", "cores": threads, "kind": "port",
               "sample": f"the same 60M-row query on the same generated tables, median of 5 runs after 1 "
                         f"warm-up ({med:.3f} s), inputs pre-materialized in host RAM",
               "single_thread": {"value": small.input_rows() / med1, "cores": 1,
                                 "sample": f"scale {args.cpu_scale:g} ({small.input_rows():,} input rows), "
                                           f"median of 3 runs after 1 warm-up ({med1:.3f} s)"}}
        log(f"cpu baseline {cpu['value'] / 1e6:.1f} M rows/s on {threads} threads (median {med:.3f}s); "
            f"1 thread {cpu['single_thread']['value'] / 1e6:.1f} M rows/s")

    out = {
        "metric": "input rows/s & % HBM peak, 60M-row join+agg; learned-sized vs plain speedup",
        "value": value,
        "unit": "input rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64/f64",
        "data": "synthetic (seeded, counter-based splitmix64)",
        "config": {"workload": f"star join Q5-shape scale {args.scale:g}: lineitem {fact_rows:,}, orders "
                               f"{sj.orders.num_rows:,}, customer {sj.customer.num_rows:,}, supplier "
                               f"{sj.supplier.num_rows:,}, nation 25, region 5; GROUP BY n_name SUM(price*(1-disc))",
                   "input_rows": input_rows, "seed": args.seed},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": profiled_traffic(),
                     "traffic_source": "profiles/flow_kernel_star_join_q5.md (rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                       "per dispatch, scaled by a 2 GiB copy kernel's counts)",
                     "kernel": "flow_kernel<SINK_DIRECT_AGG, FLOW_FAST> (lineitem probe chain)",
                     "algorithmic_bytes": fact_bytes, "kernel_ms": fact_ms_avg,
                     "measured_copy_gbs": copy_gbs},
        "cpu_baseline": cpu,
        "learned_sizing": learned,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
