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
             query and tables, on the host threads this process may use
  parity     the GPU result against that CPU result (names exact, sums 1e-12)
  filter_groupby, join_topn   the other two query shapes (6,001,215-row
             filter + GROUP BY; the SF1 two-way join + GROUP BY + top-10):
             rows/s and their dominant kernel's roofline
  fragments  256 batched join+aggregate fragments: launches per fragment,
             one persistent launch sequence, a replayed hipGraph, default and
             autotuned tunings
  learned_sizing  learned vs plain hash-table sizing on the 60M-row Zipf(1.1)
             aggregate + join
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


def host_threads():
    """Threads for the CPU baseline: the cores this process may run on, capped
    by the host's declared thread budget (OMP_NUM_THREADS, 16 on a one-GPU
    share of the box); plus the counts they come from."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    return max(1, min(affinity, budget)), {"nproc": nproc, "affinity_cores": affinity,
                                           "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


class Stats:
    """ml_ctx_enable_stats / ml_ctx_take_stats: HIP events around the launches
    of fused pipelines over at least `min_rows` source rows."""

    def __init__(self, eng):
        import ctypes as C
        from minilake_amd.gpu import lib
        self.C, self.eng, self.L = C, eng, lib()

        class Stat(C.Structure):
            _fields_ = [("sink", C.c_int32), ("_pad", C.c_int32), ("rows", C.c_int64), ("ms", C.c_double)]

        self.L.ml_ctx_enable_stats.argtypes = [C.c_void_p, C.c_int32]
        self.L.ml_ctx_enable_stats.restype = C.c_int32
        self.L.ml_ctx_take_stats.argtypes = [C.c_void_p, C.POINTER(Stat), C.c_int32]
        self.L.ml_ctx_take_stats.restype = C.c_int32
        self.buf = (Stat * 8192)()

    def start(self, min_rows):
        self.L.ml_ctx_enable_stats(self.eng.ctx, int(min_rows))
        self.L.ml_ctx_take_stats(self.eng.ctx, self.buf, 8192)

    def take(self):
        n = self.L.ml_ctx_take_stats(self.eng.ctx, self.buf, 8192)
        self.L.ml_ctx_enable_stats(self.eng.ctx, 0)
        return [(self.buf[i].rows, self.buf[i].ms) for i in range(n)]


def time_steps(eng, torch, fn, steps, warmup, stats=None, min_rows=None):
    """ms per step over `steps` timed calls of fn (HIP events on the engine's
    stream, synchronized on both sides), and the recorded kernel launches."""
    for _ in range(warmup):
        out = fn()
    torch.cuda.synchronize()
    if stats is not None:
        stats.start(min_rows)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(eng.stream)
    for _ in range(steps):
        out = fn()
    b.record(eng.stream)
    torch.cuda.synchronize()
    launches = stats.take() if stats is not None else []
    return a.elapsed_time(b) / steps, launches, out


def kernel_avg(launches, rows):
    ms = [m for r, m in launches if r == rows]
    return (sum(ms) / len(ms), len(ms)) if ms else (None, 0)


def roofline(bytes_, ms):
    if not ms:
        return None
    gbs = bytes_ / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "algorithmic_bytes": bytes_, "kernel_ms": ms}


def same_q5(got, want, rtol=1e-12):
    """Nation names exact, revenues within rtol (order-independent)."""
    from minilake_amd.table import HostTable  # noqa: F401
    g = sorted(zip(got.columns[0].decoded(), got.columns[1].decoded()))
    w = sorted(zip(want.columns[0].decoded(), want.columns[1].decoded()))
    if len(g) != len(w):
        return False
    return all(a[0] == b[0] and abs(a[1] - b[1]) <= rtol * max(abs(a[1]), abs(b[1])) for a, b in zip(g, w))


def bench_filter_groupby(eng, torch, W, Q, stats, steps, warmup, seed):
    li = W.filter_groupby(seed=seed)
    dev = eng.table(li)
    n = li.num_rows
    nbytes = li.nbytes()  # 40 B/row: qty i32, price/disc/tax f64, two dict codes u32, date i32
    ms, launches, res = time_steps(eng, torch, lambda: Q.run_q1(eng, dev), steps, warmup, stats, n)
    kms, nk = kernel_avg(launches, n)
    groups = res.to_host().num_rows
    log(f"filter_groupby: {ms:.3f} ms/step, kernel {kms:.3f} ms ({nk} launches), {groups} groups")
    out = {"workload": f"filter + GROUP BY (flag1, flag2), {n:,} rows, 8 aggregates", "rows": n,
           "ms_per_step": ms, "rows_per_s": n / (ms * 1e-3), "groups": groups,
           "roofline": roofline(nbytes, kms), "step_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel": "flow_kernel<SINK_DIRECT_AGG, FLOW_GENERIC> (filter atoms + LDS cells)"}
    del dev
    return out


def bench_join_topn(eng, torch, W, Q, stats, steps, warmup, seed):
    jt = W.join_topn(seed=seed)
    dev = W.JoinTopN(eng.table(jt.customer), eng.table(jt.orders), eng.table(jt.lineitem))
    rows, nbytes = jt.input_rows(), jt.input_bytes()
    nl = jt.lineitem.num_rows
    ms, launches, res = time_steps(eng, torch, lambda: Q.run_q3(eng, dev.customer, dev.orders, dev.lineitem)[0],
                                   steps, warmup, stats, 1)
    # the pipeline kernels of a step: customer and orders builds, the lineitem
    # aggregate (one launch per 2^21-row device batch), the TopN
    kms = sum(m for _, m in launches) / steps
    log(f"join_topn: {ms:.3f} ms/step, pipeline kernels {kms:.3f} ms/step ({len(launches) // steps} launches)")
    out = {"workload": f"customer {jt.customer.num_rows:,} -> orders {jt.orders.num_rows:,} -> lineitem {nl:,}; "
                       "GROUP BY (orderkey, orderdate, shippriority) SUM(price*(1-disc)), top 10",
           "rows": rows, "ms_per_step": ms, "rows_per_s": rows / (ms * 1e-3),
           "step_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "step_bytes": nbytes,
           "roofline": roofline(nbytes, kms),
           "kernel": "the step's pipeline kernels (builds, lineitem filter + orders probe into the group table, TopN)"}
    del dev
    return out


def bench_fragments(eng, torch, seed, reps=5):
    import numpy as np
    from minilake_amd import fragments as FR, workloads as W

    shapes = W.fragment_shapes(seed=seed, n=256)
    frags = FR.generate(eng, shapes, seed=seed)
    probe_rows = sum(s.probe_rows for s in shapes)
    build_rows = sum(s.build_rows for s in shapes)
    nbytes = FR.algorithmic_bytes(frags)

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()  # every run waits for its kernels (it reads back the error words)
            ts.append(time.perf_counter() - t)
        return sorted(ts)[len(ts) // 2] * 1e3

    fs = FR.FragmentSet(eng, frags)
    try:
        times = {}
        for mode in FR.MODES:
            times[mode + "_default_ms"] = med(lambda: fs.run(mode))
        ref = fs.results()
        t0 = time.perf_counter()
        best = fs.autotune()
        tune_s = time.perf_counter() - t0
        for mode in FR.MODES:
            times[mode + "_ms"] = med(lambda: fs.run(mode))
        got = fs.results()
        same = all(np.array_equal(gc, rc) and np.allclose(gs, rs, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(rs).max()))
                   for (gs, gc), (rs, rc) in zip(got, ref))
    finally:
        fs.close()
    best_ms = min(times["persistent_ms"], times["graph_ms"])
    out = {"workload": "256 fragments SELECT g, SUM(v), COUNT(*) FROM probe JOIN build GROUP BY g: probe rows "
                       "floor(2^U[10,24]), build rows probe/U[4,40], groups floor(2^U[0,20]), Zipf s in {0,0.8,1.2}",
           "fragments": len(shapes), "probe_rows": probe_rows, "build_rows": build_rows, **times,
           "tuning": {k: getattr(best, k) for k, _ in best._fields_}, "autotune_s": tune_s,
           "speedup_vs_per_fragment": times["per_fragment_ms"] / best_ms,
           "rows_per_s": (probe_rows + build_rows) / (best_ms * 1e-3),
           "roofline": roofline(nbytes, best_ms), "modes_agree": bool(same),
           "oracle_parity": "tests/test_gpu_fragments.py::test_batched_fragment_workload_matches_oracle"}
    log(f"fragments: {probe_rows:,} probe rows; per-fragment {times['per_fragment_ms']:.2f} ms, persistent "
        f"{times['persistent_ms']:.2f} ms, graph {times['graph_ms']:.2f} ms (default {times['persistent_default_ms']:.2f}), "
        f"autotune {tune_s:.1f}s")
    del frags
    return out


# Key distributions for the sizing ablation: (rows, Zipf s (0 = uniform), key domain).
# The first is the learned-sizing workload; in the others MiniLake's input/10
# estimate is far off (too high, too low, too high, too low).
ABLATION_WORKLOADS = [(1.0, 1.1, 100_000_000), (1.0, 0.0, 1_000_000), (0.5, 0.6, 1_000_000_000),
                      (1.0, 1.4, 100_000_000), (1.0, 0.0, 100_000_000)]


def bench_sizing_ablation(eng, torch, sizing, mlp, rows):
    """Each sizing estimator (MiniLake's planner heuristic, the sample's GEE and
    raw KMV estimates, the MLP, the exact counts) pre-sizing the same aggregate
    and join on the same path, over key distributions where input/10 is right
    and where it is far off."""
    out = []
    for frac, s, domain in ABLATION_WORKLOADS:
        n = int(rows * frac)
        nb = min(2_000_000, domain // 4)  # distinct build keys drawn from the same domain
        fact, build = sizing.zipf_workload(eng, rows=n, s=s, domain=domain, build_keys=nb)
        r = sizing.ablation(eng, mlp, fact, build, build_distinct=nb)
        r["workload"] = (f"{n:,} int64 keys {'uniform' if s == 0 else f'Zipf({s:g})'} over {domain:.0e}; "
                         f"join with {nb:,} keys x 2 rows")
        legs = r["legs"]
        log(f"ablation {r['workload']}: {r['true_groups']:,} groups; " +
            ", ".join(f"{k} {v['ms']:.1f} ms (x{v['groups_ratio']:.2f})" for k, v in legs.items()))
        out.append(r)
        del fact, build
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=10.0, help="TPC-H-like scale; 10 = 60M lineitem rows")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bloom", type=int, default=-1)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain sizing comparison")
    ap.add_argument("--no-extra", action="store_true", help="only the star join (no other query shapes, fragments)")
    ap.add_argument("--extra", default="filter_groupby,join_topn,fragments",
                    help="comma-separated extra workloads to run after the star join")
    ap.add_argument("--learned-rows", type=int, default=60_000_000)
    ap.add_argument("--no-ablation", action="store_true", help="skip the sizing-estimator ablation")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the star-join benchmark runs on one GPU; sharding over several GPUs "
                                   "(replicated build sides, row-sharded fact table, a gather of the partial "
                                   "aggregates) is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from minilake_amd import queries as Q, workloads as W
    from minilake_amd.gpu import GpuEngine

    t0 = time.time()
    sj = W.star_join(seed=args.seed, scale=args.scale)
    log(f"generated {sj.input_rows():,} rows in {time.time() - t0:.1f}s")
    eng = GpuEngine(device=0, use_bloom=args.bloom)
    stats = Stats(eng)
    dev = W.StarJoin(*[eng.table(t) for t in sj.tables()])
    torch.cuda.synchronize()
    input_rows = sj.input_rows()
    fact_rows = sj.lineitem.num_rows
    fact_bytes = sj.lineitem.nbytes()  # l_orderkey i64, l_suppkey i64, price f64, disc f64 = 32 B/row

    t_wall = time.time()
    ms_per_step, launches, top = time_steps(eng, torch, lambda: Q.run_q5(eng, dev)[0], args.steps, args.warmup,
                                            stats, fact_rows)
    wall = time.time() - t_wall
    fact_ms_avg, nfact = kernel_avg(launches, fact_rows)
    achieved = fact_bytes / (fact_ms_avg * 1e-3) / 1e9 if fact_ms_avg else None
    value = input_rows / (ms_per_step * 1e-3)
    result = top.to_host()
    log(f"{args.steps} steps: {ms_per_step:.3f} ms/step (wall incl. warm-up {wall:.2f} s), "
        f"lineitem kernel {fact_ms_avg:.3f} ms over {nfact} launches; groups={result.num_rows}")

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
    del dev
    torch.cuda.empty_cache()

    extra = {}
    wanted = set() if args.no_extra else {x for x in args.extra.split(",") if x}
    if "filter_groupby" in wanted:
        extra["filter_groupby"] = bench_filter_groupby(eng, torch, W, Q, stats, args.steps, args.warmup, args.seed)
    if "join_topn" in wanted:
        extra["join_topn"] = bench_join_topn(eng, torch, W, Q, stats, args.steps, args.warmup, args.seed)
        torch.cuda.empty_cache()
    if "fragments" in wanted:
        extra["fragments"] = bench_fragments(eng, torch, args.seed)
        torch.cuda.empty_cache()

    # learned vs plain hash-table sizing on the 60M-row Zipf(1.1) aggregate + join
    learned = None
    if not args.no_learned:
        from minilake_amd import sizing
        params = sizing.load_params()
        if params is None:
            learned = {"error": "no trained estimator weights (python -m minilake_amd.train_estimator)"}
        else:
            mlp = sizing.Mlp(eng, params)
            fact, build = sizing.zipf_workload(eng, rows=args.learned_rows)
            learned = sizing.compare_plain_learned(eng, mlp, fact, build, build_distinct=2_000_000, log=log)
            learned["workload"] = (f"{args.learned_rows:,} int64 keys Zipf(1.1) over 1e8, GROUP BY key COUNT(*), "
                                   f"SUM(key mod 1009); join with 2,000,000 keys x 2 rows")
            del fact, build
            torch.cuda.empty_cache()
            if not args.no_ablation:
                learned["ablation"] = bench_sizing_ablation(eng, torch, sizing, mlp, args.learned_rows)
            del mlp

    cpu, parity = None, None
    if not args.no_cpu:
        from oracle.oracle import OracleEngine
        threads, host = host_threads()

        def timed(oe, data, runs):
            out = Q.run_q5(oe, data)[0]  # warm-up, as the reference's bench does (plan once, 1 warm-up, timed runs)
            ts = []
            for _ in range(runs):
                t = time.time()
                Q.run_q5(oe, data)
                ts.append(time.time() - t)
            return sorted(ts)[len(ts) // 2], out

        med, want = timed(OracleEngine(threads=threads), sj, 5)
        parity = same_q5(result, want.to_host())
        small = W.star_join(seed=args.seed, scale=args.cpu_scale)
        med1, _ = timed(OracleEngine(threads=1), small, 3)
        cpu = {"value": input_rows / med, "unit": "input rows/s", "cores": threads, "kind": "port",
               "sample": f"the same 60M-row query on the same generated tables, median of 5 runs after 1 "
                         f"warm-up ({med:.3f} s), inputs pre-materialized in host RAM; {threads} threads = the "
                         f"cores this process may use within the host's thread budget",
               "host": host,
               "single_thread": {"value": small.input_rows() / med1, "cores": 1,
                                 "sample": f"scale {args.cpu_scale:g} ({small.input_rows():,} input rows), "
                                           f"median of 3 runs after 1 warm-up ({med1:.3f} s)"}}
        log(f"cpu baseline {cpu['value'] / 1e6:.1f} M rows/s on {threads} threads (median {med:.3f}s, host {host}); "
            f"1 thread {cpu['single_thread']['value'] / 1e6:.1f} M rows/s; parity {parity}")

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
        "parity": parity,
        **extra,
        "learned_sizing": learned,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
