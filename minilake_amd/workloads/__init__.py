"""Seeded synthetic tables for the benchmark and test workloads.

Each generator is deterministic in ``seed`` (counter-based splitmix64, see
``gen.cpp``). Shapes follow TPC-H's lineitem/orders/customer/supplier/nation/
region at the stated scales, restricted to the columns each query reads.

Workloads:
  * ``filter_groupby``   — 6,001,215-row filter + low-cardinality GROUP BY (TPC-H Q1 shape)
  * ``join_topn``        — 2-way join + high-cardinality GROUP BY + top-10 (Q3 shape, SF1)
  * ``star_join``        — 6-table star join at 10x (Q5 shape, 60M fact rows)
  * ``fragments``        — many independent join+aggregate fragments, Zipf-skewed keys
  * ``zipf_agg``         — Zipf(1.1) int64-key aggregation + join (learned-sizing evaluation)
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Dict, List

import numpy as np

from .. import abi
from ..table import HostColumn, HostTable

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libminilake_workloads.so")
_lib = None


_GPU_LIB_PATH = os.path.join(_HERE, "libminilake_workloads_gpu.so")
_gpu_lib = None


def build_library() -> str:
    src = os.path.join(_HERE, "gen.cpp")
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", _LIB_PATH + ".tmp", src]
    subprocess.run(cmd, check=True)
    os.replace(_LIB_PATH + ".tmp", _LIB_PATH)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    arch = os.environ.get("MINILAKE_GPU_ARCH", "gfx950")
    cmd = [hipcc, f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-o", _GPU_LIB_PATH + ".tmp", os.path.join(_HERE, "gen_gpu.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for gen_gpu.hip:\n" + r.stderr)
    os.replace(_GPU_LIB_PATH + ".tmp", _GPU_LIB_PATH)
    return _LIB_PATH


def gpu_lib():
    """Device generator (libminilake_workloads_gpu.so)."""
    global _gpu_lib
    if _gpu_lib is None:
        if not os.path.exists(_GPU_LIB_PATH):
            build_library()
        L = C.CDLL(_GPU_LIB_PATH)
        i64, u64, vp = C.c_int64, C.c_uint64, C.c_void_p
        L.mlw_gpu_zipf_keys.argtypes = [vp, i64, u64, u64, C.c_double, i64, u64, vp]
        L.mlw_gpu_zipf_keys.restype = C.c_int
        L.mlw_gpu_build_keys.argtypes = [vp, i64, i64, u64, u64, i64, vp]
        L.mlw_gpu_build_keys.restype = C.c_int
        L.mlw_gpu_fragment.argtypes = [vp, vp, i64, vp, vp, i64, i64, C.c_double, u64, u64, vp]
        L.mlw_gpu_fragment.restype = C.c_int
        _gpu_lib = L
    return _gpu_lib


def gpu_zipf_keys(out_ptr, n, seed, stream, s, domain, perm_stream, hip_stream=None):
    if gpu_lib().mlw_gpu_zipf_keys(out_ptr, n, seed, stream, float(s), int(domain), perm_stream, hip_stream) != 0:
        raise RuntimeError("device key generation failed")


def gpu_build_keys(out_ptr, nkeys, mult, seed, perm_stream, domain, hip_stream=None):
    if gpu_lib().mlw_gpu_build_keys(out_ptr, nkeys, mult, seed, perm_stream, int(domain), hip_stream) != 0:
        raise RuntimeError("device key generation failed")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_library()
        L = C.CDLL(_LIB_PATH)
        i64, u64, f64p, i64p = C.c_int64, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.mlw_uniform_int.argtypes = [u64, u64, i64, i64, i64, i64, i64p]
        L.mlw_uniform_f64.argtypes = [u64, u64, i64, i64, C.c_double, C.c_double, f64p]
        L.mlw_pick_f64.argtypes = [u64, u64, i64, i64, f64p, i64, f64p]
        L.mlw_zipf.argtypes = [u64, u64, i64, i64, C.c_double, i64, i64p]
        L.mlw_permute.argtypes = [u64, u64, i64, i64, i64, i64p]
        L.mlw_gather8.argtypes = [i64p, i64p, i64, i64p]
        L.mlw_permute_values.argtypes = [u64, u64, i64, i64, i64p, i64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def uniform_int(seed, stream, n, lo, hi, offset=0) -> np.ndarray:
    out = np.empty(n, dtype=np.int64)
    lib().mlw_uniform_int(seed, stream, offset, n, lo, hi, _p(out, C.c_int64))
    return out


def uniform_f64(seed, stream, n, lo, hi, offset=0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().mlw_uniform_f64(seed, stream, offset, n, lo, hi, _p(out, C.c_double))
    return out


def pick_f64(seed, stream, n, table, offset=0) -> np.ndarray:
    t = np.ascontiguousarray(table, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    lib().mlw_pick_f64(seed, stream, offset, n, _p(t, C.c_double), len(t), _p(out, C.c_double))
    return out


def zipf(seed, stream, n, s, domain, offset=0) -> np.ndarray:
    out = np.empty(n, dtype=np.int64)
    lib().mlw_zipf(seed, stream, offset, n, float(s), int(domain), _p(out, C.c_int64))
    return out


def permutation(seed, stream, n, count=None, offset=0) -> np.ndarray:
    count = n if count is None else count
    out = np.empty(count, dtype=np.int64)
    lib().mlw_permute(seed, stream, offset, count, n, _p(out, C.c_int64))
    return out


def permute_values(seed, stream, n, values) -> np.ndarray:
    v = np.ascontiguousarray(values, dtype=np.int64)
    out = np.empty(len(v), dtype=np.int64)
    lib().mlw_permute_values(seed, stream, len(v), n, _p(v, C.c_int64), _p(out, C.c_int64))
    return out


def zipf_keys(seed, stream, n, s, domain, perm_stream) -> np.ndarray:
    """Host twin of the device generator: scrambled Zipf(s) keys in [0, domain)."""
    ranks = zipf(seed, stream, n, s, domain) if s > 0 else uniform_int(seed, stream, n, 1, domain)
    return permute_values(seed, perm_stream, domain, ranks - 1)


def gather8(src: np.ndarray, idx: np.ndarray) -> np.ndarray:
    out = np.empty(len(idx), dtype=src.dtype)
    lib().mlw_gather8(_p(src, C.c_int64), _p(np.ascontiguousarray(idx, dtype=np.int64), C.c_int64), len(idx),
                      _p(out, C.c_int64))
    return out


# Discount and tax values k/100 (computed in double, as the literal 0.0k would be).
DISC_TABLE = np.array([k / 100.0 for k in range(11)])
TAX_TABLE = np.array([k / 100.0 for k in range(9)])

# days since 1970-01-01
D_1992_01_01 = 8035
D_1992_01_02 = 8036
D_1994_01_01 = 8766
D_1995_01_01 = 9131
D_1995_03_15 = 9204

FILTER_GROUPBY_ROWS = 6_001_215
FILTER_GROUPBY_CUTOFF = 10510  # keeps 2475 of 2526 days (97.98%)


def filter_groupby(seed: int = 42, n: int = FILTER_GROUPBY_ROWS) -> HostTable:
    """lineitem(qty i32, price f64, disc f64, tax f64, flag1 dict{A,N,R}, flag2 dict{F,O}, date)."""
    qty = uniform_int(seed, 101, n, 1, 50).astype(np.int32)
    price = uniform_f64(seed, 102, n, 900.0, 105000.0)
    disc = pick_f64(seed, 103, n, DISC_TABLE)
    tax = pick_f64(seed, 104, n, TAX_TABLE)
    f1 = uniform_int(seed, 105, n, 0, 2).astype(np.uint32)
    f2 = uniform_int(seed, 106, n, 0, 1).astype(np.uint32)
    date = uniform_int(seed, 107, n, D_1992_01_02, D_1992_01_02 + 2525).astype(np.int32)
    return HostTable([
        HostColumn(abi.INT32, qty), HostColumn(abi.FLOAT64, price), HostColumn(abi.FLOAT64, disc),
        HostColumn(abi.FLOAT64, tax), HostColumn(abi.UTF8, f1, ["A", "N", "R"]),
        HostColumn(abi.UTF8, f2, ["F", "O"]), HostColumn(abi.DATE, date),
    ], names=["l_quantity", "l_extendedprice", "l_discount", "l_tax", "l_returnflag", "l_linestatus",
              "l_shipdate"])


SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY"]


@dataclass
class JoinTopN:
    customer: HostTable  # c_custkey i64, c_mktsegment dict
    orders: HostTable    # o_orderkey i64, o_custkey i64, o_orderdate date, o_shippriority i32
    lineitem: HostTable  # l_orderkey i64, l_extendedprice f64, l_discount f64, l_shipdate date

    def input_rows(self):
        return self.customer.num_rows + self.orders.num_rows + self.lineitem.num_rows

    def input_bytes(self):
        return self.customer.nbytes() + self.orders.nbytes() + self.lineitem.nbytes()


def join_topn(seed: int = 42, scale: float = 1.0) -> JoinTopN:
    """Q3 shape: 150k customers, 1.5M orders, 1..7 lineitems per order (~6M)."""
    n_c = max(1, int(150_000 * scale))
    n_o = max(1, int(1_500_000 * scale))
    c_key = np.arange(1, n_c + 1, dtype=np.int64)
    c_seg = uniform_int(seed, 201, n_c, 0, 4).astype(np.uint32)
    o_key = np.arange(1, n_o + 1, dtype=np.int64)
    o_cust = uniform_int(seed, 202, n_o, 1, n_c)
    o_date = uniform_int(seed, 203, n_o, D_1992_01_01, D_1992_01_01 + 2405).astype(np.int32)
    o_prio = uniform_int(seed, 204, n_o, 0, 4).astype(np.int32)
    cnt = uniform_int(seed, 205, n_o, 1, 7)
    owner = np.repeat(np.arange(n_o, dtype=np.int64), cnt)
    n_l = len(owner)
    l_key = o_key[owner]
    l_price = uniform_f64(seed, 206, n_l, 900.0, 105000.0)
    l_disc = pick_f64(seed, 207, n_l, DISC_TABLE)
    l_ship = (o_date[owner].astype(np.int64) + uniform_int(seed, 208, n_l, 1, 121)).astype(np.int32)
    return JoinTopN(
        HostTable([HostColumn(abi.INT64, c_key), HostColumn(abi.UTF8, c_seg, SEGMENTS)],
                  names=["c_custkey", "c_mktsegment"]),
        HostTable([HostColumn(abi.INT64, o_key), HostColumn(abi.INT64, o_cust), HostColumn(abi.DATE, o_date),
                   HostColumn(abi.INT32, o_prio)],
                  names=["o_orderkey", "o_custkey", "o_orderdate", "o_shippriority"]),
        HostTable([HostColumn(abi.INT64, l_key), HostColumn(abi.FLOAT64, l_price), HostColumn(abi.FLOAT64, l_disc),
                   HostColumn(abi.DATE, l_ship)],
                  names=["l_orderkey", "l_extendedprice", "l_discount", "l_shipdate"]),
    )


REGIONS = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
NATIONS = ["ALGERIA", "ARGENTINA", "BRAZIL", "CANADA", "EGYPT", "ETHIOPIA", "FRANCE", "GERMANY", "INDIA",
           "INDONESIA", "IRAN", "IRAQ", "JAPAN", "JORDAN", "KENYA", "MOROCCO", "MOZAMBIQUE", "PERU", "CHINA",
           "ROMANIA", "SAUDI ARABIA", "VIETNAM", "RUSSIA", "UNITED KINGDOM", "UNITED STATES"]
NATION_REGION = [0, 1, 1, 1, 4, 0, 3, 3, 2, 2, 4, 4, 2, 4, 0, 0, 0, 1, 2, 3, 4, 2, 3, 3, 1]


@dataclass
class StarJoin:
    region: HostTable    # r_regionkey i32, r_name dict
    nation: HostTable    # n_nationkey i32, n_name dict, n_regionkey i32
    customer: HostTable  # c_custkey i64, c_nationkey i32
    orders: HostTable    # o_orderkey i64, o_custkey i64, o_orderdate date
    supplier: HostTable  # s_suppkey i64, s_nationkey i32
    lineitem: HostTable  # l_orderkey i64, l_suppkey i64, l_extendedprice f64, l_discount f64

    def tables(self):
        return [self.region, self.nation, self.customer, self.orders, self.supplier, self.lineitem]

    def input_rows(self):
        return sum(t.num_rows for t in self.tables())

    def input_bytes(self):
        return sum(t.nbytes() for t in self.tables())


def star_join(seed: int = 42, scale: float = 10.0) -> StarJoin:
    """Q5 shape: lineitem 6M*scale, orders 1.5M*scale, customer 150k*scale, supplier 100k*scale."""
    n_c = max(1, int(150_000 * scale))
    n_o = max(1, int(1_500_000 * scale))
    n_s = max(1, int(100_000 * scale))
    n_l = max(1, int(6_000_000 * scale))
    region = HostTable([HostColumn(abi.INT32, np.arange(5, dtype=np.int32)),
                        HostColumn(abi.UTF8, np.arange(5, dtype=np.uint32), REGIONS)],
                       names=["r_regionkey", "r_name"])
    nation = HostTable([HostColumn(abi.INT32, np.arange(25, dtype=np.int32)),
                        HostColumn(abi.UTF8, np.arange(25, dtype=np.uint32), NATIONS),
                        HostColumn(abi.INT32, np.array(NATION_REGION, dtype=np.int32))],
                       names=["n_nationkey", "n_name", "n_regionkey"])
    customer = HostTable([HostColumn(abi.INT64, np.arange(1, n_c + 1, dtype=np.int64)),
                          HostColumn(abi.INT32, uniform_int(seed, 301, n_c, 0, 24).astype(np.int32))],
                         names=["c_custkey", "c_nationkey"])
    orders = HostTable([HostColumn(abi.INT64, np.arange(1, n_o + 1, dtype=np.int64)),
                        HostColumn(abi.INT64, uniform_int(seed, 302, n_o, 1, n_c)),
                        HostColumn(abi.DATE, uniform_int(seed, 303, n_o, D_1992_01_01,
                                                         D_1992_01_01 + 2405).astype(np.int32))],
                       names=["o_orderkey", "o_custkey", "o_orderdate"])
    supplier = HostTable([HostColumn(abi.INT64, np.arange(1, n_s + 1, dtype=np.int64)),
                          HostColumn(abi.INT32, uniform_int(seed, 304, n_s, 0, 24).astype(np.int32))],
                         names=["s_suppkey", "s_nationkey"])
    lineitem = HostTable([HostColumn(abi.INT64, uniform_int(seed, 305, n_l, 1, n_o)),
                          HostColumn(abi.INT64, uniform_int(seed, 306, n_l, 1, n_s)),
                          HostColumn(abi.FLOAT64, uniform_f64(seed, 307, n_l, 900.0, 105000.0)),
                          HostColumn(abi.FLOAT64, pick_f64(seed, 308, n_l, DISC_TABLE))],
                         names=["l_orderkey", "l_suppkey", "l_extendedprice", "l_discount"])
    return StarJoin(region, nation, customer, orders, supplier, lineitem)


# ------------------------------------------------------------ batched fragments
@dataclass
class FragmentShape:
    probe_rows: int
    build_rows: int
    groups: int
    zipf_s: float


def fragment_shapes(seed: int = 42, n: int = 256, max_log2_probe: float = 24.0) -> List[FragmentShape]:
    """Sizes of the batched-fragment workload: probe rows = floor(2^U[10, 24]),
    build rows = probe rows / U[4, 40], groups = floor(2^U[0, 20]), Zipf
    s = (0, 0.8, 1.2)[i % 3]. `max_log2_probe` < 24 shrinks the probe range's
    upper end (the other draws are unchanged)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        lp = rng.uniform(10.0, 24.0)
        ratio = rng.uniform(4.0, 40.0)
        lg = rng.uniform(0.0, 20.0)
        lp = 10.0 + (lp - 10.0) * (max_log2_probe - 10.0) / 14.0
        prow = int(2.0 ** lp)
        out.append(FragmentShape(prow, max(1, int(prow / ratio)), max(1, int(2.0 ** lg)), (0.0, 0.8, 1.2)[i % 3]))
    return out


def gpu_fragment(shape: FragmentShape, index: int, seed: int, bkeys, bgroups, pkeys, pvals, hip_stream=None):
    """Fill device buffers (pointers) with fragment `index`: unique build keys,
    uniform build groups, Zipf(s) probe keys over the build keys, U[0,1) values."""
    rc = gpu_lib().mlw_gpu_fragment(bkeys, bgroups, shape.build_rows, pkeys, pvals, shape.probe_rows, shape.groups,
                                    float(shape.zipf_s), seed, index, hip_stream)
    if rc != 0:
        raise RuntimeError("device fragment generation failed")
