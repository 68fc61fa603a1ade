"""Host-side mirror of MiniLake's operator interface over the gfx950 C ABI.

``GpuEngine`` exposes the operators and sinks of the reference executor
(crates/minilake-exec/src/pipeline.rs:32-79) on device-resident tables:

  join_build   JoinBuildSink + build_table        (operators/join.rs:72-203)
  hashagg      HashAggregateSink                  (operators/aggregate/hash.rs:285-484)
  topn         SortSink with a limit              (operators/sort.rs:145-268)
  collect      CollectSink                        (operators/collect.rs:11-53)
  filter       FilterOperator / select            (operators/filter.rs:19-25)
  project      ProjectionOperator                 (operators/projection.rs:21-32)
  hash_columns kernels/hash.rs:16-23
  join_probe   HashJoinProbe::execute             (operators/join.rs:263-315)

Every pipeline (source -> filters/probes -> sink) runs as one fused kernel.
PyTorch only provides device memory and the stream; all compute is in
``libminilake_gpu.so``. There is no CPU fallback: without the library or a
GPU, constructing an engine raises.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from . import abi
from .table import NP_DTYPE, HostColumn, HostTable, bitmap_from_bools, make_dict

_HERE = os.path.dirname(os.path.abspath(__file__))
# MINILAKE_GPU_LIB selects another build of the library (e.g. one compiled with
# different tuning constants by minilake_amd.build.build_gpu_library)
LIB_PATH = os.environ.get("MINILAKE_GPU_LIB") or os.path.join(_HERE, "libminilake_gpu.so")
_lib = None

# Every symbol include/minilake_gpu.h declares.
EXPORTED = [
    "ml_default_config", "ml_ctx_create", "ml_ctx_destroy", "ml_last_error", "ml_ctx_synchronize", "ml_ctx_stream",
    "ml_result_num_rows", "ml_result_num_columns", "ml_result_column", "ml_result_copy_column",
    "ml_result_copy_validity", "ml_result_free", "ml_expr_type", "ml_filter", "ml_project", "ml_hash_columns",
    "ml_hashagg_create", "ml_hashagg_sink", "ml_hashagg_finalize", "ml_hashagg_num_groups",
    "ml_hashagg_table_stats", "ml_hashagg_free", "ml_join_build_create", "ml_join_build_sink",
    "ml_join_build_finalize", "ml_join_table_num_rows", "ml_join_table_unique", "ml_join_table_free",
    "ml_join_table_num_columns", "ml_join_table_column", "ml_pipeline_collect_sized", "ml_pipeline_count",
    "ml_join_probe", "ml_pipeline_hashagg", "ml_pipeline_join_build", "ml_pipeline_topn", "ml_pipeline_collect",
    "ml_pipeline_limit", "ml_mlb1_write", "ml_mlb1_read", "ml_parquet_open", "ml_parquet_num_columns",
    "ml_parquet_column_type", "ml_parquet_column_name", "ml_parquet_column_nullable", "ml_parquet_num_row_groups",
    "ml_parquet_row_group_rows", "ml_parquet_column_stats", "ml_parquet_prune", "ml_parquet_read_row_group",
    "ml_parquet_free", "ml_ctx_set_memory_limit", "ml_ctx_memory_usage", "ml_join_build_finalize_many",
    "ml_topn_create", "ml_topn_sink", "ml_topn_finalize", "ml_topn_free", "ml_ctx_enable_stats",
    "ml_ctx_take_stats", "ml_fragments_create", "ml_fragments_run", "ml_fragments_run_graph",
    "ml_fragments_run_per_fragment", "ml_fragments_autotune", "ml_fragments_free", "ml_mlp_create",
    "ml_mlp_get_params", "ml_mlp_forward", "ml_mlp_loss_grad", "ml_mlp_train_step", "ml_mlp_free",
    "ml_sizing_features", "ml_ctx_set_estimator", "ml_filter_noalloc", "ml_hashagg_create_local",
    "ml_hashagg_combine", "ml_join_build_create_local", "ml_join_build_combine", "ml_topn_create_local",
    "ml_topn_combine", "ml_limit_create", "ml_limit_sink", "ml_limit_is_finished", "ml_limit_finalize",
    "ml_limit_free",
]


class GpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load libminilake_gpu.so (building it first if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build_gpu_library()
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        P = C.POINTER
        sig = {
            "ml_default_config": ([P(abi.MlExecConfig)], None),
            "ml_ctx_create": ([P(abi.MlExecConfig), P(vp)], i32),
            "ml_ctx_destroy": ([vp], None),
            "ml_last_error": ([vp], C.c_char_p),
            "ml_ctx_synchronize": ([vp], i32),
            "ml_ctx_stream": ([vp], vp),
            "ml_result_num_rows": ([vp], i64),
            "ml_result_num_columns": ([vp], i32),
            "ml_result_column": ([vp, i32], P(abi.MlColumn)),
            "ml_result_copy_column": ([vp, vp, i32, vp], i32),
            "ml_result_copy_validity": ([vp, vp, i32, vp], i32),
            "ml_result_free": ([vp], None),
            "ml_expr_type": ([vp, P(abi.MlExpr), P(abi.MlBatch), P(i32)], i32),
            "ml_filter": ([vp, P(abi.MlExpr), P(abi.MlBatch), vp, P(i64)], i32),
            "ml_filter_noalloc": ([vp, P(abi.MlExpr), P(abi.MlBatch), vp, P(i64)], i32),
            "ml_hashagg_create_local": ([vp, vp, P(vp)], i32),
            "ml_hashagg_combine": ([vp, vp], i32),
            "ml_join_build_create_local": ([vp, vp, P(vp)], i32),
            "ml_join_build_combine": ([vp, vp], i32),
            "ml_topn_create_local": ([vp, vp, P(vp)], i32),
            "ml_topn_combine": ([vp, vp], i32),
            "ml_limit_create": ([vp, i64, i64, P(vp)], i32),
            "ml_limit_sink": ([vp, vp, P(abi.MlBatch)], i32),
            "ml_limit_is_finished": ([vp], i32),
            "ml_limit_finalize": ([vp, P(vp)], i32),
            "ml_limit_free": ([vp], None),
            "ml_project": ([vp, P(abi.MlExpr), i32, P(abi.MlBatch), P(vp)], i32),
            "ml_hash_columns": ([vp, P(abi.MlExpr), i32, P(abi.MlBatch), vp], i32),
            "ml_hashagg_create": ([vp, P(abi.MlAggSpec), P(vp)], i32),
            "ml_hashagg_sink": ([vp, P(abi.MlBatch)], i32),
            "ml_hashagg_finalize": ([vp, P(vp)], i32),
            "ml_hashagg_num_groups": ([vp, P(i64)], i32),
            "ml_hashagg_table_stats": ([vp, P(i64), P(i32)], i32),
            "ml_hashagg_free": ([vp], None),
            "ml_join_build_create": ([vp, P(abi.MlJoinBuildSpec), P(vp)], i32),
            "ml_join_build_sink": ([vp, P(abi.MlBatch)], i32),
            "ml_join_build_finalize": ([vp, P(vp)], i32),
            "ml_join_table_num_rows": ([vp], i64),
            "ml_join_table_unique": ([vp], i32),
            "ml_join_table_free": ([vp], None),
            "ml_join_table_num_columns": ([vp], i32),
            "ml_join_table_column": ([vp, i32], P(abi.MlColumn)),
            "ml_join_probe": ([vp, vp, P(abi.MlExpr), i32, P(abi.MlBatch), P(vp)], i32),
            "ml_pipeline_hashagg": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, vp], i32),
            "ml_pipeline_join_build": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, vp], i32),
            "ml_pipeline_topn": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, vp], i32),
            "ml_pipeline_collect": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, P(vp)], i32),
            "ml_pipeline_limit": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, i64, i64, P(vp)], i32),
            "ml_mlb1_write": ([vp, P(abi.MlBatch), vp, i64, P(i64)], i32),
            "ml_mlb1_read": ([vp, C.c_char_p, i64, P(i64), P(vp)], i32),
            "ml_parquet_open": ([vp, C.c_char_p, i64, P(vp)], i32),
            "ml_parquet_num_columns": ([vp], i32),
            "ml_parquet_column_type": ([vp, i32], i32),
            "ml_parquet_column_name": ([vp, i32], C.c_char_p),
            "ml_parquet_column_nullable": ([vp, i32], i32),
            "ml_parquet_num_row_groups": ([vp], i32),
            "ml_parquet_row_group_rows": ([vp, i32], i64),
            "ml_parquet_column_stats": ([vp, i32, i32, P(abi.MlColumnStats)], i32),
            "ml_parquet_prune": ([vp, P(abi.MlPrunePredicate), i32, P(i32), P(i32)], i32),
            "ml_parquet_read_row_group": ([vp, i32, P(i32), i32, P(vp)], i32),
            "ml_parquet_free": ([vp], None),
            "ml_ctx_set_memory_limit": ([vp, i64, i32], i32),
            "ml_join_build_finalize_many": ([vp, i32, vp], i32),
            "ml_ctx_memory_usage": ([vp, P(i64), P(i64), P(i64)], None),
            "ml_pipeline_collect_sized": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, i64, P(i32), P(vp)], i32),
            "ml_pipeline_count": ([vp, P(abi.MlBatch), P(abi.MlOperator), i32, P(i64)], i32),
            "ml_topn_create": ([vp, P(abi.MlSortSpec), P(vp)], i32),
            "ml_topn_sink": ([vp, P(abi.MlBatch)], i32),
            "ml_topn_finalize": ([vp, P(vp)], i32),
            "ml_topn_free": ([vp], None),
            "ml_ctx_enable_stats": ([vp, i32], i32),
            "ml_ctx_take_stats": ([vp, P(abi.MlLaunchStat), i32], i32),
            "ml_fragments_create": ([vp, P(abi.MlFragment), i32, P(abi.MlFragmentTuning), P(vp)], i32),
            "ml_fragments_run": ([vp], i32),
            "ml_fragments_run_graph": ([vp], i32),
            "ml_fragments_run_per_fragment": ([vp], i32),
            "ml_fragments_autotune": ([vp, P(abi.MlFragmentTuning)], i32),
            "ml_fragments_free": ([vp], None),
            "ml_mlp_create": ([vp, vp, P(vp)], i32),
            "ml_mlp_get_params": ([vp, vp], i32),
            "ml_mlp_forward": ([vp, vp, i64, vp], i32),
            "ml_mlp_loss_grad": ([vp, vp, vp, i64, P(C.c_float), vp], i32),
            "ml_mlp_train_step": ([vp, vp, vp, i64, C.c_float, P(C.c_float)], i32),
            "ml_mlp_free": ([vp], None),
            "ml_sizing_features": ([vp, vp, i64, i64, i32, C.c_uint64, vp], i32),
            "ml_ctx_set_estimator": ([vp, vp], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols(path: str = LIB_PATH):
    """Names from EXPORTED that the built library actually exports (no GPU needed)."""
    L = C.CDLL(path)
    return [n for n in EXPORTED if hasattr(L, n)]


_hip = None


def device_to_host(dst: int, src: int, nbytes: int) -> None:
    """hipMemcpy device -> host (synchronous), for inspecting library-owned buffers."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    rc = _hip.hipMemcpy(dst, src, nbytes, 2)
    if rc != 0:
        raise GpuError(abi.ML_ERR_INTERNAL, f"hipMemcpy failed ({rc})")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the gfx950 engine needs a visible GPU (torch.cuda.is_available() is False)")
    return torch


_TORCH_DT = {abi.BOOLEAN: "uint8", abi.INT32: "int32", abi.INT64: "int64", abi.FLOAT64: "float64",
             abi.DATE: "int32", abi.UTF8: "int32"}


class _Selection:
    """A device selection vector: pointer and length (an empty one keeps its pointer)."""

    def __init__(self, ptr: int, n: int):
        self.ptr, self.n = ptr, n

    def data_ptr(self) -> int:
        return self.ptr

    def numel(self) -> int:
        return self.n


class DeviceTable:
    """A HostTable uploaded to HBM (one tensor per column, dictionaries kept on the host)."""

    def __init__(self, host: HostTable, device="cuda"):
        torch = _torch()
        self.host = host
        self.num_rows = host.num_rows
        self.types = host.types
        self.tensors = []
        self.bitmaps = []  # device validity bitmaps (None = all valid)
        for c in host.columns:
            v = c.values.view(np.int32) if c.type == abi.UTF8 else c.values
            self.tensors.append(torch.from_numpy(np.ascontiguousarray(v)).to(device))
            bm = None
            if c.validity is not None and not np.all(c.validity) and len(c) > 0:
                bm = torch.from_numpy(bitmap_from_bools(np.asarray(c.validity, bool)).view(np.int64)).to(device)
            self.bitmaps.append(bm)
        self.dicts = [c.dictionary for c in host.columns]

    @classmethod
    def from_tensors(cls, tensors, types, dicts=None, num_rows=None) -> "DeviceTable":
        """Wrap device tensors (already in HBM) as a table without a host copy."""
        t = cls.__new__(cls)
        t.host = None
        t.tensors = list(tensors)
        t.types = list(types)
        t.dicts = list(dicts) if dicts else [None] * len(t.tensors)
        t.bitmaps = [None] * len(t.tensors)
        t.num_rows = int(num_rows if num_rows is not None else (t.tensors[0].numel() if t.tensors else 0))
        return t

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in self.tensors)

    def ml_batch(self, keep: abi.Keep, selection=None) -> abi.MlBatch:
        cols = keep((abi.MlColumn * max(1, len(self.tensors)))())
        dcache = {}
        for i, (t, ty) in enumerate(zip(self.tensors, self.types)):
            cols[i].type = ty
            cols[i].data = t.data_ptr()
            cols[i].len = self.num_rows
            if self.bitmaps[i] is not None:
                cols[i].validity = C.cast(C.c_void_p(self.bitmaps[i].data_ptr()), C.POINTER(C.c_uint64))
            if ty == abi.UTF8:
                key = id(self.dicts[i])
                if key not in dcache:
                    dcache[key] = make_dict(self.dicts[i], keep)
                cols[i].dict = C.pointer(dcache[key])
        sel_ptr, sel_len = None, 0
        if selection is not None:
            sel_ptr, sel_len = selection.data_ptr(), selection.numel()
        return keep(abi.MlBatch(cols, len(self.tensors), 0, self.num_rows, sel_ptr, sel_len))


class GpuResult:
    """A library-owned device result (ml_result): usable as a source or copied to the host."""

    def __init__(self, engine: "GpuEngine", handle):
        self.engine, self.handle = engine, handle
        L = lib()
        self.num_rows = L.ml_result_num_rows(handle)
        self.num_columns = L.ml_result_num_columns(handle)
        self.types = [L.ml_result_column(handle, i).contents.type for i in range(self.num_columns)]

    def ml_batch(self, keep: abi.Keep, selection=None) -> abi.MlBatch:
        L = lib()
        cols = keep((abi.MlColumn * max(1, self.num_columns))())
        for i in range(self.num_columns):
            cols[i] = L.ml_result_column(self.handle, i).contents
        sel_ptr, sel_len = None, 0
        if selection is not None:
            sel_ptr, sel_len = selection.data_ptr(), selection.numel()
        return keep(abi.MlBatch(cols, self.num_columns, 0, self.num_rows, sel_ptr, sel_len))

    def to_tensors(self):
        """The result's columns as device tensors (copies), by column type."""
        L, torch = lib(), self.engine.torch
        out = []
        for i in range(self.num_columns):
            dt = {abi.BOOLEAN: torch.uint8, abi.INT32: torch.int32, abi.DATE: torch.int32, abi.UTF8: torch.int32,
                  abi.INT64: torch.int64, abi.FLOAT64: torch.float64}[self.types[i]]
            t = torch.empty(max(1, self.num_rows), dtype=dt, device=f"cuda:{self.engine.device}")
            self.engine._check(L.ml_result_copy_column(self.engine.ctx, self.handle, i, t.data_ptr()))
            out.append(t[: self.num_rows])
        return out

    def fingerprint(self):
        """Sorted 64-bit per-row fingerprints over every column: two results are
        the same multiset of rows exactly when these tensors are equal (up to a
        fingerprint collision)."""
        torch = self.engine.torch
        h = torch.full((self.num_rows,), 0x632BE59BD9B4E019, dtype=torch.int64,
                       device=f"cuda:{self.engine.device}")
        for c in self.to_tensors():
            v = c.view(torch.int64) if c.dtype in (torch.int64, torch.float64) else c.to(torch.int64)
            h = (h ^ v) * -7046029254386353131  # 0x9E3779B97F4A7C15 as a signed 64-bit value
            h = h ^ (h >> 29)
        return torch.sort(h).values

    def to_host(self) -> HostTable:
        L = lib()
        cols = []
        for i in range(self.num_columns):
            mc = L.ml_result_column(self.handle, i).contents
            dt = NP_DTYPE[mc.type]
            vals = np.zeros(max(1, self.num_rows), dtype=dt)
            valid = np.zeros(max(1, self.num_rows), dtype=np.uint8)
            self.engine._check(L.ml_result_copy_column(self.engine.ctx, self.handle, i, vals.ctypes.data))
            self.engine._check(L.ml_result_copy_validity(self.engine.ctx, self.handle, i, valid.ctypes.data))
            vals, valid = vals[:self.num_rows], valid[:self.num_rows].astype(bool)
            d = None
            if mc.type == abi.UTF8:
                md = mc.dict.contents
                offs = [md.offsets[k] for k in range(md.len + 1)]
                raw = C.string_at(md.bytes, offs[-1]) if offs[-1] else b""
                d = [raw[offs[k]:offs[k + 1]].decode() for k in range(md.len)]
            cols.append(HostColumn(mc.type, vals, d, None if valid.all() else valid))
        return HostTable(cols, num_rows=self.num_rows)

    def __del__(self):
        if self.handle and _lib is not None:
            _lib.ml_result_free(self.handle)
            self.handle = None


class ParquetTable:
    """Table::open (crates/minilake-storage/src/table.rs:59-134) over one
    Parquet file held in host memory. Metadata, statistics and pruning are
    host work (usable without a GPU, ctx = None); read_row_group decodes on
    the device."""

    def __init__(self, data: bytes, engine: Optional["GpuEngine"] = None):
        self.data, self.engine = bytes(data), engine
        ctx = engine.ctx if engine is not None else None
        h = C.c_void_p()
        rc = lib().ml_parquet_open(ctx, self.data, len(self.data), C.byref(h))
        if rc != 0:
            msg = lib().ml_last_error(ctx).decode() if ctx else ""
            raise GpuError(rc, msg)
        self.handle = h.value
        L = lib()
        self.num_columns = L.ml_parquet_num_columns(self.handle)
        self.types = [L.ml_parquet_column_type(self.handle, i) for i in range(self.num_columns)]
        self.names = [L.ml_parquet_column_name(self.handle, i).decode() for i in range(self.num_columns)]
        self.nullable = [bool(L.ml_parquet_column_nullable(self.handle, i)) for i in range(self.num_columns)]
        self.num_row_groups = L.ml_parquet_num_row_groups(self.handle)
        self.row_group_rows = [L.ml_parquet_row_group_rows(self.handle, i) for i in range(self.num_row_groups)]

    def _rc(self, rc):
        if rc != 0:
            ctx = self.engine.ctx if self.engine is not None else None
            raise GpuError(rc, lib().ml_last_error(ctx).decode() if ctx else "")

    def stats(self, rg: int, c: int) -> dict:
        """ColumnStats of a row group's column: min / max (None when unknown), null_count."""
        st = abi.MlColumnStats()
        self._rc(lib().ml_parquet_column_stats(self.handle, rg, c, C.byref(st)))

        def val(has, i, f, p, n):
            if not has:
                return None
            if st.type == abi.FLOAT64:
                return f
            if st.type == abi.UTF8:
                return C.string_at(p, n).decode("utf-8")
            return i
        return {"min": val(st.has_min, st.min_i64, st.min_f64, st.min_str, st.min_len),
                "max": val(st.has_max, st.max_i64, st.max_f64, st.max_str, st.max_len),
                "null_count": st.null_count if st.has_null_count else None}

    def prune(self, preds) -> List[int]:
        """Table::prune (table.rs:165-172); preds: (column, op, value), the value of the column's type."""
        keep = abi.Keep()
        arr = keep((abi.MlPrunePredicate * max(1, len(preds)))())
        for k, (c, op, v) in enumerate(preds):
            arr[k].column, arr[k].op = c, abi.PRUNE_OPS[op]
            t = self.types[c]
            if isinstance(v, str):
                raw = keep(v.encode())
                arr[k].type, arr[k].str, arr[k].str_len = abi.UTF8, raw, len(raw)
            elif isinstance(v, float):
                arr[k].type, arr[k].f64 = abi.FLOAT64, v
            else:
                arr[k].type, arr[k].i64 = (abi.DATE if t == abi.DATE else abi.INT64), int(v)
        out = (C.c_int32 * max(1, self.num_row_groups))()
        n = C.c_int32()
        self._rc(lib().ml_parquet_prune(self.handle, arr, len(preds), out, C.byref(n)))
        return list(out[: n.value])

    def read_row_group(self, rg: int, projection: Optional[Sequence[int]] = None) -> "GpuResult":
        """Table::read_row_group (table.rs:204-240), decoded on the device."""
        if self.engine is None:
            raise GpuError(abi.ML_ERR_INTERNAL, "reading row groups needs a GpuEngine")
        proj = list(range(self.num_columns)) if projection is None else list(projection)
        arr = (C.c_int32 * max(1, len(proj)))(*proj)
        out = C.c_void_p()
        self._rc(lib().ml_parquet_read_row_group(self.handle, rg, arr, len(proj), C.byref(out)))
        return GpuResult(self.engine, out.value)

    def close(self):
        if self.handle:
            lib().ml_parquet_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuJoin:
    def __init__(self, handle):
        self.handle = handle

    def num_rows(self):
        return lib().ml_join_table_num_rows(self.handle)

    def unique(self):
        return bool(lib().ml_join_table_unique(self.handle))

    def column(self, i: int) -> np.ndarray:
        """Build column i in table row order, copied to the host."""
        mc = lib().ml_join_table_column(self.handle, i).contents
        n = self.num_rows()
        out = np.zeros(max(n, 1), dtype=NP_DTYPE[mc.type])
        if n:
            device_to_host(out.ctypes.data, mc.data, n * out.itemsize)
        return out[:n]

    def close(self):
        if self.handle and _lib is not None:
            _lib.ml_join_table_free(self.handle)
        self.handle = None

    def __del__(self):
        self.close()


class GpuEngine:
    """The gfx950 executor behind the engine interface of minilake_amd.queries."""

    name = "gpu"

    def __init__(self, device: int = 0, batch_rows: int = 1 << 21, use_bloom: int = -1, block_size: int = 256,
                 grid_per_cu: int = 8):
        torch = _torch()
        L = lib()
        self.torch = torch
        torch.cuda.set_device(device)
        # one stream for torch's copies/allocations and the library's kernels,
        # so everything the engine does is stream-ordered
        self.stream = torch.cuda.Stream(device)
        self._prev_stream = torch.cuda.current_stream(device)
        torch.cuda.set_stream(self.stream)
        cfg = abi.MlExecConfig()
        L.ml_default_config(C.byref(cfg))
        cfg.batch_rows, cfg.device, cfg.use_bloom = batch_rows, device, use_bloom
        cfg.block_size, cfg.grid_per_cu = block_size, grid_per_cu
        cfg.stream = self.stream.cuda_stream
        ctx = C.c_void_p()
        rc = L.ml_ctx_create(C.byref(cfg), C.byref(ctx))
        if rc != 0:
            raise GpuError(rc, "ml_ctx_create failed")
        self.ctx = ctx.value
        self.device = device

    def close(self):
        if self.ctx:
            lib().ml_ctx_destroy(self.ctx)
            self.ctx = None
            if self.torch.cuda.current_stream(self.device) == self.stream:
                self.torch.cuda.set_stream(self._prev_stream)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise GpuError(rc, lib().ml_last_error(self.ctx).decode())

    def set_memory_limit(self, nbytes: int, spill: bool = False):
        """MemoryPool (memory.rs:22-157) for this context; 0 = unlimited."""
        self._check(lib().ml_ctx_set_memory_limit(self.ctx, int(nbytes), int(spill)))

    def memory_usage(self) -> dict:
        used, peak, limit = C.c_int64(), C.c_int64(), C.c_int64()
        lib().ml_ctx_memory_usage(self.ctx, C.byref(used), C.byref(peak), C.byref(limit))
        return {"used": used.value, "peak": peak.value, "limit": limit.value}

    def synchronize(self):
        self._check(lib().ml_ctx_synchronize(self.ctx))

    # ------------------------------------------------------------ tables
    def table(self, t) -> DeviceTable:
        if isinstance(t, (DeviceTable, GpuResult)):
            return t
        dt = DeviceTable(t, device=f"cuda:{self.device}")
        self.torch.cuda.synchronize()
        return dt

    def _src(self, src, keep, selection=None):
        src = self.table(src)
        sel = None
        if selection is not None:
            arr = np.ascontiguousarray(selection, dtype=np.uint32).view(np.int32)
            # an empty selection still needs a non-NULL pointer (NULL means "no selection")
            dev = self.torch.from_numpy(arr if len(arr) else np.zeros(1, np.int32)).to(f"cuda:{self.device}")
            keep(dev)
            sel = _Selection(dev.data_ptr(), len(arr))
        return src, src.ml_batch(keep, sel)

    def _ops(self, ops, keep):
        return abi.op_array(ops, keep, table_handle=lambda j: j.handle)

    # ------------------------------------------------------------ sinks
    def join_build(self, src, ops, keys, key_types) -> GpuJoin:
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        spec = abi.join_build_spec(keys, key_types, keep)
        jb = C.c_void_p()
        self._check(L.ml_join_build_create(self.ctx, C.byref(spec), C.byref(jb)))
        rc = L.ml_pipeline_join_build(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), jb)
        if rc != 0:
            msg = L.ml_last_error(self.ctx).decode()
            t = C.c_void_p()
            L.ml_join_build_finalize(jb, C.byref(t))  # frees the builder
            if t.value:
                L.ml_join_table_free(t)
            raise GpuError(rc, msg)
        t = C.c_void_p()
        self._check(L.ml_join_build_finalize(jb, C.byref(t)))
        return GpuJoin(t.value)

    def join_build_many(self, specs) -> List[GpuJoin]:
        """Independent join builds, (src, ops, keys, key_types) each, finalized
        together (ml_join_build_finalize_many: one host round trip)."""
        L = lib()
        builders = []
        try:
            for src, ops, keys, key_types in specs:
                keep = abi.Keep()
                _, b = self._src(src, keep)
                spec = abi.join_build_spec(keys, key_types, keep)
                jb = C.c_void_p()
                self._check(L.ml_join_build_create(self.ctx, C.byref(spec), C.byref(jb)))
                builders.append(jb.value)
                self._check(L.ml_pipeline_join_build(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), jb))
        except Exception:
            for jb in builders:  # finalizing frees each builder
                t = C.c_void_p()
                L.ml_join_build_finalize(jb, C.byref(t))
                if t.value:
                    L.ml_join_table_free(t)
            raise
        arr = (C.c_void_p * len(builders))(*builders)
        out = (C.c_void_p * len(builders))()
        self._check(L.ml_join_build_finalize_many(arr, len(builders), out))
        return [GpuJoin(out[i]) for i in range(len(builders))]

    def hashagg(self, src, ops, group_exprs, key_types, aggs, sizing=abi.SIZING_PLAIN, capacity_hint=0,
                stats: Optional[dict] = None, mode: int = abi.AGG_MODE_SINGLE) -> GpuResult:
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        spec = abi.agg_spec(group_exprs, key_types, aggs, keep, sizing, capacity_hint, mode)
        h = C.c_void_p()
        self._check(L.ml_hashagg_create(self.ctx, C.byref(spec), C.byref(h)))
        try:
            self._check(L.ml_pipeline_hashagg(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), h))
            out = C.c_void_p()
            self._check(L.ml_hashagg_finalize(h, C.byref(out)))
            if stats is not None:
                cap, res = C.c_int64(), C.c_int32()
                self._check(L.ml_hashagg_table_stats(h, C.byref(cap), C.byref(res)))
                stats["capacity"], stats["resizes"] = cap.value, res.value
        finally:
            L.ml_hashagg_free(h)
        return GpuResult(self, out.value)

    def topn(self, src, ops, sort_keys, limit) -> GpuResult:
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        spec = abi.sort_spec(sort_keys, limit, keep)
        h = C.c_void_p()
        self._check(L.ml_topn_create(self.ctx, C.byref(spec), C.byref(h)))
        try:
            self._check(L.ml_pipeline_topn(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), h))
            out = C.c_void_p()
            self._check(L.ml_topn_finalize(h, C.byref(out)))
        finally:
            L.ml_topn_free(h)
        return GpuResult(self, out.value)

    def collect(self, src, ops) -> GpuResult:
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        out = C.c_void_p()
        self._check(L.ml_pipeline_collect(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), C.byref(out)))
        return GpuResult(self, out.value)

    def limit(self, src, ops, limit: int, offset: int = 0) -> GpuResult:
        """LimitSink (operators/limit.rs:26-117): rows [offset, offset + limit) in source order."""
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        out = C.c_void_p()
        self._check(L.ml_pipeline_limit(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), limit, offset,
                                        C.byref(out)))
        return GpuResult(self, out.value)

    def mlb1_write(self, src, selection: Optional[np.ndarray] = None) -> bytes:
        """write_batch (ipc.rs:44-87) on the device; the bytes are copied to the host."""
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep, selection)
        size = C.c_int64()
        self._check(L.ml_mlb1_write(self.ctx, C.byref(b), None, 0, C.byref(size)))
        buf = self.torch.empty(max(1, size.value), dtype=self.torch.uint8, device=f"cuda:{self.device}")
        self._check(L.ml_mlb1_write(self.ctx, C.byref(b), buf.data_ptr(), buf.numel(), C.byref(size)))
        return bytes(buf[: size.value].cpu().numpy())

    def mlb1_read(self, data: bytes):
        """read_batch (ipc.rs:111-170): (device result, or None at a clean end of stream; bytes consumed)."""
        out, used = C.c_void_p(), C.c_int64()
        self._check(lib().ml_mlb1_read(self.ctx, data, len(data), C.byref(used), C.byref(out)))
        return (GpuResult(self, out.value) if out.value else None), used.value

    def count_distinct(self, src, ops, group_exprs, key_types, sizing=abi.SIZING_PLAIN, capacity_hint=0,
                       stats: Optional[dict] = None) -> int:
        """Number of groups of a GROUP BY (the aggregate's table, no result rows emitted)."""
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        spec = abi.agg_spec(group_exprs, key_types, [abi.count_star()], keep, sizing, capacity_hint)
        h = C.c_void_p()
        self._check(L.ml_hashagg_create(self.ctx, C.byref(spec), C.byref(h)))
        try:
            self._check(L.ml_pipeline_hashagg(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), h))
            n = C.c_int64()
            self._check(L.ml_hashagg_num_groups(h, C.byref(n)))
            if stats is not None:
                cap, res = C.c_int64(), C.c_int32()
                self._check(L.ml_hashagg_table_stats(h, C.byref(cap), C.byref(res)))
                stats["capacity"], stats["resizes"] = cap.value, res.value
        finally:
            L.ml_hashagg_free(h)
        return n.value

    def collect_sized(self, src, ops, capacity_hint: int):
        """Collect with a pre-sized output; returns (result, used_fallback)."""
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        out, fb = C.c_void_p(), C.c_int32()
        self._check(L.ml_pipeline_collect_sized(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), int(capacity_hint),
                                                C.byref(fb), C.byref(out)))
        return GpuResult(self, out.value), bool(fb.value)

    def count(self, src, ops) -> int:
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        n = C.c_int64()
        self._check(L.ml_pipeline_count(self.ctx, C.byref(b), self._ops(ops, keep), len(ops), C.byref(n)))
        return n.value

    # -------------------------------------------------------- operators
    def expr_type(self, e, src):
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        me = abi.to_ml_expr(e, keep)
        t = C.c_int32()
        self._check(L.ml_expr_type(self.ctx, C.byref(me), C.byref(b), C.byref(t)))
        return t.value

    def filter(self, pred, src, selection: Optional[np.ndarray] = None) -> np.ndarray:
        L, keep = lib(), abi.Keep()
        s, b = self._src(src, keep, selection)
        n = s.num_rows if selection is None else len(selection)
        out = self.torch.zeros(max(1, n), dtype=self.torch.int32, device=f"cuda:{self.device}")
        me = abi.to_ml_expr(pred, keep)
        cnt = C.c_int64()
        self._check(L.ml_filter(self.ctx, C.byref(me), C.byref(b), out.data_ptr(), C.byref(cnt)))
        return out[:cnt.value].cpu().numpy().view(np.uint32).copy()

    def project(self, exprs: Sequence, src) -> HostTable:
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep)
        arr = abi.expr_array(exprs, keep)
        out = C.c_void_p()
        self._check(L.ml_project(self.ctx, arr, len(exprs), C.byref(b), C.byref(out)))
        return GpuResult(self, out.value).to_host()

    def hash_columns(self, keys: Sequence, src, selection: Optional[np.ndarray] = None) -> np.ndarray:
        L, keep = lib(), abi.Keep()
        s, b = self._src(src, keep, selection)
        n = s.num_rows if selection is None else len(selection)
        out = self.torch.zeros(max(1, n), dtype=self.torch.int64, device=f"cuda:{self.device}")
        arr = abi.expr_array(keys, keep)
        self._check(L.ml_hash_columns(self.ctx, arr, len(keys), C.byref(b), out.data_ptr()))
        return out[:n].cpu().numpy().view(np.uint64).copy()

    def join_probe(self, jt: GpuJoin, keys: Sequence, src, selection: Optional[np.ndarray] = None):
        L, keep = lib(), abi.Keep()
        _, b = self._src(src, keep, selection)
        arr = abi.expr_array(keys, keep)
        out = C.c_void_p()
        self._check(L.ml_join_probe(self.ctx, jt.handle, arr, len(keys), C.byref(b), C.byref(out)))
        t = GpuResult(self, out.value).to_host()
        return t.columns[0].values, t.columns[1].values
