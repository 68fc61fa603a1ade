"""ctypes mirror of ``include/minilake_gpu.h`` plus a small plan-building DSL.

The structs below match the C header field for field. ``Expr`` builds
PhysicalExpr trees (crates/minilake-exec/src/expr/mod.rs:109-173) and flattens
them to the postfix programs the C ABI takes. Literal typing follows the
reference binder: integers are Int64 unless coerced to the other operand's
Int32/Date type, decimal literals become Float64 at run time
(crates/minilake-sql/src/binder.rs:787-802,934-953; expr/eval.rs:52-57).
"""

from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence

# ---------------------------------------------------------------- constants
ML_OK = 0
ML_ERR_PARSE = 1
ML_ERR_PLAN = 2
ML_ERR_EXECUTION = 3
ML_ERR_UNSUPPORTED = 4
ML_ERR_IO = 5
ML_ERR_PARQUET = 6
ML_ERR_RESOURCES_EXHAUSTED = 7
ML_ERR_CANCELLED = 8
ML_ERR_INTERNAL = 9
ML_ERR_CAPACITY = 10

STATUS_NAMES = {
    0: "Ok", 1: "Parse", 2: "Plan", 3: "Execution", 4: "Unsupported", 5: "Io",
    6: "Parquet", 7: "ResourcesExhausted", 8: "Cancelled", 9: "Internal", 10: "Capacity",
}

BOOLEAN, INT32, INT64, FLOAT64, DATE, UTF8, NULL = 0, 1, 2, 3, 4, 5, -1
TYPE_NAMES = {0: "BOOLEAN", 1: "INT32", 2: "INT64", 3: "FLOAT64", 4: "DATE", 5: "UTF8", -1: "NULL"}
TYPE_WIDTH = {BOOLEAN: 1, INT32: 4, INT64: 8, FLOAT64: 8, DATE: 4, UTF8: 4}

EXPR_COLUMN, EXPR_LITERAL, EXPR_BINARY, EXPR_NOT, EXPR_NEGATE, EXPR_CAST = range(6)
(OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_EQ, OP_NOTEQ, OP_LT, OP_LTEQ, OP_GT, OP_GTEQ,
 OP_AND, OP_OR) = range(13)

AGG_COUNT_STAR, AGG_COUNT, AGG_SUM, AGG_AVG, AGG_MIN, AGG_MAX = range(6)
AGG_MODE_SINGLE, AGG_MODE_PARTIAL, AGG_MODE_FINAL = range(3)
SIZING_PLAIN, SIZING_HINT, SIZING_LEARNED = range(3)
PIPE_FILTER, PIPE_PROBE, PIPE_PROJECT = range(3)

MLP_IN, MLP_H1, MLP_H2, MLP_OUT = 32, 64, 64, 2
MLP_NUM_PARAMS = MLP_IN * MLP_H1 + MLP_H1 + MLP_H1 * MLP_H2 + MLP_H2 + MLP_H2 * MLP_OUT + MLP_OUT


# ------------------------------------------------------------------ structs
class MlDict(C.Structure):
    _fields_ = [("len", C.c_int32), ("_pad", C.c_int32),
                ("offsets", C.POINTER(C.c_uint32)), ("bytes", C.POINTER(C.c_uint8))]


class MlColumnStats(C.Structure):
    _fields_ = [("type", C.c_int32), ("has_min", C.c_int32), ("has_max", C.c_int32), ("has_null_count", C.c_int32),
                ("null_count", C.c_int64), ("min_i64", C.c_int64), ("max_i64", C.c_int64),
                ("min_f64", C.c_double), ("max_f64", C.c_double), ("min_str", C.c_void_p), ("min_len", C.c_int64),
                ("max_str", C.c_void_p), ("max_len", C.c_int64)]


class MlPrunePredicate(C.Structure):
    _fields_ = [("column", C.c_int32), ("op", C.c_int32), ("type", C.c_int32), ("_pad", C.c_int32),
                ("i64", C.c_int64), ("f64", C.c_double), ("str", C.c_char_p), ("str_len", C.c_int64)]


PRUNE_OPS = {"=": 0, "<": 1, "<=": 2, ">": 3, ">=": 4}


class MlColumn(C.Structure):
    _fields_ = [("type", C.c_int32), ("_pad", C.c_int32), ("data", C.c_void_p), ("len", C.c_int64),
                ("validity", C.POINTER(C.c_uint64)), ("dict", C.POINTER(MlDict))]


class MlBatch(C.Structure):
    _fields_ = [("columns", C.POINTER(MlColumn)), ("num_columns", C.c_int32), ("_pad", C.c_int32),
                ("num_rows", C.c_int64), ("selection", C.c_void_p), ("selection_len", C.c_int64)]


class MlExprNode(C.Structure):
    _fields_ = [("kind", C.c_int32), ("op", C.c_int32), ("type", C.c_int32), ("column", C.c_int32),
                ("i64", C.c_int64), ("f64", C.c_double), ("str", C.c_char_p), ("str_len", C.c_int32),
                ("_pad", C.c_int32)]


class MlExpr(C.Structure):
    _fields_ = [("nodes", C.POINTER(MlExprNode)), ("num_nodes", C.c_int32), ("_pad", C.c_int32)]


class MlExecConfig(C.Structure):
    _fields_ = [("batch_rows", C.c_int64), ("device", C.c_int32), ("block_size", C.c_int32),
                ("grid_per_cu", C.c_int32), ("use_bloom", C.c_int32), ("stream", C.c_void_p)]


class MlAggregate(C.Structure):
    _fields_ = [("func", C.c_int32), ("arg_type", C.c_int32), ("arg", MlExpr)]


class MlAggSpec(C.Structure):
    _fields_ = [("group_exprs", C.POINTER(MlExpr)), ("key_types", C.POINTER(C.c_int32)),
                ("num_keys", C.c_int32), ("num_aggs", C.c_int32), ("aggs", C.POINTER(MlAggregate)),
                ("mode", C.c_int32), ("sizing", C.c_int32), ("capacity_hint", C.c_int64)]


class MlJoinBuildSpec(C.Structure):
    _fields_ = [("keys", C.POINTER(MlExpr)), ("key_types", C.POINTER(C.c_int32)),
                ("num_keys", C.c_int32), ("_pad", C.c_int32)]


class MlOperator(C.Structure):
    _fields_ = [("kind", C.c_int32), ("build_is_left", C.c_int32), ("predicate", MlExpr),
                ("table", C.c_void_p), ("keys", C.POINTER(MlExpr)), ("num_keys", C.c_int32),
                ("num_exprs", C.c_int32), ("exprs", C.POINTER(MlExpr))]


class MlSortKey(C.Structure):
    _fields_ = [("expr", MlExpr), ("asc", C.c_int32), ("nulls_first", C.c_int32)]


class MlSortSpec(C.Structure):
    _fields_ = [("keys", C.POINTER(MlSortKey)), ("num_keys", C.c_int32), ("_pad", C.c_int32),
                ("limit", C.c_int64)]


class MlFragment(C.Structure):
    _fields_ = [("build_keys", C.c_void_p), ("build_groups", C.c_void_p), ("build_validity", C.c_void_p),
                ("build_rows", C.c_int64), ("probe_keys", C.c_void_p), ("probe_values", C.c_void_p),
                ("probe_validity", C.c_void_p), ("probe_rows", C.c_int64), ("num_groups", C.c_int64),
                ("out_sums", C.c_void_p), ("out_counts", C.c_void_p)]


class MlLaunchStat(C.Structure):
    _fields_ = [("sink", C.c_int32), ("_pad", C.c_int32), ("rows", C.c_int64), ("ms", C.c_double)]


class MlFragmentTuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("tile_rows", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("lds_groups", C.c_int32), ("probe_width", C.c_int32), ("lds_cache", C.c_int32)]


# --------------------------------------------------------------- expressions
class Expr:
    """A PhysicalExpr tree node; ``postfix()`` flattens it for the C ABI."""

    __slots__ = ("kind", "op", "type", "column", "value", "children")

    def __init__(self, kind, op=0, type=0, column=0, value=None, children=()):
        self.kind, self.op, self.type, self.column = kind, op, type, column
        self.value, self.children = value, tuple(children)

    # arithmetic
    def __add__(self, o): return binary(self, OP_ADD, o)
    def __radd__(self, o): return binary(o, OP_ADD, self)
    def __sub__(self, o): return binary(self, OP_SUB, o)
    def __rsub__(self, o): return binary(o, OP_SUB, self)
    def __mul__(self, o): return binary(self, OP_MUL, o)
    def __rmul__(self, o): return binary(o, OP_MUL, self)
    def __truediv__(self, o): return binary(self, OP_DIV, o)
    def __rtruediv__(self, o): return binary(o, OP_DIV, self)
    def __mod__(self, o): return binary(self, OP_MOD, o)
    def __neg__(self): return Expr(EXPR_NEGATE, children=(self,))
    def __and__(self, o): return binary(self, OP_AND, o)
    def __or__(self, o): return binary(self, OP_OR, o)
    def __invert__(self): return Expr(EXPR_NOT, children=(self,))

    # comparisons are methods so Expr stays hashable and usable in lists
    def eq(self, o): return binary(self, OP_EQ, o)
    def ne(self, o): return binary(self, OP_NOTEQ, o)
    def lt(self, o): return binary(self, OP_LT, o)
    def le(self, o): return binary(self, OP_LTEQ, o)
    def gt(self, o): return binary(self, OP_GT, o)
    def ge(self, o): return binary(self, OP_GTEQ, o)

    def cast(self, to): return Expr(EXPR_CAST, type=to, children=(self,))

    def postfix(self) -> List["Expr"]:
        out: List[Expr] = []
        for c in self.children:
            out.extend(c.postfix())
        out.append(self)
        return out

    def __repr__(self):
        if self.kind == EXPR_COLUMN:
            return f"#{self.column}"
        if self.kind == EXPR_LITERAL:
            return f"{self.value!r}:{TYPE_NAMES[self.type]}"
        if self.kind == EXPR_BINARY:
            sym = ["+", "-", "*", "/", "%", "=", "<>", "<", "<=", ">", ">=", "AND", "OR"][self.op]
            return f"({self.children[0]!r} {sym} {self.children[1]!r})"
        if self.kind == EXPR_CAST:
            return f"CAST({self.children[0]!r} AS {TYPE_NAMES[self.type]})"
        return f"{['', '', '', 'NOT', '-', ''][self.kind]}({self.children[0]!r})"


def col(i: int) -> Expr:
    return Expr(EXPR_COLUMN, column=i)


def lit(v, type: Optional[int] = None) -> Expr:
    """Literal; default typing: bool->BOOLEAN, int->INT64, float->FLOAT64, str->UTF8, None->NULL."""
    if type is None:
        if v is None:
            type = NULL
        elif isinstance(v, bool):
            type = BOOLEAN
        elif isinstance(v, int):
            type = INT64
        elif isinstance(v, float):
            type = FLOAT64
        elif isinstance(v, str):
            type = UTF8
        else:
            raise TypeError(f"unsupported literal {v!r}")
    return Expr(EXPR_LITERAL, type=type, value=v)


def date_lit(days: int) -> Expr:
    return lit(int(days), DATE)


def _as_expr(x) -> Expr:
    return x if isinstance(x, Expr) else lit(x)


def binary(l, op, r) -> Expr:
    return Expr(EXPR_BINARY, op=op, children=(_as_expr(l), _as_expr(r)))


class Keep:
    """Holds every ctypes object a plan points into, so nothing is freed early."""

    def __init__(self):
        self.objs = []

    def __call__(self, o):
        self.objs.append(o)
        return o


def to_ml_expr(e: Optional[Expr], keep: Keep) -> MlExpr:
    if e is None:
        return MlExpr(None, 0, 0)
    nodes = e.postfix()
    arr = keep((MlExprNode * len(nodes))())
    for i, n in enumerate(nodes):
        a = arr[i]
        a.kind, a.op, a.type, a.column = n.kind, n.op, n.type, n.column
        if n.kind == EXPR_LITERAL and n.value is not None:
            if n.type == FLOAT64:
                a.f64 = float(n.value)
            elif n.type == UTF8:
                b = keep(C.create_string_buffer(n.value.encode()))
                a.str = C.cast(b, C.c_char_p)
                a.str_len = len(n.value.encode())
            else:
                a.i64 = int(n.value)
    return MlExpr(arr, len(nodes), 0)


def expr_array(exprs: Sequence[Expr], keep: Keep):
    arr = keep((MlExpr * max(1, len(exprs)))())
    for i, e in enumerate(exprs):
        arr[i] = to_ml_expr(e, keep)
    return arr


def int_array(vals: Iterable[int], keep: Keep, ctype=C.c_int32):
    vals = list(vals)
    arr = keep((ctype * max(1, len(vals)))(*vals))
    return arr


# ------------------------------------------------------------ plan pieces
class Agg:
    """One aggregate (AggregateExpr, crates/minilake-exec/src/operators/aggregate/mod.rs:22-46)."""

    def __init__(self, func: int, arg: Optional[Expr] = None, arg_type: int = INT64):
        self.func, self.arg, self.arg_type = func, arg, arg_type


def count_star() -> Agg:
    return Agg(AGG_COUNT_STAR, None, INT64)


def agg_spec(group_exprs: Sequence[Expr], key_types: Sequence[int], aggs: Sequence[Agg], keep: Keep,
             sizing: int = SIZING_PLAIN, capacity_hint: int = 0, mode: int = AGG_MODE_SINGLE) -> MlAggSpec:
    """AggMode (aggregate/mod.rs:50-57): SINGLE and PARTIAL evaluate group_exprs
    and the aggregate arguments over the input; FINAL reads the input by
    position instead, the group keys then each aggregate's state columns (what
    a PARTIAL aggregate with the same aggregates emits)."""
    ga = expr_array(group_exprs, keep)
    kt = int_array(key_types, keep)
    aa = keep((MlAggregate * max(1, len(aggs)))())
    for i, a in enumerate(aggs):
        aa[i].func = a.func
        aa[i].arg_type = a.arg_type
        aa[i].arg = to_ml_expr(a.arg, keep)
    return MlAggSpec(ga, kt, len(group_exprs), len(aggs), aa, mode, sizing, capacity_hint)


def join_build_spec(keys: Sequence[Expr], key_types: Sequence[int], keep: Keep) -> MlJoinBuildSpec:
    return MlJoinBuildSpec(expr_array(keys, keep), int_array(key_types, keep), len(keys), 0)


def sort_spec(keys: Sequence[tuple], limit: int, keep: Keep) -> MlSortSpec:
    """keys: (expr, asc, nulls_first) triples."""
    arr = keep((MlSortKey * max(1, len(keys)))())
    for i, (e, asc, nf) in enumerate(keys):
        arr[i].expr = to_ml_expr(e, keep)
        arr[i].asc = int(asc)
        arr[i].nulls_first = int(nf)
    return MlSortSpec(arr, len(keys), 0, limit)


class Op:
    """A pipeline operator: filter, probe (of a built join table) or project."""

    def __init__(self, kind, predicate=None, table=None, keys=(), build_is_left=False, exprs=()):
        self.kind, self.predicate, self.table = kind, predicate, table
        self.keys, self.build_is_left, self.exprs = list(keys), build_is_left, list(exprs)


def filter_op(pred: Expr) -> Op:
    return Op(PIPE_FILTER, predicate=pred)


def probe_op(table, keys: Sequence[Expr], build_is_left: bool = False) -> Op:
    return Op(PIPE_PROBE, table=table, keys=keys, build_is_left=build_is_left)


def project_op(exprs: Sequence[Expr]) -> Op:
    return Op(PIPE_PROJECT, exprs=exprs)


def op_array(ops: Sequence[Op], keep: Keep, table_handle=lambda t: t):
    arr = keep((MlOperator * max(1, len(ops)))())
    for i, o in enumerate(ops):
        a = arr[i]
        a.kind = o.kind
        a.build_is_left = int(o.build_is_left)
        a.predicate = to_ml_expr(o.predicate, keep)
        a.table = table_handle(o.table) if o.table is not None else None
        a.keys = expr_array(o.keys, keep)
        a.num_keys = len(o.keys)
        a.exprs = expr_array(o.exprs, keep)
        a.num_exprs = len(o.exprs)
    return arr
