"""Order-independent result comparison.

Integer, date, boolean and string (decoded) columns must match bit-exactly;
FLOAT64 columns within a relative tolerance (1e-12 by default), with NULLs
matching exactly. Rows are matched by sorting on the exact columns.
"""

from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

from minilake_amd import abi
from minilake_amd.table import HostTable


def rows_of(t: HostTable) -> List[tuple]:
    cols = [c.decoded() for c in t.columns]
    return list(zip(*cols)) if cols else []


def _key(row, float_cols):
    return tuple((v is None, 0 if v is None else v) if i not in float_cols else (False, 0)
                 for i, v in enumerate(row))


def float_close(a, b, rtol=1e-12, atol=0.0):
    if a is None or b is None:
        return a is None and b is None
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return abs(a - b) <= max(rtol * max(abs(a), abs(b)), atol)


def assert_tables_equal(got: HostTable, want: HostTable, rtol=1e-12, atol=0.0, ordered=False):
    assert got.num_rows == want.num_rows, f"row count {got.num_rows} != {want.num_rows}"
    assert len(got.columns) == len(want.columns), "column count differs"
    for i, (g, w) in enumerate(zip(got.columns, want.columns)):
        assert g.type == w.type, f"column {i}: type {abi.TYPE_NAMES[g.type]} != {abi.TYPE_NAMES[w.type]}"
    float_cols = {i for i, c in enumerate(want.columns) if c.type == abi.FLOAT64}
    gr, wr = rows_of(got), rows_of(want)
    if rtol == 0 and atol == 0:
        float_cols_key = set()  # exact comparison: order rows by every column
    else:
        float_cols_key = float_cols
    if not ordered:
        gr = sorted(gr, key=lambda r: _key(r, float_cols_key))
        wr = sorted(wr, key=lambda r: _key(r, float_cols_key))
    for k, (a, b) in enumerate(zip(gr, wr)):
        for i, (x, y) in enumerate(zip(a, b)):
            if i in float_cols:
                assert float_close(x, y, rtol, atol), f"row {k} col {i}: {x!r} != {y!r} (rel tol {rtol})"
            else:
                assert x == y, f"row {k} col {i}: {x!r} != {y!r}"


def concat_tables(parts: Sequence[HostTable]) -> HostTable:
    """Row-wise concatenation of tables of one schema; string columns are
    re-encoded over one dictionary of distinct entries."""
    from minilake_amd.table import HostColumn
    cols = []
    for i, c0 in enumerate(parts[0].columns):
        cs = [p.columns[i] for p in parts]
        valid = None
        if any(c.validity is not None for c in cs):
            valid = np.concatenate([np.ones(len(c), bool) if c.validity is None else np.asarray(c.validity, bool)
                                    for c in cs])
        if c0.type == abi.UTF8:
            strings = [c.dictionary[k] for c in cs for k in c.values]
            entries = sorted(set(strings))
            code = {e: k for k, e in enumerate(entries)}
            cols.append(HostColumn(c0.type, np.array([code[x] for x in strings], np.uint32), entries, valid))
        else:
            cols.append(HostColumn(c0.type, np.concatenate([c.values for c in cs]), None, valid))
    return HostTable(cols)
