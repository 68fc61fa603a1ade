"""Host-side columnar tables (numpy) and their C-ABI views.

A ``HostTable`` mirrors MiniLake's ``Batch`` (crates/minilake-core/src/batch.rs:25-29):
equal-length typed columns, strings dictionary-encoded as u32 codes plus a
dictionary (``DictVec``, crates/minilake-core/src/column.rs:126-131).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import abi

NP_DTYPE = {abi.BOOLEAN: np.uint8, abi.INT32: np.int32, abi.INT64: np.int64, abi.FLOAT64: np.float64,
            abi.DATE: np.int32, abi.UTF8: np.uint32}


@dataclass
class HostColumn:
    type: int
    values: np.ndarray
    dictionary: Optional[List[str]] = None
    validity: Optional[np.ndarray] = None  # bool per row, None = all valid

    def __post_init__(self):
        self.values = np.ascontiguousarray(self.values, dtype=NP_DTYPE[self.type])
        if self.type == abi.UTF8 and self.dictionary is None:
            raise ValueError("UTF8 columns must carry a dictionary")

    def __len__(self):
        return len(self.values)

    def decoded(self):
        """Python values (strings decoded, NULL as None)."""
        if self.type == abi.UTF8:
            out = [self.dictionary[c] for c in self.values]
        else:
            out = self.values.tolist()
        if self.validity is not None:
            out = [v if ok else None for v, ok in zip(out, self.validity)]
        return out


@dataclass
class HostTable:
    columns: List[HostColumn]
    num_rows: int = -1
    names: List[str] = field(default_factory=list)

    def __post_init__(self):
        if self.num_rows < 0:
            self.num_rows = len(self.columns[0]) if self.columns else 0
        for c in self.columns:
            if len(c) != self.num_rows:
                raise ValueError("column length differs from table rows")

    @property
    def types(self):
        return [c.type for c in self.columns]

    def nbytes(self) -> int:
        return sum(c.values.nbytes for c in self.columns)

    def take(self, idx) -> "HostTable":
        return HostTable([HostColumn(c.type, c.values[idx], c.dictionary,
                                     None if c.validity is None else c.validity[idx]) for c in self.columns],
                         names=list(self.names))

    def head(self, n: int) -> "HostTable":
        return self.take(np.arange(min(n, self.num_rows)))

    def select(self, cols: Sequence[int]) -> "HostTable":
        names = [self.names[i] for i in cols] if self.names else []
        return HostTable([self.columns[i] for i in cols], names=names)


def make_dict(entries: Sequence[str], keep: abi.Keep) -> abi.MlDict:
    data = b"".join(e.encode() for e in entries)
    offs = [0]
    for e in entries:
        offs.append(offs[-1] + len(e.encode()))
    off_arr = keep((C.c_uint32 * len(offs))(*offs))
    byte_arr = keep((C.c_uint8 * max(1, len(data))).from_buffer_copy(data + b"\0"))
    return keep(abi.MlDict(len(entries), 0, off_arr, byte_arr))


def bitmap_from_bools(valid: np.ndarray) -> np.ndarray:
    """LSB-first u64 words (Bitmap, crates/minilake-core/src/bitmap.rs:9-12)."""
    n = len(valid)
    words = np.zeros((n + 63) // 64, dtype=np.uint64)
    idx = np.nonzero(valid)[0]
    np.bitwise_or.at(words, idx // 64, (np.uint64(1) << (idx % 64).astype(np.uint64)))
    return words


def bools_from_bitmap(ptr, n: int) -> np.ndarray:
    words = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint64)), shape=((n + 63) // 64,)).copy()
    bits = (words[np.arange(n) // 64] >> (np.arange(n) % 64).astype(np.uint64)) & np.uint64(1)
    return bits.astype(bool)


def host_batch(t: HostTable, keep: abi.Keep, selection: Optional[np.ndarray] = None) -> abi.MlBatch:
    """MlBatch over host memory (the oracle's input)."""
    cols = keep((abi.MlColumn * max(1, len(t.columns)))())
    dict_cache = {}
    for i, c in enumerate(t.columns):
        keep(c.values)
        cols[i].type = c.type
        cols[i].data = c.values.ctypes.data
        cols[i].len = len(c.values)
        if c.validity is not None and not np.all(c.validity):
            words = keep(bitmap_from_bools(c.validity))
            cols[i].validity = words.ctypes.data_as(C.POINTER(C.c_uint64))
        if c.type == abi.UTF8:
            key = id(c.dictionary)
            if key not in dict_cache:
                dict_cache[key] = make_dict(c.dictionary, keep)
            cols[i].dict = C.pointer(dict_cache[key])
    sel_ptr, sel_len = None, 0
    if selection is not None:
        sel = keep(np.ascontiguousarray(selection, dtype=np.uint32))
        sel_ptr, sel_len = sel.ctypes.data, len(sel)
    return keep(abi.MlBatch(cols, len(t.columns), 0, t.num_rows, sel_ptr, sel_len))


def column_from_ml(mc: abi.MlColumn, n: int) -> HostColumn:
    """Copy a host-memory MlColumn into a HostColumn."""
    dt = NP_DTYPE[mc.type]
    if n:
        vals = np.ctypeslib.as_array(C.cast(mc.data, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,)).copy()
    else:
        vals = np.zeros(0, dtype=dt)
    valid = bools_from_bitmap(mc.validity, n) if mc.validity else None
    d = None
    if mc.type == abi.UTF8:
        md = mc.dict.contents
        offs = [md.offsets[i] for i in range(md.len + 1)]
        raw = C.string_at(md.bytes, offs[-1]) if offs[-1] else b""
        d = [raw[offs[i]:offs[i + 1]].decode() for i in range(md.len)]
    return HostColumn(mc.type, vals, d, valid)
