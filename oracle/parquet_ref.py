"""CPU ORACLE for the device Parquet scan (test infrastructure only: loaded by
tests/, never by the product path).

MiniLake reads Parquet through the `parquet` crate, version 60.0.0 (Cargo.lock),
which is not part of the reference tree. This module restates the published
Apache Parquet format for the subset the scan supports (flat schema,
UNCOMPRESSED and SNAPPY chunks, data pages v1 and v2, the PLAIN, dictionary, RLE
(booleans), DELTA_BINARY_PACKED, DELTA_LENGTH_BYTE_ARRAY, DELTA_BYTE_ARRAY and
BYTE_STREAM_SPLIT encodings, RLE / bit-packed hybrid definition levels and
indices) and MiniLake's own code on top of it:

- Conversion::for_column and the value conversions of read_column_chunk
  (crates/minilake-storage/src/parquet_reader.rs:55-80, 210-287): DATE stays
  days, DECIMAL -> f64 as unscaled / 10f64.powi(scale) (decimal bytes through
  be_i128, parquet_reader.rs:156-167), FLOAT -> f64, strings
  as Python str (encode_strings' dictionary choice is a representation detail);
- convert_stats (parquet_reader.rs:290-358);
- PrunePredicate::may_match with ScalarValue::compare (stats.rs:53-70,
  scalar.rs:115-131) and Table::prune (table.rs:165-172).

Plain Python loops: meant for files of up to a few hundred thousand values.
"""

from __future__ import annotations

import struct
from typing import List, Optional

T_BOOLEAN, T_INT32, T_INT64, T_INT96, T_FLOAT, T_DOUBLE, T_BYTE_ARRAY, T_FLBA = range(8)
CT_DECIMAL, CT_DATE = 5, 6


class ParquetError(Exception):
    pass


class Unsupported(Exception):
    pass


class _Thrift:
    """Thrift compact protocol reader."""

    def __init__(self, b: bytes, pos: int = 0):
        self.b, self.pos = b, pos

    def byte(self) -> int:
        if self.pos >= len(self.b):
            raise ParquetError("metadata ends early")
        x = self.b[self.pos]
        self.pos += 1
        return x

    def varint(self) -> int:
        v, s = 0, 0
        while True:
            x = self.byte()
            v |= (x & 0x7F) << s
            if not x & 0x80:
                return v
            s += 7

    def zz(self) -> int:
        v = self.varint()
        return (v >> 1) ^ -(v & 1)

    def binary(self) -> bytes:
        n = self.varint()
        out = self.b[self.pos:self.pos + n]
        if len(out) != n:
            raise ParquetError("binary past the end")
        self.pos += n
        return out

    def read(self, t: int):
        """Value of a field or element of type t (structs as {id: value})."""
        if t in (1, 2):
            return t == 1
        if t == 3:
            return struct.unpack("<b", bytes([self.byte()]))[0]
        if t in (4, 5, 6):
            return self.zz()
        if t == 7:
            v = struct.unpack_from("<d", self.b, self.pos)[0]
            self.pos += 8
            return v
        if t == 8:
            return self.binary()
        if t in (9, 10):
            h = self.byte()
            n, et = h >> 4, h & 0x0F
            if n == 15:
                n = self.varint()
            return [(self.byte() == 1) if et in (1, 2) else self.read(et) for _ in range(n)]
        if t == 11:
            n = self.varint()
            if n == 0:
                return {}
            kv = self.byte()
            return {self.read(kv >> 4): self.read(kv & 0x0F) for _ in range(n)}
        if t == 12:
            out, last = {}, 0
            while True:
                h = self.byte()
                if h == 0:
                    return out
                d = h >> 4
                fid = last + d if d else self.zz()
                last = fid
                out[fid] = self.read(h & 0x0F)
        raise ParquetError(f"thrift type {t}")


def _pow10i(s: int) -> float:
    """10f64.powi(s) by square-and-multiply (compiler-rt __powidf2)."""
    recip, b, a, r = s < 0, abs(s), 10.0, 1.0
    while True:
        if b & 1:
            r *= a
        b //= 2
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


class Column:
    def __init__(self, el: dict):
        self.name = el.get(4, b"").decode()
        self.phys = el.get(1)
        self.type_length = el.get(2, 0)
        self.optional = el.get(3, 0) == 1
        self.converted = el.get(6)
        self.scale = el.get(7, 0)
        if el.get(3) == 2 or el.get(5, 0) > 0:
            raise Unsupported(f"nested parquet column '{self.name}'")
        c, p = self.converted, self.phys
        # Conversion::for_column (parquet_reader.rs:55-80)
        if p == T_BOOLEAN:
            self.conv = "bool"
        elif p == T_INT32:
            self.conv = "date" if c == CT_DATE else ("dec_i32" if c == CT_DECIMAL else "int32")
        elif p == T_INT64:
            self.conv = "dec_i64" if c == CT_DECIMAL else "int64"
        elif p == T_FLOAT:
            self.conv = "float"
        elif p == T_DOUBLE:
            self.conv = "double"
        elif p in (T_BYTE_ARRAY, T_FLBA) and c == CT_DECIMAL:
            self.conv = "dec_bytes"
        elif p == T_BYTE_ARRAY:
            self.conv = "utf8"
        else:
            raise Unsupported(f"parquet column '{self.name}' type")

    @property
    def ml_type(self) -> str:
        return {"bool": "bool", "int32": "int32", "date": "date", "int64": "int64", "utf8": "utf8"}.get(self.conv,
                                                                                                     "float64")


def snappy_decompress(b: bytes) -> bytes:
    """Snappy, from the format's published description: varint length, then
    literals (tag & 3 == 0) and back-references with 1-, 2- or 4-byte offsets
    (tag & 3 == 1, 2, 3), copied forward byte by byte."""
    i, n, shift = 0, 0, 0
    while True:
        x = b[i]
        i += 1
        n |= (x & 0x7F) << shift
        if not x & 0x80:
            break
        shift += 7
    out = bytearray()
    while i < len(b):
        tag = b[i]
        i += 1
        kind = tag & 3
        if kind == 0:
            ln = tag >> 2
            if ln >= 60:
                nb = ln - 59
                ln = int.from_bytes(b[i:i + nb], "little")
                i += nb
            ln += 1
            out += b[i:i + ln]
            i += ln
            continue
        if kind == 1:
            ln, off = 4 + ((tag >> 2) & 7), ((tag >> 5) << 8) | b[i]
            i += 1
        elif kind == 2:
            ln, off = (tag >> 2) + 1, int.from_bytes(b[i:i + 2], "little")
            i += 2
        else:
            ln, off = (tag >> 2) + 1, int.from_bytes(b[i:i + 4], "little")
            i += 4
        if off == 0 or off > len(out):
            raise ParquetError("snappy back-reference")
        for _ in range(ln):
            out.append(out[-off])
    if len(out) != n:
        raise ParquetError("snappy length")
    return bytes(out)


def _hybrid(b: bytes, pos: int, end: int, bw: int, want: int) -> List[int]:
    """RLE / bit-packed hybrid values (at most `want`)."""
    out: List[int] = []
    vbytes = (bw + 7) // 8
    t = _Thrift(b[:end], pos)
    while len(out) < want and t.pos < end:
        h = t.varint()
        if h & 1:
            groups = h >> 1
            nbytes = groups * bw
            chunk = int.from_bytes(b[t.pos:t.pos + nbytes], "little")
            t.pos += nbytes
            mask = (1 << bw) - 1
            for j in range(groups * 8):
                out.append((chunk >> (j * bw)) & mask if bw else 0)
        else:
            v = int.from_bytes(b[t.pos:t.pos + vbytes], "little")
            t.pos += vbytes
            out.extend([v] * (h >> 1))
    return out[:want]


def be_i128(r: bytes) -> int:
    """be_i128 (parquet_reader.rs:156-167): big-endian two's complement of the
    first 16 bytes at most, the sign taken from the first byte; no bytes -> 0."""
    v = -1 if r and r[0] & 0x80 else 0
    for x in r[:16]:
        v = (v << 8) | x
    v &= (1 << 128) - 1
    return v - (1 << 128) if v >> 127 else v


def delta_binary_packed(b: bytes, pos: int, end: int, bits: int):
    """DELTA_BINARY_PACKED, from the format's published description: a header
    <block size> <miniblocks per block> <total count> <first value (zigzag)>,
    then blocks of <min delta (zigzag)> <a bit width byte per miniblock>
    <miniblocks, bit-packed LSB first>; each value is the previous one + min
    delta + the unpacked number, wrapping at `bits` bits. Miniblocks the last
    block does not need take no bytes. Returns the values as unsigned `bits`-bit
    integers and the position after the last miniblock read."""
    t = _Thrift(b[:end], pos)
    block, nmb, total = t.varint(), t.varint(), t.varint()
    first = t.zz()
    if block == 0 or block % 128 or nmb == 0 or block % nmb or (block // nmb) % 32:
        raise ParquetError("delta header")
    per, mask = block // nmb, (1 << bits) - 1
    out: List[int] = []
    if total == 0:
        return out, t.pos
    v = first & mask
    out.append(v)
    while len(out) < total:
        md = t.zz()
        widths = [t.byte() for _ in range(nmb)]
        for w in widths:
            if len(out) >= total:
                break
            if w > bits:
                raise ParquetError("delta bit width")
            nbytes = per * w // 8
            if t.pos + nbytes > end:
                raise ParquetError("delta miniblock past the page")
            chunk = int.from_bytes(b[t.pos:t.pos + nbytes], "little")
            t.pos += nbytes
            for j in range(min(per, total - len(out))):
                v = (v + md + ((chunk >> (j * w)) & ((1 << w) - 1))) & mask
                out.append(v)
    return out, t.pos


def delta_length_byte_array(b: bytes, pos: int, end: int):
    """DELTA_LENGTH_BYTE_ARRAY: the lengths as DELTA_BINARY_PACKED (32-bit),
    then the values' bytes back to back."""
    lens, p = delta_binary_packed(b, pos, end, 32)
    out = []
    for ln in lens:
        if ln >> 31 or p + ln > end:
            raise ParquetError("delta length")
        out.append(b[p:p + ln])
        p += ln
    return out, p


def delta_byte_array(b: bytes, pos: int, end: int):
    """DELTA_BYTE_ARRAY: prefix lengths as DELTA_BINARY_PACKED, then the
    suffixes as DELTA_LENGTH_BYTE_ARRAY; value k = the first prefix[k] bytes of
    value k-1, then suffix k."""
    pre, p = delta_binary_packed(b, pos, end, 32)
    suf, p = delta_length_byte_array(b, p, end)
    if len(pre) != len(suf):
        raise ParquetError("delta prefix count")
    out, prev = [], b""
    for k, s in zip(pre, suf):
        if k > len(prev):
            raise ParquetError("delta prefix")
        prev = prev[:k] + s
        out.append(prev)
    return out, p


class ParquetFile:
    """Table::open (table.rs:59-134) over one file's bytes."""

    def __init__(self, data: bytes):
        if len(data) < 12 or data[:4] != b"PAR1" or data[-4:] != b"PAR1":
            raise ParquetError("not a parquet file")
        flen = struct.unpack_from("<I", data, len(data) - 8)[0]
        meta = _Thrift(data, len(data) - 8 - flen).read(12)
        self.data = data
        schema = meta[2]
        self.columns = [Column(el) for el in schema[1:]]
        if schema[0].get(5, 0) != len(self.columns):
            raise Unsupported("nested parquet schema")
        self.row_groups = []
        for rg in meta.get(4, []):
            chunks = []
            for cc in rg[1]:
                if 1 in cc:
                    raise Unsupported("column chunks in other files")
                chunks.append(cc[3])
            self.row_groups.append((rg[3], chunks))

    # convert_stats (parquet_reader.rs:290-358)
    def stats(self, rg: int, c: int) -> dict:
        col = self.columns[c]
        md = self.row_groups[rg][1][c]
        st = md.get(12)
        out = {"min": None, "max": None, "null_count": None}
        if st is None:
            return out
        out["null_count"] = st.get(3)

        def conv(raw: Optional[bytes], exact: bool):
            if raw is None or not exact:
                return None
            if col.conv in ("int32", "date", "dec_i32") and md[1] == T_INT32 and len(raw) == 4:
                x = struct.unpack("<i", raw)[0]
                return x / _pow10i(col.scale) if col.conv == "dec_i32" else x
            if col.conv in ("int64", "dec_i64") and md[1] == T_INT64 and len(raw) == 8:
                x = struct.unpack("<q", raw)[0]
                return x / _pow10i(col.scale) if col.conv == "dec_i64" else x
            if col.conv == "double" and len(raw) == 8:
                return struct.unpack("<d", raw)[0]
            if col.conv == "float" and len(raw) == 4:
                return float(struct.unpack("<f", raw)[0])
            if col.conv == "utf8":
                try:
                    return raw.decode("utf-8")
                except UnicodeDecodeError:
                    return None
            return None

        out["min"] = conv(st.get(6), st.get(8, True))
        out["max"] = conv(st.get(5), st.get(7, True))
        return out

    # read_column_chunk (parquet_reader.rs:210-287): Python values, None = NULL
    def read_column(self, rg: int, c: int) -> list:
        col = self.columns[c]
        nrows, chunks = self.row_groups[rg]
        md = chunks[c]
        codec = md.get(4, 0)
        if codec not in (0, 1):
            raise Unsupported(f"codec {codec}")
        start = md.get(11) if md.get(11) is not None and md.get(11) < md[9] else md[9]
        raw = self.data[start:start + md[7]]
        # the pages' bodies, uncompressed, one after another (headers dropped);
        # a v2 page keeps its levels uncompressed in front of its values
        b, bodies, pos = bytearray(), [], 0
        while pos < len(raw):
            t = _Thrift(raw, pos)
            h = t.read(12)
            body = raw[t.pos:t.pos + h[3]]
            if h[1] == 3:
                lv = h[8].get(5, 0) + h[8].get(6, 0)
                vals = body[lv:]
                if codec == 1 and h[8].get(7, True):
                    vals = snappy_decompress(vals)
                body = body[:lv] + vals
            elif codec == 1:
                body = snappy_decompress(body)
            bodies.append((h, len(b), len(b) + len(body)))
            b += body
            pos = t.pos + h[3]
        b = bytes(b)
        width = col.type_length if col.conv == "dec_bytes" else (8 if col.phys in (T_INT64, T_DOUBLE) else 4)
        dictionary = None
        out: list = []
        for h, body, end in bodies:
            if len(out) >= nrows:
                break
            if h[1] == 2:
                dictionary = self._plain(col, b, body, end, h[7][1], width)
            elif h[1] in (0, 3):
                if h[1] == 0:
                    rows, enc, v = h[5][1], h[5][2], body
                    if col.optional:
                        dl = struct.unpack_from("<I", b, v)[0]
                        levels = _hybrid(b, v + 4, v + 4 + dl, 1, rows)
                        v += 4 + dl
                else:  # v2: repetition levels, definition levels (no length prefix), values
                    v2 = h[8]
                    rows, enc = v2[3], v2[4]
                    if v2[1] != rows:
                        raise Unsupported("repeated values")
                    v = body + v2.get(6, 0)
                    if col.optional:
                        levels = _hybrid(b, v, v + v2.get(5, 0), 1, rows)
                    v += v2.get(5, 0)
                valid = [lv == 1 for lv in levels] if col.optional else [True] * rows
                vals = self._values(col, enc, b, v, end, sum(valid), width, dictionary)
                it = iter(vals)
                out.extend(next(it) if ok else None for ok in valid)
        if len(out) != nrows:
            raise ParquetError(f"expected {nrows} rows, decoded {len(out)}")
        return out

    def _values(self, col: Column, enc: int, b: bytes, v: int, end: int, n: int, width: int, dictionary) -> list:
        """The `n` non-NULL values of a data page in encoding `enc`."""
        conv = col.conv
        if enc == 0:
            return self._plain(col, b, v, end, n, width)
        if enc in (2, 8):
            return [dictionary[k] for k in _hybrid(b, v + 1, end, b[v], n)]
        if enc == 3 and conv == "bool":  # RLE: u32 length, then the hybrid of bit width 1
            ln = struct.unpack_from("<I", b, v)[0]
            return [x == 1 for x in _hybrid(b, v + 4, v + 4 + ln, 1, n)]
        if enc == 5 and col.phys in (T_INT32, T_INT64):
            bits = 32 if col.phys == T_INT32 else 64
            ints, _ = delta_binary_packed(b, v, end, bits)
            if len(ints) != n:
                raise ParquetError("delta value count")
            return self._ints(col, [x - (1 << bits) if x >> (bits - 1) else x for x in ints])
        if (enc == 6 and col.phys == T_BYTE_ARRAY) or (enc == 7 and col.phys in (T_BYTE_ARRAY, T_FLBA)):
            raws, _ = (delta_length_byte_array if enc == 6 else delta_byte_array)(b, v, end)
            if len(raws) != n:
                raise ParquetError("delta value count")
            if conv == "utf8":
                return [r.decode("utf-8", errors="surrogateescape") for r in raws]
            if col.phys == T_FLBA and any(len(r) != width for r in raws):
                raise ParquetError("fixed-length value of another length")
            return [float(be_i128(r)) / _pow10i(col.scale) for r in raws]
        if enc == 9 and conv not in ("bool", "utf8") and col.phys != T_BYTE_ARRAY:
            if (end - v) % width:
                raise ParquetError("byte stream split size")
            m = (end - v) // width
            if m < n:
                raise ParquetError("byte stream split count")
            return self._fixed(col, [bytes(b[v + k * m + j] for k in range(width)) for j in range(n)])
        raise Unsupported(f"encoding {enc}")

    def _plain(self, col: Column, b: bytes, pos: int, end: int, n: int, width: int) -> list:
        conv = col.conv
        if conv == "bool":
            return [bool((b[pos + j // 8] >> (j % 8)) & 1) for j in range(n)]
        if conv == "utf8" or col.phys == T_BYTE_ARRAY:
            vals = []
            for _ in range(n):
                ln = struct.unpack_from("<I", b, pos)[0]
                vals.append(b[pos + 4:pos + 4 + ln])
                pos += 4 + ln
            if conv == "utf8":
                return [r.decode("utf-8", errors="surrogateescape") for r in vals]
            return [float(be_i128(r)) / _pow10i(col.scale) for r in vals]
        return self._fixed(col, [b[pos + j * width:pos + (j + 1) * width] for j in range(n)])

    def _fixed(self, col: Column, raw: list) -> list:
        """Fixed-width little-endian values (big-endian for decimal bytes)."""
        conv = col.conv
        if conv == "double":
            return [struct.unpack("<d", r)[0] for r in raw]
        if conv == "float":
            return [float(struct.unpack("<f", r)[0]) for r in raw]
        if conv == "dec_bytes":
            d = _pow10i(col.scale)
            return [float(be_i128(r)) / d for r in raw]
        fmt = "<i" if col.phys == T_INT32 else "<q"
        return self._ints(col, [struct.unpack(fmt, r)[0] for r in raw])

    def _ints(self, col: Column, xs: list) -> list:
        if col.conv in ("dec_i32", "dec_i64"):
            d = _pow10i(col.scale)
            return [x / d for x in xs]
        return xs

    # PrunePredicate::may_match (stats.rs:53-70) / Table::prune (table.rs:165-172)
    def prune(self, preds) -> List[int]:
        """preds: (column, op, value) with op in '=', '<', '<=', '>', '>='."""
        return [rg for rg in range(len(self.row_groups))
                if all(may_match(self.stats(rg, c), op, v) for c, op, v in preds)]


def _compare(bound, v) -> Optional[int]:
    """ScalarValue::compare (scalar.rs:115-131) for same-family values."""
    if bound is None or v is None:
        return None
    if isinstance(bound, str) != isinstance(v, str):
        return None
    if isinstance(bound, str):
        a, b = bound.encode(), v.encode()
    else:
        a, b = bound, v
        if a != a or b != b:
            return None
    return (a > b) - (a < b)


def may_match(stats: dict, op: str, v) -> bool:
    lo, hi = _compare(stats["min"], v), _compare(stats["max"], v)
    if op == "=":
        return lo != 1 and hi != -1
    if op == "<":
        return lo not in (1, 0)
    if op == "<=":
        return lo != 1
    if op == ">":
        return hi not in (-1, 0)
    if op == ">=":
        return hi != -1
    raise ValueError(op)
