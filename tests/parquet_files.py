"""Parquet files for the scan tests, written with pyarrow (an independent
writer of the format) in memory."""

import datetime
import decimal
import io
import struct

import numpy as np


def sample_table(n=5000, seed=0, nulls=True):
    import pyarrow as pa
    rng = np.random.default_rng(seed)

    def maybe_null(vals, frac=0.15):
        if not nulls:
            return vals
        mask = rng.random(len(vals)) < frac
        return [None if m else v for v, m in zip(vals, mask)]

    words = ["AIR", "MAIL", "RAIL", "SHIP", "TRUCK", "REG AIR", "FOB", "", "naïve"]
    cols = {
        "i64": pa.array(maybe_null(rng.integers(-(1 << 40), 1 << 40, n).tolist()), pa.int64()),
        "i32": pa.array(rng.integers(-1000, 1000, n).astype(np.int32)),  # no NULLs
        "f64": pa.array(maybe_null(np.round(rng.normal(0, 100, n), 3).tolist()), pa.float64()),
        "f32": pa.array(maybe_null(rng.normal(0, 10, n).astype(np.float32).tolist()), pa.float32()),
        "day": pa.array(maybe_null(rng.integers(8000, 9000, n).tolist()), pa.int32()).cast(pa.date32()),
        "mode": pa.array(maybe_null([words[k] for k in rng.integers(0, len(words), n)]), pa.string()),
        "name": pa.array(maybe_null([f"Customer#{k:09d}" for k in rng.integers(0, n * 4, n)]), pa.string()),
        "flag": pa.array(maybe_null([bool(x) for x in rng.integers(0, 2, n)]), pa.bool_()),
        "price": pa.array(maybe_null([decimal.Decimal(int(x)).scaleb(-2) for x in rng.integers(-10**8, 10**8, n)]),
                          pa.decimal128(12, 2)),
        "qty": pa.array(maybe_null([decimal.Decimal(int(x)).scaleb(-1) for x in rng.integers(0, 5000, n)]),
                        pa.decimal128(5, 1)),
    }
    return pa.table(cols)


def write(table, **kw) -> bytes:
    import pyarrow.parquet as pq
    opts = dict(compression="NONE", row_group_size=2000, data_page_size=4096)
    opts.update(kw)
    buf = io.BytesIO()
    pq.write_table(table, buf, **opts)
    return buf.getvalue()


def byte_array_decimals(values, precision, scale, **kw) -> bytes:
    """A file of one OPTIONAL BYTE_ARRAY DECIMAL(precision, scale) column "dec"
    holding `values` (big-endian two's complement bytes, or None). pyarrow
    writes DECIMAL only as fixed-length bytes, so it writes a binary column and
    the footer's schema element of it gets converted_type DECIMAL, scale and
    precision (Thrift compact fields 6, 7, 8) spliced in after its name."""
    import pyarrow as pa
    assert 0 <= scale < 64 and 0 < precision < 64
    data = write(pa.table({"dec": pa.array(values, pa.binary())}), **kw)
    flen = struct.unpack_from("<I", data, len(data) - 8)[0]
    start = len(data) - 8 - flen
    foot = data[start:len(data) - 8]
    name = b"\x18\x03dec"  # field 4 (name) after field 3, 3 bytes
    at = foot.index(name) + len(name)
    assert foot[at] == 0  # the element ends after its name
    add = bytes([0x25, 2 * 5, 0x15, 2 * scale, 0x15, 2 * precision])  # i32 fields, zigzag values
    foot = foot[:at] + add + foot[at:]
    return data[:start] + foot + struct.pack("<I", len(foot)) + b"PAR1"


def decimal_bytes_sample(n, seed, min_len=1, max_len=16):
    """Big-endian two's complement values of every length in [min_len,
    max_len], some NULL."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        if k % 13 == 5:
            out.append(None)
            continue
        ln = int(rng.integers(min_len, max_len + 1))
        out.append(bytes(rng.integers(0, 256, ln).astype(np.uint8).tobytes()))
    return out


def as_engine_value(v):
    """pyarrow's Python value in the engine's terms (DATE as days, DECIMAL as f64)."""
    if isinstance(v, datetime.date):
        return (v - datetime.date(1970, 1, 1)).days
    if isinstance(v, decimal.Decimal):
        return float(v)
    return v


def pyarrow_rows(data: bytes, rg: int):
    import pyarrow.parquet as pq
    t = pq.ParquetFile(io.BytesIO(data)).read_row_group(rg)
    return [[as_engine_value(v) for v in t.column(i).to_pylist()] for i in range(t.num_columns)]
