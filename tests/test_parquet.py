"""Parquet scan, host side (no GPU): the oracle's decoder and statistics
against pyarrow's reading of the same files; the product's footer parser,
statistics (convert_stats, parquet_reader.rs:290-358) and pruning
(stats.rs:53-70, table.rs:165-172) against the oracle; the reference's own
pruning test (stats.rs:80-99) as known answers."""

import io
import math

import numpy as np
import pytest

from minilake_amd import abi
from oracle import parquet_ref as R
from tests import parquet_files as F

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

# one encoding per column, each other than PLAIN and dictionary
ENCODINGS = {"i64": "DELTA_BINARY_PACKED", "i32": "DELTA_BINARY_PACKED", "day": "DELTA_BINARY_PACKED",
             "f64": "BYTE_STREAM_SPLIT", "f32": "BYTE_STREAM_SPLIT", "mode": "DELTA_LENGTH_BYTE_ARRAY",
             "name": "DELTA_BYTE_ARRAY", "flag": "RLE", "price": "DELTA_BYTE_ARRAY", "qty": "BYTE_STREAM_SPLIT"}

WRITES = {
    "dictionary": {},
    "plain": {"use_dictionary": False},
    "small_pages": {"data_page_size": 300},
    "decimals_as_integers": {"store_decimal_as_integer": True},
    "snappy": {"compression": "SNAPPY"},
    "dictionary_fallback": {"dictionary_pagesize_limit": 512},  # later pages PLAIN
    "pages_v2": {"data_page_version": "2.0"},  # booleans RLE-encoded
    "pages_v2_snappy": {"data_page_version": "2.0", "compression": "SNAPPY", "data_page_size": 1024},
    "encodings": {"use_dictionary": False, "column_encoding": ENCODINGS},
    "encodings_v2_snappy": {"use_dictionary": False, "column_encoding": ENCODINGS, "data_page_version": "2.0",
                            "compression": "SNAPPY", "data_page_size": 300},
    "decimals_as_integers_delta": {"use_dictionary": False, "store_decimal_as_integer": True,
                                   "column_encoding": {"price": "DELTA_BINARY_PACKED", "qty": "BYTE_STREAM_SPLIT"}},
}


def wide_table(n=3000, seed=5):
    """Decimals past 64 bits (FLBA of 16 and 9 bytes) and integers at their
    extremes, whose deltas wrap around."""
    import decimal
    rng = np.random.default_rng(seed)
    big = [decimal.Decimal(int(a) * 10**20 + int(b)).scaleb(-3) for a, b in
           zip(rng.integers(-10**14, 10**14, n), rng.integers(0, 10**18, n))]
    big[:4] = [decimal.Decimal(2**63).scaleb(-3), decimal.Decimal(-2**63 - 1).scaleb(-3),
               decimal.Decimal(2**64 + 1).scaleb(-3), decimal.Decimal(10**37 - 1).scaleb(-3)]
    mid = [decimal.Decimal(int(x) * 1000 + 7) for x in rng.integers(-10**16, 10**16, n)]
    i64 = rng.choice(np.array([-2**63, 2**63 - 1, 0, -1, 1], np.int64), n)
    i32 = rng.choice(np.array([-2**31, 2**31 - 1, 0, -7], np.int32), n)
    return pa.table({"big": pa.array(big, pa.decimal128(38, 3)), "mid": pa.array(mid, pa.decimal128(20, 0)),
                     "i64": pa.array(i64), "i32": pa.array(i32),
                     "seq": pa.array(np.arange(n, dtype=np.int64) * 3 - 1000)})


def _files():
    t = F.sample_table(5000, seed=1)
    out = {k: F.write(t, **kw) for k, kw in WRITES.items()}
    req = pa.table({"k": pa.array(np.arange(3000, dtype=np.int64)), "s": pa.array([f"v{i % 11}" for i in range(3000)])},
                   schema=pa.schema([pa.field("k", pa.int64(), nullable=False),
                                     pa.field("s", pa.string(), nullable=False)]))
    out["required"] = F.write(req)
    n = 3000
    edge = pa.table({"nulls": pa.array([None] * n, pa.int64()),
                     "empty": pa.array(["" if i % 3 else None for i in range(n)], pa.string()),
                     "sparse": pa.array([float(i) if i % 97 == 0 else None for i in range(n)], pa.float64()),
                     "flag": pa.array([None if i % 5 else bool(i % 2) for i in range(n)], pa.bool_())})
    out["edge"] = F.write(edge, data_page_size=256)
    out["edge_v2"] = F.write(edge, data_page_size=256, data_page_version="2.0")
    w = wide_table()
    out["wide"] = F.write(w, use_dictionary=False)
    out["wide_encodings"] = F.write(w, use_dictionary=False, data_page_size=512,
                                    column_encoding={"big": "BYTE_STREAM_SPLIT", "mid": "DELTA_BYTE_ARRAY",
                                                     "i64": "DELTA_BINARY_PACKED", "i32": "DELTA_BINARY_PACKED",
                                                     "seq": "DELTA_BINARY_PACKED"})
    dec = F.decimal_bytes_sample(3000, 3)
    out["byte_decimals"] = F.byte_array_decimals(dec, 38, 4, use_dictionary=False)
    out["byte_decimals_dict"] = F.byte_array_decimals(dec, 38, 4, data_page_size=1024)
    out["byte_decimals_delta"] = F.byte_array_decimals(dec, 38, 4, use_dictionary=False, data_page_version="2.0",
                                                       compression="SNAPPY", column_encoding={"dec": "DELTA_BYTE_ARRAY"})
    out["byte_decimals_long"] = F.byte_array_decimals(LONG_DECIMALS, 30, 2, use_dictionary=False,
                                                      column_encoding={"dec": "DELTA_LENGTH_BYTE_ARRAY"})
    return out


# BYTE_ARRAY decimals of 0 to 20 bytes: pyarrow refuses lengths 0 and past 16
LONG_DECIMALS = F.decimal_bytes_sample(3000, 7, min_len=0, max_len=20)
EXTRA_FILES = ["required", "edge", "edge_v2", "wide", "wide_encodings", "byte_decimals", "byte_decimals_dict",
               "byte_decimals_delta"]
ORACLE_ONLY = ["byte_decimals_long"]


@pytest.fixture(scope="module")
def files():
    return _files()


@pytest.mark.parametrize("kind", list(WRITES) + EXTRA_FILES)
def test_oracle_decodes_like_pyarrow(files, kind):
    data = files[kind]
    f = R.ParquetFile(data)
    md = pq.ParquetFile(io.BytesIO(data)).metadata
    assert len(f.row_groups) == md.num_row_groups
    for rg in range(md.num_row_groups):
        want = F.pyarrow_rows(data, rg)
        for c in range(len(f.columns)):
            got = f.read_column(rg, c)
            if f.columns[c].conv == "dec_bytes":
                # the reference rounds twice (i128 -> f64, then / 10^scale); pyarrow's Decimal -> float
                # once: they may differ in the last bit past 2^53
                assert len(got) == len(want[c]) and all(
                    (a is None and b is None) or (a is not None and b is not None and math.isclose(a, b, rel_tol=1e-15))
                    for a, b in zip(got, want[c])), (rg, f.columns[c].name)
            else:
                assert got == want[c], (rg, f.columns[c].name)
            st = md.row_group(rg).column(c).statistics
            got = f.stats(rg, c)
            assert got["null_count"] == st.null_count
            if f.columns[c].conv in ("bool", "dec_bytes"):  # convert_stats keeps no bounds for these
                assert got["min"] is None and got["max"] is None
            else:
                assert got["min"] == F.as_engine_value(st.min) and got["max"] == F.as_engine_value(st.max)


def test_prune_known_answers():
    """stats.rs:80-99: stats (10, 20) of a DATE column."""
    s = {"min": 10, "max": 20, "null_count": 0}
    assert not R.may_match(s, "<", 10) and R.may_match(s, "<=", 10)
    assert not R.may_match(s, ">", 20) and R.may_match(s, ">=", 20)
    assert not R.may_match(s, "=", 21) and R.may_match(s, "=", 15)
    assert R.may_match({"min": None, "max": None, "null_count": None}, "=", 99)
    # the same through the library's host code, on a file whose only row group spans days 10..20
    from minilake_amd.gpu import ParquetTable
    data = F.write(pa.table({"d": pa.array(np.arange(10, 21, dtype=np.int32)).cast(pa.date32())}))
    t = ParquetTable(data)
    assert t.stats(0, 0) == {"min": 10, "max": 20, "null_count": 0}
    for op, v, keep in (("<", 10, False), ("<=", 10, True), (">", 20, False), (">=", 20, True), ("=", 21, False),
                        ("=", 15, True)):
        assert t.prune([(0, op, v)]) == ([0] if keep else []), (op, v)


@pytest.mark.parametrize("kind", list(WRITES) + EXTRA_FILES + ORACLE_ONLY)
def test_library_metadata_stats_and_prune_match_oracle(files, kind):
    from minilake_amd.gpu import ParquetTable
    data = files[kind]
    f = R.ParquetFile(data)
    t = ParquetTable(data)
    assert t.num_columns == len(f.columns) and t.num_row_groups == len(f.row_groups)
    assert t.names == [c.name for c in f.columns]
    assert t.nullable == [c.optional for c in f.columns]
    types = {"bool": abi.BOOLEAN, "int32": abi.INT32, "date": abi.DATE, "int64": abi.INT64, "utf8": abi.UTF8,
             "float64": abi.FLOAT64}
    assert t.types == [types[c.ml_type] for c in f.columns]
    assert t.row_group_rows == [n for n, _ in f.row_groups]
    for rg in range(t.num_row_groups):
        for c in range(t.num_columns):
            assert t.stats(rg, c) == f.stats(rg, c), (rg, c)
    preds = []
    for c, col in enumerate(f.columns):
        st = f.stats(1 if len(f.row_groups) > 1 else 0, c)
        if st["min"] is not None:
            for op in ("=", "<", "<=", ">", ">="):
                preds.append((c, op, st["min"]))
                preds.append((c, op, st["max"]))
    for p in preds:
        assert t.prune([p]) == f.prune([p]), p
    assert t.prune(preds[:3]) == f.prune(preds[:3])


def test_unsupported_files_are_refused():
    from minilake_amd.gpu import GpuError, ParquetTable
    nested = F.write(pa.table({"l": pa.array([[1, 2], [3]], pa.list_(pa.int64()))}))
    with pytest.raises(GpuError) as e:
        ParquetTable(nested)
    assert e.value.code == abi.ML_ERR_UNSUPPORTED
    with pytest.raises(R.Unsupported):
        R.ParquetFile(nested)
    with pytest.raises(GpuError) as e:
        ParquetTable(b"PAR1" + b"\0" * 8)
    assert e.value.code == abi.ML_ERR_PARQUET


def test_snappy_known_answers():
    """Snappy: a hand-encoded stream (literal "abc", then a 1-byte-offset copy of
    6 bytes at offset 3 that overlaps its own output) and pyarrow's Snappy
    codec round trips over literal-heavy and repetitive data."""
    assert R.snappy_decompress(b"\x09\x08abc\x09\x03") == b"abcabcabc"
    rng = np.random.default_rng(2)
    for data in (rng.integers(0, 256, 5000).astype(np.uint8).tobytes(), b"MiniLake " * 3000,
                 bytes(70_000), rng.integers(0, 4, 100_000).astype(np.uint8).tobytes()):
        assert R.snappy_decompress(pa.compress(data, codec="snappy", asbytes=True)) == data


def test_delta_encodings_known_answers():
    """The Parquet format's own examples of the DELTA encodings, encoded by
    hand (block of 128 values in 4 miniblocks): DELTA_BINARY_PACKED of 1..5
    (min delta 1, bit widths 0) and of 7 5 3 1 2 3 4 5 (min delta -2, bit width
    2); DELTA_LENGTH_BYTE_ARRAY of Hello World Foobar ABCDEF; DELTA_BYTE_ARRAY
    of axis axle babble babyhood (prefixes 0 2 0 3)."""
    def hdr(total, first_zigzag):
        return bytes([0x80, 0x01, 0x04, total, first_zigzag])

    a = hdr(5, 2) + bytes([0x02, 0, 0, 0, 0])
    assert R.delta_binary_packed(a, 0, len(a), 64) == ([1, 2, 3, 4, 5], len(a))
    b = hdr(8, 14) + bytes([0x03, 2, 0, 0, 0, 0xC0, 0xFF, 0, 0, 0, 0, 0, 0])
    assert R.delta_binary_packed(b, 0, len(b), 32) == ([7, 5, 3, 1, 2, 3, 4, 5], len(b))
    c = hdr(4, 10) + bytes([0x00, 1, 0, 0, 0, 0x02, 0, 0, 0]) + b"HelloWorldFoobarABCDEF"
    assert R.delta_length_byte_array(c, 0, len(c)) == ([b"Hello", b"World", b"Foobar", b"ABCDEF"], len(c))
    pre = hdr(4, 0) + bytes([0x03, 3, 0, 0, 0, 0x44, 0x01] + [0] * 10)
    suf = hdr(4, 8) + bytes([0x03, 3, 0, 0, 0, 0x70] + [0] * 11)
    d = pre + suf + b"axislebabbleyhood"
    assert R.delta_byte_array(d, 0, len(d)) == ([b"axis", b"axle", b"babble", b"babyhood"], len(d))
    # a delta that wraps: INT64 max then min (delta 1 mod 2^64)
    w = bytes([0x80, 0x01, 0x04, 2]) + _varint(2 * (2**63 - 1)) + bytes([0x02, 0, 0, 0, 0])
    assert R.delta_binary_packed(w, 0, len(w), 64)[0] == [2**63 - 1, 2**63]


def _varint(v):
    out = bytearray()
    while True:
        out.append((v & 0x7F) | (0x80 if v > 0x7F else 0))
        v >>= 7
        if not v:
            return bytes(out)


def test_be_i128_known_answers():
    """be_i128 (parquet_reader.rs:156-167): sign from the first byte, at most 16 bytes."""
    assert R.be_i128(b"") == 0
    assert R.be_i128(b"\xff") == -1 and R.be_i128(b"\x80") == -128 and R.be_i128(b"\x00\x80") == 128
    assert R.be_i128(b"\x7f" + b"\xff" * 15) == 2**127 - 1
    assert R.be_i128(b"\x80" + b"\x00" * 15) == -2**127
    assert R.be_i128(b"\x01" * 17) == int.from_bytes(b"\x01" * 16, "big")
    assert R.be_i128(b"\xfe" + b"\x00" * 16) == -2**121


def test_oracle_byte_array_decimals_of_any_length(files):
    """BYTE_ARRAY decimals of 0 to 20 bytes decode to be_i128(bytes) / 10^scale
    (parquet_reader.rs:156-167, 268-281), NULLs kept."""
    f = R.ParquetFile(files["byte_decimals_long"])
    got = [v for rg in range(len(f.row_groups)) for v in f.read_column(rg, 0)]
    assert got == [None if v is None else float(R.be_i128(v)) / 100.0 for v in LONG_DECIMALS]
