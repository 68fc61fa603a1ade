"""Parquet scan, host side (no GPU): the oracle's decoder and statistics
against pyarrow's reading of the same files; the product's footer parser,
statistics (convert_stats, parquet_reader.rs:290-358) and pruning
(stats.rs:53-70, table.rs:165-172) against the oracle; the reference's own
pruning test (stats.rs:80-99) as known answers."""

import io

import numpy as np
import pytest

from minilake_amd import abi
from oracle import parquet_ref as R
from tests import parquet_files as F

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

WRITES = {
    "dictionary": {},
    "plain": {"use_dictionary": False},
    "small_pages": {"data_page_size": 300},
    "decimals_as_integers": {"store_decimal_as_integer": True},
    "snappy": {"compression": "SNAPPY"},
    "dictionary_fallback": {"dictionary_pagesize_limit": 512},  # later pages PLAIN
}


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
    return out


@pytest.fixture(scope="module")
def files():
    return _files()


@pytest.mark.parametrize("kind", list(WRITES) + ["required", "edge"])
def test_oracle_decodes_like_pyarrow(files, kind):
    data = files[kind]
    f = R.ParquetFile(data)
    md = pq.ParquetFile(io.BytesIO(data)).metadata
    assert len(f.row_groups) == md.num_row_groups
    for rg in range(md.num_row_groups):
        want = F.pyarrow_rows(data, rg)
        for c in range(len(f.columns)):
            assert f.read_column(rg, c) == want[c], (rg, f.columns[c].name)
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


@pytest.mark.parametrize("kind", list(WRITES) + ["required", "edge"])
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
