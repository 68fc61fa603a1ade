"""Device Parquet scan (read_row_group, table.rs:204-240; read_column_chunk,
parquet_reader.rs:210-287) against the oracle's decoder on files written by
pyarrow: every column type, NULLs, dictionary and PLAIN pages, many small
pages, data pages v2, the RLE / DELTA / BYTE_STREAM_SPLIT encodings, decimals
as fixed-length bytes (up to 16), as BYTE_ARRAY (0 to 20 bytes) and as
integers, REQUIRED columns, SNAPPY chunks, projections, a scan result as the
source of an aggregate, a 300,000-row row group of DELTA pages against
pyarrow, and other codecs refused."""

import numpy as np
import pytest

from minilake_amd import abi
from minilake_amd.abi import AGG_COUNT_STAR, AGG_MAX, AGG_SUM, FLOAT64, INT64, UTF8, Agg, col
from oracle import parquet_ref as R
from tests import parquet_files as F
from tests.compare import assert_tables_equal
from tests.test_parquet import ENCODINGS, EXTRA_FILES, ORACLE_ONLY, WRITES, _files

pytestmark = pytest.mark.gpu
pa = pytest.importorskip("pyarrow")


@pytest.fixture(scope="module")
def files():
    return _files()


@pytest.mark.parametrize("kind", list(WRITES) + EXTRA_FILES + ORACLE_ONLY)
def test_device_scan_matches_oracle(gpu_engine, files, kind):
    from minilake_amd.gpu import ParquetTable
    data = files[kind]
    f = R.ParquetFile(data)
    t = ParquetTable(data, gpu_engine)
    for rg in range(t.num_row_groups):
        got = t.read_row_group(rg).to_host()
        assert got.num_rows == f.row_groups[rg][0]
        assert got.types == t.types
        for c in range(t.num_columns):
            assert got.columns[c].decoded() == f.read_column(rg, c), (kind, rg, f.columns[c].name)
    proj = [t.num_columns - 1, 0]
    got = t.read_row_group(0, proj).to_host()
    assert [c.decoded() for c in got.columns] == [f.read_column(0, c) for c in proj]


@pytest.mark.parametrize("version", ["1.0", "2.0"])
def test_device_scan_of_one_large_row_group(gpu_engine, version):
    """300,000 rows in one row group (thousands of DELTA miniblocks per column
    chunk, so the prefix sums cross many 1024-number rounds), against
    pyarrow's reading of the same file."""
    from minilake_amd.gpu import ParquetTable
    import pyarrow.parquet as pq
    import io
    n = 300_000
    t = F.sample_table(n, seed=9)
    data = F.write(t, use_dictionary=False, column_encoding=ENCODINGS, row_group_size=n, data_page_size=1 << 20,
                   data_page_version=version, compression="SNAPPY")
    pt = ParquetTable(data, gpu_engine)
    assert pt.num_row_groups == 1
    got = pt.read_row_group(0).to_host()
    want = pq.ParquetFile(io.BytesIO(data)).read_row_group(0)
    for c, name in enumerate(pt.names):
        if name in ("price", "qty"):
            continue  # decimals: the oracle-sized tests cover them
        exp = [F.as_engine_value(v) for v in want.column(c).to_pylist()]
        assert got.columns[c].decoded() == exp, name


def test_scan_result_feeds_an_aggregate(gpu_engine, oracle_engine, files):
    """A decoded row group (NULLs included) as the source of a GROUP BY."""
    from minilake_amd.gpu import ParquetTable
    t = ParquetTable(files["dictionary"], gpu_engine)
    res = t.read_row_group(1)
    names = t.names
    mode, price, i64 = names.index("mode"), names.index("price"), names.index("i64")
    aggs = [Agg(AGG_COUNT_STAR), Agg(AGG_SUM, col(i64), INT64), Agg(AGG_MAX, col(price), FLOAT64)]
    want = oracle_engine.hashagg(res.to_host(), [], [col(mode)], [UTF8], aggs).to_host()
    got = gpu_engine.hashagg(res, [], [col(mode)], [UTF8], aggs).to_host()
    assert_tables_equal(got, want, rtol=1e-12)


@pytest.mark.parametrize("codec", ["ZSTD", "GZIP"])
def test_other_codecs_refused(gpu_engine, codec):
    from minilake_amd.gpu import GpuError, ParquetTable
    data = F.write(F.sample_table(500, seed=4), compression=codec)
    t = ParquetTable(data, gpu_engine)
    with pytest.raises(GpuError) as e:
        t.read_row_group(0)
    assert e.value.code == abi.ML_ERR_UNSUPPORTED
