"""The allocation-free filter (ml_filter_noalloc) against ml_filter, through
the C ABI: identical selections on two calls in a row (the second over a
smaller input, so stale contents of the shared scratch array would show),
with validity bitmaps, a dictionary predicate and an incoming selection; and
a call over the row limit is refused before any launch."""

import ctypes as C

import numpy as np
import pytest

from minilake_amd import abi
from minilake_amd.abi import BOOLEAN, DATE, FLOAT64, INT32, INT64, UTF8, col, lit
from minilake_amd.gpu import GpuEngine, GpuError, lib
from minilake_amd.table import HostColumn, HostTable

pytestmark = pytest.mark.gpu

ML_ERR_RESOURCES_EXHAUSTED = 7
ML_FILTER_NOALLOC_MAX_ROWS = 1 << 24


def table(n, seed):
    rng = np.random.default_rng(seed)
    valid = rng.random(n) > 0.1
    return HostTable([
        HostColumn(INT32, rng.integers(-1000, 1000, n).astype(np.int32)),
        HostColumn(FLOAT64, rng.normal(0, 100, n), None, valid),
        HostColumn(UTF8, rng.integers(0, 4, n).astype(np.uint32), ["apple", "banana", "cherry", "date"]),
        HostColumn(INT64, rng.integers(-3, 4, n)),
    ])


PRED = ((col(0).lt(lit(200)) & col(1).gt(lit(-50.0))) | col(2).eq(lit("cherry"))) & ~(col(3).eq(lit(0)))


def noalloc_filter(eng, pred, t, selection=None):
    L, keep = lib(), abi.Keep()
    s, b = eng._src(t, keep, selection)
    n = s.num_rows if selection is None else len(selection)
    out = eng.torch.full((max(1, n),), -1, dtype=eng.torch.int32, device=f"cuda:{eng.device}")
    me = abi.to_ml_expr(pred, keep)
    cnt = C.c_int64()
    rc = L.ml_filter_noalloc(eng.ctx, C.byref(me), C.byref(b), out.data_ptr(), C.byref(cnt))
    eng._check(rc)
    return out[:cnt.value].cpu().numpy().view(np.uint32).copy()


def test_noalloc_filter_matches_filter_twice():
    eng = GpuEngine()
    try:
        for n, seed in ((300_000, 1), (5_000, 2)):  # the second call is smaller
            t = table(n, seed)
            want = eng.filter(PRED, t)
            got = noalloc_filter(eng, PRED, t)
            assert len(want) > 0
            np.testing.assert_array_equal(got, want)
            sel = np.arange(0, n, 3, dtype=np.uint32)
            np.testing.assert_array_equal(noalloc_filter(eng, PRED, t, sel), eng.filter(PRED, t, sel))
        # an empty input
        t0 = table(0, 3)
        assert len(noalloc_filter(eng, PRED, t0)) == 0
    finally:
        eng.close()


def test_noalloc_filter_refuses_too_many_rows():
    eng = GpuEngine()
    try:
        b = abi.MlBatch()
        b.columns = None
        b.num_columns = 0
        b.num_rows = ML_FILTER_NOALLOC_MAX_ROWS + 1
        keep = abi.Keep()
        me = abi.to_ml_expr(lit(True), keep)
        cnt = C.c_int64(-1)
        rc = lib().ml_filter_noalloc(eng.ctx, C.byref(me), C.byref(b), None, C.byref(cnt))
        assert rc == ML_ERR_RESOURCES_EXHAUSTED
        assert b"at most" in lib().ml_last_error(eng.ctx)
        assert cnt.value == -1
    finally:
        eng.close()
