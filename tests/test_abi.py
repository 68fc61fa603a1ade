"""The C ABI without a GPU: the built library loads and exports every function
include/minilake_gpu.h declares, and the ctypes mirror matches the C struct
layouts byte for byte (checked against gcc's own sizeof/offsetof)."""

import ctypes as C
import os
import re
import subprocess

import pytest

from minilake_amd import abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "minilake_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ml_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    import ctypes
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, f"declared but not exported: {missing}"
    assert set(gpu.EXPORTED) <= set(declared_functions())


STRUCTS = {
    "ml_dict": abi.MlDict, "ml_column": abi.MlColumn, "ml_batch": abi.MlBatch, "ml_expr_node": abi.MlExprNode,
    "ml_expr": abi.MlExpr, "ml_exec_config": abi.MlExecConfig, "ml_aggregate": abi.MlAggregate,
    "ml_agg_spec": abi.MlAggSpec, "ml_join_build_spec": abi.MlJoinBuildSpec, "ml_operator": abi.MlOperator,
    "ml_sort_key": abi.MlSortKey, "ml_sort_spec": abi.MlSortSpec, "ml_fragment": abi.MlFragment,
    "ml_fragment_tuning": abi.MlFragmentTuning, "ml_launch_stat": abi.MlLaunchStat,
}


def test_ctypes_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l)
    for cname, py in STRUCTS.items():
        assert int(got[f"{cname} size"]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


def test_constants_match_header():
    text = open(HEADER).read()
    assert abi.MLP_NUM_PARAMS == 32 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2 == 6402
    for name, val in [("ML_ERR_CAPACITY", 10), ("ML_TYPE_UTF8", 5)]:
        assert re.search(rf"{name}\s*=\s*{val}", text)


@pytest.mark.parametrize("block", [32, 100, 320])
def test_ctx_rejects_partial_wave_block_sizes(block):
    """Streaming kernels reduce across whole 64-lane waves and are compiled for
    at most 256 threads: other block sizes are a plan error, before any device
    is touched."""
    L = gpu.lib()
    cfg = abi.MlExecConfig()
    L.ml_default_config(C.byref(cfg))
    cfg.block_size = block
    ctx = C.c_void_p()
    assert L.ml_ctx_create(C.byref(cfg), C.byref(ctx)) == abi.ML_ERR_PLAN
