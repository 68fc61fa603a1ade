"""Build the gfx950 library (``libminilake_gpu.so``) and the workload generator.

Every HIP source is compiled by ``hipcc --offload-arch=gfx950`` in parallel into
``minilake_amd/_build/`` and linked in-tree, so the shared library travels with
the source tree. Flags:
  -ffp-contract=off      every row value is an IEEE evaluation in source order
  -munsafe-fp-atomics    hardware FP64 atomics (global/ds_add_f64) instead of CAS loops
  -packed-fp32-ops off   no packed FP32 instructions
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libminilake_gpu.so")
SOURCES = ["kernels.hip", "compile.hip", "engine.hip", "fragments.hip", "mlp.hip", "parquet.hip", "noalloc.hip"]
ARCH = os.environ.get("MINILAKE_GPU_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-I", os.path.join(HERE, "..", "include"),
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def _compile(src: str, build_dir: str = BUILD, defines=()) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj


def build_gpu_library(jobs: int = 8, out: str = LIB, build_dir: str = BUILD, defines=()) -> str:
    """Compile and link the library; `defines` (NAME=VALUE) override the
    tuning constants of the sources (ML_FAST_U, ML_BLOOM_KEYS_PER_WORD, ML_NT_KEYS)."""
    os.makedirs(build_dir, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, defines), sources))
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, out)
    return out


def build_all() -> None:
    from . import workloads
    workloads.build_library()
    build_gpu_library()


if __name__ == "__main__":
    sys.path.insert(0, os.path.join(HERE, ".."))
    build_all()
    print(LIB)
