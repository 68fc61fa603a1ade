"""Batched fragments through the C ABI (ml_fragments_*): the Python mirror the
tests and the bench drive. Each fragment is
    SELECT g, SUM(v), COUNT(*) FROM probe JOIN build USING (key) GROUP BY g
over device columns; results land in per-fragment device outputs."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import abi
from . import workloads as W
from .gpu import GpuEngine, lib

MODES = ("persistent", "graph", "per_fragment")


@dataclass
class DeviceFragment:
    build_keys: object      # torch int64 (device)
    build_groups: object    # torch int64
    probe_keys: object      # torch int64
    probe_values: object    # torch float64
    groups: int
    build_validity: object = None  # torch int64 bitmap words or None
    probe_validity: object = None
    out_sums: object = None
    out_counts: object = None


def device_fragment(eng: GpuEngine, bk, bg, pk, pv, groups, bvalid=None, pvalid=None) -> DeviceFragment:
    """A fragment from host arrays (validity: one bool per row, or None)."""
    t = eng.torch
    dev = f"cuda:{eng.device}"

    def bitmap(v):
        if v is None:
            return None
        bits = np.packbits(np.asarray(v, bool), bitorder="little")
        words = np.zeros((len(bits) + 7) // 8 * 8, np.uint8)
        words[:len(bits)] = bits
        return t.from_numpy(words.view(np.int64).copy()).to(dev)

    as_dev = lambda a, dt: t.from_numpy(np.ascontiguousarray(a, dt)).to(dev)
    return DeviceFragment(as_dev(bk, np.int64), as_dev(bg, np.int64), as_dev(pk, np.int64), as_dev(pv, np.float64),
                          int(groups), bitmap(bvalid), bitmap(pvalid))


def generate(eng: GpuEngine, shapes: Sequence[W.FragmentShape], seed: int = 42) -> List[DeviceFragment]:
    """The batched-fragment workload, generated on the device."""
    t = eng.torch
    dev = f"cuda:{eng.device}"
    out = []
    for i, sh in enumerate(shapes):
        bk = t.empty(sh.build_rows, dtype=t.int64, device=dev)
        bg = t.empty(sh.build_rows, dtype=t.int64, device=dev)
        pk = t.empty(max(1, sh.probe_rows), dtype=t.int64, device=dev)
        pv = t.empty(max(1, sh.probe_rows), dtype=t.float64, device=dev)
        W.gpu_fragment(sh, i, seed, bk.data_ptr(), bg.data_ptr(), pk.data_ptr(), pv.data_ptr(), eng.stream.cuda_stream)
        out.append(DeviceFragment(bk, bg, pk[:sh.probe_rows], pv[:sh.probe_rows], sh.groups))
    eng.torch.cuda.synchronize()
    return out


class FragmentSet:
    """ml_fragments_create over device fragments; run / autotune / free."""

    def __init__(self, eng: GpuEngine, frags: Sequence[DeviceFragment], tuning: Optional[abi.MlFragmentTuning] = None):
        t = eng.torch
        self.eng, self.frags = eng, list(frags)
        self.arr = (abi.MlFragment * len(frags))()
        ptr = lambda x: 0 if x is None else x.data_ptr()
        for i, f in enumerate(self.frags):
            if f.out_sums is None:
                f.out_sums = t.zeros(f.groups, dtype=t.float64, device=f"cuda:{eng.device}")
                f.out_counts = t.zeros(f.groups, dtype=t.int64, device=f"cuda:{eng.device}")
            self.arr[i] = abi.MlFragment(ptr(f.build_keys), ptr(f.build_groups), ptr(f.build_validity),
                                         len(f.build_keys), ptr(f.probe_keys), ptr(f.probe_values),
                                         ptr(f.probe_validity), len(f.probe_keys), f.groups, ptr(f.out_sums),
                                         ptr(f.out_counts))
        t.cuda.synchronize()
        h = C.c_void_p()
        tune = C.byref(tuning) if tuning is not None else None
        eng._check(lib().ml_fragments_create(eng.ctx, self.arr, len(frags), tune, C.byref(h)))
        self.handle = h.value

    def run(self, mode: str = "persistent"):
        fn = {"persistent": lib().ml_fragments_run, "graph": lib().ml_fragments_run_graph,
              "per_fragment": lib().ml_fragments_run_per_fragment}[mode]
        self.eng._check(fn(self.handle))

    def autotune(self) -> abi.MlFragmentTuning:
        best = abi.MlFragmentTuning()
        self.eng._check(lib().ml_fragments_autotune(self.handle, C.byref(best)))
        return best

    def results(self):
        return [(f.out_sums.cpu().numpy().copy(), f.out_counts.cpu().numpy().copy()) for f in self.frags]

    def close(self):
        if self.handle:
            lib().ml_fragments_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def algorithmic_bytes(frags: Sequence[DeviceFragment]) -> int:
    """16 B per probe row (key, value) + 16 B per build row (key, group) +
    16 B per output group (sum, count)."""
    return sum(16 * (len(f.probe_keys) + len(f.build_keys) + f.groups) for f in frags)
