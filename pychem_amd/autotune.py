"""Autotuner for the J/K digestion launch shapes.

For every ERI class with stored quartets it times the single-class digest
(pq_jk_class) over a grid of (work-item length, blocks launched) candidates
with HIP events on the working stream, keeps the fastest, applies it to the
handle (pq_jk_set_params) and returns the table, which callers may persist
as JSON and re-apply with ``apply``.
"""
import ctypes
import json

import numpy as np
import torch

from . import _lib
from ._lib import check
from .engine import _ptr, _stream

RUN_LENGTHS = (0, 64, 128, 256, 512, 1024)      # 0 = automatic
GRIDS = (0, 256, 512, 1024, 2048, 4096)         # 0 = one block per work item


def _time_class(h, cls, Dj, Dk, J, K, reps=3):
    L = _lib.lib()
    nitems = ctypes.c_int(0)
    check(L.pq_jk_class(h, cls, 3, _ptr(Dj), _ptr(Dk), _ptr(J), _ptr(K), ctypes.byref(nitems), _stream()), "pq_jk_class")
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        check(L.pq_jk_class(h, cls, 3, _ptr(Dj), _ptr(Dk), _ptr(J), _ptr(K), None, _stream()), "pq_jk_class")
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / reps, nitems.value


def tune_jk(engine, Dj, Dk, J, K, verbose=False):
    """engine: Engine or BatchEngine with ERIs built.  Dj, Dk, J, K: device
    buffers of the handle's packed matrix size.  Returns {class: {...}}."""
    h = engine.h
    L = _lib.lib()
    stats = engine.stats()
    nq = [s["nq64"] for s in stats] if isinstance(stats, list) else [int(r[0]) for r in stats]
    table = {}
    for cls in range(L.pq_eri_nclass()):
        if nq[cls] == 0:
            continue
        best = None
        base_ms = None
        for run in RUN_LENGTHS:
            check(L.pq_jk_set_params(h, cls, run, 0, _stream()), "pq_jk_set_params")
            ms0, nitems = _time_class(h, cls, Dj, Dk, J, K)
            if run == 0:
                base_ms = ms0
            for grid in GRIDS:
                if grid and grid >= nitems:
                    continue
                if grid:
                    check(L.pq_jk_set_params(h, cls, run, grid, _stream()), "pq_jk_set_params")
                    ms, _ = _time_class(h, cls, Dj, Dk, J, K)
                else:
                    ms = ms0
                if best is None or ms < best[0]:
                    best = (ms, run, grid, nitems)
        ms, run, grid, nitems = best
        check(L.pq_jk_set_params(h, cls, run, grid, _stream()), "pq_jk_set_params")
        table[cls] = {"run_len": run, "grid": grid, "ms": ms, "default_ms": base_ms, "nitems": nitems}
        if verbose:
            print("class %2d: run %4d grid %4d  %.4f ms (default %.4f)" % (cls, run, grid, ms, base_ms))
    return table


def apply(engine, table):
    L = _lib.lib()
    for cls, p in table.items():
        check(L.pq_jk_set_params(engine.h, int(cls), int(p["run_len"]), int(p["grid"]), _stream()),
              "pq_jk_set_params")


def save(table, path):
    with open(path, "w") as fh:
        json.dump({str(k): v for k, v in table.items()}, fh, indent=1)


def load(path):
    with open(path) as fh:
        return {int(k): v for k, v in json.load(fh).items()}
