"""Autotuner for the device solver's leaf bound.

The leaf bound (rows per leaf, solved by one wavefront's Jacobi sweep in LDS)
trades leaf work (O(m^3) per leaf, a few waves per CU) against tree depth
(one launch sequence per level).  The best value depends on the size mix and
the spectrum, so it is measured: each candidate is timed as a captured-graph
replay of the actual workload, and the fastest wins.  Results can be cached
in a JSON file keyed by the device name and a signature of the size list.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from . import BatchSolver, GraphSolve, Solver

CANDIDATES = (8, 12, 16, 24, 32, 48, 64)


def _signature(kind: str, sizes, device_name: str) -> str:
    h = hashlib.sha1(np.ascontiguousarray(sizes, dtype=np.int64).tobytes()).hexdigest()[:16]
    return f"{kind}:{device_name}:{len(sizes)}:{h}"


def _time_graph(gs: GraphSolve, reps: int) -> float:
    import torch
    for _ in range(2):
        gs.run()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        gs.run()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _store(path, cache):
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def tune_batched(d_cat, e_cat, sizes, candidates=CANDIDATES, reps: int = 5,
                 cache_path: str | None = None) -> dict:
    """Pick the leaf bound for a batch.  d_cat, e_cat: the batch's
    concatenated rows as float64 device tensors (read only).  Returns
    {"leaf_max": best, "ms": {leaf: ms per batch}, "cached": bool}."""
    import torch
    dev = d_cat.device
    key = _signature("batch", sizes, torch.cuda.get_device_name(dev))
    cache = _load(cache_path)
    if key in cache:
        return dict(cache[key], cached=True)
    timings = {}
    for leaf in candidates:
        gs = GraphSolve(BatchSolver(sizes, device=dev, leaf_max=leaf), d_cat, e_cat)
        timings[int(leaf)] = _time_graph(gs, reps)
        del gs
    best = min(timings, key=timings.get)
    res = {"leaf_max": int(best), "ms": timings}
    cache[key] = res
    _store(cache_path, cache)
    return dict(res, cached=False)


def tune_single(d, e, candidates=CANDIDATES, reps: int = 5, cache_path: str | None = None) -> dict:
    """Pick the leaf bound for one matrix (d, e: float64 device tensors)."""
    import torch
    n = d.numel()
    key = _signature("single", [n], torch.cuda.get_device_name(d.device))
    cache = _load(cache_path)
    if key in cache:
        return dict(cache[key], cached=True)
    timings = {}
    for leaf in candidates:
        gs = GraphSolve(Solver(n, device=d.device, leaf_max=leaf), d, e)
        timings[int(leaf)] = _time_graph(gs, reps)
        del gs
    best = min(timings, key=timings.get)
    res = {"leaf_max": int(best), "ms": timings}
    cache[key] = res
    _store(cache_path, cache)
    return dict(res, cached=False)
