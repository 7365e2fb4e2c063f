"""Autotuner of the device solver's launch shape: the leaf bound and the
tile/launch knobs of brdc_options.

* leaf_max: rows per leaf (solved by the leaf kernel, implicit QL with one
  lane per leaf, O(m^2) per leaf) against tree depth (one launch sequence per
  level);
* stream_min_k: the merge size from which the streamed secular kernel (a
  persistent grid sharing LDS-staged pole chunks) takes over from the
  lanes-per-root kernel;
* secular_lanes_log2: the secular lanes a level aims to keep in flight, from
  which the lanes per root follow (more lanes: shorter passes, more waves);
* wide_max_waves: the widest group of the product-form and boundary-row
  kernels (waves splitting one 64-item group's stream);
* subtree_max_log2: up to how many subtree roots (nodes of at most 128 rows)
  the bottom of the trees runs in the one-launch subtree kernel (latency
  bound: it wins while the subtrees are few, the level launches win when
  thousands of subtrees fill the device anyway); -1 never, 30 always.

The best values depend on the size mix and the spectrum, so they are
measured: each candidate is timed as a captured-graph replay of the actual
workload (all its matrices), one knob at a time in the order above, each
keeping the best value found so far for the others (a coordinate search; the
defaults start it).  Results can be cached in a JSON file keyed by the device
name and a signature of the sizes.  The knobs other than leaf_max move results
within rounding only (they change which kernel solves a merge and how its
sums are split).
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from . import BatchSolver, GraphSolve, Solver

CANDIDATES = (8, 12, 16, 24, 32, 48, 64)
#: candidate values of each launch-shape knob, the default first
KNOB_CANDIDATES = {
    "stream_min_k": (2048, 1024, 4096),
    "secular_lanes_log2": (18, 17, 19),
    "wide_max_waves": (8, 4, 2),
    "subtree_max_log2": (0, -1, 30),
}


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


# A candidate replaces the configuration kept so far only when it is faster by
# more than this fraction: replays of one configuration spread by about half a
# per cent, so smaller wins are noise, and taking them made the chosen launch
# shape (and with it the per-kernel launch mix) change from run to run.
MARGIN = 0.01


def _search(time_of, leaf_candidates, knobs: bool, margin: float = MARGIN):
    """Coordinate search: time_of(config) -> ms for a dict of options.  The
    first value of each candidate list is the starting point; a later value
    is taken only if it beats the kept configuration by more than `margin`
    (a fraction of the kept configuration's time)."""
    timings = {}

    def run(cfg):
        key = ",".join(f"{k}={cfg[k]}" for k in sorted(cfg))
        if key not in timings:
            timings[key] = time_of(cfg)
        return timings[key]

    def better(cfg, kept):
        return run(cfg) < run(kept) * (1.0 - margin)

    best = {"leaf_max": int(leaf_candidates[0])}
    if knobs:
        best.update({k: v[0] for k, v in KNOB_CANDIDATES.items()})
    for leaf in leaf_candidates:
        cfg = dict(best, leaf_max=int(leaf))
        if better(cfg, best):
            best = cfg
    if knobs:
        for name, values in KNOB_CANDIDATES.items():
            for v in values:
                cfg = dict(best, **{name: int(v)})
                if better(cfg, best):
                    best = cfg
    return best, timings


def tune_batched(d_cat, e_cat, sizes, candidates=CANDIDATES, reps: int = 5,
                 cache_path: str | None = None, knobs: bool = True) -> dict:
    """Pick the launch shape for a batch.  d_cat, e_cat: the batch's
    concatenated rows as float64 device tensors (read only).  Returns
    {"leaf_max": best leaf, "options": {every tuned option}, "ms": {config:
    ms per batch}, "cached": bool}."""
    import torch
    dev = d_cat.device
    key = _signature("batch" + ("+knobs" if knobs else ""), sizes,
                     torch.cuda.get_device_name(dev))
    cache = _load(cache_path)
    if key in cache:
        return dict(cache[key], cached=True)

    def time_of(cfg):
        gs = GraphSolve(BatchSolver(sizes, device=dev, **cfg), d_cat, e_cat)
        ms = _time_graph(gs, reps)
        del gs
        return ms

    best, timings = _search(time_of, candidates, knobs)
    res = {"leaf_max": int(best["leaf_max"]), "options": best, "ms": timings}
    cache[key] = res
    _store(cache_path, cache)
    return dict(res, cached=False)


def tune_single(problems, candidates=(8, 12, 16, 24, 32), reps: int = 5,
                cache_path: str | None = None, knobs: bool = True) -> dict:
    """Pick the launch shape for single solves of one order n: problems is a
    list of (d, e) float64 device tensors of equal length (for instance one
    matrix of each kind the caller solves); the summed graph-replay time of
    all of them is minimised.  Returns {"options": {...}, "ms": {config: summed
    ms}, "cached": bool}."""
    import torch
    n = problems[0][0].numel()
    dev = problems[0][0].device
    key = _signature("single" + ("+knobs" if knobs else ""), [n, len(problems)],
                     torch.cuda.get_device_name(dev))
    cache = _load(cache_path)
    if key in cache:
        return dict(cache[key], cached=True)

    def time_of(cfg):
        total = 0.0
        for d, e in problems:
            gs = GraphSolve(Solver(n, device=dev, **cfg), d, e)
            total += _time_graph(gs, reps)
            del gs
        return total

    best, timings = _search(time_of, candidates, knobs)
    res = {"options": best, "ms": timings}
    cache[key] = res
    _store(cache_path, cache)
    return dict(res, cached=False)
