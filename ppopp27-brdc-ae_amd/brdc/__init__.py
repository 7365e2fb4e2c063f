"""brdc -- MI355X boundary-row values-only D&C tridiagonal eigensolver.

Host-side mirror of the reference's entry points for this path:

* :func:`dstebr` -- DSTEDC_DC_BR semantics (cpu/src/dstedc_dc_br.f:7-151) on
  host arrays: returns (eigenvalues ascending, INFO).
* :func:`dstebr_device` -- the device entry (the role of
  tridiag_solver_run_from_device_tridiagonal, gpu/src/tridiag_solver.cu:2541)
  on torch tensors already resident in HBM, enqueued on the current stream.
* :func:`workspace_bytes`, :func:`merge_stats` -- workspace query and per-merge
  diagnostics.

All compute runs in libbrdc.so (hand-written gfx950 HIP kernels); torch only
provides device memory and the stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BrdcError, options

__all__ = ["dstebr", "dstebr_device", "workspace_bytes", "merge_stats", "Solver", "BrdcError",
           "profile_kernels", "profile_levels", "GraphSolve", "BatchSolver", "options", "query"]


def query(n: int) -> tuple[int, int, int]:
    """Workspace query of the host entry (LWORK = -1): (info, lwmin, liwmin)."""
    L = _lib.load()
    w = (ctypes.c_double * 1)()
    iw = (ctypes.c_int * 1)()
    info = L.brdc_dstebr(int(n), None, None, w, -1, iw, -1)
    return info, int(w[0]), int(iw[0])


def dstebr(d, e):
    """Eigenvalues of the symmetric tridiagonal (d, e) via the host-array entry.

    Returns (w, info) with w ascending; d and e are not modified.
    """
    L = _lib.load()
    d = np.array(d, dtype=np.float64, copy=True)
    n = d.shape[0]
    e = np.ascontiguousarray(np.asarray(e, dtype=np.float64).reshape(-1)[: max(n - 1, 0)]).copy()
    if e.shape[0] != max(n - 1, 0):
        raise ValueError("e must hold n-1 entries")
    info, lw, liw = query(n)
    work = np.zeros(max(lw, 1))
    iwork = np.zeros(max(liw, 1), dtype=np.int32)
    e_buf = np.zeros(max(n, 1))
    e_buf[: e.shape[0]] = e
    info = L.brdc_dstebr(n, d.ctypes.data_as(_lib._dp), e_buf.ctypes.data_as(_lib._dp),
                         work.ctypes.data_as(_lib._dp), lw, iwork.ctypes.data_as(_lib._ip), liw)
    return d, int(info)


def workspace_bytes(n: int, leaf_max: int = 0, **tuning) -> int:
    """Device workspace bytes of an order-n solve with these options (the
    leaf bound, and the subtree knob that picks the default one)."""
    L = _lib.load()
    out = ctypes.c_size_t(0)
    o = options(leaf_max, **tuning)
    _lib.check(L.brdc_dstebr_device_workspace_ex(int(n), ctypes.byref(o), ctypes.byref(out)),
               "workspace query")
    return int(out.value)


class Solver:
    """A reusable device solve for one order n: owns the workspace and info word.

    ``solve(d, e)`` enqueues on the current torch stream and overwrites the
    device tensor ``d`` with the eigenvalues; nothing synchronises.
    """

    def __init__(self, n: int, device="cuda", leaf_max: int = 0, **tuning):
        import torch
        self.n = int(n)
        self.opts = options(leaf_max, **tuning)
        self.bytes = workspace_bytes(self.n, leaf_max, **tuning)
        self.ws = torch.empty(max(self.bytes, 256), dtype=torch.uint8, device=device)
        self.info = torch.zeros(1, dtype=torch.int32, device=device)
        if self.ws.data_ptr() % 256:
            raise BrdcError("workspace allocation is not 256-byte aligned")

    def solve(self, d, e, stream=None):
        import torch
        if d.dtype != torch.float64 or e.dtype != torch.float64:
            raise TypeError("d and e must be float64")
        if not (d.is_cuda and e.is_cuda and d.is_contiguous() and e.is_contiguous()):
            raise ValueError("d and e must be contiguous device tensors")
        if d.numel() != self.n or e.numel() < max(self.n - 1, 0):
            raise ValueError("shape mismatch with the solver's n")
        st = stream if stream is not None else torch.cuda.current_stream()
        L = _lib.load()
        _lib.check(L.brdc_dstebr_device_ex(self.n, d.data_ptr(), e.data_ptr() if self.n > 1 else None,
                                           self.ws.data_ptr(), self.bytes, st.cuda_stream,
                                           self.info.data_ptr(), ctypes.byref(self.opts)),
                   "brdc_dstebr_device")
        return d

    def merge_stats(self, stream=None):
        """Per-merge records of the last solve: level, start, size, K, G
        (Givens deflations), evals (secular evaluations, all roots), tiny
        (tiny-z deflations), evals_max (most evaluations of one root),
        failed (roots that did not converge)."""
        import torch
        L = _lib.load()
        st = stream if stream is not None else torch.cuda.current_stream()
        total = max(1, self.n)
        buf = np.zeros(_lib.MERGE_STAT_FIELDS * total, dtype=np.int32)
        cnt = L.brdc_merge_stats(self.n, self.ws.data_ptr(), self.bytes, st.cuda_stream,
                                 ctypes.byref(self.opts), buf.ctypes.data_as(_lib._ip), total)
        if cnt < 0:
            _lib.check(-cnt, "merge stats")
        return _merge_records(buf, cnt)


def _merge_records(buf, cnt):
    F = _lib.MERGE_STAT_FIELDS
    rec = buf[: F * cnt].reshape(cnt, F)
    out = np.zeros(cnt, dtype=[(name, "i4") for name in _lib.MERGE_STAT_NAMES])
    for i, name in enumerate(_lib.MERGE_STAT_NAMES):
        out[name] = rec[:, i]
    return out


class BatchSolver:
    """Batched solve of many tridiagonals of uneven sizes (one flattened merge
    forest; no counterpart in the reference, which solves one matrix a call).

    ``sizes`` fixes the batch shape: the plan (every matrix's tree and the
    per-height launch shapes) is built once and copied into this object's
    workspace.  ``solve(d, e)`` takes the concatenated rows (see
    :func:`brdc.generators.pack_batch`), overwrites ``d`` with each matrix's
    eigenvalues ascending in its own rows and leaves per-matrix INFO in
    ``self.info``; it enqueues on the current torch stream and does not
    synchronise.
    """

    def __init__(self, sizes, device="cuda", leaf_max: int = 0, **tuning):
        import torch
        L = _lib.load()
        self.sizes = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32).reshape(-1))
        self.nmat = int(self.sizes.shape[0])
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes, dtype=np.int64)])
        self.rows = int(self.offsets[-1])
        self.opts = options(leaf_max, **tuning)
        self._plan = ctypes.c_void_p()
        _lib.check(L.brdc_batch_plan_create(self.nmat, self.sizes.ctypes.data_as(_lib._ip),
                                            ctypes.byref(self.opts), ctypes.byref(self._plan)),
                   "batch plan")
        nbytes = ctypes.c_size_t(0)
        _lib.check(L.brdc_batch_workspace(self._plan, ctypes.byref(nbytes)), "batch workspace")
        self.bytes = int(nbytes.value)
        self.ws = torch.empty(max(self.bytes, 256), dtype=torch.uint8, device=device)
        if self.ws.data_ptr() % 256:
            raise BrdcError("workspace allocation is not 256-byte aligned")
        self.info = torch.zeros(max(self.nmat, 1), dtype=torch.int32, device=device)
        st = torch.cuda.current_stream(self.ws.device)
        _lib.check(L.brdc_batch_upload(self._plan, self.ws.data_ptr(), self.bytes, st.cuda_stream),
                   "batch upload")

    def __del__(self):
        # at interpreter exit the module globals may already be cleared; the
        # process is ending then and the host-side plan goes with it
        plan = getattr(self, "_plan", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if plan and lib is not None:
            lib.brdc_batch_plan_destroy(plan)
            self._plan = None

    def stats(self) -> dict:
        out = np.zeros(7, dtype=np.int32)
        k = _lib.load().brdc_batch_plan_stats(self._plan, out.ctypes.data_as(_lib._ip), 7)
        if k < 0:
            _lib.check(-k, "batch plan stats")
        keys = ("rows", "matrices", "leaves", "heights", "merges", "largest_leaf", "leaf_bound")
        return {kk: int(v) for kk, v in zip(keys, out[:k])}

    def solve(self, d, e, stream=None):
        import torch
        if d.dtype != torch.float64 or e.dtype != torch.float64:
            raise TypeError("d and e must be float64")
        if not (d.is_cuda and e.is_cuda and d.is_contiguous() and e.is_contiguous()):
            raise ValueError("d and e must be contiguous device tensors")
        if d.numel() != self.rows or e.numel() < self.rows:
            raise ValueError("d and e must hold the batch's concatenated rows")
        st = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(_lib.load().brdc_batch_solve(self._plan, d.data_ptr(), e.data_ptr(),
                                                self.ws.data_ptr(), self.bytes, st.cuda_stream,
                                                self.info.data_ptr()),
                   "brdc_batch_solve")
        return d

    def merge_stats(self, stream=None):
        """Per-merge records of the last batched solve (fields as
        Solver.merge_stats; level = forest height, start in the concatenated
        rows)."""
        import torch
        L = _lib.load()
        st = stream if stream is not None else torch.cuda.current_stream()
        total = max(1, self.stats()["merges"])
        buf = np.zeros(_lib.MERGE_STAT_FIELDS * total, dtype=np.int32)
        cnt = L.brdc_batch_merge_stats(self._plan, self.ws.data_ptr(), self.bytes,
                                       st.cuda_stream, buf.ctypes.data_as(_lib._ip), total)
        if cnt < 0:
            _lib.check(-cnt, "batch merge stats")
        return _merge_records(buf, cnt)

    def split(self, w):
        """Per-matrix views of concatenated rows (numpy or torch)."""
        return [w[int(self.offsets[b]):int(self.offsets[b + 1])] for b in range(self.nmat)]


def dstebr_device(d, e, leaf_max: int = 0):
    """Eigenvalues of device tensors (d, e); returns (w, info) with w a new tensor."""
    import torch
    s = Solver(d.numel(), device=d.device, leaf_max=leaf_max)
    w = d.clone()
    s.solve(w, e.contiguous())
    torch.cuda.synchronize(d.device)
    return w, int(s.info.item())


def merge_stats(solver: Solver):
    return solver.merge_stats()


#: algorithmic HBM bytes per matrix row for the memory-bound kernels, per
#: merge the row takes part in (prepare: read d, blo, bhi + write d, z, q1,
#: q2; deflate: read 4 + write 4 doubles plus flags and indices; finalize:
#: read 3 + write 3 doubles) or once per solve (leaf: read d, e + write d,
#: blo, bhi; norm: read d, e; prep_e: read e + write the scaled e)
_BYTES_PER_ROW = {"prepare": 56, "deflate": 112, "finalize": 48, "leaf": 40, "norm": 16,
                  "prep_e": 16, "subtree": 40}
# classes whose bytes are paid once per row (the others once per row of
# every launched merge)
_ROW_BYTE_CLASSES = ("leaf", "norm", "prep_e", "subtree")

# leaf bounds when brdc_options.leaf_max is 0 (kDefaultLeaf, kSubtreeLeaf)
# and the default of subtree_max_log2 (kSubMaxLog2), as in the HIP driver
_DEFAULT_LEAF = 16
_SUBTREE_LEAF = 8
_SUB_MAX_LOG2 = 10


def _tree_levels(n: int, leaf: int) -> int:
    k = 0
    while -(-n // (1 << k)) > leaf:
        k += 1
    return k


def _subtree_depth(n: int, leaf: int) -> int:
    """Depth of the subtree roots of an order-n tree (-1: none); mirrors
    subtree_depth in the HIP driver."""
    if n <= 1:
        return -1
    levels = _tree_levels(n, leaf)
    for depth in range(levels):
        if -(-n // (1 << depth)) <= _lib.SUBTREE_ROWS:
            return depth if levels - depth <= 6 else -1
    return -1


def _subtree_limit(opts) -> int:
    lg = opts.subtree_max_log2 or _SUB_MAX_LOG2
    return 0 if lg < 0 else 1 << lg


def effective_leaf(sizes, opts) -> int:
    """The leaf bound a solve over matrices of these orders uses: the option
    if set, else 8 when the subtree kernel takes the bottom of every tree at
    that bound (at most 2^subtree_max_log2 subtree roots in all), else 16."""
    if opts.leaf_max:
        return int(opts.leaf_max)
    roots = sum(1 << d for d in (_subtree_depth(int(n), _SUBTREE_LEAF) for n in sizes) if d >= 0)
    return _SUBTREE_LEAF if 0 < roots <= _subtree_limit(opts) else _DEFAULT_LEAF


def in_subtree(solver, stats) -> np.ndarray:
    """Mask of the merge records (merge_stats) solved inside the subtree
    kernel rather than by the level launches: the subtrees of every matrix
    when their roots number at most 2^subtree_max_log2 (over a whole batch),
    none otherwise."""
    sizes = solver.sizes if isinstance(solver, BatchSolver) else np.array([solver.n])
    leaf = effective_leaf(sizes, solver.opts)
    depth = np.array([_subtree_depth(int(n), leaf) for n in sizes], dtype=np.int64)
    lev = np.array([_tree_levels(int(n), leaf) if n > 0 else 0 for n in sizes], dtype=np.int64)
    roots = int(np.where(depth >= 0, 2 ** np.maximum(depth, 0), 0).sum())
    hs = np.where(depth >= 0, lev - depth, 0)
    if roots > _subtree_limit(solver.opts):
        hs[:] = 0
    if isinstance(solver, BatchSolver):
        mat = np.searchsorted(solver.offsets, stats["start"], side="right") - 1
        return stats["level"] <= hs[mat]
    return (hs[0] > 0) & (stats["level"] >= lev[0] - hs[0])


def _level_max_size(solver, stats) -> np.ndarray:
    """Per merge record, the largest merge of its level launch (the maxsz the
    driver sizes that level's launches by): the largest record of the same
    level outside the subtree kernel."""
    sub = in_subtree(solver, stats)
    out = np.zeros(len(stats), dtype=np.int64)
    for lv in np.unique(stats["level"]):
        m = (stats["level"] == lv) & ~sub
        if m.any():
            out[stats["level"] == lv] = int(stats["size"][m].max())
    return out


def _root_merges(solver, stats) -> np.ndarray:
    """Mask of the records that are a matrix's root merge (no boundary rows)."""
    if isinstance(solver, BatchSolver):
        return np.isin(stats["start"], solver.offsets[:-1]) & \
            (stats["size"] == solver.sizes[np.searchsorted(solver.offsets, stats["start"],
                                                           side="right") - 1])
    return stats["level"] == 0


def profile_kernels(solver: Solver, problems, stream=None):
    """Run one solve per (w, d, e) in problems with kernel-class timing on
    (solver: a Solver, or a BatchSolver whose problems hold the concatenated
    rows).

    Returns {"kernels": {class: {"ms", "launches", "avg_us"}}, "flops": {class:
    FP64 flops}, "bytes": {class: bytes}} summed over the problems, where the
    flop counts use each solve's own per-merge K and secular evaluation
    counts: 7 flops per pole per evaluation (secular for merges below
    the solver's streaming threshold (its stream_min_k option, else
    STREAM_MIN_K); secular_stream for the others' evaluations after the
    first), 4 per pole for the first, midpoint evaluation of the merges of
    the threshold or more (secular_mid: the sum alone), 3 per pole pair
    (product-form w) and 9 per pole pair (boundary rows) of the non-root
    merges that k_wprod and k_rows solve (more than SMALL_MERGE_TOP poles, at
    levels of nodes of more than SMALL_MERGE rows; the finalize launch solves
    the others).
    """
    import torch
    L = _lib.load()
    nc = len(_lib.KERNEL_CLASSES)
    tot_ms = np.zeros(nc)
    tot_n = np.zeros(nc, dtype=np.int64)
    flops = {"secular": 0.0, "secular_stream": 0.0, "secular_mid": 0.0, "wprod": 0.0,
             "rows": 0.0, "subtree": 0.0}
    nbytes = {k: 0.0 for k in _BYTES_PER_ROW}
    st = stream if stream is not None else torch.cuda.current_stream()
    for (w, d, e) in problems:
        w.copy_(d)
        torch.cuda.synchronize()
        _lib.check(L.brdc_profile_begin(), "profile begin")
        solver.solve(w, e, st)
        ms = np.zeros(nc)
        cnt = np.zeros(nc, dtype=np.int32)
        _lib.check(L.brdc_profile_end(ms.ctypes.data_as(_lib._dp), cnt.ctypes.data_as(_lib._ip), nc),
                   "profile end")
        tot_ms += ms
        tot_n += cnt
        stats = solver.merge_stats(st)
        sub = in_subtree(solver, stats)
        k = stats["k"].astype(np.float64)
        ev = stats["evals"].astype(np.float64)
        kstream = solver.opts.stream_min_k or _lib.STREAM_MIN_K
        big = (stats["k"] >= kstream) & ~sub
        small = (stats["k"] < kstream) & ~sub
        flops["secular"] += float(np.sum(7.0 * k[small] * ev[small]))
        # every root of a streamed merge has its first evaluation in secular_mid
        flops["secular_stream"] += float(np.sum(7.0 * k[big] * (ev[big] - k[big])))
        flops["secular_mid"] += float(np.sum(4.0 * k[big] * k[big]))
        root = _root_merges(solver, stats)
        nonroot = ~root & (stats["k"] > 2)
        # merges of at most SMALL_MERGE poles at levels of nodes of at most
        # SMALL_MERGE rows, else of at most SMALL_MERGE_TOP poles, get their
        # product form and rows in the finalize launch (an HBM class, whose
        # time then includes them); the larger ones in wprod and rows
        kmax = np.where(_level_max_size(solver, stats) <= _lib.SMALL_MERGE, _lib.SMALL_MERGE,
                        _lib.SMALL_MERGE_TOP)
        wide = stats["k"] > kmax
        flops["wprod"] += float(np.sum(3.0 * k[nonroot & ~sub & wide] ** 2))
        flops["rows"] += float(np.sum(9.0 * k[nonroot & ~sub & wide] ** 2))
        flops["subtree"] += float(np.sum(7.0 * k[sub] * ev[sub]) +
                                  np.sum(12.0 * k[nonroot & sub] ** 2))
        # every launched merge moves its rows once; the subtree kernel reads
        # and writes each of its rows once
        merged_rows = float(stats["size"][~sub].sum())
        rows = float(getattr(solver, "rows", getattr(solver, "n", 0)))
        for kname, b in _BYTES_PER_ROW.items():
            nbytes[kname] += b * (rows if kname in _ROW_BYTE_CLASSES else merged_rows)
    kern = {}
    for i, name in enumerate(_lib.KERNEL_CLASSES):
        if tot_n[i]:
            kern[name] = {"ms": float(tot_ms[i]), "launches": int(tot_n[i]),
                          "avg_us": float(tot_ms[i] * 1e3 / tot_n[i])}
    return {"kernels": kern, "flops": flops, "bytes": nbytes}


def profile_levels(solver, w, d, e, stream=None, nlevels=64):
    """Kernel-class milliseconds of one solve split by merge level (the
    per-level timers of the reference's TridiagMergeLevelDiagnostics):
    returns {level: {class: ms}} with level -1 for the launches before the
    merges (norm, prep_e, leaf) and l >= 0 for the merges of tree depth l
    (0 = root merge), or forest height l for a BatchSolver."""
    import torch
    L = _lib.load()
    nc = len(_lib.KERNEL_CLASSES)
    st = stream if stream is not None else torch.cuda.current_stream()
    w.copy_(d)
    torch.cuda.synchronize()
    _lib.check(L.brdc_profile_begin(), "profile begin")
    solver.solve(w, e, st)
    ms = np.zeros((nlevels + 1) * nc)
    cnt = np.zeros((nlevels + 1) * nc, dtype=np.int32)
    _lib.check(L.brdc_profile_end_levels(ms.ctypes.data_as(_lib._dp),
                                         cnt.ctypes.data_as(_lib._ip), nc, nlevels),
               "profile end")
    ms = ms.reshape(nlevels + 1, nc)
    cnt = cnt.reshape(nlevels + 1, nc)
    out = {}
    for row in range(nlevels + 1):
        if cnt[row].sum():
            out[row - 1] = {name: float(ms[row, i]) for i, name in
                            enumerate(_lib.KERNEL_CLASSES) if cnt[row, i]}
    return out


class GraphSolve:
    """One solve captured into a HIP graph (via torch's graph capture).

    The device entry point does no host synchronisation and no allocation, so
    the whole launch sequence (copy of d into the solver's output buffer, then
    every level's kernels) records into one graph; ``run()`` replays it on the
    current stream.  Inputs are read from the tensors given at capture time.
    """

    def __init__(self, solver: Solver, d, e):
        import torch
        self.solver = solver
        self.d, self.e = d, e
        self.w = torch.empty_like(d)
        self.graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=d.device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up outside capture
            self.w.copy_(d)
            solver.solve(self.w, e)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize(d.device)
        with torch.cuda.graph(self.graph):
            self.w.copy_(d)
            solver.solve(self.w, e)
        torch.cuda.synchronize(d.device)

    def run(self):
        self.graph.replay()
        return self.w
