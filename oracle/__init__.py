"""ctypes bindings for the CPU oracle (libbrdc_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference's BR values-only
D&C (cpu/src/dstedc_dc_br.f, dlaed0_br.f, dlaed7_br.f, dlaed9_br.f,
dlaed4_br.f) and the LAPACK routines it calls.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product library never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbrdc_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


class MergeStat(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int), ("start", ctypes.c_int),
                ("size", ctypes.c_int), ("cut", ctypes.c_int),
                ("k", ctypes.c_int), ("givens", ctypes.c_int)]


def build() -> str:
    """Compile the oracle in place (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        L.brdc_oracle_dstebr_ex.restype = ctypes.c_int
        L.brdc_oracle_dstebr_ex.argtypes = [
            ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.POINTER(MergeStat), ctypes.c_int, _ip]
        L.brdc_oracle_laed4.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                        ctypes.c_double, _dp]
        L.brdc_oracle_laed4_br.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                           ctypes.c_double, _dp, _ip, _dp, _dp, _dp]
        L.brdc_oracle_laed5.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.brdc_oracle_laed6.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        _dp, _dp, ctypes.c_double, _dp]
        L.brdc_oracle_leaf.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.brdc_oracle_sterf.argtypes = [ctypes.c_int, _dp, _dp]
        L.brdc_oracle_lamrg.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                        ctypes.c_int, _ip]
        L.brdc_oracle_lamrg.restype = None
        L.brdc_oracle_wprod_nthreads.argtypes = []
        L.brdc_oracle_eig_quad.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    return a.ctypes.data_as(_ip)


def query(n: int, smlsiz: int = 0) -> tuple[int, int, int]:
    """Workspace query (LWORK = -1) for leaves of at most smlsiz rows (0: the
    default): returns (info, lwmin, liwmin)."""
    w = np.zeros(1)
    iw = np.zeros(1, dtype=np.int32)
    dd = np.zeros(max(n, 1))
    ee = np.zeros(max(n, 1))
    info = lib().brdc_oracle_dstebr_ex(n, _p(dd), _p(ee), _p(w), -1, _p(iw), -1, smlsiz, 0,
                                       None, 0, None)
    return info, int(w[0]), int(iw[0])


def dstebr(d, e, smlsiz: int = 0, nthreads: int = 0, stats: bool = False):
    """Eigenvalues (ascending) by the BR D&C restatement.

    Returns (w, info) or (w, info, merge_stats) when stats=True, where
    merge_stats is a structured numpy array (level, start, size, cut, k, givens).
    """
    d = np.array(d, dtype=np.float64, copy=True)
    n = d.shape[0]
    e = np.array(e, dtype=np.float64, copy=True).reshape(-1)
    e_buf = np.zeros(max(n, 1))
    e_buf[: e.shape[0]] = e
    info, lw, liw = query(n, smlsiz)
    work = np.zeros(max(lw, 1))
    iwork = np.zeros(max(liw, 1), dtype=np.int32)
    if stats:
        cap = 4 * max(n, 1) + 16
        st = (MergeStat * cap)()
        cnt = ctypes.c_int(0)
        info = lib().brdc_oracle_dstebr_ex(n, _p(d), _p(e_buf), _p(work), lw, _p(iwork), liw,
                                           smlsiz, nthreads, st, cap, ctypes.byref(cnt))
        rec = np.array([(s.level, s.start, s.size, s.cut, s.k, s.givens)
                        for s in st[: cnt.value]],
                       dtype=[("level", "i4"), ("start", "i4"), ("size", "i4"),
                              ("cut", "i4"), ("k", "i4"), ("givens", "i4")])
        return d, info, rec
    info = lib().brdc_oracle_dstebr_ex(n, _p(d), _p(e_buf), _p(work), lw, _p(iwork), liw,
                                       smlsiz, nthreads, None, 0, None)
    return d, info


def laed4(d, z, rho: float, i: int):
    """Root i (0-based): (lambda, delta, info)."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    delta = np.zeros(d.shape[0])
    lam = ctypes.c_double(0.0)
    info = lib().brdc_oracle_laed4(d.shape[0], i, _p(d), _p(z), _p(delta), rho, ctypes.byref(lam))
    return lam.value, delta, info


def laed4_br(d, z, rho: float, i: int):
    """Root i plus compact delta: (lambda, org, tau, near1, near2, delta, info)."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    delta = np.zeros(d.shape[0])
    lam, tau, n1, n2 = (ctypes.c_double(0.0) for _ in range(4))
    org = ctypes.c_int(0)
    info = lib().brdc_oracle_laed4_br(d.shape[0], i, _p(d), _p(z), _p(delta), rho,
                                      ctypes.byref(lam), ctypes.byref(org), ctypes.byref(tau),
                                      ctypes.byref(n1), ctypes.byref(n2))
    return lam.value, org.value, tau.value, n1.value, n2.value, delta, info


def laed5(d, z, rho: float, i: int):
    d = np.ascontiguousarray(d, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    delta = np.zeros(2)
    lam = ctypes.c_double(0.0)
    lib().brdc_oracle_laed5(i, _p(d), _p(z), _p(delta), rho, ctypes.byref(lam))
    return lam.value, delta


def laed6(kniter: int, orgati: bool, rho: float, d, z, finit: float):
    d = np.ascontiguousarray(d, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    tau = ctypes.c_double(0.0)
    info = lib().brdc_oracle_laed6(kniter, int(bool(orgati)), rho, _p(d), _p(z), finit,
                                   ctypes.byref(tau))
    return tau.value, info


def leaf(d, e):
    """Leaf eigensolve: (w ascending, first row, last row, info)."""
    d = np.array(d, dtype=np.float64, copy=True)
    n = d.shape[0]
    e = np.array(e, dtype=np.float64, copy=True).reshape(-1)
    e_buf = np.zeros(max(n, 1))
    e_buf[: e.shape[0]] = e
    blo = np.zeros(n)
    bhi = np.zeros(n)
    info = lib().brdc_oracle_leaf(n, _p(d), _p(e_buf), _p(blo), _p(bhi))
    return d, blo, bhi, info


def sterf(d, e):
    d = np.array(d, dtype=np.float64, copy=True)
    n = d.shape[0]
    e_buf = np.zeros(max(n, 1))
    e_buf[: max(n - 1, 0)] = np.asarray(e, dtype=np.float64)[: max(n - 1, 0)]
    info = lib().brdc_oracle_sterf(n, _p(d), _p(e_buf))
    return d, info


def eig_quad(d, e, k: int, lo: float, hi: float, steps: int = 120):
    """Eigenvalue k (0-based, ascending) by binary128 Sturm bisection from the
    bracket [lo, hi]; returns (hi, lo) of a double-double, or None when the
    bracket does not hold eigenvalue k."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    n = d.shape[0]
    e_buf = np.zeros(max(n, 1))
    e_buf[: max(n - 1, 0)] = np.asarray(e, dtype=np.float64)[: max(n - 1, 0)]
    out = np.zeros(2)
    rc = lib().brdc_oracle_eig_quad(n, _p(d), _p(e_buf), int(k), float(lo), float(hi),
                                    int(steps), _p(out))
    return None if rc else (float(out[0]), float(out[1]))


def eig_quad_probe(d, e, k: int, approx: float, steps: int = 70):
    """Eigenvalue k near approx: a bracket approx -/+ r with r doubled from
    1e-12 |T| until it holds the eigenvalue (the whole Gershgorin interval at
    most), then steps binary128 halvings."""
    d = np.asarray(d, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    ea = np.abs(e)
    rad = np.zeros(d.size)
    rad[:-1] += ea
    rad[1:] += ea
    glo, ghi = float(np.min(d - rad)), float(np.max(d + rad))
    tn = max(abs(glo), abs(ghi), 1e-300)
    r = 1e-12 * tn
    while True:
        lo, hi = max(approx - r, glo - r), min(approx + r, ghi + r)
        res = eig_quad(d, e, k, lo, hi, steps)
        if res is not None or (lo <= glo - r / 2 and hi >= ghi + r / 2):
            return res
        r *= 4.0


def lamrg(a, n1: int, n2: int, s1: int = 1, s2: int = 1):
    a = np.ascontiguousarray(a, dtype=np.float64)
    idx = np.zeros(n1 + n2, dtype=np.int32)
    lib().brdc_oracle_lamrg(n1, n2, _p(a), s1, s2, _p(idx))
    return idx


MERGE_CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            _dp, _dp)


def merge_problems(d, e, smlsiz: int = 0):
    """Every merge's secular problem of the BR restatement on (d, e):
    a list of (level, start, rho, dlambda[K], w[K]) in the order solved."""
    out = []

    def cb(level, start, k, rho, dl, w):
        out.append((level, start, rho, np.ctypeslib.as_array(dl, shape=(k,)).copy(),
                    np.ctypeslib.as_array(w, shape=(k,)).copy()))

    f = MERGE_CB(cb)
    L = lib()
    L.brdc_oracle_set_merge_callback.argtypes = [MERGE_CB]
    L.brdc_oracle_set_merge_callback.restype = None
    L.brdc_oracle_set_merge_callback(f)
    try:
        dstebr(d, e, smlsiz=smlsiz, nthreads=1)
    finally:
        L.brdc_oracle_set_merge_callback(MERGE_CB())
    return out
