"""Writes the golden fixtures under tests/golden/ (run once, in a container with scipy).

Sources (third-party LAPACK as bundled by scipy 1.15 / OpenBLAS 0.3.28, the
library family the reference links, cpu/README.md):
  * scipy.linalg.lapack.dsterf            -> eigenvalue fixtures (DSTERF is the
    reference's acceptance oracle, scripts/cpu/smoke_br.f90:30)
  * scipy_dlaed4_/scipy_dlaed5_/scipy_dlaed6_ via ctypes -> per-root secular
    vectors (the reference calls DLAED4 at dlaed4_br.f:59 and dlaed9_br.f:87-127)

Inputs come from brdc.generators, so the fixtures also pin the generators.
Usage:  python tests/golden/make_golden.py [reduced | config2]
        (``reduced``: only the DSTERF fixture of reduced_dense_49152.de.bin;
        ``config2``: only the DSTERF fixtures of the N >= 16384 plain solves)
"""
from __future__ import annotations

import ctypes
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ppopp27-brdc-ae_amd"))

from brdc.generators import GENERATORS, generate, seed_for  # noqa: E402


def openblas():
    import scipy
    cands = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                   "libscipy_openblas*.so"))
    if not cands:
        raise RuntimeError("scipy's bundled OpenBLAS not found")
    return ctypes.CDLL(cands[0])


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lapack_dlaed4(L, d, z, rho, i):
    n = ctypes.c_int(d.shape[0])
    ii = ctypes.c_int(i + 1)
    delta = np.zeros(d.shape[0])
    r = ctypes.c_double(rho)
    lam = ctypes.c_double(0.0)
    info = ctypes.c_int(0)
    L.scipy_dlaed4_(ctypes.byref(n), ctypes.byref(ii), d.ctypes.data_as(_dp),
                    z.ctypes.data_as(_dp), delta.ctypes.data_as(_dp), ctypes.byref(r),
                    ctypes.byref(lam), ctypes.byref(info))
    return lam.value, delta, info.value


def lapack_dlaed5(L, d, z, rho, i):
    ii = ctypes.c_int(i + 1)
    delta = np.zeros(2)
    r = ctypes.c_double(rho)
    lam = ctypes.c_double(0.0)
    L.scipy_dlaed5_(ctypes.byref(ii), d.ctypes.data_as(_dp), z.ctypes.data_as(_dp),
                    delta.ctypes.data_as(_dp), ctypes.byref(r), ctypes.byref(lam))
    return lam.value, delta


def lapack_dlaed6(L, kniter, orgati, rho, d, z, finit):
    k = ctypes.c_int(kniter)
    o = ctypes.c_int(1 if orgati else 0)
    r = ctypes.c_double(rho)
    f = ctypes.c_double(finit)
    tau = ctypes.c_double(0.0)
    info = ctypes.c_int(0)
    L.scipy_dlaed6_(ctypes.byref(k), ctypes.byref(o), ctypes.byref(r), d.ctypes.data_as(_dp),
                    z.ctypes.data_as(_dp), ctypes.byref(f), ctypes.byref(tau), ctypes.byref(info))
    return tau.value, info.value


def secular_cases(rng):
    """(d, z, rho) cases: random, clustered poles, tiny weights, extreme rho."""
    cases = []
    for n in (1, 2, 3, 4, 5, 7, 8, 16, 33, 64, 128):
        d = np.sort(rng.standard_normal(n))
        z = rng.standard_normal(n)
        z /= np.linalg.norm(z)
        cases.append((d, z, 1.0 + rng.random()))
    for n in (6, 40):
        base = np.sort(rng.standard_normal(n))
        d = np.sort(base + np.arange(n) * 1e-9)  # close but distinct poles
        z = rng.standard_normal(n)
        z /= np.linalg.norm(z)
        cases.append((d, z, 0.7))
    for n in (10, 50):
        d = np.sort(rng.uniform(-1, 1, n))
        z = rng.standard_normal(n) * np.where(rng.random(n) < 0.3, 1e-6, 1.0)
        z /= np.linalg.norm(z)
        cases.append((d, z, 2.0))
    for rho in (1e-8, 1e4):
        d = np.sort(rng.uniform(-1, 1, 20))
        z = rng.standard_normal(20)
        z /= np.linalg.norm(z)
        cases.append((d, z, rho))
    # equally spaced poles with equal weights (Toeplitz-like merges)
    d = np.linspace(-1.0, 1.0, 30)
    z = np.full(30, 1.0 / np.sqrt(30))
    cases.append((d, z, 1.0))
    return cases


def reduced_dense_fixture():
    """DSTERF eigenvalues of the reference's shipped input
    data/inputs/reduced_dense_49152.de.bin (kept here as
    reduced_dense_49152.de.bin: the compact tridiagonal of a reduced dense
    matrix, int64 n then D then E)."""
    from scipy.linalg.lapack import dsterf
    from brdc.generators import read_de_bin
    d, e = read_de_bin(os.path.join(HERE, "reduced_dense_49152.de.bin"))
    w, info = dsterf(d.copy(), e.copy())[:2]
    assert info == 0
    np.savez_compressed(os.path.join(HERE, "reduced_dense_49152_dsterf.npz"), w=w)


CONFIG2_SIZES = (16384, 32768, 49152)


def config2_fixture():
    """DSTERF eigenvalues of the twelve plain-solve matrices: N in {16384,
    32768, 49152} x {uniform, normal, toeplitz, clustered}, drawn with
    ``generate(kind, n, seed_for(2, n))`` (the inputs the GPU tests and the
    bench solve).  DSTERF is O(N^2) and sequential: about 4 minutes for all
    twelve in one core."""
    from scipy.linalg.lapack import dsterf
    ev = {}
    for n in CONFIG2_SIZES:
        for kind in GENERATORS:
            d, e = generate(kind, n, seed_for(2, n))
            w, info = dsterf(d.copy(), e.copy())[:2]
            assert info == 0
            ev[f"{kind}_{n}"] = w
            print(f"dsterf {kind} n={n} done", flush=True)
    np.savez_compressed(os.path.join(HERE, "dsterf_config2.npz"), **ev)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "reduced":
        reduced_dense_fixture()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "config2":
        config2_fixture()
        return
    L = openblas()
    from scipy.linalg.lapack import dsterf

    # 1. DSTERF eigenvalue fixtures: config-1 matrix and small matrices of every generator
    ev = {}
    d, e = generate("normal", 4096, seed_for(1, 4096))
    w, info = dsterf(d.copy(), e.copy())[:2]
    assert info == 0
    ev["normal_4096"] = w
    for kind in GENERATORS:
        for n in (26, 64, 65, 300, 1000):
            d, e = generate(kind, n, seed_for(1, n))
            w, info = dsterf(d.copy(), e.copy())[:2]
            assert info == 0
            ev[f"{kind}_{n}"] = w
    # the raw generator outputs at one size pin the generator bits themselves
    for kind in GENERATORS:
        d, e = generate(kind, 64, seed_for(1, 64))
        ev[f"gen_{kind}_64_d"] = d
        ev[f"gen_{kind}_64_e"] = e
    np.savez_compressed(os.path.join(HERE, "dsterf_eigs.npz"), **ev)

    # 2. DLAED4 per-root vectors
    rng = np.random.default_rng(20260101)
    arrays = {}
    for ci, (d, z, rho) in enumerate(secular_cases(rng)):
        n = d.shape[0]
        lams = np.zeros(n)
        deltas = np.zeros((n, n))
        infos = np.zeros(n, dtype=np.int32)
        for i in range(n):
            lams[i], deltas[i], infos[i] = lapack_dlaed4(L, d, z, rho, i)
        arrays[f"c{ci}_d"] = d
        arrays[f"c{ci}_z"] = z
        arrays[f"c{ci}_rho"] = np.array([rho])
        arrays[f"c{ci}_lam"] = lams
        arrays[f"c{ci}_delta"] = deltas
        arrays[f"c{ci}_info"] = infos
    arrays["ncases"] = np.array([len(secular_cases(np.random.default_rng(20260101)))])
    np.savez_compressed(os.path.join(HERE, "dlaed4_vectors.npz"), **arrays)

    # 3. DLAED5 and DLAED6 vectors
    a5 = []
    for _ in range(40):
        d = np.sort(rng.standard_normal(2))
        z = rng.standard_normal(2)
        z /= np.linalg.norm(z)
        rho = float(rng.uniform(0.1, 3.0))
        for i in range(2):
            lam, delta = lapack_dlaed5(L, d, z, rho, i)
            a5.append([d[0], d[1], z[0], z[1], rho, i, lam, delta[0], delta[1]])
    a6 = []
    for _ in range(200):
        # three poles around an origin at 0 (shifted deltas), a root bracketed
        # between the two poles on the orgati side
        orgati = bool(rng.random() < 0.5)
        gaps = rng.uniform(0.05, 2.0, 2)
        if orgati:  # current point 0 lies between the 2nd and 3rd pole
            dd = np.array([-gaps[0], -0.1 * gaps[0], gaps[1]])
        else:       # current point 0 lies between the 1st and 2nd pole
            dd = np.array([-gaps[0], 0.1 * gaps[1], gaps[1]])
        zz = rng.uniform(0.01, 1.0, 3)
        rho = float(rng.uniform(-0.5, 0.5))
        finit = rho + float(np.sum(zz / dd))
        for kniter in (2, 3):
            tau, info = lapack_dlaed6(L, kniter, orgati, rho, dd, zz, finit)
            a6.append([kniter, int(orgati), rho, dd[0], dd[1], dd[2], zz[0], zz[1], zz[2],
                       finit, tau, info])
    np.savez_compressed(os.path.join(HERE, "dlaed56_vectors.npz"), laed5=np.array(a5),
                        laed6=np.array(a6))
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
