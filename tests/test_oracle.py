"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Every oracle routine is checked against something that a plausible mistake
would fail: LAPACK golden vectors from scipy's OpenBLAS (tests/golden/), dense
eigh, closed-form spectra, and invariants of the spectrum.
"""
import os

import numpy as np
import pytest
import scipy.special

from brdc.generators import GENERATORS, generate, seed_for

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_err(w, ref):
    """Normwise max|dw| / max|ref| (scripts/cpu/smoke_br.f90:32-33)."""
    return float(np.max(np.abs(w - ref)) / max(np.max(np.abs(ref)), np.finfo(float).tiny))


# ---------------------------------------------------------------- generators
def test_generator_bits_match_fixture():
    g = np.load(os.path.join(GOLD, "dsterf_eigs.npz"))
    for kind in GENERATORS:
        d, e = generate(kind, 64, seed_for(1, 64))
        np.testing.assert_array_equal(d, g[f"gen_{kind}_64_d"])
        np.testing.assert_array_equal(e, g[f"gen_{kind}_64_e"])


def test_generator_statistics():
    d, e = generate("normal", 200000, 12345)
    assert abs(d.mean()) < 0.01 and abs(d.std() - 1.0) < 0.01
    d, e = generate("uniform", 200000, 12345)
    assert d.min() >= -1.0 and d.max() < 1.0 and abs(d.mean()) < 0.01
    d, e = generate("clustered", 1000, 7)
    centres = d[:8]
    assert np.all(np.abs(d - centres[np.arange(1000) % 8]) < 1e-6)
    assert np.all((e >= 0) & (e < 1e-6))


# ----------------------------------------------------------- secular solvers
def test_laed4_matches_lapack_golden(orc):
    g = np.load(os.path.join(GOLD, "dlaed4_vectors.npz"))
    for c in range(int(g["ncases"][0])):
        d, z, rho = g[f"c{c}_d"], g[f"c{c}_z"], float(g[f"c{c}_rho"][0])
        for i in range(d.shape[0]):
            lam, delta, info = orc.laed4(d, z, rho, i)
            assert info == g[f"c{c}_info"][i]
            ref_lam, ref_delta = g[f"c{c}_lam"][i], g[f"c{c}_delta"][i]
            assert abs(lam - ref_lam) <= 4e-16 * max(abs(ref_lam), 1.0), (c, i)
            # delta carries high relative accuracy: compare entrywise-relative
            np.testing.assert_allclose(delta, ref_delta, rtol=1e-14, atol=0, err_msg=f"{c},{i}")


def test_laed5_laed6_match_lapack_golden(orc):
    g = np.load(os.path.join(GOLD, "dlaed56_vectors.npz"))
    for row in g["laed5"]:
        lam, delta = orc.laed5(row[0:2], row[2:4], row[4], int(row[5]))
        assert abs(lam - row[6]) <= 4e-16 * max(abs(row[6]), 1.0)
        np.testing.assert_allclose(delta, row[7:9], rtol=1e-14, atol=1e-16)
    for row in g["laed6"]:
        tau, info = orc.laed6(int(row[0]), bool(row[1]), row[2], row[3:6], row[6:9], row[9])
        assert info == row[11]
        assert abs(tau - row[10]) <= 1e-15 * max(abs(row[10]), 1e-300) + 1e-300


def test_laed4_closed_forms_and_interlacing(orc):
    rng = np.random.default_rng(5)
    # K = 1: lambda = d + rho z^2
    lam, delta, info = orc.laed4(np.array([0.3]), np.array([0.8]), 2.0, 0)
    assert info == 0 and abs(lam - (0.3 + 2.0 * 0.64)) < 1e-15 and delta[0] == 1.0
    # K = 2: roots of the 2x2 matrix diag(d) + rho z z^T
    d = np.array([-0.4, 0.9]); z = np.array([0.6, 0.8]); rho = 1.3
    ref = np.linalg.eigvalsh(np.diag(d) + rho * np.outer(z, z))
    for i in range(2):
        lam, _, _ = orc.laed4(d, z, rho, i)
        assert abs(lam - ref[i]) < 1e-15
    # general K: interlacing d_i < lam_i < d_{i+1}, lam_K < d_K + rho|z|^2,
    # eigenvalues of diag(d) + rho z z^T, and delta_j = d_j - lam
    for n in (3, 9, 40):
        d = np.sort(rng.standard_normal(n)); z = rng.standard_normal(n); z /= np.linalg.norm(z)
        rho = 0.9
        ref = np.linalg.eigvalsh(np.diag(d) + rho * np.outer(z, z))
        for i in range(n):
            lam, delta, info = orc.laed4(d, z, rho, i)
            assert info == 0
            assert d[i] < lam and (lam < d[i + 1] if i < n - 1 else lam <= d[-1] + rho)
            assert abs(lam - ref[i]) < 1e-14
            np.testing.assert_allclose(delta, d - lam, rtol=1e-12, atol=1e-15)


def test_laed4_br_compact_delta_reconstructs(orc):
    """dlaed4_br.f:79-103: delta_j = (d_j - d_org) - tau, near = pole-adjacent deltas."""
    rng = np.random.default_rng(9)
    n = 30
    d = np.sort(rng.standard_normal(n)); z = rng.standard_normal(n); z /= np.linalg.norm(z)
    for i in range(n):
        lam, org, tau, n1, n2, delta, info = orc.laed4_br(d, z, 1.1, i)
        assert info == 0
        if i < n - 1:
            assert org == (i if lam < 0.5 * (d[i] + d[i + 1]) else i + 1)
            assert (tau >= 0) == (org == i)
            assert n1 == delta[i] and n2 == delta[i + 1]
        else:
            assert org == n - 1 and n1 == delta[n - 2] and n2 == delta[n - 1]
        assert tau == -delta[org]
        rec = (d - d[org]) - tau
        np.testing.assert_allclose(rec, delta, rtol=4e-15, atol=0)


# --------------------------------------------------------------------- leaf
@pytest.mark.parametrize("n", [1, 2, 3, 7, 25, 48, 64])
def test_leaf_matches_dense_eigh(orc, n):
    d, e = generate("normal", n, seed_for(9, n))
    w, blo, bhi, info = orc.leaf(d, e)
    assert info == 0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ref_w, V = np.linalg.eigh(T)
    assert rel_err(w, ref_w) < 1e-14
    np.testing.assert_allclose(np.abs(blo), np.abs(V[0, :]), atol=1e-12)
    np.testing.assert_allclose(np.abs(bhi), np.abs(V[-1, :]), atol=1e-12)
    # rows of an orthogonal matrix: unit norm; first vs last row orthogonal
    assert abs(np.dot(blo, blo) - 1) < 1e-13 and abs(np.dot(bhi, bhi) - 1) < 1e-13
    if n > 1:
        assert abs(np.dot(blo, bhi)) < 1e-13


# ------------------------------------------------------------- end to end
def test_dstebr_matches_dsterf_fixtures(orc):
    g = np.load(os.path.join(GOLD, "dsterf_eigs.npz"))
    d, e = generate("normal", 4096, seed_for(1, 4096))
    w, info = orc.dstebr(d, e)
    assert info == 0 and rel_err(w, g["normal_4096"]) < 1e-13
    for kind in GENERATORS:
        for n in (26, 64, 65, 300, 1000):
            d, e = generate(kind, n, seed_for(1, n))
            w, info = orc.dstebr(d, e)
            assert info == 0
            assert np.all(np.diff(w) >= 0)
            assert rel_err(w, g[f"{kind}_{n}"]) < 1e-13, (kind, n)


@pytest.mark.parametrize("n", [16384, 32768, 49152])
def test_dstebr_matches_dsterf_at_full_size(orc, n):
    """The plain-solve matrices at full size (four generators,
    seed_for(2, n)) against LAPACK DSTERF (tests/golden/dsterf_config2.npz)."""
    g = np.load(os.path.join(GOLD, "dsterf_config2.npz"))
    for kind in GENERATORS:
        d, e = generate(kind, n, seed_for(2, n))
        w, info = orc.dstebr(d, e)
        assert info == 0 and np.all(np.diff(w) >= 0)
        assert rel_err(w, g[f"{kind}_{n}"]) < 1e-13, (kind, n)


@pytest.mark.parametrize("smlsiz", [0, 8, 32, 64])
def test_dstebr_leaf_size_independent(orc, smlsiz):
    g = np.load(os.path.join(GOLD, "dsterf_eigs.npz"))
    for kind in GENERATORS:
        d, e = generate(kind, 1000, seed_for(1, 1000))
        w, info = orc.dstebr(d, e, smlsiz=smlsiz)
        assert info == 0 and rel_err(w, g[f"{kind}_1000"]) < 1e-13


def test_closed_form_spectra(orc):
    n = 777
    # Toeplitz a, b: a + 2b cos(k pi / (n+1))
    w, info = orc.dstebr(np.full(n, 2.0), np.full(n - 1, 1.0))
    ref = np.sort(2.0 + 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    assert info == 0 and rel_err(w, ref) < 1e-14
    # Kac-Sylvester-Clement: D = 0, E_i = sqrt(i (n-i)) -> -(n-1), -(n-3), ..., n-1
    i = np.arange(1, n)
    w, info = orc.dstebr(np.zeros(n), np.sqrt(i * (n - i)))
    assert info == 0 and rel_err(w, np.arange(-(n - 1), n, 2, dtype=float)) < 1e-13
    # Gauss-Legendre Jacobi matrix: nodes of P_n
    w, info = orc.dstebr(np.zeros(n), i / np.sqrt(4.0 * i * i - 1.0))
    assert info == 0 and rel_err(w, scipy.special.roots_legendre(n)[0]) < 1e-13
    # Gauss-Hermite: E_i = sqrt(i/2) -> roots of H_n
    m = 200
    j = np.arange(1, m)
    w, info = orc.dstebr(np.zeros(m), np.sqrt(j / 2.0))
    assert info == 0 and rel_err(w, scipy.special.roots_hermite(m)[0]) < 1e-13


def test_trace_and_frobenius_invariants(orc):
    for kind in GENERATORS:
        d, e = generate(kind, 3000, seed_for(2, 3000))
        w, info = orc.dstebr(d, e)
        assert info == 0
        scale = np.max(np.abs(w))
        assert abs(w.sum() - d.sum()) < 1e-11 * scale * 3000
        assert abs((w * w).sum() - (d * d).sum() - 2 * (e * e).sum()) < 1e-11 * scale ** 2 * 3000


def test_merge_stats_deflation_bookkeeping(orc):
    # Toeplitz halves are mirror images: every pole is doubled, so the root
    # merge deflates at least half of its poles (K <= N/2)
    n = 2048
    w, info, st = orc.dstebr(np.full(n, 2.0), np.full(n - 1, 1.0), stats=True)
    assert info == 0
    root = st[st["level"] == 0]
    assert root.shape[0] == 1 and root["size"][0] == n and root["k"][0] <= n // 2
    assert np.all(st["k"] <= st["size"]) and np.all(st["k"] >= 0)
    # every level of merges covers the matrix exactly once
    for lev in np.unique(st["level"]):
        assert st[st["level"] == lev]["size"].sum() == n
    # localised spectra deflate heavily
    d, e = generate("clustered", 4096, seed_for(2, 4096))
    w, info, st = orc.dstebr(d, e, stats=True)
    assert info == 0 and st[st["level"] == 0]["k"][0] < 4096 // 4


def test_query_and_argument_errors(orc):
    info, lw, liw = orc.query(4096)
    nthr = orc.lib().brdc_oracle_wprod_nthreads()
    assert info == 0 and lw == (16 + 2 * nthr) * 4096 and liw == 7 * 4096
    info, lw, liw = orc.query(25)
    assert (info, lw, liw) == (0, 1, 1)
    assert orc.query(-1)[0] == -1


# BRDC_WPROD_NTHREADS (dstedc_dc_br.f:153-213), written out from the
# reference's text: with no usable override the buffer count is a step of the
# OpenMP team size NT (NT <= 1 -> 1, <= 8 -> 2, <= 28 -> 4, else 16, capped at
# NT); a numeric OPENBLAS_BRDC_WPROD_THREADS gives max(1, min(value, NT));
# blanks are skipped, any other character voids the override, and 0 counts as
# no override.  Each case: (OMP_NUM_THREADS, override or None, expected).
WPROD_CASES = [(1, None, 1), (2, None, 2), (8, None, 2), (9, None, 4), (28, None, 4),
               (29, None, 16), (56, None, 16),
               (8, "3", 3), (8, "99", 8), (8, "3x", 2), (8, " 5 ", 5), (8, "0", 2),
               (1, "4", 1), (29, "7", 7), (56, "", 16)]

_WPROD_PROBE = """
import ctypes, os, sys
lib = ctypes.CDLL(sys.argv[1], mode=os.RTLD_LAZY)
f = getattr(lib, sys.argv[2])
f.restype = ctypes.c_int
print(f())
"""


def _wprod_in_subprocess(path, symbol, threads, override):
    """OpenMP reads OMP_NUM_THREADS when its runtime starts, so each setting
    runs in a process of its own."""
    import subprocess
    import sys
    env = {k: v for k, v in os.environ.items() if k != "OPENBLAS_BRDC_WPROD_THREADS"}
    env["OMP_NUM_THREADS"] = str(threads)
    if override is not None:
        env["OPENBLAS_BRDC_WPROD_THREADS"] = override
    out = subprocess.run([sys.executable, "-c", _WPROD_PROBE, path, symbol], env=env,
                         check=True, capture_output=True, text=True).stdout
    return int(out.strip())


@pytest.mark.parametrize("threads,override,expected", WPROD_CASES)
def test_wprod_nthreads_table(orc, threads, override, expected):
    assert _wprod_in_subprocess(orc.LIB_PATH, "brdc_oracle_wprod_nthreads", threads,
                                override) == expected


REF_DRIVER = os.path.join(os.path.dirname(GOLD), "..", "oracle", "_ref",
                          "libref_dstedc_dc_br.so")


@pytest.mark.skipif(not os.path.exists(REF_DRIVER),
                    reason="oracle/_ref not built (needs the reference tree and amdflang)")
@pytest.mark.parametrize("threads,override,expected", WPROD_CASES)
def test_wprod_nthreads_matches_reference_build(threads, override, expected):
    """The reference's own BRDC_WPROD_NTHREADS, compiled from
    cpu/src/dstedc_dc_br.f by oracle/Makefile's `ref` target, gives the same
    table."""
    assert _wprod_in_subprocess(REF_DRIVER, "brdc_wprod_nthreads_", threads,
                                override) == expected


def test_query_follows_wprod_threads(orc):
    """The workspace query's LWORK = (16 + 2*NTHR)*N follows the buffer
    count (dstedc_dc_br.f:51-59): at 29 threads NTHR is 16, so N = 4096 needs
    48*4096 doubles; an override of 3 gives 22*4096."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, sys.argv[1]); import oracle; "
            "print(*oracle.query(4096))")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for threads, override, nthr in ((29, None, 16), (8, "3", 3), (1, None, 1)):
        env = {k: v for k, v in os.environ.items() if k != "OPENBLAS_BRDC_WPROD_THREADS"}
        env["OMP_NUM_THREADS"] = str(threads)
        if override is not None:
            env["OPENBLAS_BRDC_WPROD_THREADS"] = override
        out = subprocess.run([sys.executable, "-c", code, root], env=env, check=True,
                             capture_output=True, text=True).stdout.split()
        assert [int(x) for x in out] == [0, (16 + 2 * nthr) * 4096, 7 * 4096]


def test_splits_zero_matrix_and_small(orc):
    # exact zeros in E split the matrix into independent blocks
    d, e = generate("normal", 500, 3)
    e[[40, 41, 300]] = 0.0
    w, info = orc.dstebr(d, e)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert info == 0 and rel_err(w, np.linalg.eigvalsh(T)) < 1e-13
    w, info = orc.dstebr(np.zeros(100), np.zeros(99))
    assert info == 0 and np.all(w == 0)
    for n in (1, 2, 5, 25):
        d, e = generate("uniform", n, 4)
        w, info = orc.dstebr(d, e)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert info == 0 and rel_err(w, np.linalg.eigvalsh(T)) < 1e-14


def test_split_blocks_keep_their_own_scale(orc):
    """dstedc_dc_br.f:86-140 scales each split block by its own norm (and
    sends blocks of 25 rows or fewer to DSTERF): a block of size 1e-200 next
    to one of size 1e150 keeps full relative accuracy in its eigenvalues,
    checked componentwise against the block alone (scaled to 1 and back, which
    is exact in binary floating point for these powers of ten up to rounding
    of one multiply)."""
    from scipy.linalg import eigh_tridiagonal
    big_d, big_e = generate("normal", 300, 21)
    for m in (20, 200):  # a DSTERF-sized block and a D&C-sized one
        sm_d, sm_e = generate("uniform", m, 22)
        d = np.concatenate([big_d * 1e150, sm_d * 1e-200])
        e = np.concatenate([big_e * 1e150, [0.0], sm_e * 1e-200])
        w, info = orc.dstebr(d, e)
        assert info == 0 and np.all(np.diff(w) >= 0)
        small = np.sort(w[np.abs(w) < 1e-100])
        ref = np.sort(eigh_tridiagonal(sm_d, sm_e, eigvals_only=True)) * 1e-200
        assert small.shape == (m,)
        assert np.max(np.abs(small - ref) / np.max(np.abs(ref))) < 1e-13
        large = np.sort(w[np.abs(w) >= 1e-100])
        refl = np.sort(eigh_tridiagonal(big_d, big_e, eigvals_only=True)) * 1e150
        assert rel_err(large, refl) < 1e-13


def test_sterf_matches_dsterf_fixtures(orc):
    """The oracle's values-only QL (the DSTERF role) against LAPACK DSTERF."""
    g = np.load(os.path.join(GOLD, "dsterf_eigs.npz"))
    for kind in GENERATORS:
        for n in (26, 300, 1000):
            d, e = generate(kind, n, seed_for(1, n))
            w, info = orc.sterf(d, e)
            assert info == 0 and rel_err(w, g[f"{kind}_{n}"]) <= 1e-13, (kind, n)


def test_lamrg_brute_force(orc):
    """DLAMRG: two sorted runs (ascending for stride +1, descending for -1)
    merged into one ascending index list; ties go to the first run."""
    rng = np.random.default_rng(5)
    for n1, n2 in ((0, 3), (3, 0), (1, 1), (5, 7), (40, 33)):
        for s1 in (1, -1):
            for s2 in (1, -1):
                a1 = np.sort(rng.integers(0, 6, n1).astype(float))
                a2 = np.sort(rng.integers(0, 6, n2).astype(float))
                a = np.concatenate([a1 if s1 > 0 else a1[::-1], a2 if s2 > 0 else a2[::-1]])
                idx = orc.lamrg(a, n1, n2, s1, s2)
                assert sorted(idx.tolist()) == list(range(n1 + n2))
                assert np.all(np.diff(a[idx]) >= 0)
                # ties: within equal values, every first-run index precedes
                # every second-run index
                for v in np.unique(a):
                    pos = [p for p in idx if a[p] == v]
                    side = [p >= n1 for p in pos]
                    assert side == sorted(side)


def test_eig_quad_closed_forms(orc):
    """The binary128 Sturm bisection against closed forms evaluated in
    40-digit arithmetic (mpmath): tridiag(1, 0, 1), eigenvalues 2 cos(k pi /
    (n + 1)), exactly representable entries, so the only error is the
    bisection's own (about 1e-33); and the Kac matrix D = 0, E_i = sqrt(i (n -
    i)) (eigenvalues -(n-1), -(n-3), ..., n-1) to the rounding of its E."""
    import mpmath
    mpmath.mp.dps = 40
    n = 300
    d, e = np.zeros(n), np.ones(n - 1)
    for k in (0, 1, 77, 150, 298, 299):
        exact = 2 * mpmath.cos((n - k) * mpmath.pi / (n + 1))
        hi, lo = orc.eig_quad_probe(d, e, k, float(exact) + 1e-9)
        assert abs(mpmath.mpf(hi) + mpmath.mpf(lo) - exact) < 1e-30
    m = 64
    ek = np.sqrt(np.arange(1, m) * (m - np.arange(1, m)))
    for k in (0, 5, 31, 63):
        hi, lo = orc.eig_quad_probe(np.zeros(m), ek, k, -(m - 1) + 2 * k + 0.3)
        assert abs(hi - (-(m - 1) + 2 * k)) <= 1e-13 * m


def test_eig_quad_agrees_with_oracle_and_rejects_bad_brackets(orc):
    """A random normal matrix: binary128 probes agree with the double BR
    oracle to its rounding, and a bracket that misses eigenvalue k is
    reported (None) rather than bisected."""
    n = 500
    d, e = generate("normal", n, seed_for(9, n))
    w, info = orc.dstebr(d, e)
    assert info == 0
    for k in (0, 123, 250, 499):
        hi, lo = orc.eig_quad_probe(d, e, k, w[k])
        assert abs(hi - w[k]) <= 1e-14 * np.max(np.abs(w))
    assert orc.eig_quad(d, e, 250, w[300] - 1e-9, w[300] + 1e-9) is None
