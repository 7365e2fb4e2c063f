"""The device secular root finder (ppopp27-brdc-ae_amd/csrc/brdc_secular.h),
compiled for the host, against LAPACK DLAED4 golden vectors and the oracle.
No GPU needed: the same source the gfx950 kernels inline is run on the CPU."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
LIB = os.path.join(HERE, "cpu_harness", "libsecular_host.so")
_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def sec():
    src = os.path.join(HERE, "cpu_harness", "secular_host.hip")
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        import subprocess
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-host-only", "-O2", "-fPIC", "-shared",
                        "-std=c++17", "-o", LIB, src], check=True)
    L = ctypes.CDLL(LIB)
    L.secular_host.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, _dp]
    L.secular_host.restype = ctypes.c_int
    return L


def run(L, d, z, rho, i):
    out = np.zeros(6)
    info = L.secular_host(d.shape[0], i, d.ctypes.data_as(_dp), z.ctypes.data_as(_dp), rho,
                          out.ctypes.data_as(_dp))
    return info, out


def test_device_secular_matches_lapack(sec):
    g = np.load(os.path.join(GOLD, "dlaed4_vectors.npz"))
    for c in range(int(g["ncases"][0])):
        d, z, rho = g[f"c{c}_d"], g[f"c{c}_z"], float(g[f"c{c}_rho"][0])
        K = d.shape[0]
        for i in range(K):
            info, out = run(sec, d, z, rho, i)
            assert info == 0
            ref = g[f"c{c}_lam"][i]
            assert abs(out[0] - ref) <= 4e-16 * max(1.0, abs(ref)), (c, i)
            if K > 2:
                # compact delta (org, tau, near1, near2) reproduces LAPACK's delta
                org = int(out[4])
                rec = (d - d[org]) - out[1]
                a = K - 2 if i == K - 1 else i
                rec[a], rec[a + 1] = out[2], out[3]
                np.testing.assert_allclose(rec, g[f"c{c}_delta"][i], rtol=2e-14, atol=0)
                assert (out[1] >= 0) == (org == i or i == K - 1)


def test_device_secular_matches_oracle_on_merge_like_inputs(sec, orc):
    rng = np.random.default_rng(77)
    for K in (3, 4, 17, 100, 600):
        d = np.sort(rng.standard_normal(K))
        z = rng.standard_normal(K) * np.where(rng.random(K) < 0.2, 1e-7, 1.0)
        z /= np.linalg.norm(z)
        rho = float(rng.uniform(0.1, 4.0))
        for i in range(K):
            info, out = run(sec, d, z, rho, i)
            lam, org, tau, n1, n2, delta, info2 = orc.laed4_br(d, z, rho, i)
            assert info == 0 and info2 == 0
            # both stop at |w| <= eps*erretm; the device recomputes delta from
            # (org, tau) instead of accumulating it, so roots agree to a few ulps
            assert abs(out[0] - lam) <= 1e-14 * max(1.0, abs(lam))


def test_secular_launch_slots_cover_every_merge(sec):
    """The secular launch gives each merge as many wave slots as the largest
    need of any merge size up to the level's maximum (the need is not monotone
    in the size), at every level of several tree shapes."""
    sec.secular_slot_violations.argtypes = [ctypes.c_int, ctypes.c_int]
    sec.secular_slot_violations.restype = ctypes.c_int
    for n in (1000, 3000, 16384, 49152, 100000, 262144):
        for leaf in (8, 16, 32, 64):
            assert sec.secular_slot_violations(n, leaf) == 0, (n, leaf)


@pytest.mark.parametrize("kind", ["normal", "uniform", "clustered", "toeplitz"])
def test_device_secular_on_every_merge_of_a_solve(sec, orc, kind):
    """Every secular problem the oracle's tree meets on a seeded n = 3000
    matrix (leaf 16): the device root finder converges
    on each root and agrees with the oracle's DLAED4."""
    from brdc.generators import generate, seed_for
    n = 3000
    d, e = generate(kind, n, seed_for(2, n))
    probs = orc.merge_problems(d, e, smlsiz=16)
    assert len(probs) == 2 ** 8 - 1
    checked = 0
    rng = np.random.default_rng(3)
    for (level, start, rho, dl, w) in probs:
        K = dl.shape[0]
        if K == 0:
            continue
        # every root of the big merges would be O(K^2) ctypes calls: sample
        roots = range(K) if K <= 64 else sorted(set(rng.integers(0, K, 48).tolist()) | {0, K - 1})
        for i in roots:
            info, out = run(sec, dl, w, rho, i)
            assert info == 0, (level, start, K, i)
            lam, _, info_ref = orc.laed4(dl, w, rho, i)
            assert info_ref == 0
            scale = max(abs(lam), np.max(np.abs(dl)))
            assert abs(out[0] - lam) <= 1e-14 * scale, (level, start, K, i)
            checked += 1
    assert checked > 100
