"""The plain solves at full size: N in {16384, 32768, 49152} x {uniform,
normal, toeplitz, clustered} (inputs ``generate(kind, n, seed_for(2, n))``),
through the C ABI on the GPU.

Each solve is held to the reference's acceptance bar, max|w - ref| /
max|ref| <= 1e-10 (scripts/cpu/smoke_br.f90:30-38), against three references:
LAPACK DSTERF (tests/golden/dsterf_config2.npz, made by make_golden.py with
scipy's OpenBLAS), the CPU oracle on the same input, and for Toeplitz D=2,
E=1 the closed form 2 + 2 cos(k pi / (n + 1)).  The eigenvalues must come out
ascending, with the trace and Frobenius norm of the input.
"""
import os

import numpy as np
import pytest

from brdc.generators import GENERATORS, generate, seed_for

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


def rel_err(w, ref):
    return float(np.max(np.abs(w - ref)) / max(np.max(np.abs(ref)), np.finfo(float).tiny))


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brdc
    brdc._lib.load()
    return torch


@pytest.fixture(scope="module")
def dsterf():
    return np.load(os.path.join(GOLD, "dsterf_config2.npz"))


@pytest.mark.parametrize("n", [16384, 32768, 49152])
@pytest.mark.parametrize("kind", GENERATORS)
def test_full_size_against_dsterf_oracle_and_invariants(gpu, orc, dsterf, kind, n):
    import brdc
    torch = gpu
    print(f"config2 {kind} n={n}", flush=True)
    d, e = generate(kind, n, seed_for(2, n))
    dev = torch.device("cuda:0")
    w = torch.from_numpy(d).to(dev)
    s = brdc.Solver(n, device=dev)
    s.solve(w, torch.from_numpy(e).to(dev))
    torch.cuda.synchronize()
    w = w.cpu().numpy()
    assert int(s.info.item()) == 0
    assert np.all(np.diff(w) >= 0)
    assert rel_err(w, dsterf[f"{kind}_{n}"]) <= TOL
    ref, info_ref = orc.dstebr(d, e)
    assert info_ref == 0 and rel_err(w, ref) <= TOL
    if kind == "toeplitz":
        exact = np.sort(2.0 + 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
        assert rel_err(w, exact) <= TOL
    # trace and Frobenius norm: sum w = sum D, sum w^2 = sum D^2 + 2 sum E^2,
    # to a few units of rounding per term
    scale = max(np.max(np.abs(w)), np.max(np.abs(d)))
    assert abs(w.sum() - d.sum()) <= 64 * np.finfo(float).eps * scale * n
    assert abs((w * w).sum() - (d * d).sum() - 2 * (e * e).sum()) <= \
        64 * np.finfo(float).eps * scale ** 2 * n
