"""A seeded sweep of matrix families, orders and leaf bounds through the C ABI
on the GPU, each solve against the CPU oracle at the acceptance bar
max|w - ref| / max|ref| <= 1e-10 (scripts/cpu/smoke_br.f90:30-38), with
sortedness and the trace.

Beyond the four measured generators the families stress the deflation and
root-finding paths the reference's DLAED8_BR and DLAED4 take on structured
input: integer entries (exact ties between children's eigenvalues, so
close-pole Givens deflation on every merge, and exact zeros in E), Wilkinson
blocks glued by tiny off-diagonals (pairs of eigenvalues agreeing to many
digits), graded diagonals over eight decades, and matrices split by exact
zeros at random rows.
"""
import numpy as np
import pytest

from brdc.generators import generate

pytestmark = pytest.mark.gpu
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


def family(kind, n, rng):
    if kind in ("uniform", "normal", "toeplitz", "clustered"):
        return generate(kind, n, int(rng.integers(1 << 30)))
    if kind == "integer":
        d = rng.integers(-2, 3, n).astype(np.float64)
        e = rng.integers(0, 2, n - 1).astype(np.float64)
        return d, e
    if kind == "wilkinson":
        m = 21
        blk = np.abs(np.arange(m) - (m - 1) / 2.0)
        reps = (n + m - 1) // m
        d = np.tile(blk, reps)[:n]
        e = np.ones(n - 1)
        e[m - 1::m] = 1e-10
        return d, e
    if kind == "graded":
        g = 10.0 ** (-8.0 * np.arange(n) / max(n - 1, 1))
        return g * rng.uniform(0.5, 1.5, n), 0.5 * g[:-1] * rng.uniform(0.5, 1.5, n - 1)
    if kind == "split":
        d = rng.standard_normal(n)
        e = rng.standard_normal(n - 1)
        e[rng.integers(0, n - 1, max(1, n // 50))] = 0.0
        return d, e
    raise ValueError(kind)


KINDS = ("uniform", "normal", "toeplitz", "clustered", "integer", "wilkinson", "graded", "split")
LEAVES = (8, 16, 24, 32, 64)


def cases():
    rng = np.random.default_rng(20261204)
    out = []
    for i in range(32):
        n = int(np.exp(rng.uniform(np.log(2), np.log(6000))))
        out.append((KINDS[i % len(KINDS)], max(n, 2), LEAVES[i % len(LEAVES)], int(rng.integers(1 << 30))))
    return out


@pytest.mark.parametrize("kind,n,leaf,seed", cases())
def test_sweep_against_oracle(gpu, orc, kind, n, leaf, seed):
    import brdc
    torch = gpu
    rng = np.random.default_rng(seed)
    d, e = family(kind, n, rng)
    ref, info_ref = orc.dstebr(d, e, smlsiz=leaf)
    dev = torch.device("cuda:0")
    dt = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
    et = torch.from_numpy(np.ascontiguousarray(e)).to(dev)
    s = brdc.Solver(n, device=dev, leaf_max=leaf)
    w = dt.clone()
    s.solve(w, et)
    torch.cuda.synchronize()
    w = w.cpu().numpy()
    assert int(s.info.item()) == 0 and info_ref == 0
    assert np.all(np.diff(w) >= 0.0)
    assert rel_err(w, ref) <= TOL, (kind, n, leaf, rel_err(w, ref))
    scale = max(np.max(np.abs(ref)), np.finfo(float).tiny)
    assert abs(w.sum() - d.sum()) <= 1e-12 * n * scale
