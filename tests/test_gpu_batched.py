"""GPU parity of the batched entry (one flattened merge forest over many
matrices of uneven sizes) through the C ABI: every matrix against the CPU
oracle (the reference BR restatement) at the 1e-10 bar of
scripts/cpu/smoke_br.f90:32-38, against the single-matrix GPU entry, and
per-matrix INFO.  The batched form has no counterpart in the reference."""
import numpy as np
import pytest

from brdc.generators import batched_workload, generate, pack_batch, seed_for

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


def solve_batch(torch, mats, leaf=0):
    import brdc
    d_cat, e_cat, off = pack_batch(mats)
    dev = torch.device("cuda:0")
    bs = brdc.BatchSolver([d.shape[0] for d, _ in mats], device=dev, leaf_max=leaf)
    w = torch.from_numpy(d_cat).to(dev)
    e = torch.from_numpy(e_cat).to(dev) if e_cat.size else torch.zeros(1, dtype=torch.float64,
                                                                       device=dev)
    if bs.rows:
        bs.solve(w, e)
    torch.cuda.synchronize()
    return bs.split(w.cpu().numpy()), bs.info.cpu().numpy()[:bs.nmat], bs


KINDS = ("normal", "uniform", "toeplitz", "clustered")


@pytest.mark.parametrize("leaf", [0, 8, 32, 64])
def test_mixed_batch_matches_oracle(gpu, orc, leaf):
    sizes = [1, 2, 3, 7, 16, 17, 31, 33, 64, 65, 100, 257, 1000, 2500]
    mats = [generate(KINDS[b % 4], n, seed_for(6, n, b)) for b, n in enumerate(sizes)]
    ws, info, _ = solve_batch(gpu, mats, leaf)
    assert np.all(info == 0)
    for (d, e), w in zip(mats, ws):
        ref, info_ref = orc.dstebr(d, e)
        assert info_ref == 0
        assert np.all(np.diff(w) >= 0)
        assert rel_err(w, ref) <= TOL, (d.shape[0], rel_err(w, ref))


def test_batch_matches_single_matrix_entry(gpu):
    import brdc
    sizes = [40, 300, 1111, 4096]
    mats = [generate("normal", n, seed_for(7, n, b)) for b, n in enumerate(sizes)]
    ws, info, _ = solve_batch(gpu, mats)
    assert np.all(info == 0)
    dev = gpu.device("cuda:0")
    for (d, e), w in zip(mats, ws):
        single, info1 = brdc.dstebr_device(gpu.from_numpy(d).to(dev), gpu.from_numpy(e).to(dev))
        assert info1 == 0
        # same trees and kernels; only the secular lanes-per-root choice (which
        # follows the level's merge count) may change the summation order
        assert rel_err(w, single.cpu().numpy()) <= 1e-13


def test_batched_workload_2048(gpu, orc):
    """The batched configuration: 2048 matrices, sizes log-uniform in
    [32, 4096], D, E ~ N(0,1).  Closed-form invariants for every matrix, the
    oracle for a spread sample."""
    sizes, mats = batched_workload()
    ws, info, bs = solve_batch(gpu, mats)
    assert bs.nmat == 2048 and np.all(info == 0)
    for (d, e), w in zip(mats, ws):
        scale = max(np.max(np.abs(w)), 1e-300)
        n = d.shape[0]
        assert np.all(np.diff(w) >= 0)
        assert abs(w.sum() - d.sum()) <= 1e-12 * scale * n
        assert abs((w * w).sum() - (d * d).sum() - 2 * (e * e).sum()) <= 1e-12 * scale ** 2 * n
    for b in list(range(0, 2048, 97)) + [int(np.argmax(sizes)), int(np.argmin(sizes))]:
        d, e = mats[b]
        ref, _ = orc.dstebr(d, e)
        assert rel_err(ws[b], ref) <= TOL, b


def test_many_small_matrices_packed_roots(gpu, orc):
    """6000 matrices of 20..60 rows: the bottom heights hold thousands of small
    merges, so their secular roots run packed across merges (one wave spans
    several merges' roots).  Every matrix against the oracle."""
    rng = np.random.default_rng(12)
    sizes = [int(x) for x in rng.integers(20, 61, 6000)]
    mats = [generate(KINDS[b % 4], n, seed_for(12, n, b)) for b, n in enumerate(sizes)]
    ws, info, bs = solve_batch(gpu, mats)
    assert np.all(info == 0)
    for b, ((d, e), w) in enumerate(zip(mats, ws)):
        ref, _ = orc.dstebr(d, e)
        assert rel_err(w, ref) <= TOL, b


def test_per_matrix_info_and_empty_matrices(gpu, orc):
    mats = [generate("normal", n, seed_for(8, n, b)) for b, n in enumerate([50, 0, 300, 0, 20])]
    bad = mats[2][0].copy()
    bad[5] = np.nan
    mats[2] = (bad, mats[2][1])
    ws, info, _ = solve_batch(gpu, mats)
    assert list(info) == [0, 0, 301, 0, 0]
    for b in (0, 4):
        ref, _ = orc.dstebr(*mats[b])
        assert rel_err(ws[b], ref) <= TOL
    assert ws[1].size == 0 and ws[3].size == 0


def test_batch_graph_replay_matches_eager(gpu):
    import brdc
    torch = gpu
    dev = torch.device("cuda:0")
    sizes = [33, 200, 900, 2048, 77]
    mats = [generate("uniform", n, seed_for(9, n, b)) for b, n in enumerate(sizes)]
    eager, info, _ = solve_batch(torch, mats)
    d_cat, e_cat, _ = pack_batch(mats)
    dt = torch.from_numpy(d_cat).to(dev)
    et = torch.from_numpy(e_cat).to(dev)
    gs = brdc.GraphSolve(brdc.BatchSolver(sizes, device=dev), dt, et)
    for _ in range(3):
        gs.w.fill_(-3.0)
        torch.cuda.synchronize()
        out = gs.run()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), np.concatenate(eager))
        assert int(gs.solver.info.abs().sum().item()) == 0
        torch.cuda.synchronize()
