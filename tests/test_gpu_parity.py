"""GPU parity of the HIP BR path (through the C ABI) against the CPU oracle,
the LAPACK DSTERF fixtures, closed forms and spectrum invariants.

Bar: max|w - ref| / max|ref| <= 1e-10 (scripts/cpu/smoke_br.f90:32-38); the
observed differences are ~1e-14.  Same seeded inputs as the oracle tests.
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


def solve(torch, d, e, leaf=0):
    import brdc
    n = d.shape[0]
    dev = torch.device("cuda:0")
    dt = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
    et = torch.from_numpy(np.ascontiguousarray(e if n > 1 else np.zeros(1))).to(dev)
    s = brdc.Solver(n, device=dev, leaf_max=leaf)
    w = dt.clone()
    s.solve(w, et)
    torch.cuda.synchronize()
    return w.cpu().numpy(), int(s.info.item()), s


@pytest.mark.parametrize("kind", GENERATORS)
@pytest.mark.parametrize("n", [26, 64, 65, 300, 1000])
def test_matches_dsterf_fixtures(gpu, kind, n):
    g = np.load(os.path.join(GOLD, "dsterf_eigs.npz"))
    d, e = generate(kind, n, seed_for(1, n))
    w, info, _ = solve(gpu, d, e)
    assert info == 0
    assert np.all(np.diff(w) >= 0)
    assert rel_err(w, g[f"{kind}_{n}"]) <= TOL


def test_config1_n4096_vs_dsterf(gpu):
    g = np.load(os.path.join(GOLD, "dsterf_eigs.npz"))
    d, e = generate("normal", 4096, seed_for(1, 4096))
    w, info, _ = solve(gpu, d, e)
    assert info == 0 and rel_err(w, g["normal_4096"]) <= TOL


@pytest.mark.parametrize("kind", GENERATORS)
@pytest.mark.parametrize("leaf", [8, 32, 64])
def test_matches_oracle_and_deflation_decisions(gpu, orc, kind, leaf):
    n = 3000
    d, e = generate(kind, n, seed_for(2, n))
    ref, info_ref, st_ref = orc.dstebr(d, e, smlsiz=leaf, stats=True)
    w, info, s = solve(gpu, d, e, leaf=leaf)
    assert info == 0 and info_ref == 0
    assert rel_err(w, ref) <= TOL
    st = s.merge_stats()
    # same tree, and the same number of non-deflated poles per merge except at
    # the tolerance edge (both leaf solvers run QL, but their rotations round
    # differently, so a pole at the deflation tolerance may go either way)
    ko = {(int(r["level"]), int(r["start"])): int(r["k"]) for r in st_ref}
    kg = {(int(r["level"]), int(r["start"])): int(r["k"]) for r in st}
    assert set(ko) == set(kg)
    diffs = [abs(ko[k] - kg[k]) for k in ko]
    assert sum(1 for x in diffs if x) <= max(2, len(diffs) // 20)
    assert max(diffs) <= 3


def test_n65536_normal_vs_oracle_and_workspace(gpu, orc):
    """The largest configured size, N = 65536 normal: against the oracle to
    1e-10, the trace and Frobenius invariants, and its device workspace
    (reported by the query) a small fraction of one GPU's HBM."""
    import brdc
    n = 65536
    d, e = generate("normal", n, seed_for(3, n))
    ref, info_ref = orc.dstebr(d, e)
    w, info, _ = solve(gpu, d, e)
    assert info == 0 and info_ref == 0 and rel_err(w, ref) <= TOL
    scale = np.max(np.abs(w))
    assert abs(w.sum() - d.sum()) <= 1e-12 * scale * n
    assert abs((w * w).sum() - (d * d).sum() - 2 * (e * e).sum()) <= 1e-12 * scale ** 2 * n
    assert brdc.workspace_bytes(n) < 1 << 30


def test_n262144_invariants_and_closed_form(gpu):
    # four times the largest order the reference measures: the Toeplitz matrix
    # (root merge of 131072 poles) against its closed form, a normal matrix
    # through sortedness and the trace and Frobenius invariants
    n = 262144
    d, e = generate("toeplitz", n, 7)
    w, info, _ = solve(gpu, d, e)
    k = np.arange(1, n + 1)
    cf = np.sort(2.0 + 2.0 * np.cos(k * np.pi / (n + 1)))
    assert info == 0 and rel_err(w, cf) <= 1e-14
    d, e = generate("normal", n, 7)
    w, info, _ = solve(gpu, d, e)
    assert info == 0 and np.all(np.diff(w) >= 0.0)
    scale = np.max(np.abs(w))
    assert abs(w.sum() - d.sum()) <= 1e-13 * n * scale
    assert abs((w ** 2).sum() - (d ** 2).sum() - 2.0 * (e ** 2).sum()) <= 1e-13 * (w ** 2).sum()


def test_closed_form_families(gpu):
    import scipy.special
    n = 2049
    i = np.arange(1, n)
    w, info, _ = solve(gpu, np.zeros(n), np.sqrt(i * (n - i)))
    assert info == 0 and rel_err(w, np.arange(-(n - 1), n, 2, dtype=float)) <= TOL
    w, info, _ = solve(gpu, np.zeros(n), i / np.sqrt(4.0 * i * i - 1.0))
    assert info == 0 and rel_err(w, scipy.special.roots_legendre(n)[0]) <= TOL


def test_edge_cases(gpu, orc):
    # n = 1 and n = 2, tiny leaves, exact zeros (splits), zero matrix, scaling
    for n in (1, 2, 3):
        d, e = generate("uniform", n, 11)
        w, info, _ = solve(gpu, d, e)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert info == 0 and rel_err(w, np.linalg.eigvalsh(T)) <= 1e-14
    d, e = generate("normal", 700, 5)
    e[[10, 11, 350, 351, 352, 698]] = 0.0
    w, info, _ = solve(gpu, d, e)
    ref, _ = orc.dstebr(d, e)
    assert info == 0 and rel_err(w, ref) <= TOL
    w, info, _ = solve(gpu, np.zeros(500), np.zeros(499))
    assert info == 0 and np.all(w == 0)
    d, e = generate("normal", 900, 6)
    for scale in (1e-150, 1e150):
        w, info, _ = solve(gpu, d * scale, e * scale)
        ref, _ = orc.dstebr(d * scale, e * scale)
        assert info == 0 and rel_err(w, ref) <= TOL


def test_nonfinite_input_reports_info(gpu):
    d, e = generate("normal", 300, 7)
    d[17] = np.nan
    _, info, _ = solve(gpu, d, e)
    assert info == 301


def test_host_entry_matches_device(gpu):
    import brdc
    d, e = generate("normal", 2500, seed_for(3, 2500))
    w_host, info = brdc.dstebr(d, e)
    w_dev, info2, _ = solve(gpu, d, e)
    assert info == 0 and info2 == 0
    np.testing.assert_array_equal(w_host, w_dev)


def test_graph_replay_matches_eager(gpu):
    """The device entry records into a HIP graph; every replay (with eager
    work and synchronisations in between) reproduces the eager result."""
    import brdc
    torch = gpu
    dev = torch.device("cuda:0")
    n = 5000
    d, e = generate("normal", n, seed_for(4, n))
    dt = torch.from_numpy(d).to(dev)
    et = torch.from_numpy(e).to(dev)
    w_eager, info, _ = solve(torch, d, e)
    gs = brdc.GraphSolve(brdc.Solver(n, device=dev), dt, et)
    for rep in range(4):
        gs.w.fill_(-7.0)
        torch.cuda.synchronize()
        out = gs.run()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), w_eager)
        assert int(gs.solver.info.item()) == 0
        x = torch.randn(1000, device=dev).sum()  # unrelated eager work
        torch.cuda.synchronize()
        assert torch.isfinite(x)


def test_streamed_merges_are_reproducible(gpu, orc):
    """Merges of 2048 poles or more go through the streamed secular kernel,
    whose workgroups take roots from a shared counter in whatever order they
    arrive.  A root's arithmetic does not depend on that order or on which
    roots share its wave, so repeated eager solves and graph replays agree
    bit for bit.  Cases: Toeplitz n = 16384 (closed form) and an evenly
    spread diagonal with a constant coupling, n = 6000 (oracle); both keep
    K >= 2048 at their top merges."""
    import brdc
    torch = gpu
    dev = torch.device("cuda:0")
    n1, n2 = 16384, 6000
    cases = [(np.full(n1, 2.0), np.full(n1 - 1, 1.0),
              np.sort(2.0 + 2.0 * np.cos(np.arange(1, n1 + 1) * np.pi / (n1 + 1)))),
             (np.linspace(-1.0, 1.0, n2), np.full(n2 - 1, 0.3), None)]
    for d, e, ref in cases:
        n = d.shape[0]
        if ref is None:
            ref, info_ref = orc.dstebr(d, e)
            assert info_ref == 0
        w0, info, s = solve(torch, d, e)
        assert info == 0 and rel_err(w0, ref) <= TOL
        assert int(s.merge_stats()["k"].max()) >= 2048  # the streamed kernel ran
        for _ in range(2):
            w1, _, _ = solve(torch, d, e)
            np.testing.assert_array_equal(w1, w0)
        dt = torch.from_numpy(d).to(dev)
        et = torch.from_numpy(e).to(dev)
        gs = brdc.GraphSolve(brdc.Solver(n, device=dev), dt, et)
        for _ in range(3):
            gs.w.fill_(-5.0)
            torch.cuda.synchronize()
            out = gs.run()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(out.cpu().numpy(), w0)


@pytest.mark.parametrize("n", [4098, 4226, 6146, 8194])
def test_streamed_merge_edge_sizes(gpu, n):
    """Toeplitz D=2, E=1 keeps K = n/2 at its top merge, so these orders put
    K just above the streaming threshold (2049), one root past a 64-root tile
    (2113 = 33*64 + 1: a tile holding the last root alone, whose excluded
    poles K-2, K-1 start below the tile), one pole past a 128-pole chunk
    (3073) and 4097.  Closed form and repeat-exactness."""
    d, e = np.full(n, 2.0), np.full(n - 1, 1.0)
    ref = np.sort(2.0 + 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    w0, info, s = solve(gpu, d, e)
    assert info == 0 and rel_err(w0, ref) <= TOL
    st = s.merge_stats()
    assert int(st["k"].max()) == n // 2 and n // 2 >= 2048
    w1, _, _ = solve(gpu, d, e)
    np.testing.assert_array_equal(w1, w0)


def test_matches_rocsolver_on_the_same_gpu(gpu):
    """rocSOLVER's STERF and STEDC (values only) as an independent GPU
    reference at a size they finish quickly."""
    from brdc.comparators import rocsolver_eigvals
    torch = gpu
    dev = torch.device("cuda:0")
    for kind in ("normal", "toeplitz"):
        d, e = generate(kind, 400, seed_for(1, 400))
        w, info, _ = solve(torch, d, e)
        dt, et = torch.from_numpy(d).to(dev), torch.from_numpy(e).to(dev)
        for m in ("sterf", "stedc"):
            ref, rinfo = rocsolver_eigvals(dt, et, m)
            torch.cuda.synchronize()
            assert int(rinfo.item()) == 0 and info == 0
            assert rel_err(w, ref.cpu().numpy()) <= TOL


def test_shipped_reduced_dense_input(gpu):
    """The reference's shipped input (data/inputs/reduced_dense_49152.de.bin:
    a reduced dense matrix; root merge K ~ 40000) against its DSTERF
    fixture."""
    from brdc.generators import read_de_bin
    d, e = read_de_bin(os.path.join(GOLD, "reduced_dense_49152.de.bin"))
    ref = np.load(os.path.join(GOLD, "reduced_dense_49152_dsterf.npz"))["w"]
    w, info, _ = solve(gpu, d, e)
    assert info == 0 and rel_err(w, ref) <= TOL
    scale = np.max(np.abs(w))
    n = d.shape[0]
    assert abs(w.sum() - d.sum()) <= 1e-12 * scale * n


def test_harness_gpu_columns(gpu):
    import io
    import brdc
    from brdc import harness
    out = io.StringIO()
    rows = harness.main(["--sizes=300,5000", "--data=normal,toeplitz,clustered,splitmix-normal",
                         "--iters=3", "--warmup=1", "--run-rocsolver=1",
                         "--tridiag-diagnostics=1"], out=out)
    assert len(rows) == 8
    for r in rows:
        assert r["self_info"] == 0 and r["self_err"] <= TOL and r["self_maxerr"] <= TOL
        assert r["rslv_err"] <= TOL and r["self_ms"] > 0 and r["rslv_ms"] > 0
        st = r["merges"]
        # every root converged within LAPACK's iteration bound (30 steps
        # after the midpoint and initial evaluations), the deflation counts
        # add up per merge, and every root of a merge of three or more poles
        # took at least one evaluation (merges of one or two poles take the
        # closed form, counted once)
        assert int(st["failed"].sum()) == 0 and int(st["evals_max"].max()) <= 32
        assert np.all(st["k"] + st["givens"] + st["tiny"] == st["size"])
        big = st["k"] > 2
        assert np.all(st["evals"][big] >= st["k"][big])
        # levels of larger nodes run as level launches; the subtree launch
        # (level -1) holds the rest
        launched = {int(x) for x in st["level"][st["size"] > brdc._lib.SUBTREE_ROWS]}
        assert set(r["level_ms"]) == {-1} | launched
        assert "subtree" in r["level_ms"][-1]
    assert "iter_max" in out.getvalue()


@pytest.mark.parametrize("knobs", [{"stream_min_k": 1024}, {"stream_min_k": 4096},
                                   {"secular_lanes_log2": 16}, {"secular_lanes_log2": 20},
                                   {"wide_max_waves": 1}, {"wide_max_waves": 2},
                                   {"leaf_max": 32, "stream_min_k": 512, "secular_lanes_log2": 17,
                                    "wide_max_waves": 4}])
def test_launch_shape_knobs_keep_parity(gpu, orc, knobs):
    """Every launch-shape knob the autotuner sets keeps the results within the
    bar: an evenly spread diagonal with a constant coupling (n = 6000, top
    merges K >= 2048), Toeplitz n = 8194 (closed form) and a batch."""
    import brdc
    torch = gpu
    dev = torch.device("cuda:0")
    cases = [(np.linspace(-1.0, 1.0, 6000), np.full(5999, 0.3), None),
             (np.full(8194, 2.0), np.full(8193, 1.0),
              np.sort(2.0 + 2.0 * np.cos(np.arange(1, 8195) * np.pi / 8195)))]
    for d, e, ref in cases:
        n = d.shape[0]
        if ref is None:
            ref, info_ref = orc.dstebr(d, e)
            assert info_ref == 0
        s = brdc.Solver(n, device=dev, **knobs)
        w = torch.from_numpy(d).to(dev)
        s.solve(w, torch.from_numpy(e).to(dev))
        torch.cuda.synchronize()
        assert int(s.info.item()) == 0 and rel_err(w.cpu().numpy(), ref) <= TOL
    from brdc.generators import pack_batch
    mats = [generate("normal", m, seed_for(7, m, b)) for b, m in enumerate((300, 2500, 4100))]
    d_cat, e_cat, off = pack_batch(mats)
    bs = brdc.BatchSolver([300, 2500, 4100], device=dev, **knobs)
    w = torch.from_numpy(d_cat).to(dev)
    bs.solve(w, torch.from_numpy(e_cat).to(dev))
    torch.cuda.synchronize()
    ws = bs.split(w.cpu().numpy())
    for b, (d, e) in enumerate(mats):
        ref, _ = orc.dstebr(d, e)
        assert rel_err(ws[b], ref) <= TOL
