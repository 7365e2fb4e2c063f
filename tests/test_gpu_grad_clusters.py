"""GPU checks of the backward on numerically degenerate eigenvalues.

Twisted factorization alone returns one vector for every copy of a repeated
eigenvalue, so the backward recomputes the vectors of such clusters by inverse
iteration with Gram-Schmidt.  Only cluster-invariant losses have a defined
gradient there, so the checks use them:
  L = sum(w):    dL/dD = 1, dL/dE = 0           (V V^T = I)
  L = sum(w^2):  dL/dD = 2 D, dL/dE = 4 E        (V diag(w) V^T = T)
on matrices split into identical blocks (exact repeats) or glued by a
coupling of 1e-300; and the clustered generator at n = 1024 (its eight
clusters of close but localised eigenvalues stay with twisted factorization)
against a dense FP64 eigh gradient with g constant over each cluster.

The cluster path takes eigenvalues within 2^-44 ||T|| of a neighbour.  Pairs
whose vectors overlap and whose gap lies between that and about 1e-8 ||T||
(Wilkinson's W21+ below its top pair) keep twisted-factorization vectors,
accurate to about eps ||T|| / gap, as the module documentation states."""
import numpy as np
import pytest

from brdc.generators import generate, pack_batch, seed_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brdc
    brdc._lib.load()
    return torch


def grads(torch, d, e, loss):
    from brdc.autograd import eigvalsh_tridiagonal
    dev = torch.device("cuda:0")
    dt = torch.tensor(d, device=dev, requires_grad=True)
    et = torch.tensor(e, device=dev, requires_grad=True)
    w = eigvalsh_tridiagonal(dt, et)
    loss(w).backward()
    torch.cuda.synchronize()
    return dt.grad.cpu().numpy(), et.grad.cpu().numpy(), w.detach().cpu().numpy()


def split_blocks(kind, m, copies, glue=0.0):
    d0, e0 = generate(kind, m, seed_for(6, m))
    d = np.tile(d0, copies)
    e = np.concatenate([np.concatenate([e0, [glue]]) for _ in range(copies)])[:-1]
    return d, e


CASES = {
    "two_identical_blocks": lambda: split_blocks("normal", 500, 2),
    "three_identical_blocks": lambda: split_blocks("uniform", 300, 3),
    "identical_blocks_glued_1e-300": lambda: split_blocks("normal", 200, 2, glue=1e-300),
    "four_identical_toeplitz_blocks": lambda: split_blocks("toeplitz", 250, 4),
    "zero_matrix": lambda: (np.zeros(130), np.zeros(129)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_sum_of_eigenvalues_gives_identity(gpu, case):
    d, e = CASES[case]()
    gd, ge, w = grads(gpu, d, e, lambda w: w.sum())
    assert np.max(np.abs(gd - 1.0)) <= 1e-10, np.max(np.abs(gd - 1.0))
    assert np.max(np.abs(ge)) <= 1e-10, np.max(np.abs(ge))


@pytest.mark.parametrize("case", sorted(CASES))
def test_sum_of_squares_gives_the_matrix(gpu, case):
    d, e = CASES[case]()
    gd, ge, w = grads(gpu, d, e, lambda w: (w * w).sum())
    scale = np.max(np.abs(w))
    assert np.max(np.abs(gd - 2.0 * d)) <= 1e-10 * scale, np.max(np.abs(gd - 2.0 * d))
    assert np.max(np.abs(ge - 4.0 * e)) <= 1e-10 * scale, np.max(np.abs(ge - 4.0 * e))


def test_repeated_eigenvalues_are_really_repeated(gpu):
    """The identical-block case exercises the cluster path: its spectrum
    holds every eigenvalue twice (to rounding)."""
    d, e = split_blocks("normal", 500, 2)
    _, _, w = grads(gpu, d, e, lambda w: w.sum())
    gaps = np.diff(w)[0::2]
    assert np.max(np.abs(gaps)) <= 64 * np.finfo(float).eps * np.max(np.abs(w))


def test_clustered_generator_against_dense_eigh(gpu):
    """D_i = c_(i mod 8) + 1e-8 N(0,1), E_i = 1e-6 U(0,1), n = 1024: g constant
    over each centre's cluster of 128 eigenvalues."""
    torch = gpu
    n = 1024
    d, e = generate("clustered", n, seed_for(5, n))
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    lam, V = np.linalg.eigh(T)
    centres = d[:8]  # D_k = c_k + 1e-8 N(0,1) for k < 8: each centre to 1e-7
    assert np.min(np.diff(np.sort(centres))) > 1e-4
    which = np.argmin(np.abs(lam[:, None] - centres[None, :]), axis=1)
    assert np.all(np.bincount(which, minlength=8) == n // 8)
    gval = np.random.default_rng(3).standard_normal(8)
    g = gval[which]
    rd, re = (V * V) @ g, 2.0 * (V[:-1] * V[1:]) @ g
    gt = torch.tensor(g, device="cuda:0")
    gd, ge, w = grads(torch, d, e, lambda w: (w * gt).sum())
    assert np.max(np.abs(w - lam)) <= 1e-12
    assert np.max(np.abs(gd - rd)) <= 1e-10 * np.max(np.abs(g)), np.max(np.abs(gd - rd))
    assert np.max(np.abs(ge - re)) <= 1e-10 * np.max(np.abs(g)), np.max(np.abs(ge - re))


def test_batched_grad_with_clusters_equals_single(gpu):
    """A batch holding a split matrix between two ordinary ones: every
    matrix's gradient equals its single-matrix gradient bit for bit."""
    import brdc
    from brdc.autograd import grad_batched
    torch = gpu
    dev = torch.device("cuda:0")
    mats = [generate("normal", 300, seed_for(6, 300, 1)), split_blocks("normal", 400, 2),
            generate("uniform", 129, seed_for(6, 129, 2))]
    sizes = [m[0].shape[0] for m in mats]
    d_cat, e_cat, off = pack_batch(mats)
    bs = brdc.BatchSolver(sizes, device=dev)
    d0, e0 = torch.from_numpy(d_cat).to(dev), torch.from_numpy(e_cat).to(dev)
    w = d0.clone()
    bs.solve(w, e0)
    g = torch.from_numpy(np.random.default_rng(9).standard_normal(d_cat.size)).to(dev)
    gd, ge = grad_batched(bs, d0, e0, w, g)
    gd, ge = gd.cpu().numpy(), ge.cpu().numpy()
    for b, (d, e) in enumerate(mats):
        o, n = int(off[b]), sizes[b]
        gb = g[o:o + n].cpu().numpy()
        sd, se, _ = grads(torch, d, e, lambda ww: (ww * torch.tensor(gb, device=dev)).sum())
        np.testing.assert_array_equal(gd[o:o + n], sd)
        np.testing.assert_array_equal(ge[o:o + n - 1], se)
