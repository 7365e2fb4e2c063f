"""GPU checks of the hand-written backward (dL/dD_j = sum_i g_i v_ij^2,
dL/dE_j = 2 sum_i g_i v_ij v_i,j+1, eigenvectors recomputed by twisted
factorization) through the C ABI and torch.autograd: against a dense FP64
eigh gradient, central finite differences (h = 1e-6), the Euler identity of
the degree-1 homogeneous eigenvalues, and between the batched and
single-matrix entries.  The reference is values-only; this path is new."""
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


def dense_grad(d, e, g):
    """dL/dD, dL/dE from a dense FP64 eigendecomposition (numpy LAPACK)."""
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    lam, V = np.linalg.eigh(T)
    return (V * V) @ g, 2.0 * (V[:-1] * V[1:]) @ g, lam


def gpu_grad(torch, d, e, g):
    from brdc.autograd import eigvalsh_tridiagonal
    dev = torch.device("cuda:0")
    dt = torch.tensor(d, device=dev, requires_grad=True)
    et = torch.tensor(e, device=dev, requires_grad=True)
    w = eigvalsh_tridiagonal(dt, et)
    (w * torch.tensor(g, device=dev)).sum().backward()
    return dt.grad.cpu().numpy(), et.grad.cpu().numpy(), w.detach().cpu().numpy()


@pytest.mark.parametrize("kind,n", [("normal", 2), ("normal", 3), ("normal", 64), ("normal", 65),
                                    ("normal", 500), ("uniform", 1024), ("toeplitz", 300),
                                    ("normal", 2000)])
def test_grad_matches_dense_eigh(gpu, kind, n):
    d, e = generate(kind, n, seed_for(5, n))
    g = np.random.default_rng(n).standard_normal(n)
    gd, ge, w = gpu_grad(gpu, d, e, g)
    rd, re, lam = dense_grad(d, e, g)
    scale = np.max(np.abs(g))
    assert np.max(np.abs(w - lam)) <= 1e-12 * np.max(np.abs(lam))
    assert np.max(np.abs(gd - rd)) <= 1e-10 * scale, np.max(np.abs(gd - rd))
    assert np.max(np.abs(ge - re)) <= 1e-10 * scale, np.max(np.abs(ge - re))


def test_grad_matches_central_differences(gpu):
    """L = g . lambda(D, E); central differences with h = 1e-6 on every entry
    of a 40 x 40 matrix, eigenvalues from the GPU solve itself."""
    import brdc
    torch = gpu
    n = 40
    d, e = generate("normal", n, seed_for(5, n, 1))
    g = np.random.default_rng(1).standard_normal(n)
    gd, ge, _ = gpu_grad(torch, d, e, g)
    h = 1e-6

    def loss(dd, ee):
        w, info = brdc.dstebr(dd, ee)
        assert info == 0
        return float(g @ w)
    fd_d = np.array([(loss(d + h * (np.arange(n) == j), e) - loss(d - h * (np.arange(n) == j), e))
                     / (2 * h) for j in range(n)])
    fd_e = np.array([(loss(d, e + h * (np.arange(n - 1) == j)) -
                      loss(d, e - h * (np.arange(n - 1) == j))) / (2 * h) for j in range(n - 1)])
    assert np.max(np.abs(fd_d - gd)) <= 1e-7 * np.max(np.abs(g))
    assert np.max(np.abs(fd_e - ge)) <= 1e-7 * np.max(np.abs(g))


def test_grad_identities_at_n16384(gpu):
    """n = 16384 (the differentiated single matrix): sum_j dD_j = sum_i g_i
    (unit eigenvectors) and sum_j D_j dD_j + E_j dE_j = sum_i g_i lambda_i
    (eigenvalues are homogeneous of degree 1 in (D, E))."""
    n = 16384
    d, e = generate("normal", n, seed_for(5, n))
    g = np.random.default_rng(2).standard_normal(n)
    gd, ge, w = gpu_grad(gpu, d, e, g)
    s = np.sum(np.abs(g))
    assert abs(gd.sum() - g.sum()) <= 1e-11 * s
    assert abs(d @ gd + e @ ge - g @ w) <= 1e-11 * s * np.max(np.abs(w))
    # a spread of single eigenvalues: g = unit vector -> dD = v_i^2 sums to 1,
    # is nonnegative, and dE = 2 v_i v_i+1
    from brdc.autograd import grad_single
    torch = gpu
    dev = torch.device("cuda:0")
    dt, et = torch.tensor(d, device=dev), torch.tensor(e, device=dev)
    wt = torch.tensor(w, device=dev)
    for i in (0, 4097, n // 2, n - 1):
        gi = torch.zeros(n, dtype=torch.float64, device=dev)
        gi[i] = 1.0
        dd, de = grad_single(dt, et, wt, gi)
        dd = dd.cpu().numpy()
        de = de.cpu().numpy()
        assert abs(dd.sum() - 1.0) <= 1e-12 and dd.min() >= 0.0
        v = np.sqrt(dd)
        assert np.max(np.abs(np.abs(de) - 2 * v[:-1] * v[1:])) <= 1e-12


def test_batched_grad_equals_single_grads(gpu):
    import brdc
    from brdc.autograd import eigvalsh_tridiagonal_batched, grad_single
    torch = gpu
    dev = torch.device("cuda:0")
    sizes = [1, 2, 63, 64, 65, 1024, 1024, 3000]
    mats = [generate("normal", n, seed_for(10, n, b)) for b, n in enumerate(sizes)]
    d_cat, e_cat, off = pack_batch(mats)
    bs = brdc.BatchSolver(sizes, device=dev)
    dt = torch.tensor(d_cat, device=dev, requires_grad=True)
    et = torch.tensor(e_cat, device=dev, requires_grad=True)
    g = torch.tensor(np.random.default_rng(3).standard_normal(d_cat.size), device=dev)
    w = eigvalsh_tridiagonal_batched(dt, et, bs)
    (w * g).sum().backward()
    gd, ge = dt.grad.cpu().numpy(), et.grad.cpu().numpy()
    wn = w.detach()
    for b, (d, e) in enumerate(mats):
        o, n = int(off[b]), sizes[b]
        sd, se = grad_single(torch.tensor(d, device=dev), torch.tensor(e, device=dev),
                             wn[o:o + n].contiguous(), g[o:o + n].contiguous())
        # same tasks and the same summation order per matrix: identical bits
        np.testing.assert_array_equal(gd[o:o + n], sd.cpu().numpy())
        np.testing.assert_array_equal(ge[o:o + n - 1], se.cpu().numpy())
        assert ge[o + n - 1] == 0.0
        rd, re, _ = dense_grad(d, e, g[o:o + n].cpu().numpy())
        assert np.max(np.abs(gd[o:o + n] - rd)) <= 1e-10 * np.max(np.abs(g.cpu().numpy()))


def test_differentiated_workload_shapes(gpu):
    """The differentiated configuration: 256 matrices at n = 1024 plus one at
    n = 16384 in one batch, upstream g ~ N(0,1): every matrix obeys the two
    identities; two of the small ones against dense eigh."""
    import brdc
    from brdc.autograd import eigvalsh_tridiagonal_batched
    torch = gpu
    dev = torch.device("cuda:0")
    sizes = [1024] * 256 + [16384]
    mats = [generate("normal", n, seed_for(5, n, b)) for b, n in enumerate(sizes)]
    d_cat, e_cat, off = pack_batch(mats)
    bs = brdc.BatchSolver(sizes, device=dev)
    dt = torch.tensor(d_cat, device=dev, requires_grad=True)
    et = torch.tensor(e_cat, device=dev, requires_grad=True)
    g = np.random.default_rng(4).standard_normal(d_cat.size)
    w = eigvalsh_tridiagonal_batched(dt, et, bs)
    (w * torch.tensor(g, device=dev)).sum().backward()
    gd, ge, wv = dt.grad.cpu().numpy(), et.grad.cpu().numpy(), w.detach().cpu().numpy()
    for b, (d, e) in enumerate(mats):
        o, n = int(off[b]), sizes[b]
        gb = g[o:o + n]
        s = np.sum(np.abs(gb))
        assert abs(gd[o:o + n].sum() - gb.sum()) <= 1e-11 * s
        assert abs(d @ gd[o:o + n] + e @ ge[o:o + n - 1] - gb @ wv[o:o + n]) <= \
            1e-11 * s * np.max(np.abs(wv[o:o + n]))
    for b in (0, 131):
        d, e = mats[b]
        o, n = int(off[b]), sizes[b]
        rd, re, _ = dense_grad(d, e, g[o:o + n])
        assert np.max(np.abs(gd[o:o + n] - rd)) <= 1e-10 * np.max(np.abs(g))
        assert np.max(np.abs(ge[o:o + n - 1] - re)) <= 1e-10 * np.max(np.abs(g))


def unit_grads(torch, d, e, idx):
    """dD, dE of lambda_i alone (g = unit vector e_i) for each i in idx."""
    from brdc.autograd import grad_single
    dev = torch.device("cuda:0")
    dt, et = torch.tensor(d, device=dev), torch.tensor(e, device=dev)
    import brdc
    w, info = brdc.dstebr(d, e)
    assert info == 0
    wt = torch.tensor(w, device=dev)
    out = []
    for i in idx:
        gi = torch.zeros(d.size, dtype=torch.float64, device=dev)
        gi[i] = 1.0
        dd, de = grad_single(dt, et, wt, gi)
        out.append((dd.cpu().numpy(), de.cpu().numpy()))
    return w, out


@pytest.mark.parametrize("n", [9000, 16384])
def test_long_matrix_unit_grads_toeplitz_closed_form(gpu, n):
    """Matrices long enough for the multi-launch backward (twist search over
    row ranges, z crossed by segment transfers).  Toeplitz (2, 1): the k-th
    eigenvector is v(j) = sqrt(2/(n+1)) sin(j k pi/(n+1)), spread over every
    row, so g = e_i must give dD = v^2 and dE = 2 v(j) v(j+1) everywhere."""
    d, e = np.full(n, 2.0), np.full(n - 1, 1.0)
    idx = [100, n // 3, n // 2, n - 101]
    _, res = unit_grads(gpu, d, e, idx)
    j = np.arange(1, n + 1)
    for i, (dd, de) in zip(idx, res):
        k = n - i  # ascending order: the i-th smallest has k = n - i
        v = np.sqrt(2.0 / (n + 1)) * np.sin(j * k * np.pi / (n + 1))
        assert np.all(np.isfinite(dd)) and np.all(np.isfinite(de))
        assert np.max(np.abs(dd - v * v)) <= 1e-9 * np.max(v * v), i
        assert np.max(np.abs(de - 2 * v[:-1] * v[1:])) <= 1e-9 * np.max(v * v), i


def test_long_matrix_unit_grads_match_mrrr_vectors(gpu):
    """A normal random n = 12000 matrix (long path): v_i^2 from the backward
    against the eigenvector of scipy's eigh_tridiagonal (LAPACK MRRR) for the
    same index, to a bound of 50 eps |T| / gap_i."""
    from scipy.linalg import eigh_tridiagonal
    n = 12000
    d, e = generate("normal", n, seed_for(5, n, 7))
    idx = [0, 17, n // 4, n // 2 + 3, n - 1]
    w, res = unit_grads(gpu, d, e, idx)
    tn = np.max(np.abs(d)) + 2 * np.max(np.abs(e))
    for i, (dd, de) in zip(idx, res):
        lam, V = eigh_tridiagonal(d, e, select="i", select_range=(i, i))
        v = V[:, 0]
        gap = min(abs(w[i] - w[i - 1]) if i > 0 else np.inf,
                  abs(w[i + 1] - w[i]) if i + 1 < n else np.inf)
        tol = max(1e-11, 50 * np.finfo(float).eps * tn / gap)
        assert abs(lam[0] - w[i]) <= 1e-12 * tn
        assert np.max(np.abs(dd - v * v)) <= tol, (i, gap)
        assert np.max(np.abs(de - 2 * v[:-1] * v[1:])) <= 2 * tol, (i, gap)


def test_batched_long_matrix_grad_equals_single(gpu):
    """A batch holding two long matrices (n = 9000 and 12000, multi-launch
    backward) beside short ones: each long matrix's batched gradient equals its
    single-matrix gradient bit for bit, and every matrix obeys sum_j dD_j =
    sum_i g_i."""
    import brdc
    from brdc.autograd import eigvalsh_tridiagonal_batched, grad_single
    torch = gpu
    dev = torch.device("cuda:0")
    sizes = [300, 9000, 1024, 65, 12000]
    mats = [generate("uniform", n, seed_for(11, n, b)) for b, n in enumerate(sizes)]
    d_cat, e_cat, off = pack_batch(mats)
    bs = brdc.BatchSolver(sizes, device=dev)
    dt = torch.tensor(d_cat, device=dev, requires_grad=True)
    et = torch.tensor(e_cat, device=dev, requires_grad=True)
    g = torch.tensor(np.random.default_rng(9).standard_normal(d_cat.size), device=dev)
    w = eigvalsh_tridiagonal_batched(dt, et, bs)
    (w * g).sum().backward()
    gd, ge, wn = dt.grad.cpu().numpy(), et.grad.cpu().numpy(), w.detach()
    for b, (d, e) in enumerate(mats):
        o, n = int(off[b]), sizes[b]
        gb = g[o:o + n].cpu().numpy()
        assert abs(gd[o:o + n].sum() - gb.sum()) <= 1e-11 * np.sum(np.abs(gb))
    for b in (1, 4):  # the two long matrices, of different sizes
        o, n = int(off[b]), sizes[b]
        d, e = mats[b]
        sd, se = grad_single(torch.tensor(d, device=dev), torch.tensor(e, device=dev),
                             wn[o:o + n].contiguous(), g[o:o + n].contiguous())
        np.testing.assert_array_equal(gd[o:o + n], sd.cpu().numpy())
        np.testing.assert_array_equal(ge[o:o + n - 1], se.cpu().numpy())


@pytest.mark.parametrize("kind,n", [("normal", 9000), ("toeplitz", 9000), ("uniform", 8192)])
def test_long_matrix_transfer_scan_equals_row_walk(gpu, kind, n, monkeypatch):
    """The long-matrix z scan crosses segments with stored transfers and walks
    rows only where a transfer is unusable.  BRDC_GRAD_SCAN_LIMIT = 0 marks
    every transfer unusable, so the same gradient comes from walking every
    row: the two forms agree to rounding (n = 9000, and n = 8192, the first
    size on the long path; random g)."""
    d, e = generate(kind, n, seed_for(5, n, 3))
    g = np.random.default_rng(12).standard_normal(n)
    gd, ge, _ = gpu_grad(gpu, d, e, g)
    monkeypatch.setenv("BRDC_GRAD_SCAN_LIMIT", "0")
    wd, we, _ = gpu_grad(gpu, d, e, g)
    monkeypatch.delenv("BRDC_GRAD_SCAN_LIMIT")
    scale = np.max(np.abs(g))
    assert np.all(np.isfinite(wd)) and np.all(np.isfinite(we))
    assert np.max(np.abs(gd - wd)) <= 1e-12 * scale
    assert np.max(np.abs(ge - we)) <= 1e-12 * scale
    # the two forms round differently: identical bits would mean the limit
    # never reached the kernel
    assert not np.array_equal(gd, wd)
