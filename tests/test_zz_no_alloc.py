"""The allocation-free device entry (brdc_dstebr_device_noalloc) against the
workspace entry (brdc_dstebr_device_ex) through the C ABI: the same kernels
run on the same inputs, so the eigenvalues must agree bit for bit.  Two calls
in a row share the GPU's static workspace, the second on a smaller input, so
anything the first call left behind that the second read would show up."""
import ctypes

import numpy as np
import pytest

from brdc.generators import generate, seed_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import brdc
    brdc._lib.load()
    return torch


def solve_workspace(torch, d, e, leaf):
    """The existing device entry with a caller-owned workspace."""
    from brdc import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    n = d.shape[0]
    opts = _lib.options(leaf)
    nb = ctypes.c_size_t(0)
    _lib.check(L.brdc_dstebr_device_workspace_ex(n, ctypes.byref(opts), ctypes.byref(nb)),
               "workspace query")
    ws = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
    w = torch.from_numpy(d).to(dev)
    et = torch.from_numpy(e).to(dev)
    info = torch.full((1,), -99, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(L.brdc_dstebr_device_ex(n, w.data_ptr(), et.data_ptr(), ws.data_ptr(),
                                       int(nb.value), st, info.data_ptr(), ctypes.byref(opts)),
               "device entry")
    torch.cuda.synchronize()
    return w.cpu().numpy(), int(info.item())


def solve_noalloc(torch, d, e, leaf):
    from brdc import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    n = d.shape[0]
    opts = _lib.options(leaf)
    w = torch.from_numpy(d).to(dev)
    et = torch.from_numpy(e).to(dev)
    info = torch.full((1,), -99, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(L.brdc_dstebr_device_noalloc(n, w.data_ptr(), et.data_ptr(), st,
                                            info.data_ptr(), ctypes.byref(opts)),
               "allocation-free entry")
    torch.cuda.synchronize()
    return w.cpu().numpy(), int(info.item())


def test_noalloc_matches_workspace_entry_on_two_calls(gpu, orc):
    from brdc import _lib
    torch = gpu
    # the largest order first (Toeplitz keeps K = n/2 at the top merges, so
    # the streamed and chunked kernels run), then a smaller one of another
    # generator and leaf size
    cases = [("toeplitz", _lib.NOALLOC_MAX_N, 0), ("normal", 3000, 32)]
    for kind, n, leaf in cases:
        print(f"noalloc {kind} n={n} leaf={leaf}", flush=True)
        d, e = generate(kind, n, seed_for(2, n))
        w_ref, info_ref = solve_workspace(torch, d, e, leaf)
        w, info = solve_noalloc(torch, d, e, leaf)
        assert info == 0 and info_ref == 0
        np.testing.assert_array_equal(w, w_ref)
        assert np.all(np.diff(w) >= 0)
        if kind == "toeplitz":
            exact = np.sort(2.0 + 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
            assert np.max(np.abs(w - exact)) / np.max(np.abs(exact)) <= 1e-10
        else:
            ref, oinfo = orc.dstebr(d, e)
            assert oinfo == 0 and np.max(np.abs(w - ref)) / np.max(np.abs(ref)) <= 1e-10


def test_noalloc_rejects_orders_beyond_its_workspace(gpu):
    from brdc import _lib
    torch = gpu
    L = _lib.load()
    n = _lib.NOALLOC_MAX_N + 1
    w = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    assert L.brdc_dstebr_device_noalloc(n, w.data_ptr(), w.data_ptr(), st, None, None) == 3
