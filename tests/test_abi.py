"""C-ABI checks that need no GPU: the library loads, exports every symbol that
include/brdc.h declares, and its host-side logic (workspace queries, argument
errors) behaves like DSTEDC_DC_BR's (cpu/src/dstedc_dc_br.f:44-75)."""
import ctypes
import os
import re

import pytest

from brdc import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "brdc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(brdc_\w+|dstedc_dc_br_)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = _lib.load()
    names = header_functions()
    assert set(names) == set(_lib.EXPORTS)
    for name in names:
        assert hasattr(L, name), name


def test_profile_classes_match_the_library():
    """The Python mirror of the kernel classes and of the streamed-merge
    threshold agrees with the header and the HIP source."""
    hdr = open(os.path.join(ROOT, "include", "brdc.h")).read()
    assert int(re.search(r"#define BRDC_KERNEL_CLASSES (\d+)", hdr).group(1)) == \
        len(_lib.KERNEL_CLASSES)
    src = open(os.path.join(ROOT, "ppopp27-brdc-ae_amd", "csrc", "brdc.hip")).read()
    enum = re.search(r"enum KernelClass \{([^}]*)\}", src).group(1)
    names = [x.strip() for x in enum.split(",") if x.strip() and x.strip() != "kNumClasses"]
    camel = ["k" + "".join(p.capitalize() for p in c.split("_")) for c in _lib.KERNEL_CLASSES]
    assert names == camel
    assert int(re.search(r"constexpr int kStreamMinK = (\d+);", src).group(1)) == \
        _lib.STREAM_MIN_K
    kern = open(os.path.join(ROOT, "ppopp27-brdc-ae_amd", "csrc", "brdc_kernels.h")).read()
    assert int(re.search(r"constexpr int kSmallMaxK = (\d+);", kern).group(1)) == \
        _lib.SMALL_MERGE
    assert int(re.search(r"constexpr int kSmallTopK = (\d+);", kern).group(1)) == \
        _lib.SMALL_MERGE_TOP
    import brdc
    assert int(re.search(r"constexpr int kDefaultLeaf = (\d+);", src).group(1)) == \
        brdc._DEFAULT_LEAF
    assert int(re.search(r"constexpr int kSubtreeLeaf = (\d+);", src).group(1)) == \
        brdc._SUBTREE_LEAF
    sub = open(os.path.join(ROOT, "ppopp27-brdc-ae_amd", "csrc", "brdc_subtree.h")).read()
    assert int(re.search(r"constexpr int kSubRows = (\d+);", sub).group(1)) == _lib.SUBTREE_ROWS
    assert int(re.search(r"constexpr int kSubMaxLevels = (\d+);", sub).group(1)) == 6
    assert int(re.search(r"constexpr int kSubMaxLog2 = (\d+);", src).group(1)) == \
        brdc._SUB_MAX_LOG2


def test_status_strings():
    L = _lib.load()
    assert L.brdc_status_string(0) == b"success"
    assert L.brdc_status_string(4).startswith(b"workspace")


def test_host_query_matches_reference_convention():
    import brdc
    assert brdc.query(4096) == (0, 18 * 4096, 7 * 4096)
    assert brdc.query(25) == (0, 1, 1)
    assert brdc.query(26) == (0, 18 * 26, 7 * 26)
    assert brdc.query(-3)[0] == -1


def test_host_argument_errors():
    L = _lib.load()
    d = (ctypes.c_double * 100)()
    e = (ctypes.c_double * 100)()
    w = (ctypes.c_double * 10)()
    iw = (ctypes.c_int * 10000)()
    assert L.brdc_dstebr(100, d, e, w, 10, iw, 10000) == -5
    big = (ctypes.c_double * 2000)()
    assert L.brdc_dstebr(100, d, e, big, 2000, iw, 10) == -7
    # n <= 1 returns before any device work
    assert L.brdc_dstebr(1, d, e, w, 1, iw, 1) == 0
    assert L.brdc_dstebr(0, d, e, w, 1, iw, 1) == 0


def test_device_workspace_query_is_linear_in_n():
    import brdc
    b1 = brdc.workspace_bytes(16384)
    b2 = brdc.workspace_bytes(65536)
    assert 0 < b1 < b2 < 5 * b1                     # O(n), no n^2 slab
    assert b2 < 400 * 65536                         # < 400 bytes per row
    L = _lib.load()
    out = ctypes.c_size_t(0)
    assert L.brdc_dstebr_device_workspace(-1, ctypes.byref(out)) == 3
    o = _lib.options(65)
    assert L.brdc_dstebr_device_workspace_ex(100, ctypes.byref(o), ctypes.byref(out)) == 3
    o = _lib.options(64)
    assert L.brdc_dstebr_device_workspace_ex(100, ctypes.byref(o), ctypes.byref(out)) == 0


def test_device_entry_rejects_bad_arguments_without_launching():
    L = _lib.load()
    o = _lib.Options()
    # null workspace / pointers and undersized workspace fail before any launch
    assert L.brdc_dstebr_device(100, 0x1000, 0x2000, None, 0, None, None) == 2
    assert L.brdc_dstebr_device(100, None, 0x2000, 0x1000, 1 << 30, None, None) == 2
    assert L.brdc_dstebr_device(100, 0x1000, 0x2000, 0x100000, 16, None, None) == 4
    assert L.brdc_dstebr_device(100, 0x1000, 0x2000, 0x100010, 1 << 30, None, None) == 4
    assert L.brdc_dstebr_device(-2, 0x1000, 0x2000, 0x100000, 1 << 30, None, None) == 3
    o.reserved[0] = 1
    assert L.brdc_dstebr_device_ex(100, 0x1000, 0x2000, 0x100000, 1 << 30, None, None,
                                   ctypes.byref(o)) == 3
    assert L.brdc_dstebr_device(0, None, None, None, 0, None, None) == 0


def _plan(L, sizes, leaf=0):
    import numpy as np
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    p = ctypes.c_void_p()
    o = _lib.options(leaf)
    st = L.brdc_batch_plan_create(len(sizes), sizes.ctypes.data_as(_lib._ip), ctypes.byref(o),
                                  ctypes.byref(p))
    return st, p


def test_batch_plan_is_every_matrix_tree():
    """The flattened forest holds each matrix's own tree (the tree of the
    single-matrix entry): leaves = sum 2^levels_b, merges = sum (2^levels_b - 1),
    heights = max levels_b, over the same floor/ceil halving."""
    import numpy as np
    from brdc.generators import batched_sizes
    L = _lib.load()
    for leaf in (8, 16, 32, 64):
        sizes = np.concatenate([[0, 1, 2, 3, leaf, leaf + 1, 2 * leaf + 1, 1000],
                                batched_sizes(300, 32, 4096, 99)]).astype(np.int32)
        st, p = _plan(L, sizes, leaf)
        assert st == 0
        out = np.zeros(7, dtype=np.int32)
        assert L.brdc_batch_plan_stats(p, out.ctypes.data_as(_lib._ip), 7) == 7
        lev = []
        for n in sizes:
            if n == 0:
                continue
            k = 0
            while -(-n // (1 << k)) > leaf:
                k += 1
            lev.append(k)
        lev = np.array(lev)
        assert out[0] == sizes.sum() and out[1] == len(sizes)
        assert out[2] == np.sum(2 ** lev) and out[4] == np.sum(2 ** lev - 1)
        assert out[3] == lev.max() and out[5] <= leaf and out[6] == leaf
        nb = ctypes.c_size_t(0)
        assert L.brdc_batch_workspace(p, ctypes.byref(nb)) == 0
        assert 0 < nb.value < 400 * int(sizes.sum()) + 64 * out[2] + 4096 * len(sizes)
        L.brdc_batch_plan_destroy(p)


def test_batch_entry_rejects_bad_arguments_without_launching():
    import numpy as np
    L = _lib.load()
    st, p = _plan(L, [10, -1, 5])
    assert st == 3 and not p.value
    st, p = _plan(L, [10, 20], leaf=65)
    assert st == 3 and not p.value
    st, p = _plan(L, [10, 20])
    assert st == 0 and p.value
    nb = ctypes.c_size_t(0)
    L.brdc_batch_workspace(p, ctypes.byref(nb))
    assert L.brdc_batch_upload(p, None, nb.value, None) == 2
    assert L.brdc_batch_upload(p, 0x100000, 16, None) == 4
    assert L.brdc_batch_solve(p, 0x1000, 0x2000, 0x100000, nb.value, None, None) == 2
    assert L.brdc_batch_solve(p, 0x1000, 0x2000, 0x100010, nb.value, None, 0x3000) == 4
    assert L.brdc_batch_solve(None, 0x1000, 0x2000, 0x100000, nb.value, None, 0x3000) == 2
    L.brdc_batch_plan_destroy(p)
    st, p = _plan(L, np.zeros(0, dtype=np.int32))
    assert st == 0
    assert L.brdc_batch_solve(p, None, None, None, 0, None, None) == 0  # empty batch
    L.brdc_batch_plan_destroy(p)


def test_noalloc_workspace_bound_covers_every_leaf():
    """The allocation-free entry's static array holds the workspace of its
    largest order at every leaf bound, and the Python mirror of its limits
    matches the header."""
    hdr = open(os.path.join(ROOT, "include", "brdc.h")).read()
    assert int(re.search(r"#define BRDC_NOALLOC_MAX_N (\d+)", hdr).group(1)) == \
        _lib.NOALLOC_MAX_N
    assert re.search(r"#define BRDC_NOALLOC_BYTES \(20u << 20\)", hdr)
    L = _lib.load()
    for leaf in range(2, 65):
        opts = _lib.options(leaf)
        out = ctypes.c_size_t(0)
        assert L.brdc_dstebr_device_workspace_ex(_lib.NOALLOC_MAX_N, ctypes.byref(opts),
                                                 ctypes.byref(out)) == 0
        assert out.value <= _lib.NOALLOC_BYTES, leaf


def test_noalloc_argument_errors_before_any_device_work():
    L = _lib.load()
    assert L.brdc_dstebr_device_noalloc(-1, None, None, None, None, None) == 3
    assert L.brdc_dstebr_device_noalloc(_lib.NOALLOC_MAX_N + 1, None, None, None, None,
                                        None) == 3
    assert L.brdc_dstebr_device_noalloc(0, None, None, None, None, None) == 0
    assert L.brdc_dstebr_device_noalloc(10, None, None, None, None, None) == 2


AMDFLANG = "/opt/rocm/llvm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(AMDFLANG), reason="amdflang not available")
@pytest.mark.parametrize("route", ["reference_symbol", "renamed_reference"])
def test_fortran_caller_links_and_runs(tmp_path, route):
    """An unchanged Fortran caller of DSTEDC_DC_BR (tests/cpu_harness/
    fortran_caller.f90, LP64, arguments by reference like
    scripts/cpu/smoke_br.f90:24-27) links against libbrdc.so in place of the
    reference's objects -- through the library's dstedc_dc_br_ symbol, or
    with its reference renamed to brdc_dstebr_ by objcopy -- and the
    workspace queries and argument errors behave as dstedc_dc_br.f:44-75
    defines them."""
    import shutil
    import subprocess
    src = os.path.join(ROOT, "tests", "cpu_harness", "fortran_caller.f90")
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = str(tmp_path / "caller")
    if route == "reference_symbol":
        subprocess.run([AMDFLANG, "-O1", src, "-o", exe, f"-L{libdir}", "-lbrdc",
                        f"-Wl,-rpath,{libdir}"], check=True)
    else:
        if shutil.which("objcopy") is None:
            pytest.skip("objcopy not available")
        obj = str(tmp_path / "caller.o")
        subprocess.run([AMDFLANG, "-O1", "-c", src, "-o", obj], check=True)
        subprocess.run(["objcopy", "--redefine-sym", "dstedc_dc_br_=brdc_dstebr_", obj],
                       check=True)
        subprocess.run([AMDFLANG, obj, "-o", exe, f"-L{libdir}", "-lbrdc",
                        f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    assert got["query4096"] == ["0", str(18 * 4096), str(7 * 4096)]
    assert got["query25"] == ["0", "1", "1"]
    assert got["badn"] == ["-1"]
    assert got["smalllwork"] == ["-5"]
    assert got["smallliwork"] == ["-7"]
    assert got["n1"] == ["0", "3.500"]


def test_tuning_options_are_validated_without_a_gpu():
    """brdc_options' launch-shape knobs: 0 keeps the default, in-range values
    are accepted, out-of-range ones rejected with BRDC_STATUS_INVALID_VALUE
    (checked by the workspace query, which touches no device)."""
    L = _lib.load()
    out = ctypes.c_size_t(0)

    def status(**kw):
        o = _lib.options(**kw)
        return L.brdc_dstebr_device_workspace_ex(4096, ctypes.byref(o), ctypes.byref(out))
    assert status() == 0
    for kw in ({"stream_min_k": 256}, {"stream_min_k": 1 << 20}, {"secular_lanes_log2": 12},
               {"secular_lanes_log2": 24}, {"wide_max_waves": 1}, {"wide_max_waves": 4},
               {"leaf_max": 24, "stream_min_k": 1024, "secular_lanes_log2": 17,
                "wide_max_waves": 2}, {"subtree_max_log2": -1}, {"subtree_max_log2": 1},
               {"subtree_max_log2": 30}):
        assert status(**kw) == 0, kw
    for kw in ({"stream_min_k": 100}, {"stream_min_k": -1}, {"secular_lanes_log2": 11},
               {"secular_lanes_log2": 25}, {"wide_max_waves": 3}, {"wide_max_waves": 16},
               {"subtree_max_log2": -2}, {"subtree_max_log2": 31}):
        assert status(**kw) == 3, kw
    o = _lib.options()
    o.reserved[2] = 1
    assert L.brdc_dstebr_device_workspace_ex(4096, ctypes.byref(o), ctypes.byref(out)) == 3


def test_grad_scratch_query_and_its_bound():
    """The backward's scratch (checkpointed pivot chains of every task) grows
    as n^2: about 17.6 GiB at n = 65536, the largest configured order; orders
    whose scratch would pass 2^37 bytes are refused with
    BRDC_STATUS_INVALID_VALUE before any device work."""
    L = _lib.load()
    out = ctypes.c_size_t(0)
    assert L.brdc_dstebr_grad_workspace(65536, ctypes.byref(out)) == 0
    assert 16 * 2 ** 30 < out.value < 20 * 2 ** 30
    small = ctypes.c_size_t(0)
    assert L.brdc_dstebr_grad_workspace(16384, ctypes.byref(small)) == 0
    assert 10 < out.value / small.value < 20  # quadratic growth
    assert L.brdc_dstebr_grad_workspace(200000, ctypes.byref(out)) == 3
    assert L.brdc_dstebr_grad_device(200000, None, None, None, None, None, None, None, 0,
                                     None) == 2  # pointers are checked first
