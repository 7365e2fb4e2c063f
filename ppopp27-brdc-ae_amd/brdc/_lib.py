"""ctypes binding of libbrdc.so (the C ABI declared in include/brdc.h).

The library is built in place by ``make -C ppopp27-brdc-ae_amd`` (or
``__graft_entry__.build()``).  Nothing here falls back to a CPU path: if the
library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libbrdc.so")

# brdc_status codes (include/brdc.h)
STATUS = {0: "success", 2: "invalid pointer", 3: "invalid value", 4: "invalid workspace",
          5: "alloc failed", 6: "hip error", 7: "internal error"}

#: every symbol include/brdc.h declares
EXPORTS = ("brdc_dstebr", "brdc_dstebr_", "dstedc_dc_br_", "brdc_dstebr_device_workspace",
           "brdc_dstebr_device_workspace_ex", "brdc_dstebr_device", "brdc_dstebr_device_ex",
           "brdc_merge_stats", "brdc_status_string", "brdc_profile_begin", "brdc_profile_end",
           "brdc_batch_plan_create", "brdc_batch_plan_destroy", "brdc_batch_workspace",
           "brdc_batch_upload", "brdc_batch_solve", "brdc_batch_plan_stats",
           "brdc_dstebr_grad_workspace", "brdc_dstebr_grad_device", "brdc_batch_grad_workspace",
           "brdc_batch_grad", "brdc_dstebr_device_noalloc", "brdc_profile_end_levels",
           "brdc_batch_merge_stats")

KERNEL_CLASSES = ("norm", "prep_e", "leaf", "prepare", "deflate", "secular", "wprod", "rows",
                  "finalize", "secular_stream", "secular_mid", "subtree")
# ints per merge record of brdc_merge_stats / brdc_batch_merge_stats
MERGE_STAT_FIELDS = 9
MERGE_STAT_NAMES = ("level", "start", "size", "k", "givens", "evals", "tiny", "evals_max",
                    "failed")
# the allocation-free entry's static workspace (BRDC_NOALLOC_* in include/brdc.h)
NOALLOC_MAX_N = 65536
NOALLOC_BYTES = 20 << 20
# merges of at least this many poles go to the streamed secular kernel
# (kStreamMinK in csrc/brdc.hip)
STREAM_MIN_K = 2048
# merges of at most SMALL_MERGE poles (at levels of nodes of at most
# SMALL_MERGE rows; SMALL_MERGE_TOP above) get their product form and
# boundary rows in the finalize launch, one block each (kSmallMaxK,
# kSmallTopK in csrc/brdc_kernels.h)
SMALL_MERGE = 256
SMALL_MERGE_TOP = 128
# largest node (rows) solved inside the subtree kernel, with all of its leaves
SUBTREE_ROWS = 128


class Options(ctypes.Structure):
    """brdc_options: leaf_max and the launch-shape knobs (0 = default each)
    plus reserved words."""
    _fields_ = [("leaf_max", ctypes.c_int), ("stream_min_k", ctypes.c_int),
                ("secular_lanes_log2", ctypes.c_int), ("wide_max_waves", ctypes.c_int),
                ("subtree_max_log2", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


class BrdcError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BrdcError(f"HIP library {LIB_PATH} is not built; run "
                        f"`make -C {PKG_ROOT}` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    L.brdc_dstebr.restype = ctypes.c_int
    L.brdc_dstebr.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int]
    L.brdc_dstebr_.restype = None
    L.brdc_dstebr_.argtypes = [_ip, _dp, _dp, _dp, _ip, _ip, _ip, _ip]
    L.brdc_dstebr_device_workspace.restype = ctypes.c_int
    L.brdc_dstebr_device_workspace.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    L.brdc_dstebr_device_workspace_ex.restype = ctypes.c_int
    L.brdc_dstebr_device_workspace_ex.argtypes = [ctypes.c_int, ctypes.POINTER(Options),
                                                  ctypes.POINTER(ctypes.c_size_t)]
    L.brdc_dstebr_device.restype = ctypes.c_int
    L.brdc_dstebr_device.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp]
    L.brdc_dstebr_device_ex.restype = ctypes.c_int
    L.brdc_dstebr_device_ex.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp,
                                        ctypes.POINTER(Options)]
    L.brdc_dstebr_device_noalloc.restype = ctypes.c_int
    L.brdc_dstebr_device_noalloc.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp,
                                             ctypes.POINTER(Options)]
    L.brdc_merge_stats.restype = ctypes.c_int
    L.brdc_merge_stats.argtypes = [ctypes.c_int, _vp, ctypes.c_size_t, _vp,
                                   ctypes.POINTER(Options), _ip, ctypes.c_int]
    L.brdc_profile_begin.restype = ctypes.c_int
    L.brdc_profile_begin.argtypes = []
    L.brdc_profile_end.restype = ctypes.c_int
    L.brdc_profile_end.argtypes = [_dp, _ip, ctypes.c_int]
    L.brdc_profile_end_levels.restype = ctypes.c_int
    L.brdc_profile_end_levels.argtypes = [_dp, _ip, ctypes.c_int, ctypes.c_int]
    L.brdc_batch_merge_stats.restype = ctypes.c_int
    L.brdc_batch_merge_stats.argtypes = [_vp, _vp, ctypes.c_size_t, _vp, _ip, ctypes.c_int]
    L.brdc_status_string.restype = ctypes.c_char_p
    L.brdc_status_string.argtypes = [ctypes.c_int]
    L.brdc_batch_plan_create.restype = ctypes.c_int
    L.brdc_batch_plan_create.argtypes = [ctypes.c_int, _ip, ctypes.POINTER(Options),
                                         ctypes.POINTER(ctypes.c_void_p)]
    L.brdc_batch_plan_destroy.restype = None
    L.brdc_batch_plan_destroy.argtypes = [_vp]
    L.brdc_batch_workspace.restype = ctypes.c_int
    L.brdc_batch_workspace.argtypes = [_vp, ctypes.POINTER(ctypes.c_size_t)]
    L.brdc_batch_upload.restype = ctypes.c_int
    L.brdc_batch_upload.argtypes = [_vp, _vp, ctypes.c_size_t, _vp]
    L.brdc_batch_solve.restype = ctypes.c_int
    L.brdc_batch_solve.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp]
    L.brdc_batch_plan_stats.restype = ctypes.c_int
    L.brdc_batch_plan_stats.argtypes = [_vp, _ip, ctypes.c_int]
    L.brdc_dstebr_grad_workspace.restype = ctypes.c_int
    L.brdc_dstebr_grad_workspace.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    L.brdc_dstebr_grad_device.restype = ctypes.c_int
    L.brdc_dstebr_grad_device.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          ctypes.c_size_t, _vp]
    L.brdc_batch_grad_workspace.restype = ctypes.c_int
    L.brdc_batch_grad_workspace.argtypes = [_vp, ctypes.POINTER(ctypes.c_size_t)]
    L.brdc_batch_grad.restype = ctypes.c_int
    L.brdc_batch_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp,
                                  ctypes.c_size_t, _vp]
    _lib = L
    return L


#: names of the tunable fields of brdc_options besides leaf_max
TUNING_FIELDS = ("stream_min_k", "secular_lanes_log2", "wide_max_waves")


def options(leaf_max: int = 0, stream_min_k: int = 0, secular_lanes_log2: int = 0,
            wide_max_waves: int = 0, subtree_max_log2: int = 0) -> Options:
    o = Options()
    o.leaf_max = int(leaf_max)
    o.stream_min_k = int(stream_min_k)
    o.secular_lanes_log2 = int(secular_lanes_log2)
    o.wide_max_waves = int(wide_max_waves)
    o.subtree_max_log2 = int(subtree_max_log2)
    return o


def check(status: int, what: str) -> None:
    if status != 0:
        msg = load().brdc_status_string(status).decode()
        raise BrdcError(f"{what}: {msg} (status {status})")
