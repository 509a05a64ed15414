"""ctypes binding of libmtcheck_amd.so (include/mtcheck.h, include/mtcheck_workload.h).

The library must be built in-tree (``__graft_entry__.build()``); importing this
module without it raises, so nothing silently runs without the HIP code.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# MTC_LIBRARY selects another build of the library (a schedule experiment, say).
LIB_PATH = os.environ.get("MTC_LIBRARY") or os.path.join(PKG, "libmtcheck_amd.so")

MTC_OK = 0
MTC_PENDING = 9
STATUS_NAMES = {
    0: "MTC_OK", 1: "MTC_ERR_INVALID_ARG", 2: "MTC_ERR_UNSUPPORTED_APPROACH", 3: "MTC_ERR_OUT_OF_MEMORY",
    4: "MTC_ERR_TABLE_FULL", 5: "MTC_ERR_CORRUPT_DIFF", 6: "MTC_ERR_HIP", 7: "MTC_ERR_STATE",
    8: "MTC_ERR_BUFFER_TOO_SMALL", 9: "MTC_PENDING",
}

# Every symbol the public headers declare (checked by the CPU test suite).
EXPORTS = [
    "mtc_create", "mtc_destroy", "mtc_checkpoint", "mtc_diff_device", "mtc_copy_diff", "mtc_copy_diff_async",
    "mtc_copy_wait", "mtc_get_stats",
    "mtc_restart", "mtc_restart_basic", "mtc_restart_noalloc", "mtc_restart_times", "mtc_digest_host", "mtc_hash_chunks", "mtc_set_compat", "mtc_set_debug", "mtc_copy_tree", "mtc_copy_labels",
    "mtc_last_error", "mtc_status_string", "mtc_set_tuning", "mtc_set_sweep_tiles", "mtc_batch_create", "mtc_batch_destroy",
    "mtc_checkpoint_batch", "mtc_batch_last_error", "mtc_server_create", "mtc_server_destroy", "mtc_server_add_client",
    "mtc_server_client", "mtc_server_remove_client", "mtc_server_submit", "mtc_server_result", "mtc_server_stats",
    "mtc_server_last_error",
    "mtc_wl_fill_stream", "mtc_wl_fill_runs", "mtc_wl_copy_chunks",
]


class MtcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Header(ctypes.Structure):
    _fields_ = [("ref_id", ctypes.c_uint32), ("chkpt_id", ctypes.c_uint32), ("datalen", ctypes.c_uint64),
                ("chunk_size", ctypes.c_uint32), ("num_first_ocur", ctypes.c_uint32),
                ("num_prior_chkpts", ctypes.c_uint32), ("num_shift_dupl", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("current_id", ctypes.c_uint32), ("baseline_id", ctypes.c_uint32),
                ("num_first_ocur", ctypes.c_uint32), ("num_shift_dupl", ctypes.c_uint32),
                ("num_prior_chkpts", ctypes.c_uint32), ("num_chunks", ctypes.c_uint32),
                ("data_len", ctypes.c_uint64), ("diff_size", ctypes.c_uint64), ("data_chunks", ctypes.c_uint64),
                ("table_entries", ctypes.c_uint64), ("table_capacity", ctypes.c_uint64),
                ("ms_setup", ctypes.c_float), ("ms_dedup", ctypes.c_float), ("ms_collect", ctypes.c_float),
                ("ms_copy", ctypes.c_float), ("changed_chunks", ctypes.c_uint64), ("ms_leaf", ctypes.c_float),
                ("leaf_labels_written", ctypes.c_uint32), ("leaf_digests_stored", ctypes.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.mtc_create.argtypes = [ctypes.POINTER(vp), u32, i32, i32, i32]
        L.mtc_destroy.argtypes = [vp]
        L.mtc_checkpoint.argtypes = [vp, vp, u64, i32, i32, vp, ctypes.POINTER(Header), ctypes.POINTER(u64)]
        L.mtc_diff_device.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(u64)]
        L.mtc_copy_diff.argtypes = [vp, vp, u64, vp]
        L.mtc_copy_diff_async.argtypes = [vp, vp, u64]
        L.mtc_copy_wait.argtypes = [vp]
        L.mtc_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.mtc_restart.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(u64), u32, u32, i32, vp, u64, i32, vp]
        L.mtc_restart_basic.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(u64), u32, u32, i32, vp, u64, i32, vp]
        L.mtc_digest_host.argtypes = [vp, u64, i32, vp]
        L.mtc_hash_chunks.argtypes = [vp, u64, u32, i32, vp, vp]
        L.mtc_set_debug.argtypes = [vp, i32]
        L.mtc_copy_tree.argtypes = [vp, vp, u64]
        L.mtc_copy_labels.argtypes = [vp, vp, u64]
        L.mtc_last_error.argtypes = [vp]
        L.mtc_last_error.restype = ctypes.c_char_p
        L.mtc_status_string.argtypes = [i32]
        L.mtc_status_string.restype = ctypes.c_char_p
        L.mtc_set_tuning.argtypes = [vp, u32, ctypes.c_float, u32]
        L.mtc_set_sweep_tiles.argtypes = [vp, u32, u32]
        L.mtc_batch_create.argtypes = [ctypes.POINTER(vp), i32, i32]
        L.mtc_batch_destroy.argtypes = [vp]
        L.mtc_checkpoint_batch.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(u64),
                                           ctypes.POINTER(i32), u32, i32, vp, ctypes.POINTER(Header),
                                           ctypes.POINTER(u64)]
        L.mtc_batch_last_error.argtypes = [vp]
        L.mtc_batch_last_error.restype = ctypes.c_char_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.mtc_server_create.argtypes = [ctypes.POINTER(vp), i32, u64, u32, u32]
        L.mtc_server_destroy.argtypes = [vp]
        L.mtc_server_add_client.argtypes = [vp, u32, i32, u64, ctypes.POINTER(u32)]
        L.mtc_server_client.argtypes = [vp, u32, ctypes.POINTER(vp)]
        L.mtc_server_remove_client.argtypes = [vp, u32]
        L.mtc_server_submit.argtypes = [vp, u32, vp, u64, i32, vp, ctypes.POINTER(u64)]
        L.mtc_server_result.argtypes = [vp, u64, i32, ctypes.POINTER(Header), ctypes.POINTER(u64), dp, dp, dp]
        L.mtc_server_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                       ctypes.POINTER(u64), ctypes.POINTER(u32)]
        L.mtc_server_last_error.argtypes = [vp]
        L.mtc_server_last_error.restype = ctypes.c_char_p
        L.mtc_wl_fill_stream.argtypes = [vp, u64, u64, u64, vp]
        L.mtc_wl_fill_runs.argtypes = [vp, u64, u32, vp, u32, u64, i32, vp]
        L.mtc_wl_copy_chunks.argtypes = [vp, u64, u32, vp, u32, vp, vp]
        for name in EXPORTS:
            if name not in ("mtc_last_error", "mtc_status_string", "mtc_batch_last_error", "mtc_server_last_error"):
                getattr(L, name).restype = i32
        _lib = L
    return _lib


def check(rc: int, handle=None) -> None:
    if rc != MTC_OK:
        msg = lib().mtc_last_error(handle)
        raise MtcError(rc, msg.decode() if msg else "")
