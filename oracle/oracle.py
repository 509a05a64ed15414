"""ctypes view of the CPU oracle (oracle/mtc_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of bench.py -- as the checker, never as the thing
measured or shipped.  The library under test (mtcheck_amd) never imports it.

The C file restates MTCheck's serial tree low-offset deduplication
(include/reference_impl.hpp:15-132), its diff writer (src/tree_approach.cpp:511-710)
and its restart (src/tree_approach.cpp:712-1128); see the header of mtc_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmtc_oracle.so")
SRC_PATH = os.path.join(_HERE, "mtc_oracle.c")

FIRST_OCUR, FIXED_DUPL, SHIFT_DUPL, FIRST_DUPL, DONE = 0, 1, 2, 3, 4
HASH_MURMUR3, HASH_MD5 = 0, 1
HEADER = struct.Struct("<IIQIIII")  # include/utils.hpp:29-37


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (the recipe build() in __graft_entry__ runs)."""
    if force or not os.path.exists(LIB_PATH):
        tmp = LIB_PATH + ".tmp"
        subprocess.check_call(["gcc", "-O2", "-std=c11", "-Wall", "-fopenmp", "-shared", "-fPIC",
                               "-o", tmp, SRC_PATH])
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u8p, u32, u64, i32 = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.orc_murmur3_mtcheck.argtypes = [vp, u64, vp]
        L.orc_murmur3_canonical.argtypes = [vp, u64, u32, vp]
        L.orc_md5.argtypes = [vp, u64, vp]
        L.orc_hash.argtypes = [i32, vp, u64, vp]
        L.orc_leaf_runs.argtypes = [u32, u32, vp, vp]
        L.orc_leaf_runs.restype = i32
        L.orc_ref_num_leaf_descendents.argtypes = [u32, u32]
        L.orc_ref_num_leaf_descendents.restype = u32
        L.orc_ref_leftmost_leaf.argtypes = [u32, u32]
        L.orc_ref_leftmost_leaf.restype = u32
        L.orc_dedup_create.argtypes = [u32, i32, i32]
        L.orc_dedup_create.restype = vp
        L.orc_dedup_destroy.argtypes = [vp]
        L.orc_dedup_checkpoint.argtypes = [vp, vp, u64, i32]
        L.orc_dedup_checkpoint.restype = i32
        L.orc_dedup_checkpoint_par.argtypes = [vp, vp, u64, i32, i32]
        L.orc_dedup_checkpoint_par.restype = i32
        for name in ("orc_dedup_diff_size", "orc_dedup_table_size"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = u64
        for name in ("orc_dedup_num_first_ocur", "orc_dedup_num_shift_dupl", "orc_dedup_num_nodes",
                     "orc_dedup_current_id"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = u32
        for name in ("orc_dedup_copy_diff", "orc_dedup_copy_labels", "orc_dedup_copy_tree",
                     "orc_dedup_copy_first_roots", "orc_dedup_copy_shift_roots"):
            getattr(L, name).argtypes = [vp, vp]
        L.orc_dedup_table_find.argtypes = [vp, vp, vp, vp]
        L.orc_dedup_table_find.restype = i32
        L.orc_restart.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64), u32, u32,
                                  vp, u64, i32]
        L.orc_restart.restype = i32
        L.orc_basic_checkpoint.argtypes = [vp, vp, u64, i32]
        L.orc_basic_checkpoint.restype = i32
        L.orc_restart_basic.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64), u32, u32,
                                        vp, u64]
        L.orc_restart_basic.restype = i32
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _as_u8(data) -> np.ndarray:
    if isinstance(data, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(data), dtype=np.uint8)
    return np.ascontiguousarray(data).view(np.uint8).reshape(-1)


def murmur3_mtcheck(data) -> bytes:
    a = _as_u8(data)
    out = np.zeros(16, np.uint8)
    lib().orc_murmur3_mtcheck(_ptr(a) if a.size else None, a.size, _ptr(out))
    return out.tobytes()


def murmur3_canonical(data, seed: int = 0) -> bytes:
    a = _as_u8(data)
    out = np.zeros(16, np.uint8)
    lib().orc_murmur3_canonical(_ptr(a) if a.size else None, a.size, seed, _ptr(out))
    return out.tobytes()


def md5(data) -> bytes:
    a = _as_u8(data)
    out = np.zeros(16, np.uint8)
    lib().orc_md5(_ptr(a) if a.size else None, a.size, _ptr(out))
    return out.tobytes()


def hash_bytes(kind: int, data) -> bytes:
    return md5(data) if kind == HASH_MD5 else murmur3_mtcheck(data)


def leaf_runs(node: int, n: int):
    st = np.zeros(2, np.uint32)
    ct = np.zeros(2, np.uint32)
    r = lib().orc_leaf_runs(node, n, _ptr(st), _ptr(ct))
    return [(int(st[i]), int(ct[i])) for i in range(r)]


def ref_num_leaf_descendents(node: int, num_nodes: int) -> int:
    return int(lib().orc_ref_num_leaf_descendents(node, num_nodes))


def ref_leftmost_leaf(node: int, num_nodes: int) -> int:
    return int(lib().orc_ref_leftmost_leaf(node, num_nodes))


class OracleDedup:
    """Serial tree deduplicator (the reference's lowest-offset rule)."""

    def __init__(self, chunk_size: int, hash_kind: int = HASH_MURMUR3, compat_ref: bool = False):
        self._destroy = lib().orc_dedup_destroy
        self._h = lib().orc_dedup_create(chunk_size, hash_kind, 1 if compat_ref else 0)
        if not self._h:
            raise ValueError("bad chunk size")
        self.chunk_size = chunk_size

    def __del__(self):
        # the destroy function is held by the instance: module globals may
        # already be cleared when an instance dies at interpreter exit
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)
            self._h = None

    def checkpoint(self, data, make_baseline: bool = False, threads: int = 0) -> bytes:
        """threads > 0: the OpenMP checkpoint on that many threads (same bytes)."""
        a = _as_u8(data)
        if threads > 0:
            rc = lib().orc_dedup_checkpoint_par(self._h, _ptr(a), a.size, 1 if make_baseline else 0, threads)
        else:
            rc = lib().orc_dedup_checkpoint(self._h, _ptr(a), a.size, 1 if make_baseline else 0)
        if rc != 0:
            raise RuntimeError(f"oracle checkpoint failed: {rc}")
        out = np.empty(lib().orc_dedup_diff_size(self._h), np.uint8)
        lib().orc_dedup_copy_diff(self._h, _ptr(out))
        return out.tobytes()

    @property
    def current_id(self) -> int:
        return int(lib().orc_dedup_current_id(self._h))

    def num_first_ocur(self) -> int:
        return int(lib().orc_dedup_num_first_ocur(self._h))

    def num_shift_dupl(self) -> int:
        return int(lib().orc_dedup_num_shift_dupl(self._h))

    def labels(self) -> np.ndarray:
        out = np.empty(lib().orc_dedup_num_nodes(self._h), np.uint8)
        lib().orc_dedup_copy_labels(self._h, _ptr(out))
        return out

    def tree(self) -> np.ndarray:
        out = np.empty((lib().orc_dedup_num_nodes(self._h), 16), np.uint8)
        lib().orc_dedup_copy_tree(self._h, _ptr(out))
        return out

    def first_roots(self) -> np.ndarray:
        out = np.empty(self.num_first_ocur(), np.uint32)
        if out.size:
            lib().orc_dedup_copy_first_roots(self._h, _ptr(out))
        return out

    def shift_roots(self) -> np.ndarray:
        out = np.empty(self.num_shift_dupl(), np.uint32)
        if out.size:
            lib().orc_dedup_copy_shift_roots(self._h, _ptr(out))
        return out

    def table_find(self, digest: bytes):
        d = np.frombuffer(digest, np.uint8).copy()
        node = ctypes.c_uint32()
        tree = ctypes.c_uint32()
        if lib().orc_dedup_table_find(self._h, _ptr(d), ctypes.byref(node), ctypes.byref(tree)):
            return int(node.value), int(tree.value)
        return None

    def table_size(self) -> int:
        return int(lib().orc_dedup_table_size(self._h))


def restart(diffs, chkpt_idx: int, out_len: int | None = None) -> np.ndarray:
    """Rebuild checkpoint ``chkpt_idx`` from diffs indexed by checkpoint id."""
    bufs = [np.frombuffer(bytes(d), np.uint8) if d is not None else None for d in diffs]
    n = len(bufs)
    ptrs = (ctypes.c_void_p * n)(*[(_ptr(b) if b is not None else None) for b in bufs])
    sizes = (ctypes.c_uint64 * n)(*[(b.size if b is not None else 0) for b in bufs])
    hdr = parse_header(diffs[chkpt_idx])
    length = hdr["datalen"] if out_len is None else out_len
    out = np.zeros(max(length, 1), np.uint8)
    rc = lib().orc_restart(ptrs, sizes, n, chkpt_idx, _ptr(out), length, 0)
    if rc != 0:
        raise RuntimeError(f"oracle restart failed: {rc}")
    return out[:hdr["datalen"]]


class OracleBasic(OracleDedup):
    """The basic approach (src/basic_approach.cpp): per-chunk digests compared with
    the previous checkpoint's; the diff lists the changed chunks, ascending."""

    def __init__(self, chunk_size: int, hash_kind: int = HASH_MURMUR3):
        super().__init__(chunk_size, hash_kind)

    def checkpoint(self, data, make_baseline: bool = False, threads: int = 0) -> bytes:
        a = _as_u8(data)
        rc = lib().orc_basic_checkpoint(self._h, _ptr(a), a.size, 1 if make_baseline else 0)
        if rc != 0:
            raise RuntimeError(f"oracle basic checkpoint failed: {rc}")
        out = np.empty(lib().orc_dedup_diff_size(self._h), np.uint8)
        lib().orc_dedup_copy_diff(self._h, _ptr(out))
        return out.tobytes()


def restart_basic(diffs, chkpt_idx: int) -> np.ndarray:
    """Rebuild checkpoint ``chkpt_idx`` of the basic approach from its diffs."""
    bufs = [np.frombuffer(bytes(d), np.uint8) if d is not None else None for d in diffs]
    n = len(bufs)
    ptrs = (ctypes.c_void_p * n)(*[(_ptr(b) if b is not None else None) for b in bufs])
    sizes = (ctypes.c_uint64 * n)(*[(b.size if b is not None else 0) for b in bufs])
    hdr = parse_header(diffs[chkpt_idx])
    out = np.zeros(max(hdr["datalen"], 1), np.uint8)
    rc = lib().orc_restart_basic(ptrs, sizes, n, chkpt_idx, _ptr(out), hdr["datalen"])
    if rc != 0:
        raise RuntimeError(f"oracle basic restart failed: {rc}")
    return out[:hdr["datalen"]]


def parse_header(diff) -> dict:
    f = HEADER.unpack_from(bytes(diff[:32]))
    keys = ("ref_id", "chkpt_id", "datalen", "chunk_size", "num_first_ocur", "num_prior_chkpts", "num_shift_dupl")
    return dict(zip(keys, f))


def parse_diff(diff) -> dict:
    """Split a .hashtree.incr_chkpt buffer into its sections (src/tree_approach.cpp:777-784)."""
    b = bytes(diff)
    h = parse_header(b)
    F, P, S = h["num_first_ocur"], h["num_prior_chkpts"], h["num_shift_dupl"]
    o = 32
    first = np.frombuffer(b, np.uint32, F, o); o += 4 * F
    table = np.frombuffer(b, np.uint32, 2 * P, o).reshape(P, 2); o += 8 * P
    shift = np.frombuffer(b, np.uint32, 2 * S, o).reshape(S, 2); o += 8 * S
    return {"header": h, "first": first, "table": table, "shift": shift, "data": b[o:]}
