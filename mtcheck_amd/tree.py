"""Host-side mirror of MTCheck's TreeDeduplicator over the C ABI.

Method names, argument meaning and the id-advance rules follow
include/tree_approach.hpp:16-205 and src/tree_approach.cpp:1577-1915; device
buffers are torch tensors (or raw (pointer, length) pairs), diffs are bytes.
"""
from __future__ import annotations

import ctypes
import os
import struct
import time
from typing import Sequence

from . import _lib
from ._lib import Header, Stats, check, lib

# Approaches (include/utils.hpp:39-49) and hashes (include/hash_functions.hpp:16-20).
UNKNOWN, FULL, BASIC, LIST, TREE, TREE_LOW_OFFSET_REF, TREE_LOW_OFFSET, TREE_LOW_ROOT_REF, TREE_LOW_ROOT = range(9)
MURMUR3, MD5 = 0, 1
# Labels, include/utils.hpp:51-57.
FIRST_OCUR, FIXED_DUPL, SHIFT_DUPL, FIRST_DUPL, DONE = 0, 1, 2, 3, 4

HEADER = struct.Struct("<IIQIIII")
_MODES = {"--run-full-chkpt": FULL, "--run-basic-chkpt": BASIC, "--run-list-chkpt": LIST,
          "--run-tree-chkpt": TREE, "--run-tree-low-offset-ref-chkpt": TREE_LOW_OFFSET_REF,
          "--run-tree-low-offset-chkpt": TREE_LOW_OFFSET, "--run-tree-low-root-ref-chkpt": TREE_LOW_ROOT_REF,
          "--run-tree-low-root-chkpt": TREE_LOW_ROOT}


def get_mode(argv: Sequence[str]) -> int:
    """Mode flag parsing, src/utils.cpp:15-36 (unknown flags give UNKNOWN)."""
    for a in argv:
        if a in _MODES:
            return _MODES[a]
    return UNKNOWN


def _dev_ptr(data):
    """(device pointer, length in bytes) of a torch tensor or an explicit pair."""
    if isinstance(data, tuple):
        return int(data[0]), int(data[1])
    if not data.is_cuda:
        raise ValueError("checkpoint data must live in device memory")
    if not data.is_contiguous():
        raise ValueError("checkpoint data must be contiguous")
    return data.data_ptr(), data.numel() * data.element_size()


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    return int(getattr(stream, "cuda_stream", stream))


def parse_header(diff) -> dict:
    keys = ("ref_id", "chkpt_id", "datalen", "chunk_size", "num_first_ocur", "num_prior_chkpts", "num_shift_dupl")
    return dict(zip(keys, HEADER.unpack_from(bytes(diff[:32]))))


class TreeDeduplicator:
    """Tree (lowest-offset) deduplicator resident on one GPU."""

    def __init__(self, chunk_size: int, approach: int = TREE, hash_kind: int = MURMUR3, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().mtc_create(ctypes.byref(h), chunk_size, approach, hash_kind, device))
        self._h = h
        self.chunk_size = chunk_size
        self.approach = BASIC if approach == BASIC else TREE
        self.device = device
        self.hash_kind = hash_kind
        self.datasizes = (0, 0)
        self.timers = [0.0, 0.0, 0.0, 0.0]
        self.restart_timers = [0.0, 0.0]

    def close(self):
        if getattr(self, "_h", None):
            lib().mtc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------
    def stats(self) -> Stats:
        s = Stats()
        check(lib().mtc_get_stats(self._h, ctypes.byref(s)), self._h)
        return s

    @property
    def current_id(self) -> int:
        return self.stats().current_id

    def num_first_ocur(self) -> int:  # include/tree_approach.hpp:53-55
        return self.stats().num_first_ocur

    def num_shift_dupl(self) -> int:  # include/tree_approach.hpp:57-59
        return self.stats().num_shift_dupl

    def set_tuning(self, leaf_tiles_per_block: int = 1, load_factor: float = 0.125, fuse_levels: int = 8) -> None:
        """Schedule parameters; load factor and fuse depth apply from the next baseline."""
        check(lib().mtc_set_tuning(self._h, leaf_tiles_per_block, load_factor, fuse_levels), self._h)

    def set_sweep_tiles(self, tiles_a: int = 0, tiles_b: int = 0) -> None:
        """First-stage tiles per wavefront of sweep A and sweep B (0: the default)."""
        check(lib().mtc_set_sweep_tiles(self._h, tiles_a, tiles_b), self._h)

    def set_debug(self, on: bool = True) -> None:
        check(lib().mtc_set_debug(self._h, 1 if on else 0), self._h)

    def set_compat(self, on: bool = True) -> None:
        """Reference-identical bytes where the library departs from them (the
        reference's leaf ranges; no shifted root 0), from the next checkpoint on."""
        check(lib().mtc_set_compat(self._h, 1 if on else 0), self._h)

    # -- checkpoint overloads (src/tree_approach.cpp:1577-1776) -----------------
    def _run(self, data, make_baseline: bool, advance: bool, stream=None, copy: bool = True):
        ptr, n = _dev_ptr(data)
        hdr = Header()
        size = ctypes.c_uint64()
        check(lib().mtc_checkpoint(self._h, ptr, n, 1 if make_baseline else 0, 1 if advance else 0,
                                   _stream_ptr(stream), ctypes.byref(hdr), ctypes.byref(size)), self._h)
        diff = None
        copy_s = 0.0
        if copy:  # the copy of the diff to the host is the fourth timer (:1678-1688)
            buf = ctypes.create_string_buffer(size.value)
            t0 = time.perf_counter()
            check(lib().mtc_copy_diff(self._h, buf, size.value, _stream_ptr(stream)), self._h)
            copy_s = time.perf_counter() - t0
            diff = buf.raw
        s = self.stats()
        meta = 4 * hdr.num_first_ocur + 8 * hdr.num_prior_chkpts + 8 * hdr.num_shift_dupl
        self.datasizes = (size.value - meta, meta)  # (:707-708) data size includes the header
        self.timers = [s.ms_setup / 1e3, s.ms_dedup / 1e3, s.ms_collect / 1e3, copy_s]
        return hdr, diff

    def copy_last_diff_device(self, out=None):
        """The last diff copied into a (new) uint8 cuda tensor."""
        import torch
        size = self.stats().diff_size
        if out is None:
            out = torch.empty(size, dtype=torch.uint8, device=f"cuda:{self.device}")
        check(lib().mtc_copy_diff(self._h, out.data_ptr(), out.numel(), None), self._h)
        return out[:size]

    def copy_last_diff(self) -> bytes:
        """The last checkpoint's diff, copied from the device."""
        p = ctypes.c_void_p()
        size = ctypes.c_uint64()
        check(lib().mtc_diff_device(self._h, ctypes.byref(p), ctypes.byref(size)), self._h)
        buf = ctypes.create_string_buffer(size.value)
        check(lib().mtc_copy_diff(self._h, buf, size.value, None), self._h)
        return buf.raw

    def copy_last_diff_async(self, out) -> int:
        """Queues a copy of the last diff into `out` (a pinned uint8 CPU tensor of at
        least the diff's size) and returns the diff's size; `out` is complete once
        wait_copies() returns.  The next checkpoint may run meanwhile."""
        size = self.stats().diff_size
        if out.numel() < size:
            raise ValueError("buffer smaller than the diff")
        check(lib().mtc_copy_diff_async(self._h, out.data_ptr(), out.numel()), self._h)
        return size

    def wait_copies(self) -> None:
        """Waits for every copy queued by copy_last_diff_async."""
        check(lib().mtc_copy_wait(self._h), self._h)

    def checkpoint_header(self, data, make_baseline: bool = False, stream=None):
        """checkpoint(header_t&, ptr, len, diff_h, bool): does not advance the id."""
        hdr, diff = self._run(data, make_baseline, advance=False, stream=stream)
        return hdr, diff

    def checkpoint(self, data, make_baseline: bool = False, stream=None) -> bytes:
        """checkpoint(ptr, len, diff_h, bool), :1740-1748: diff to the host, id advances."""
        return self._run(data, make_baseline, advance=True, stream=stream)[1]

    def checkpoint_device(self, data, make_baseline: bool = False, stream=None):
        """Diff left in device memory: (device pointer, size); id advances."""
        hdr, _ = self._run(data, make_baseline, advance=True, stream=stream, copy=False)
        p = ctypes.c_void_p()
        size = ctypes.c_uint64()
        check(lib().mtc_diff_device(self._h, ctypes.byref(p), ctypes.byref(size)), self._h)
        return hdr, p.value, size.value

    def checkpoint_log(self, data, logname: str, make_baseline: bool = False, stream=None) -> bytes:
        """checkpoint(ptr, len, diff_h, logname, bool), :1765-1776: writes the three CSV logs."""
        hdr, diff = self._run(data, make_baseline, advance=True, stream=stream)
        self.write_chkpt_log(hdr, diff, logname)
        return diff

    def checkpoint_file(self, data, filename: str, logname: str, make_baseline: bool = False, stream=None) -> bytes:
        """checkpoint(ptr, len, filename, logname, bool), :1705-1724: also writes the diff file."""
        hdr, diff = self._run(data, make_baseline, advance=True, stream=stream)
        self.write_chkpt_log(hdr, diff, logname)
        with open(filename, "wb") as f:
            f.write(diff)
        return diff

    # -- restart overloads (src/tree_approach.cpp:1786-1840) ---------------------
    def restart(self, out, chkpts, chkpt_id: int, logname: str | None = None, stream=None) -> None:
        restart_into(out, chkpts, chkpt_id, device=self.device, stream=stream, approach=self.approach)
        self.restart_timers = last_restart_times()
        if logname:
            self.write_restart_log(chkpt_id, f"{logname}.chunk_size.{self.chunk_size}.restart_timing.csv")

    def restart_files(self, out, filenames: Sequence[str], chkpt_id: int, logname: str | None = None,
                      stream=None) -> None:
        diffs = []
        for name in filenames:
            suffix = ".basic.incr_chkpt" if self.approach == BASIC else ".hashtree.incr_chkpt"
            with open(name + suffix, "rb") as f:
                diffs.append(f.read())
        restart_into(out, diffs, chkpt_id, device=self.device, stream=stream, approach=self.approach)
        self.restart_timers = last_restart_times()
        if logname:
            self.write_restart_log(chkpt_id, logname)

    # -- logs (src/tree_approach.cpp:1853-1915, src/utils.cpp:38-124) -------------
    def write_chkpt_log(self, header, diff: bytes, logname: str) -> None:
        cid = header.chkpt_id
        base = f"{logname}.chunk_size.{self.chunk_size}"
        result, size, timing = base + ".csv", base + ".size.csv", base + ".timing.csv"
        new = not os.path.exists(result) or os.path.getsize(result) == 0
        d, m = self.datasizes
        t = self.timers
        with open(result, "a") as f:
            if new:
                f.write("Approach,Chkpt ID,Chunk Size,Uncompressed Size,Compressed Size,Data Size,Metadata Size,"
                        "Setup Time,Comparison Time,Gather Time,Write Time\n")
            f.write(f"Tree,{cid},{self.chunk_size},{header.datalen},{d + m},{d},{m},{t[0]},{t[1]},{t[2]},{t[3]}\n")
        with open(timing, "a") as f:
            f.write(f"Tree,{cid},{self.chunk_size},{t[1]},{t[2]},{t[3]}\n")
        with open(size, "a") as f:
            f.write(f"Tree,{cid},{self.chunk_size},{d},{m},")
            f.write(metadata_breakdown(header, diff, cid))

    def write_restart_log(self, select_chkpt: int, logname: str) -> None:
        with open(logname, "a") as f:
            f.write(f"Tree,{select_chkpt},{self.chunk_size},{self.restart_timers[0]},{self.restart_timers[1]}\n")


class Batch:
    """Many deduplicators checkpointed together (mtc_checkpoint_batch): one sequence
    of batched launches per step, replayed from a captured hipGraph when the step's
    shape repeats.  Each member's diff equals what its own checkpoint() returns."""

    def __init__(self, device: int = 0, use_graph: bool = True):
        b = ctypes.c_void_p()
        check(lib().mtc_batch_create(ctypes.byref(b), device, 1 if use_graph else 0))
        self._b = b
        self.device = device

    def close(self):
        if getattr(self, "_b", None):
            lib().mtc_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def checkpoint(self, dedups, datas, make_baseline=False, advance: bool = True, stream=None, copy: bool = False):
        """One checkpoint of every deduplicator.  make_baseline: bool or one bool per
        member.  Returns (headers, diff sizes) or, with copy, the host diffs."""
        n = len(dedups)
        if len(datas) != n:
            raise ValueError("one buffer per deduplicator")
        hs = (ctypes.c_void_p * n)(*[d._h.value for d in dedups])
        ptrs = (ctypes.c_void_p * n)()
        lens = (ctypes.c_uint64 * n)()
        for i, x in enumerate(datas):
            ptrs[i], lens[i] = _dev_ptr(x)
        bl = make_baseline if isinstance(make_baseline, (list, tuple)) else [make_baseline] * n
        base = (ctypes.c_int * n)(*[1 if b else 0 for b in bl])
        hdrs = (Header * n)()
        sizes = (ctypes.c_uint64 * n)()
        rc = lib().mtc_checkpoint_batch(self._b, hs, ptrs, lens, base, n, 1 if advance else 0, _stream_ptr(stream),
                                        hdrs, sizes)
        if rc != _lib.MTC_OK:
            raise _lib.MtcError(rc, (lib().mtc_batch_last_error(self._b) or b"").decode())
        if not copy:
            return list(hdrs), list(sizes)
        out = []
        for d, size in zip(dedups, sizes):
            buf = ctypes.create_string_buffer(size)
            check(lib().mtc_copy_diff(d._h, buf, size, None), d._h)
            out.append(buf.raw)
        return out


def metadata_breakdown(header, diff: bytes, num_chkpts: int) -> str:
    """write_metadata_breakdown, src/utils.cpp:38-124 (tree mode)."""
    F, P, S = header.num_first_ocur, header.num_prior_chkpts, header.num_shift_dupl
    out = [f"32,{F * 4},"]
    if header.ref_id != header.chkpt_id:
        out.append(f"{8 * P}")
        for i in range(num_chkpts):
            if i < P:
                off = 32 + 4 * F + 8 * i
                _, num = struct.unpack_from("<II", diff, off)
                out.append(f",{num * 8}")
            else:
                out.append(",0")
    else:
        out.append(f"0,{S * 8}")
        for _ in range(1, num_chkpts):
            out.append(",0")
    out.append("\n")
    return "".join(out)


def last_restart_times() -> list:
    """[upload, whole restart] of this thread's last restart, in seconds
    (restart_timers, src/tree_approach.cpp:1792-1793; upload 0 for device chains)."""
    up, total = ctypes.c_float(), ctypes.c_float()
    check(lib().mtc_restart_times(ctypes.byref(up), ctypes.byref(total)))
    return [up.value / 1e3, total.value / 1e3]


def restart_into(out, chkpts, chkpt_id: int, device: int = 0, stream=None, approach: int = TREE) -> None:
    """Rebuild checkpoint chkpt_id into the device buffer `out` from diffs indexed by id.

    chkpts: host diffs (bytes-like, or None below the chain's baseline) or device
    diffs given as (pointer, size) pairs / uint8 cuda tensors.  approach=BASIC
    reads basic diffs (mtc_restart_basic).
    """
    optr, olen = _dev_ptr(out)
    n = len(chkpts)
    on_device = any((isinstance(c, tuple) or hasattr(c, "data_ptr")) for c in chkpts if c is not None)
    ptrs = (ctypes.c_void_p * n)()
    sizes = (ctypes.c_uint64 * n)()
    keep = []
    for i, c in enumerate(chkpts):
        if c is None:
            continue
        if on_device:
            p, s = _dev_ptr(c)
        else:
            b = bytes(c)
            buf = ctypes.create_string_buffer(b, len(b))
            keep.append(buf)
            p, s = ctypes.addressof(buf), len(b)
        ptrs[i] = p
        sizes[i] = s
    fn = lib().mtc_restart_basic if approach == BASIC else lib().mtc_restart
    check(fn(ptrs, sizes, n, chkpt_id, 1 if on_device else 0, optr, olen, device, _stream_ptr(stream)))


def digest_host(data: bytes, hash_kind: int = MURMUR3) -> bytes:
    """calculate_digest_host, include/hash_functions.hpp:46-82 (raw 16 bytes)."""
    b = bytes(data)
    out = ctypes.create_string_buffer(16)
    src = ctypes.create_string_buffer(b, len(b)) if b else None
    check(lib().mtc_digest_host(src, len(b), hash_kind, out))
    return out.raw


def digest_str(digest: bytes) -> str:
    """Hex digest in groups of four bytes, as calculate_digest_host prints it."""
    h = digest.hex().upper()
    return " ".join(h[i:i + 8] for i in range(0, 32, 8)) + " "


def hash_chunks(data, chunk_size: int, hash_kind: int = MURMUR3, out=None, stream=None):
    """Leaf digests of every chunk of a device buffer (16 bytes each)."""
    import torch
    ptr, n = _dev_ptr(data)
    nch = (n + chunk_size - 1) // chunk_size
    if out is None:
        out = torch.empty(nch * 16, dtype=torch.uint8, device=data.device)
    check(lib().mtc_hash_chunks(ptr, n, chunk_size, hash_kind, out.data_ptr(), _stream_ptr(stream)))
    return out


def debug_tree(d: TreeDeduplicator):
    """Full tree of the last checkpoint (needs set_debug before it): (2n-1, 16) uint8."""
    import numpy as np
    s = d.stats()
    nodes = 2 * s.num_chunks - 1
    out = np.zeros((nodes, 16), np.uint8)
    check(lib().mtc_copy_tree(d._h, out.ctypes.data, out.nbytes), d._h)
    return out


def debug_labels(d: TreeDeduplicator):
    """Labels of every node after the last checkpoint (needs set_debug before it)."""
    import numpy as np
    s = d.stats()
    out = np.zeros(2 * s.num_chunks - 1, np.uint8)
    check(lib().mtc_copy_labels(d._h, out.ctypes.data, out.nbytes), d._h)
    return out
