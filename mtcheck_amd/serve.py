"""Served checkpoints over mtc_server_* (the library's native request loop).

A Server owns the clients' resident state; `submit` queues a checkpoint of a
client's device buffer and returns a ticket; `result` polls or waits for it.
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import Header, lib
from .tree import _dev_ptr, _stream_ptr


class ServeError(_lib.MtcError):
    pass


class ClientHandle:
    """Read access to a served client's last diff (valid until its next checkpoint runs)."""

    def __init__(self, ptr):
        self._h = ctypes.c_void_p(ptr)

    def copy_last_diff(self) -> bytes:
        p = ctypes.c_void_p()
        size = ctypes.c_uint64()
        _lib.check(lib().mtc_diff_device(self._h, ctypes.byref(p), ctypes.byref(size)), self._h)
        buf = ctypes.create_string_buffer(size.value)
        _lib.check(lib().mtc_copy_diff(self._h, buf, size.value, None), self._h)
        return buf.raw

    def copy_last_diff_device(self):
        """The last diff copied into a new uint8 cuda tensor."""
        import torch
        s = self.stats()
        out = torch.empty(s.diff_size, dtype=torch.uint8, device="cuda:0")
        _lib.check(lib().mtc_copy_diff(self._h, ctypes.c_void_p(out.data_ptr()), out.numel(), None), self._h)
        return out

    def stats(self):
        """The client's deduplicator statistics (read between its checkpoints)."""
        s = _lib.Stats()
        _lib.check(lib().mtc_get_stats(self._h, ctypes.byref(s)), self._h)
        return s


class Server:
    def __init__(self, device: int = 0, memory_budget: int = 200 << 30, lanes: int = 2, max_batch: int = 64):
        s = ctypes.c_void_p()
        self._check(lib().mtc_server_create(ctypes.byref(s), device, memory_budget, lanes, max_batch), None)
        self._s = s

    def _check(self, rc: int, s) -> None:
        if rc != _lib.MTC_OK:
            msg = lib().mtc_server_last_error(s)
            raise ServeError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_s", None):
            lib().mtc_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_client(self, chunk_size: int, max_len: int, hash_kind: int = 0) -> int:
        c = ctypes.c_uint32()
        self._check(lib().mtc_server_add_client(self._s, chunk_size, hash_kind, max_len, ctypes.byref(c)), self._s)
        return c.value

    def client(self, cid: int) -> ClientHandle:
        h = ctypes.c_void_p()
        self._check(lib().mtc_server_client(self._s, cid, ctypes.byref(h)), self._s)
        return ClientHandle(h.value)

    def remove_client(self, cid: int) -> None:
        self._check(lib().mtc_server_remove_client(self._s, cid), self._s)

    def submit(self, cid: int, data, make_baseline: bool = False, stream=None) -> int:
        ptr, n = _dev_ptr(data)
        t = ctypes.c_uint64()
        self._check(lib().mtc_server_submit(self._s, cid, ptr, n, 1 if make_baseline else 0, _stream_ptr(stream),
                                            ctypes.byref(t)), self._s)
        return t.value

    def result(self, ticket: int, wait: bool = True):
        """None while pending (wait=False); else (header, diff size, t_submit, t_start, t_done)."""
        h = Header()
        size = ctypes.c_uint64()
        ts, t0, t1 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        rc = lib().mtc_server_result(self._s, ticket, 1 if wait else 0, ctypes.byref(h), ctypes.byref(size),
                                     ctypes.byref(ts), ctypes.byref(t0), ctypes.byref(t1))
        if rc == _lib.MTC_PENDING:
            return None
        self._check(rc, self._s)
        return h, size.value, ts.value, t0.value, t1.value

    def stats(self) -> dict:
        v = [ctypes.c_uint64() for _ in range(4)]
        q = ctypes.c_uint32()
        self._check(lib().mtc_server_stats(self._s, *[ctypes.byref(x) for x in v], ctypes.byref(q)), self._s)
        return {"steps": v[0].value, "served": v[1].value, "bytes": v[2].value, "resident_bytes": v[3].value,
                "queued": q.value}
