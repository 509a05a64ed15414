"""The C-ABI library loads without a GPU, exports every symbol the public headers
declare, and reports errors through status codes -- CPU only."""
import ctypes
import os
import re

import mtcheck_amd as M
from mtcheck_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in ("mtcheck.h", "mtcheck_workload.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        syms |= set(re.findall(r"^(?:int|const char\*)\s+(mtc_\w+)\(", text, re.M))
    return syms


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 17
    for s in syms:
        assert hasattr(L, s), s
    assert syms == set(_lib.EXPORTS)


def test_status_strings_and_argument_errors():
    L = _lib.lib()
    assert L.mtc_status_string(0) == b"ok"
    assert L.mtc_status_string(5) == b"corrupt diff"
    h = ctypes.c_void_p()
    assert L.mtc_create(ctypes.byref(h), 0, M.TREE, M.MURMUR3, 0) == 1        # chunk size 0
    assert L.mtc_create(ctypes.byref(h), 128, M.TREE, 7, 0) == 1              # unknown hash
    assert L.mtc_create(ctypes.byref(h), 128, M.FULL, M.MURMUR3, 0) == 2      # unsupported approach
    assert L.mtc_checkpoint(None, None, 0, 0, 0, None, None, None) == 1
    assert L.mtc_digest_host(None, 5, 0, None) == 1


def test_header_layout():
    # include/utils.hpp:29-37: 32 bytes, datalen at offset 8
    assert ctypes.sizeof(_lib.Header) == 32
    assert _lib.Header.datalen.offset == 8
    assert _lib.Header.num_shift_dupl.offset == 28
