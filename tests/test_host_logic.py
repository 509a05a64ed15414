"""Host-side mirror of the reference interface (no GPU calls) -- CPU only."""
import struct

import mtcheck_amd as M
from mtcheck_amd import _lib
from mtcheck_amd.tree import metadata_breakdown


def test_get_mode_parsing():
    # src/utils.cpp:15-36
    assert M.get_mode(["prog", "128", "10", "--run-tree-chkpt"]) == M.TREE
    assert M.get_mode(["--run-tree-low-offset-chkpt"]) == M.TREE_LOW_OFFSET
    assert M.get_mode(["--run-basic-chkpt"]) == M.BASIC
    assert M.get_mode(["--run-naive-chkpt"]) == M.UNKNOWN


def _diff(ref, cid, F, table, S):
    h = struct.pack("<IIQIIII", ref, cid, 1024, 128, F, len(table), S)
    body = b"".join(struct.pack("<I", i) for i in range(F))
    body += b"".join(struct.pack("<II", c, m) for c, m in table)
    body += b"\0" * (8 * S)
    return h + body


def test_metadata_breakdown_incremental():
    # src/utils.cpp:38-124: header, first-occurrence bytes, repeat map, per-checkpoint bytes
    hdr = _lib.Header(0, 3, 1024, 128, 2, 2, 5)
    diff = _diff(0, 3, 2, [(0, 2), (2, 3)], 5)
    assert metadata_breakdown(hdr, diff, 3) == "32,8,16,16,24,0\n"


def test_metadata_breakdown_baseline():
    hdr = _lib.Header(0, 0, 1024, 128, 1, 1, 4)
    diff = _diff(0, 0, 1, [(0, 4)], 4)
    assert metadata_breakdown(hdr, diff, 3) == "32,4,0,32,0,0\n"


def test_parse_header_roundtrip():
    d = _diff(2, 5, 1, [], 0)
    h = M.parse_header(d)
    assert (h["ref_id"], h["chkpt_id"], h["datalen"], h["chunk_size"], h["num_first_ocur"]) == (2, 5, 1024, 128, 1)
