"""The oracle's basic approach (src/basic_approach.cpp) on the CPU: diff layout,
the changed-chunk rule, baselines and restarts.  No reference test holds basic
diffs, so the checks are invariants: every checkpoint restarts from its chain,
the listed chunks are exactly those whose bytes changed (digests are compared,
so a chunk rewritten with its old bytes is not listed), the first checkpoint
lists every chunk, and the diff size is 32 + F (4 + C)."""
import numpy as np
import pytest

from mtcheck_amd.workload import CheckpointSeries
from oracle import oracle as O


def chunks_changed(prev, cur, C):
    n = (cur.size + C - 1) // C
    return [c for c in range(n) if prev is None or not np.array_equal(prev[c * C:(c + 1) * C], cur[c * C:(c + 1) * C])]


@pytest.mark.parametrize("nbytes,C,hk", [(1 << 18, 128, O.HASH_MURMUR3), (100_003, 64, O.HASH_MD5),
                                         (5 * 128 + 1, 128, O.HASH_MURMUR3)])
def test_basic_series(nbytes, C, hk):
    series = CheckpointSeries(nbytes, C, seed=21, fresh_frac=0.05, copy_frac=0.03)
    d = O.OracleBasic(C, hk)
    buf = np.zeros(nbytes, np.uint8)
    prev, diffs, versions = None, [], []
    for k in range(6):
        series.apply_host(buf, k)
        diff = d.checkpoint(buf, make_baseline=(k == 0))
        h = O.parse_header(diff)
        ids = np.frombuffer(diff[32:32 + 4 * h["num_first_ocur"]], np.uint32).tolist()
        assert ids == chunks_changed(prev, buf, C)
        assert (h["num_prior_chkpts"], h["num_shift_dupl"], h["chkpt_id"], h["ref_id"]) == (0, 0, k, 0)
        assert len(diff) == 32 + h["num_first_ocur"] * (4 + C)
        diffs.append(diff)
        versions.append(buf.copy())
        prev = buf.copy()
    for k in range(6):
        assert np.array_equal(O.restart_basic(diffs, k), versions[k]), k


def test_basic_baseline_resets_chain():
    C = 128
    d = O.OracleBasic(C)
    a = np.arange(4096, dtype=np.uint8)
    b = a.copy()
    b[:C] ^= 1
    d0 = d.checkpoint(a, True)
    d1 = d.checkpoint(b)
    d2 = d.checkpoint(b, True)  # a new baseline lists every chunk again
    assert O.parse_header(d1)["num_first_ocur"] == 1
    h2 = O.parse_header(d2)
    assert (h2["ref_id"], h2["num_first_ocur"]) == (2, 4096 // C)
    assert np.array_equal(O.restart_basic([d0, d1, d2], 2), b)
