"""Pins the oracle's labelling, diffs and restarts -- CPU only.

* the reference's own expectations: test/test_case_01..10.cpp (counts at :100/:101)
  and restart == original (test/tree_chkpt.cpp:65-73);
* properties that follow from the definitions, checked by code written from those
  definitions rather than from the algorithm: regions are disjoint and cover exactly
  the non-fixed leaves, every first-occurrence leaf is new (not in any checkpoint
  since the baseline, lowest offset within the checkpoint), every shifted entry
  copies equal bytes, ordering of every section;
* brute force over all 8-leaf inputs of length 4 on a 3-letter alphabet;
* the two documented departures from the reference, and the switch that undoes them.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import REFERENCE_CASES


def heap_leaves(node: int, n: int):
    """Leaves (chunk indices, ascending) under `node` by walking the heap."""
    out, stack = [], [node]
    while stack:
        v = stack.pop()
        if v >= n - 1:
            out.append(v - (n - 1))
        else:
            stack += [2 * v + 1, 2 * v + 2]
    return sorted(out)


@pytest.mark.parametrize("case", sorted(REFERENCE_CASES))
def test_reference_cases(case):
    s0, s1, counts, first, shift = REFERENCE_CASES[case]
    d = O.OracleDedup(1)
    d0 = d.checkpoint(s0.encode(), True)
    assert O.restart([d0], 0).tobytes() == s0.encode()
    d1 = d.checkpoint(s1.encode(), False)
    assert O.restart([d0, d1], 1).tobytes() == s1.encode()
    assert (d.num_first_ocur(), d.num_shift_dupl()) == counts
    assert d.first_roots().tolist() == first
    assert d.shift_roots().tolist() == shift


def test_leaf_runs_against_heap_walk():
    for n in list(range(1, 70)) + [127, 128, 129, 1000]:
        for node in range(2 * n - 1):
            runs = O.leaf_runs(node, n)
            got = [s + j for s, c in runs for j in range(c)]
            assert got == heap_leaves(node, n), (n, node)


def test_reference_leaf_helpers_agree_on_single_level_nodes():
    """kokkos_merkle_tree.hpp:102-128 are exact whenever a node's leaves share a depth."""
    for n in [4, 8, 16, 5, 6, 7, 11, 100]:
        nodes = 2 * n - 1
        for node in range(nodes):
            leaves = heap_leaves(node, n)
            depths = {int(np.floor(np.log2(l + n))) for l in leaves}
            if len(depths) == 1:
                lm = O.ref_leftmost_leaf(node, nodes) - (n - 1)
                cnt = O.ref_num_leaf_descendents(node, nodes)
                assert list(range(lm, lm + cnt)) == leaves, (n, node)


def chunks_of(buf: bytes, C: int):
    return [buf[i:i + C] for i in range(0, len(buf), C)]


def check_diff(diff: bytes, history, k: int, ref: int, C: int):
    """Definition-level checks of diff k given the buffers of checkpoints ref..k."""
    p = O.parse_diff(diff)
    h = p["header"]
    cur = chunks_of(history[k], C)
    n = len(cur)
    assert (h["ref_id"], h["chkpt_id"], h["datalen"], h["chunk_size"]) == (ref, k, len(history[k]), C)
    first = p["first"].tolist()
    assert first == sorted(first) and len(set(first)) == len(first)
    table = p["table"].tolist()
    assert [c for c, _ in table] == sorted({c for c, _ in table}) and all(m > 0 for _, m in table)
    assert sum(m for _, m in table) == h["num_shift_dupl"]
    srcs = [c for c, m in table for _ in range(m)]
    shift = p["shift"].tolist()
    for a, b in zip(range(len(shift)), range(1, len(shift))):
        if srcs[a] == srcs[b]:
            assert shift[a][0] < shift[b][0]
    # fixed leaves: same bytes (digest) as at the previous checkpoint
    prev = chunks_of(history[k - 1], C) if k > ref else None
    fixed = [prev is not None and O.hash_bytes(0, cur[c]) == O.hash_bytes(0, prev[c]) for c in range(n)]
    seen = set()
    for j in range(ref, k):
        seen.update(chunks_of(history[j], C))
    covered = []
    for r in first:
        leaves = heap_leaves(r, n)
        covered += leaves
        for c in leaves:  # first occurrence: new to the baseline's history, lowest offset
            assert not fixed[c] and cur[c] not in seen, (k, r, c)
            assert cur[c] not in cur[:c], (k, r, c)
    data = p["data"]
    pos = 0
    for r in first:
        for c in heap_leaves(r, n):
            assert data[pos * C:pos * C + len(cur[c])] == cur[c]
            pos += 1
    assert len(data) == pos * C
    for (node, prv), src in zip(shift, srcs):
        assert ref <= src <= k
        a, b = heap_leaves(node, n), heap_leaves(prv, n)
        covered += a
        old = chunks_of(history[src], C)
        assert len(a) == len(b) and [cur[x] for x in a] == [old[y] for y in b], (k, node, prv, src)
    assert sorted(covered) == sorted(set(covered)), "regions overlap"
    assert sorted(covered) == [c for c in range(n) if not fixed[c]], "regions do not cover the changed leaves"


def run_history(bufs, C, baselines=(0,), compat=False):
    d = O.OracleDedup(C, compat_ref=compat)
    diffs, ref = [], 0
    for k, b in enumerate(bufs):
        if k in baselines:
            ref = k
        diffs.append(d.checkpoint(np.frombuffer(b, np.uint8), make_baseline=(k in baselines)))
        if not compat:
            check_diff(diffs[-1], bufs, k, ref, C)
            assert O.restart(diffs, k).tobytes() == b
    return diffs


def test_brute_force_all_length4_pairs():
    alphabet = b"ABC"
    words = [bytes(w) for w in itertools.product(alphabet, repeat=4)]
    for a in words:
        for b in words:
            run_history([a, b], 1)


@pytest.mark.parametrize("L", [3, 5, 6, 7, 9])
def test_brute_force_random_ragged(L):
    """Non-power-of-two chunk counts (leaves on two depths), three checkpoints."""
    rng = np.random.default_rng(L)
    for _ in range(400):
        bufs = [bytes(rng.choice(list(b"ABC"), L).tolist()) for _ in range(3)]
        run_history(bufs, 1)


def test_random_series_with_rebaseline():
    rng = np.random.default_rng(9)
    C, n = 16, 37
    cur = rng.integers(0, 4, n * C, dtype=np.uint8)  # small alphabet: many repeats
    bufs = []
    for k in range(8):
        idx = rng.integers(0, n, 5)
        for i in idx:
            cur[i * C:(i + 1) * C] = rng.integers(0, 4, C, dtype=np.uint8) if k % 2 else cur[((i + 3) % n) * C:((i + 3) % n + 1) * C]
        bufs.append(cur.tobytes())
    run_history(bufs, C, baselines=(0, 4))


def test_root_shift_departure():
    """A, B, A: node 0 is a shifted duplicate at checkpoint 2.  The reference never
    emits node 0 as a shifted root (reference_impl.hpp:105-106), so its diff restarts
    to B; the oracle emits it and restarts to A."""
    A, B = b"ABCDEFGH", b"IJKLMNOP"
    d = run_history([A, B, A], 1)
    p = O.parse_diff(d[2])
    assert p["first"].tolist() == [] and p["shift"].tolist() == [[0, 0]]
    assert p["table"].tolist() == [[0, 1]]
    c = run_history([A, B, A], 1, compat=True)
    assert O.parse_header(c[2])["num_shift_dupl"] == 0
    assert O.restart(c, 2).tobytes() == B  # what the reference's own diff restores


def test_mixed_level_departure():
    """n = 5 chunks: node 1 spans both leaf depths (chunks 0, 3, 4).  The reference's
    leaf-range helpers drop chunk 0 from its region; the oracle stores all three."""
    base = b"ABCDE"
    d = run_history([base, b"VWXYZ"], 1)
    c = run_history([base, b"VWXYZ"], 1, compat=True)
    pd, pc = O.parse_diff(d[1]), O.parse_diff(c[1])
    assert pd["first"].tolist() == pc["first"].tolist() == [0]
    assert pd["data"] == b"VWXYZ"
    # reference: num_leaf_descendents(0, 9) = 2 and leftmost_leaf(0, 9) = 7 (chunk 3)
    assert pc["data"] == b"YZ"


def test_compat_identical_on_power_of_two():
    rng = np.random.default_rng(4)
    for _ in range(100):
        bufs = [bytes(rng.choice(list(b"ABCD"), 16).tolist()) for _ in range(3)]
        if bufs[2] == bufs[0] and bufs[1] != bufs[0]:
            continue  # node 0 shifted: the one place the modes differ
        assert run_history(bufs, 1) == run_history(bufs, 1, compat=True)


def test_checker_catches_planted_mistakes():
    """The definition-level checker must reject a diff with a planted mistake."""
    bufs = [b"ABCDEFGH", b"IJABCDEK"]  # test_case_04
    d = run_history(bufs, 1)
    good = bytearray(d[1])
    p = O.parse_diff(bytes(good))
    F, P, S = (p["header"][k] for k in ("num_first_ocur", "num_prior_chkpts", "num_shift_dupl"))
    data_off = 32 + 4 * F + 8 * P + 8 * S
    mutants = []
    m = bytearray(good); m[data_off] ^= 1; mutants.append(m)                      # wrong data byte
    m = bytearray(good); m[32:36] = (4).to_bytes(4, "little"); mutants.append(m)   # wrong first root
    sh = 32 + 4 * F + 8 * P
    m = bytearray(good); m[sh + 4:sh + 8] = (9).to_bytes(4, "little"); mutants.append(m)  # wrong prev node
    for m in mutants:
        with pytest.raises(AssertionError):
            check_diff(bytes(m), bufs, 1, 0, 1)
