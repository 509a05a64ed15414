"""Shared helpers for the parity tests (CPU oracle vs. HIP library)."""
import numpy as np

from oracle import oracle as O


def describe_diff(d: bytes) -> str:
    p = O.parse_diff(d)
    h = p["header"]
    return (f"hdr={h} first={p['first'][:8].tolist()}... table={p['table'].tolist()[:8]} "
            f"shift={p['shift'][:4].tolist()}... data={len(p['data'])}B")


def first_mismatch(a: bytes, b: bytes) -> str:
    if len(a) != len(b):
        return f"sizes differ: {len(a)} vs {len(b)}; {describe_diff(a)} || {describe_diff(b)}"
    x = np.frombuffer(a, np.uint8)
    y = np.frombuffer(b, np.uint8)
    idx = np.nonzero(x != y)[0]
    if idx.size == 0:
        return "identical"
    return f"first differing byte {idx[0]} of {len(a)}; {describe_diff(a)} || {describe_diff(b)}"


REFERENCE_CASES = {
    # test/test_case_XX.cpp: step0 -> step1, expected (num_first_ocur, num_shift_dupl) at :100/:101,
    # and the region roots the serial rules give (first roots, shifted roots).
    "01": ("ABCDEFGH", "IJKLMNOP", (1, 0), [0], []),
    "02": ("ABCDEFGH", "ABCDEFGH", (0, 0), [], []),
    "03": ("ABCDEFGH", "ABCDEIJK", (2, 0), [6, 12], []),
    "04": ("ABCDEFGH", "IJABCDEK", (2, 3), [3, 14], [4, 5, 13]),
    "05": ("ABCDEFGH", "ABCIABCJ", (2, 2), [10, 14], [5, 13]),
    "06": ("ABCDEFGH", "EFGBCDIJ", (1, 4), [6], [3, 5, 9, 10]),
    "07": ("ABCDEFGH", "IJKLABIH", (1, 2), [1], [5, 13]),
    "08": ("MNXYOPQR", "IJXYIJAB", (2, 1), [3, 6], [5]),
    "09": ("MNXYOPQR", "IJXYIJIB", (2, 2), [3, 14], [5, 13]),
    "10": ("MNXYOPQR", "IJXYIJJI", (1, 3), [3], [5, 13, 14]),
}
