"""Writes tests/golden/golden.json from the CPU oracle.

The reference stores no digests, diffs or vectors (its tests only assert region
counts and restart equality), so these fixtures are produced by the oracle, whose
hashes and rules are pinned separately (tests/test_oracle_*.py).  They freeze its
output so that later changes to the oracle or to the HIP library are compared to
the same bytes.  Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from mtcheck_amd.workload import CheckpointSeries  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.helpers import REFERENCE_CASES  # noqa: E402

HASH_INPUTS = [b"", b"a", b"abc", b"ABCDEFGH", b"message digest", bytes(range(16)), bytes(range(17)),
               bytes(range(128)), bytes(range(256)) * 16, b"\x00" * 128]

SERIES = [  # (name, nbytes, chunk, seed, steps, hash)
    ("c0_1mib_128", 1 << 20, 128, 0x4D54434B, 5, O.HASH_MURMUR3),
    ("ragged_1000003_128", 1_000_003, 128, 7, 4, O.HASH_MURMUR3),
    ("md5_256k_64", 1 << 18, 64, 11, 4, O.HASH_MD5),
]


def main():
    out = {"hash": [], "reference_cases": {}, "series": {}}
    for b in HASH_INPUTS:
        out["hash"].append({"input_hex": b.hex(), "murmur3_mtcheck": O.murmur3_mtcheck(b).hex(),
                            "md5": O.md5(b).hex()})
    for case, (s0, s1, _, _, _) in sorted(REFERENCE_CASES.items()):
        d = O.OracleDedup(1)
        out["reference_cases"][case] = [d.checkpoint(s0.encode(), True).hex(), d.checkpoint(s1.encode()).hex()]
    for name, nbytes, C, seed, steps, hk in SERIES:
        s = CheckpointSeries(nbytes, C, seed=seed)
        d = O.OracleDedup(C, hk)
        buf = np.zeros(nbytes, np.uint8)
        rec = []
        for k in range(steps):
            s.apply_host(buf, k)
            diff = d.checkpoint(buf, make_baseline=(k == 0))
            rec.append({"header": O.parse_header(diff), "sha256": hashlib.sha256(diff).hexdigest(),
                        "buffer_sha256": hashlib.sha256(buf.tobytes()).hexdigest()})
        out["series"][name] = {"nbytes": nbytes, "chunk": C, "seed": seed, "hash": hk, "steps": rec}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
