"""The oracle reproduces the committed golden fixtures (tests/golden/golden.json) -- CPU."""
import hashlib
import json
import os

import numpy as np
import pytest

from mtcheck_amd.workload import CheckpointSeries
from oracle import oracle as O
from tests.helpers import REFERENCE_CASES

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def test_golden_hashes():
    for e in GOLDEN["hash"]:
        b = bytes.fromhex(e["input_hex"])
        assert O.murmur3_mtcheck(b).hex() == e["murmur3_mtcheck"]
        assert O.md5(b).hex() == e["md5"]


@pytest.mark.parametrize("case", sorted(REFERENCE_CASES))
def test_golden_reference_case_diffs(case):
    s0, s1, _, _, _ = REFERENCE_CASES[case]
    d = O.OracleDedup(1)
    got = [d.checkpoint(s0.encode(), True).hex(), d.checkpoint(s1.encode()).hex()]
    assert got == GOLDEN["reference_cases"][case]


@pytest.mark.parametrize("name", sorted(GOLDEN["series"]))
def test_golden_series(name):
    g = GOLDEN["series"][name]
    s = CheckpointSeries(g["nbytes"], g["chunk"], seed=g["seed"])
    d = O.OracleDedup(g["chunk"], g["hash"])
    buf = np.zeros(g["nbytes"], np.uint8)
    for k, rec in enumerate(g["steps"]):
        s.apply_host(buf, k)
        assert hashlib.sha256(buf.tobytes()).hexdigest() == rec["buffer_sha256"]
        diff = d.checkpoint(buf, make_baseline=(k == 0))
        assert O.parse_header(diff) == rec["header"]
        assert hashlib.sha256(diff).hexdigest() == rec["sha256"]
