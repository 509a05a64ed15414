"""The basic approach (src/basic_approach.cpp) on the device, through the C ABI.

mtc_restart_basic is checked against the oracle's basic diffs: every checkpoint
of a series must restart, from host and from device diffs, to the buffer it was
taken of.  Basic checkpoints are checked byte for byte against the oracle.  Their
first device runs faulted (a baseline's digest-table reset went through the null
table pointer of a basic handle); tests/test_host_launches.py shows that launch on
the CPU and checks every launch of basic steps there.  Every test prints its case
first.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import first_mismatch

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


SERIES = [
    (1 << 20, 128, O.HASH_MURMUR3, (0,)),
    (1 << 20, 64, O.HASH_MD5, (0, 3)),
    (300_007, 100, O.HASH_MURMUR3, (0, 2)),
    (4 << 20, 512, O.HASH_MURMUR3, (0,)),
    (70_001, 8, O.HASH_MD5, (0,)),
    (5 * 128 + 1, 128, O.HASH_MURMUR3, (0, 4)),
]


def oracle_series(nbytes, C, hk, base_at, steps=6):
    series = CheckpointSeries(nbytes, C, seed=5, fresh_frac=0.04, copy_frac=0.03)
    orc = O.OracleBasic(C, hk)
    host = np.zeros(nbytes, np.uint8)
    diffs, versions = [], []
    for k in range(steps):
        series.apply_host(host, k)
        diffs.append(orc.checkpoint(host, make_baseline=k in base_at))
        versions.append(host.copy())
    return diffs, versions


@pytest.mark.parametrize("nbytes,C,hk,base_at", SERIES)
def test_basic_restart_from_oracle_diffs(nbytes, C, hk, base_at):
    say(f"basic restart nbytes={nbytes} C={C} hk={hk} baselines={base_at}")
    diffs, versions = oracle_series(nbytes, C, hk, base_at)
    dev_diffs = [torch.frombuffer(bytearray(d), dtype=torch.uint8).to("cuda:0") for d in diffs]
    out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
    for k in range(len(diffs)):
        out.zero_()
        M.restart_into(out, diffs, k, approach=M.BASIC)
        assert np.array_equal(out.cpu().numpy(), versions[k]), ("host diffs", k)
        out.zero_()
        M.restart_into(out, dev_diffs, k, approach=M.BASIC)
        assert np.array_equal(out.cpu().numpy(), versions[k]), ("device diffs", k)


def test_basic_restart_rejects_out_of_range_ids():
    """A chunk index >= n is caught by the restart's own bounds check before any copy."""
    C, nbytes = 64, 64 * 100
    say("basic restart of corrupt chunk ids")
    diffs, _ = oracle_series(nbytes, C, O.HASH_MURMUR3, (0,), steps=2)
    out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
    for bad in (100, 0x7FFFFFFF):
        d = bytearray(diffs[0])
        d[32:36] = np.uint32(bad).tobytes()
        with pytest.raises(M.MtcError):
            M.restart_into(out, [bytes(d)], 0, approach=M.BASIC)


def test_basic_restart_rejects_tree_diffs():
    """A tree diff with shifted duplicates is refused by the basic restart (host-side header check)."""
    C = 16
    say("basic restart of a tree diff")
    a = np.random.default_rng(4).integers(0, 256, 64 * C, dtype=np.uint8)
    b = a.copy()
    b[3 * C:4 * C] = a[10 * C:11 * C]
    t = M.TreeDeduplicator(C)
    d0 = t.checkpoint(torch.from_numpy(a).to("cuda:0"), True)
    d1 = t.checkpoint(torch.from_numpy(b).to("cuda:0"))
    assert O.parse_header(d1)["num_shift_dupl"] > 0
    out = torch.zeros(a.size, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(M.MtcError):
        M.restart_into(out, [d0, d1], 1, approach=M.BASIC)


@pytest.mark.parametrize("nbytes,C,hk,base_at", SERIES)
def test_basic_series(nbytes, C, hk, base_at):
    say(f"basic series nbytes={nbytes} C={C} hk={hk} baselines={base_at}")
    series = CheckpointSeries(nbytes, C, seed=5, fresh_frac=0.04, copy_frac=0.03)
    d = M.TreeDeduplicator(C, approach=M.BASIC, hash_kind=hk)
    orc = O.OracleBasic(C, hk)
    buf = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
    host = np.zeros(nbytes, np.uint8)
    diffs, versions, dev_diffs = [], [], []
    for k in range(6):
        series.apply_host(host, k)
        buf.copy_(torch.from_numpy(host))
        base = k in base_at
        diff = d.checkpoint(buf, make_baseline=base)
        exp = orc.checkpoint(host, make_baseline=base)
        assert diff == exp, (k, first_mismatch(diff, exp))
        diffs.append(diff)
        dev_diffs.append(torch.frombuffer(bytearray(diff), dtype=torch.uint8).to("cuda:0"))
        versions.append(host.copy())
    for k in range(6):
        out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
        d.restart(out, diffs, k)
        assert np.array_equal(out.cpu().numpy(), versions[k]), ("host diffs", k)
        out.zero_()
        d.restart(out, dev_diffs, k)
        assert np.array_equal(out.cpu().numpy(), versions[k]), ("device diffs", k)


def test_basic_rewrite_with_same_bytes_is_not_listed():
    """A chunk rewritten with its previous bytes has the previous digest: not listed."""
    C = 64
    say("basic same-bytes rewrite")
    a = np.random.default_rng(2).integers(0, 256, 64 * C, dtype=np.uint8)
    d = M.TreeDeduplicator(C, approach=M.BASIC)
    orc = O.OracleBasic(C)
    buf = torch.from_numpy(a.copy()).to("cuda:0")
    assert d.checkpoint(buf, True) == orc.checkpoint(a, True)
    b = a.copy()
    b[5 * C:6 * C] = a[9 * C:10 * C]   # changed
    b[7 * C:8 * C] = a[7 * C:8 * C]    # same bytes
    buf.copy_(torch.from_numpy(b))
    diff = d.checkpoint(buf)
    assert diff == orc.checkpoint(b)
    h = O.parse_header(diff)
    assert h["num_first_ocur"] == 1
    assert np.frombuffer(diff[32:36], np.uint32)[0] == 5
