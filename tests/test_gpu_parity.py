"""Parity of the HIP deduplicator with the CPU oracle (bit-exact), through the C ABI.

Every test prints the case it runs before touching the GPU.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import REFERENCE_CASES, first_mismatch

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(a: np.ndarray):
    return torch.from_numpy(np.array(a, dtype=np.uint8, copy=True)).to("cuda:0")


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("hk", [M.MURMUR3, M.MD5])
def test_chunk_hash_all_lengths(hk):
    """Leaf digests for every chunk length 1..300 and short tails, both hashes."""
    rng = np.random.default_rng(7)
    for C in [1, 7, 16, 48, 64, 100, 128, 256, 300]:
        for L in sorted({C, 3 * C, 5 * C - 1, 4096 + 13}):
            say(f"hash hk={hk} C={C} L={L}")
            a = rng.integers(0, 256, L, dtype=np.uint8)
            got = M.hash_chunks(dev(a), C, hk).cpu().numpy().reshape(-1, 16)
            n = (L + C - 1) // C
            for c in range(n):
                exp = O.hash_bytes(hk, a[c * C:min((c + 1) * C, L)])
                assert got[c].tobytes() == exp, (C, L, c)


def test_chunk_hash_every_length_to_4096():
    """One chunk of every length 1..4096 (generic and 16-byte paths)."""
    rng = np.random.default_rng(11)
    a = rng.integers(0, 256, 4096, dtype=np.uint8)
    d = dev(a)
    for hk in (M.MURMUR3, M.MD5):
        say(f"hash single chunk lengths hk={hk}")
        for L in list(range(1, 130)) + [255, 256, 511, 512, 1000, 1024, 2048, 4095, 4096]:
            got = M.hash_chunks(d[:L], L, hk).cpu().numpy().tobytes()
            assert got == O.hash_bytes(hk, a[:L]), (hk, L)


@pytest.mark.parametrize("case", sorted(REFERENCE_CASES))
def test_reference_case(case):
    """test/test_case_XX.cpp: 8 one-byte chunks, baseline then one incremental checkpoint."""
    s0, s1, counts, first, shift = REFERENCE_CASES[case]
    say(f"reference case {case}")
    d = M.TreeDeduplicator(1)
    orc = O.OracleDedup(1)
    diffs = []
    for step, s in enumerate((s0, s1)):
        data = np.frombuffer(s.encode(), np.uint8)
        diff = d.checkpoint(dev(data), make_baseline=(step == 0))
        exp = orc.checkpoint(data, make_baseline=(step == 0))
        assert diff == exp, first_mismatch(diff, exp)
        diffs.append(diff)
        out = torch.zeros(8, dtype=torch.uint8, device="cuda:0")
        d.restart(out, diffs, step)
        assert bytes(out.cpu().numpy()) == s.encode()
    assert (d.num_first_ocur(), d.num_shift_dupl()) == counts
    p = O.parse_diff(diffs[1])
    assert p["first"].tolist() == first
    assert sorted(p["shift"][:, 0].tolist()) == shift


def run_series(nbytes, C, steps, hk=M.MURMUR3, check_tree=True, seed=3, tuning=None, handles=None, **kw):
    series = CheckpointSeries(nbytes, C, seed=seed, **kw)
    d = M.TreeDeduplicator(C, hash_kind=hk)
    if handles is not None:
        handles.append(d)
    if tuning:
        d.set_tuning(*tuning)
    d.set_debug(check_tree)
    orc = O.OracleDedup(C, hk)
    buf = np.zeros(nbytes, np.uint8)
    diffs = []
    for k in range(steps):
        series.apply_host(buf, k)
        say(f"series nbytes={nbytes} C={C} hk={hk} step={k}")
        diff = d.checkpoint(dev(buf), make_baseline=(k == 0))
        exp = orc.checkpoint(buf, make_baseline=(k == 0))
        assert diff == exp, f"step {k}: " + first_mismatch(diff, exp)
        if check_tree:
            assert np.array_equal(M.debug_labels(d), orc.labels()), f"labels differ at step {k}"
            assert np.array_equal(M.debug_tree(d), orc.tree()), f"tree differs at step {k}"
        diffs.append(diff)
        out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
        d.restart(out, diffs, k)
        assert np.array_equal(out.cpu().numpy(), buf), f"restart of step {k} differs"
    return diffs


@pytest.mark.parametrize("nbytes,C", [(1 << 20, 128), (1 << 20, 64), (1 << 19, 512), (1 << 21, 4096)])
def test_series_power_of_two(nbytes, C):
    run_series(nbytes, C, 5)


@pytest.mark.parametrize("nbytes,C", [(1_000_003, 128), (777_777, 64), (5 * 128 + 1, 128), (3 * 16, 16),
                                      (123_457, 100)])
def test_series_ragged(nbytes, C):
    """Chunk counts that are not powers of two, short last chunks, odd chunk sizes."""
    run_series(nbytes, C, 4)


def test_series_md5():
    run_series(1 << 20, 128, 4, hk=M.MD5)


def test_series_heavy_change():
    """20% change per step, many shifted duplicates and intra-checkpoint repeats."""
    run_series(1 << 20, 64, 4, fresh_frac=0.13, copy_frac=0.07)


def test_repeated_content_and_root_shift():
    """Whole-buffer repeats: A, B, A (node 0 a shifted duplicate), constant buffers."""
    rng = np.random.default_rng(5)
    C = 64
    A = rng.integers(0, 256, 64 * C, dtype=np.uint8)
    B = rng.integers(0, 256, 64 * C, dtype=np.uint8)
    Z = np.zeros(64 * C, np.uint8)
    XY = np.tile(rng.integers(0, 256, 2 * C, dtype=np.uint8), 32)
    for seq in ([A, B, A], [Z, Z, A, Z], [XY, A, XY], [A, A, A]):
        say(f"repeat sequence len={len(seq)}")
        d = M.TreeDeduplicator(C)
        d.set_debug(True)
        orc = O.OracleDedup(C)
        diffs = []
        for k, buf in enumerate(seq):
            diff = d.checkpoint(dev(buf), make_baseline=(k == 0))
            exp = orc.checkpoint(buf, make_baseline=(k == 0))
            assert diff == exp, first_mismatch(diff, exp)
            assert np.array_equal(M.debug_labels(d), orc.labels())
            diffs.append(diff)
            out = torch.zeros(buf.size, dtype=torch.uint8, device="cuda:0")
            d.restart(out, diffs, k)
            assert np.array_equal(out.cpu().numpy(), buf)


def test_rebaseline_and_device_chain():
    """A second baseline mid-series; restart from device-resident diffs."""
    C, nbytes = 128, 1 << 18
    series = CheckpointSeries(nbytes, C, seed=9)
    d = M.TreeDeduplicator(C)
    orc = O.OracleDedup(C)
    buf = np.zeros(nbytes, np.uint8)
    diffs, dev_diffs = [], []
    for k in range(6):
        series.apply_host(buf, k)
        say(f"rebaseline step {k}")
        base = k in (0, 3)
        diff = d.checkpoint(dev(buf), make_baseline=base)
        exp = orc.checkpoint(buf, make_baseline=base)
        assert diff == exp, first_mismatch(diff, exp)
        assert M.parse_header(diff)["ref_id"] == (0 if k < 3 else 3)
        diffs.append(diff)
        dev_diffs.append(dev(np.frombuffer(diff, np.uint8)))
        out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
        M.restart_into(out, dev_diffs, k)
        assert np.array_equal(out.cpu().numpy(), buf)


def test_config0_sequence():
    """64 MiB, 128 B chunks, 10 checkpoints of the config-0 generator: all diffs
    bit-exact with the oracle, restart of checkpoint 9 byte-identical."""
    nbytes, C = 64 << 20, 128
    series = CheckpointSeries(nbytes, C)
    d = M.TreeDeduplicator(C)
    orc = O.OracleDedup(C)
    gbuf = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    diffs = []
    for k in range(10):
        say(f"config0 step {k}")
        series.apply_device(gbuf, k)
        host = gbuf.cpu().numpy()
        diff = d.checkpoint(gbuf, make_baseline=(k == 0))
        exp = orc.checkpoint(host, make_baseline=(k == 0))
        assert diff == exp, f"step {k}: " + first_mismatch(diff, exp)
        diffs.append(diff)
    out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
    d.restart(out, diffs, 9)
    assert torch.equal(out, gbuf)
    assert hashlib.sha256(out.cpu().numpy().tobytes()).digest() == hashlib.sha256(host.tobytes()).digest()


def test_device_generator_matches_host():
    for nbytes, C in [(1 << 20, 128), (100_003, 64)]:
        series = CheckpointSeries(nbytes, C, seed=21)
        g = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
        h = np.zeros(nbytes, np.uint8)
        for k in range(3):
            say(f"generator nbytes={nbytes} step {k}")
            series.apply_device(g, k)
            series.apply_host(h, k)
            assert np.array_equal(g.cpu().numpy(), h)


def test_errors_are_reported():
    say("error paths")
    with pytest.raises(M.MtcError) as e:
        M.TreeDeduplicator(128, approach=M.FULL)
    assert e.value.code == 2
    with pytest.raises(M.MtcError):
        M.TreeDeduplicator(0)
    d = M.TreeDeduplicator(64)
    a = torch.zeros(4096, dtype=torch.uint8, device="cuda:0")
    d.checkpoint(a, make_baseline=True)
    with pytest.raises(M.MtcError) as e:
        d.checkpoint(torch.zeros(2048, dtype=torch.uint8, device="cuda:0"))
    assert e.value.code == 7
    good = M.TreeDeduplicator(64).checkpoint(a, make_baseline=True)
    out = torch.zeros(4096, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(M.MtcError) as e:  # truncated diff
        M.restart_into(out, [good[:40]], 0)
    assert e.value.code == 5
    bad = bytearray(good)  # a first-occurrence root beyond the tree
    bad[32:36] = (10**6).to_bytes(4, "little")
    with pytest.raises(M.MtcError) as e:
        M.restart_into(out, [bytes(bad)], 0)
    assert e.value.code == 5
    with pytest.raises(M.MtcError) as e:  # output too small
        M.restart_into(torch.zeros(100, dtype=torch.uint8, device="cuda:0"), [good], 0)
    assert e.value.code == 8


def test_golden_fixtures():
    """The library's diffs match the committed golden fixtures byte for byte
    (sha256 of each diff), independently of the oracle at run time."""
    import json
    import os
    g_all = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
    for case, exp in sorted(g_all["reference_cases"].items()):
        s0, s1 = REFERENCE_CASES[case][:2]
        say(f"golden reference case {case}")
        d = M.TreeDeduplicator(1)
        got = [d.checkpoint(dev(np.frombuffer(s.encode(), np.uint8)), make_baseline=(i == 0)).hex()
               for i, s in enumerate((s0, s1))]
        assert got == exp, case
    for name, g in sorted(g_all["series"].items()):
        series = CheckpointSeries(g["nbytes"], g["chunk"], seed=g["seed"])
        d = M.TreeDeduplicator(g["chunk"], hash_kind=g["hash"])
        gbuf = torch.empty(g["nbytes"], dtype=torch.uint8, device="cuda:0")
        for k, rec in enumerate(g["steps"]):
            say(f"golden series {name} step {k}")
            series.apply_device(gbuf, k)
            diff = d.checkpoint(gbuf, make_baseline=(k == 0))
            assert M.parse_header(diff) == rec["header"], (name, k)
            assert hashlib.sha256(diff).hexdigest() == rec["sha256"], (name, k)


@pytest.mark.parametrize("path,env", [("windowed", {}), ("leftovers", {"MTC_REHASH_PROBE_LIMIT": "1"}),
                                      ("atomic", {"MTC_REHASH": "atomic"})])
def test_table_growth_during_long_heavy_series(monkeypatch, path, env):
    """12 checkpoints at 30% change: the table outgrows its baseline capacity, so
    checkpoints pause after the leaf kernel for the exact bound and rehash; every
    diff stays bit-exact and restarts.  Each rehash path is run: windowed (the
    default), windowed with most entries left for the atomic pass, and atomic only."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    hs = []
    diffs = run_series(1 << 19, 64, 12, fresh_frac=0.2, copy_frac=0.1, tuning=(1, 0.5, 8), handles=hs)
    assert len(diffs) == 12
    assert hs[0].stats().table_capacity >= 4 * ((1 << 19) // 64)  # grew past the baseline capacity


def test_table_growth_in_batch():
    """The same growth inside a batched step, next to members that need none."""
    # the last member's table starts at the minimum 1024 slots: one rehash window
    sizes = [(1 << 19, 64, 0.3), (1 << 20, 128, 0.01), (300_001, 64, 0.25), (30_000, 64, 0.3)]
    series = [CheckpointSeries.with_change(nb, C, ch, seed=70 + i) for i, (nb, C, ch) in enumerate(sizes)]
    bufs = [torch.empty(nb, dtype=torch.uint8, device="cuda:0") for nb, _, _ in sizes]
    mates = [M.TreeDeduplicator(C) for _, C, _ in sizes]
    for m in mates:
        m.set_tuning(1, 0.5, 8)  # a small table, so it must grow
    orc = [O.OracleDedup(C) for _, C, _ in sizes]
    batch = M.Batch()
    for k in range(10):
        for s, b in zip(series, bufs):
            s.apply_device(b, k)
        say(f"batched growth step {k}")
        got = batch.checkpoint(mates, bufs, make_baseline=(k == 0), copy=True)
        for i, g in enumerate(got):
            assert g == orc[i].checkpoint(bufs[i].cpu().numpy(), make_baseline=(k == 0)), (k, i)
    caps = [m.stats().table_capacity for m in mates]
    assert caps[0] > 2 * (sizes[0][0] // 64)  # grew past the baseline capacity
    assert caps[3] > 1024
