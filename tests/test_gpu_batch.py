"""Batched checkpoints (mtc_checkpoint_batch): every member's diff is identical to
the same deduplicator checkpointed on its own, with graph replay on and off."""
import numpy as np
import pytest

from oracle import oracle as O

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu

# (bytes, chunk, hash): ragged sizes, a single chunk, an unaligned chunk size, both hashes
MEMBERS = [(1 << 20, 128, M.MURMUR3), (300_007, 64, M.MURMUR3), (2_000_000, 512, M.MD5), (77, 128, M.MURMUR3),
           (65_536, 100, M.MURMUR3), (5 << 20, 128, M.MURMUR3), (123_457, 16, M.MD5), (4096, 4096, M.MURMUR3)]


def say(msg):
    print(msg, flush=True)


def run_series(use_graph, steps=5, rebase_at=3, tuning=None):
    series = [CheckpointSeries(nb, C, seed=100 + i) for i, (nb, C, _) in enumerate(MEMBERS)]
    bufs = [torch.empty(nb, dtype=torch.uint8, device="cuda:0") for nb, _, _ in MEMBERS]
    solo = [M.TreeDeduplicator(C, hash_kind=hk) for _, C, hk in MEMBERS]
    mates = [M.TreeDeduplicator(C, hash_kind=hk) for _, C, hk in MEMBERS]
    if tuning:
        for d in mates:
            d.set_tuning(*tuning[:3])
            if len(tuning) > 3:
                d.set_sweep_tiles(*tuning[3:])
    batch = M.Batch(use_graph=use_graph)
    out = []
    for k in range(steps):
        for s, b in zip(series, bufs):
            s.apply_device(b, k)
        torch.cuda.synchronize()
        base = [k == 0 or (k == rebase_at and i % 3 == 0) for i in range(len(MEMBERS))]
        say(f"batch graph={use_graph} tuning={tuning} step {k}")
        got = batch.checkpoint(mates, bufs, make_baseline=base, copy=True)
        exp = [d.checkpoint(b, make_baseline=bl) for d, b, bl in zip(solo, bufs, base)]
        for i, (g, e) in enumerate(zip(got, exp)):
            assert g == e, f"step {k} member {i} {MEMBERS[i]}"
        out.append(got)
    return out, bufs


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_single(use_graph):
    run_series(use_graph)


def test_batch_tuning_does_not_change_results():
    a, _ = run_series(True, steps=3, tuning=(4, 0.4, 6, 1, 64))
    b, _ = run_series(False, steps=3, tuning=(2, 0.15, 8, 3, 5))
    assert a == b


def test_batch_against_oracle_and_restart():
    """A smaller mixed batch checked against the CPU oracle, and restarts of the
    batched diffs."""
    mem = [(200_003, 128, M.MURMUR3), (50_000, 64, M.MD5), (8192, 512, M.MURMUR3)]
    series = [CheckpointSeries(nb, C, seed=7 + i) for i, (nb, C, _) in enumerate(mem)]
    bufs = [torch.empty(nb, dtype=torch.uint8, device="cuda:0") for nb, _, _ in mem]
    ds = [M.TreeDeduplicator(C, hash_kind=hk) for _, C, hk in mem]
    orc = [O.OracleDedup(C, hk) for _, C, hk in mem]
    batch = M.Batch()
    chains = [[] for _ in mem]
    for k in range(4):
        for s, b in zip(series, bufs):
            s.apply_device(b, k)
        torch.cuda.synchronize()
        say(f"batch vs oracle step {k}")
        got = batch.checkpoint(ds, bufs, make_baseline=(k == 0), copy=True)
        for i, g in enumerate(got):
            assert g == orc[i].checkpoint(bufs[i].cpu().numpy(), make_baseline=(k == 0)), (k, i)
            chains[i].append(g)
    for i, b in enumerate(bufs):
        out = torch.zeros_like(b)
        M.restart_into(out, chains[i], 3)
        assert torch.equal(out, b)


def test_batch_rejects_duplicate_handle():
    d = M.TreeDeduplicator(64)
    a = torch.zeros(4096, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(M.MtcError):
        M.Batch().checkpoint([d, d], [a, a], make_baseline=True)


def test_failed_batch_step_restarts_every_member_from_a_baseline(monkeypatch):
    """A member whose step reports an error (forced with MTC_TEST_FAIL_MEMBER) fails
    the whole batch; every member's next checkpoint is a new baseline, equal to the
    oracle's, and restarts exactly."""
    mem = [(150_001, 128, M.MURMUR3), (64 * 1000, 64, M.MURMUR3), (40_000, 512, M.MD5)]
    series = [CheckpointSeries(nb, C, seed=40 + i) for i, (nb, C, _) in enumerate(mem)]
    bufs = [torch.empty(nb, dtype=torch.uint8, device="cuda:0") for nb, _, _ in mem]
    ds = [M.TreeDeduplicator(C, hash_kind=hk) for _, C, hk in mem]
    orc = [O.OracleDedup(C, hk) for _, C, hk in mem]
    batch = M.Batch()
    for k in range(2):
        for s, b in zip(series, bufs):
            s.apply_device(b, k)
        torch.cuda.synchronize()
        say(f"failure injection: step {k}")
        got = batch.checkpoint(ds, bufs, make_baseline=(k == 0), copy=True)
        for i, g in enumerate(got):
            assert g == orc[i].checkpoint(bufs[i].cpu().numpy(), make_baseline=(k == 0)), (k, i)
    for s, b in zip(series, bufs):
        s.apply_device(b, 2)
    torch.cuda.synchronize()
    say("failure injection: failing step (member 1)")
    monkeypatch.setenv("MTC_TEST_FAIL_MEMBER", "1")
    with pytest.raises(M.MtcError):
        batch.checkpoint(ds, bufs)
    monkeypatch.delenv("MTC_TEST_FAIL_MEMBER")
    say("failure injection: retried step")
    got = batch.checkpoint(ds, bufs, copy=True)
    for i, (g, (nb, C, hk)) in enumerate(zip(got, mem)):
        h = O.parse_header(g)
        assert (h["ref_id"], h["chkpt_id"]) == (2, 2), (i, h)
        assert g == orc[i].checkpoint(bufs[i].cpu().numpy(), make_baseline=True), i
        out = torch.zeros_like(bufs[i])
        M.restart_into(out, [None, None, g], 2)
        assert torch.equal(out, bufs[i]), i
