"""Parity at full size: one 8 GiB uint8 buffer through the headline generator at
points of the 8 GiB sweep -- 128 B chunks with Murmur3 and with MD5 at 2% change
per checkpoint, 64 B chunks at 20% and 4096 B chunks at 0.1% (2:1 fresh runs :
shifted copies) -- and a 1 GiB float64 simulation state restarted at every
checkpoint.

The oracle needs minutes for an 8 GiB baseline (its leaf inserts are serial), so
at this size the checks are properties that hold at any size:
  * every checkpoint restarts, from its device-resident chain, to the buffer it
    was taken of;
  * a second handle with another schedule (several leaf tiles per block,
    shallower sweep stages) writes the same bytes;
  * the first-occurrence and shifted roots cover exactly the chunks whose bytes
    changed since the previous checkpoint (all of them at the baseline), each
    chunk once: with n = 2^D chunks a node at depth d covers the 2^(D-d)
    consecutive chunks under it;
  * the diff is as long as its header says: 32 + 4F + 8P + 8S bytes of metadata
    plus C bytes per chunk under a first-occurrence root.
Bit-exactness against the oracle is covered at smaller sizes in the other tests.
Prints each step before it runs.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries, FloatStateSeries  # noqa: E402

pytestmark = pytest.mark.gpu

N = 8 << 30


def say(msg):
    print(msg, flush=True)


def root_ranges(nodes: np.ndarray, D: int):
    """[start, end) chunk ranges of heap nodes of a tree with 2^D leaves."""
    x = nodes.astype(np.int64) + 1
    depth = np.floor(np.log2(x)).astype(np.int64)
    assert np.all((1 << depth) <= x) and np.all(x < (2 << depth))
    width = np.left_shift(1, D - depth)
    start = x * width - (1 << D)
    return start, start + width


def sections(diff: bytes):
    h = np.frombuffer(diff, np.uint32, 8)
    F, P, S = int(h[5]), int(h[6]), int(h[7])
    first = np.frombuffer(diff, np.uint32, F, 32)
    shift = np.frombuffer(diff, np.uint32, 2 * S, 32 + 4 * F + 8 * P).reshape(S, 2)
    return h, F, P, S, first, shift


def changed_chunks(cur, prev, C, n):
    if prev is None:
        return np.ones(n, bool)
    out = torch.empty(n, dtype=torch.bool, device=cur.device)
    step = 1 << 22
    for i in range(0, n, step):
        a = cur[i * C:(i + step) * C].view(-1, C)
        b = prev[i * C:(i + step) * C].view(-1, C)
        out[i:i + step] = (a != b).any(dim=1)
    return out.cpu().numpy()


@pytest.mark.parametrize("C,hk,change", [(128, M.MURMUR3, 0.02), (128, M.MD5, 0.02), (64, M.MURMUR3, 0.2),
                                         (4096, M.MURMUR3, 0.001)])
def test_full_size_8gib_series_properties_and_restart(C, hk, change):
    steps = 4
    n = N // C
    D = n.bit_length() - 1
    name = f"8 GiB, C={C}, {'md5' if hk == M.MD5 else 'murmur3'}, {change * 100:g}% change"
    series = CheckpointSeries.with_change(N, C, change)
    buf = torch.empty(N, dtype=torch.uint8, device="cuda:0")
    d = M.TreeDeduplicator(C, hash_kind=hk)
    other = M.TreeDeduplicator(C, hash_kind=hk)
    other.set_tuning(4, 0.125, 6)
    versions, dev_diffs = [], []
    for k in range(steps):
        say(f"full size: checkpoint {k} ({name})")
        series.apply_device(buf, k)
        diff = d.checkpoint(buf, make_baseline=(k == 0))
        assert other.checkpoint(buf, make_baseline=(k == 0)) == diff, k
        h, F, P, S, first, shift = sections(diff)
        assert (int(h[0]), int(h[1]), int(h[2]) | (int(h[3]) << 32), int(h[4])) == (0, k, N, C)
        fs, fe = root_ranges(first, D)
        ss, se = root_ranges(shift[:, 0], D)
        assert len(diff) == 32 + 4 * F + 8 * P + 8 * S + C * int((fe - fs).sum()), k
        starts = np.concatenate([fs, ss])
        ends = np.concatenate([fe, se])
        cover = np.cumsum(np.bincount(starts, minlength=n + 1) - np.bincount(ends, minlength=n + 1))
        changed = changed_chunks(buf, versions[-1] if versions else None, C, n)
        assert cover.max() <= 1, k
        assert np.array_equal(cover[:n] == 1, changed), k
        versions.append(buf.clone())
        dev_diffs.append(torch.frombuffer(bytearray(diff), dtype=torch.uint8).to("cuda:0"))
        del cover, starts, ends, changed
    other.close()
    d.close()
    out = torch.empty(N, dtype=torch.uint8, device="cuda:0")
    for k in range(steps):
        say(f"full size: restart {k} ({name})")
        M.restart_into(out, dev_diffs, k)
        assert torch.equal(out, versions[k]), k
    del versions, dev_diffs, out, buf
    torch.cuda.empty_cache()


def test_float_state_1gib_restart_every_checkpoint():
    """A 1 GiB float64 state (2^27 values, N(0,1)), 128 B chunks; eight checkpoints
    perturbing 1% of the values in runs of 8-4096 by N(0, 1e-3); every checkpoint
    restarts from its device chain to the state it was taken of."""
    fs = FloatStateSeries(1 << 27)
    vals = fs.initial()
    d = M.TreeDeduplicator(128)
    chain, states = [], []
    for k in range(8):
        say(f"1 GiB float state: checkpoint {k}")
        if k:
            fs.perturb(vals, k)
        raw = vals.view(torch.uint8)
        d.checkpoint_device(raw, make_baseline=(k == 0))
        chain.append(d.copy_last_diff_device().clone())
        states.append(raw.clone())
    d.close()
    out = torch.empty_like(states[0])
    for k in range(8):
        say(f"1 GiB float state: restart {k}")
        out.fill_(0)
        M.restart_into(out, chain[:k + 1], k)
        assert torch.equal(out, states[k]), k
