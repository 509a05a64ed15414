"""Parity at the headline workload's full size: one 8 GiB uint8 buffer, 128 B
chunks, Murmur3, 2% change per checkpoint (2:1 fresh runs : shifted copies).

The oracle needs minutes for an 8 GiB baseline (its leaf inserts are serial), so
at this size the checks are properties that hold at any size:
  * every checkpoint restarts, from its device-resident chain, to the buffer it
    was taken of;
  * a second handle with another schedule (several leaf tiles per block,
    shallower sweep stages) writes the same bytes;
  * the first-occurrence and shifted roots cover exactly the chunks whose bytes
    changed since the previous checkpoint (all of them at the baseline), each
    chunk once: with n = 2^26 chunks a node at depth d covers the 2^(26-d)
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
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu

D = 26
N, C = (1 << D) * 128, 128


def say(msg):
    print(msg, flush=True)


def root_ranges(nodes: np.ndarray):
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


def changed_chunks(cur, prev):
    if prev is None:
        return np.ones(1 << D, bool)
    out = torch.empty(1 << D, dtype=torch.bool, device=cur.device)
    step = 1 << 22
    for i in range(0, 1 << D, step):
        a = cur[i * C:(i + step) * C].view(-1, C)
        b = prev[i * C:(i + step) * C].view(-1, C)
        out[i:i + step] = (a != b).any(dim=1)
    return out.cpu().numpy()


def test_full_size_8gib_series_properties_and_restart():
    steps = 4
    series = CheckpointSeries.with_change(N, C, 0.02)
    buf = torch.empty(N, dtype=torch.uint8, device="cuda:0")
    d = M.TreeDeduplicator(C)
    other = M.TreeDeduplicator(C)
    other.set_tuning(4, 0.125, 6)
    versions, dev_diffs = [], []
    for k in range(steps):
        say(f"full size: checkpoint {k} (8 GiB, C=128, murmur3, 2% change)")
        series.apply_device(buf, k)
        diff = d.checkpoint(buf, make_baseline=(k == 0))
        assert other.checkpoint(buf, make_baseline=(k == 0)) == diff, k
        h, F, P, S, first, shift = sections(diff)
        assert (int(h[0]), int(h[1]), int(h[2]) | (int(h[3]) << 32), int(h[4])) == (0, k, N, C)
        fs, fe = root_ranges(first)
        ss, se = root_ranges(shift[:, 0])
        assert len(diff) == 32 + 4 * F + 8 * P + 8 * S + C * int((fe - fs).sum()), k
        starts = np.concatenate([fs, ss])
        ends = np.concatenate([fe, se])
        cover = np.cumsum(np.bincount(starts, minlength=(1 << D) + 1) - np.bincount(ends, minlength=(1 << D) + 1))
        changed = changed_chunks(buf, versions[-1] if versions else None)
        assert cover.max() <= 1, k
        assert np.array_equal(cover[:1 << D] == 1, changed), k
        versions.append(buf.clone())
        dev_diffs.append(torch.frombuffer(bytearray(diff), dtype=torch.uint8).to("cuda:0"))
    other.close()
    out = torch.empty(N, dtype=torch.uint8, device="cuda:0")
    for k in range(steps):
        say(f"full size: restart {k}")
        d.restart(out, dev_diffs, k)
        assert torch.equal(out, versions[k]), k
