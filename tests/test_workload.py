"""Synthetic checkpoint-series generator (host side) -- CPU only."""
import numpy as np

from mtcheck_amd.workload import GOLDEN, CheckpointSeries, _mix, stream_bytes, stream_key


def test_stream_bytes_match_word_formula():
    key = stream_key(5, 3)
    b = stream_bytes(key, 0, 64)
    for w in range(8):
        k = np.array([(key + (w + 1) * int(GOLDEN)) & (2**64 - 1)], np.uint64)
        word = int(_mix(k)[0])
        assert b[8 * w:8 * w + 8].tobytes() == word.to_bytes(8, "little")
    # any sub-range is the same bytes
    assert np.array_equal(stream_bytes(key, 13, 51), b[13:51])


def test_plans_are_deterministic_and_sized():
    s = CheckpointSeries(1 << 22, 128, seed=42)
    n = s.n
    p0, p0b = s.plan(0), s.plan(0)
    assert np.array_equal(p0.runs, p0b.runs)
    assert p0.zero_runs and p0.runs[:, 1].sum() >= round(0.10 * n)
    assert p0.runs[:, 1].max() <= 32 and (p0.runs[:, 0] + p0.runs[:, 1] <= n).all()
    p1 = s.plan(1)
    assert not p1.zero_runs and round(0.02 * n) <= p1.runs[:, 1].sum() < round(0.02 * n) + 64
    assert p1.runs[:, 1].max() <= 64
    dst, src = p1.pairs[:, 0], p1.pairs[:, 1]
    assert len(np.unique(dst)) == len(dst) and (dst != src).all()
    assert len(dst) <= round(0.01 * n)
    assert not np.array_equal(s.plan(2).runs, p1.runs)


def test_host_series_changes_about_the_planned_fraction():
    s = CheckpointSeries(1 << 20, 128, seed=1)
    a = s.host_buffer(0)
    zero_chunks = (a.reshape(-1, 128) == 0).all(axis=1).mean()
    assert 0.07 < zero_chunks <= 0.11
    b = a.copy()
    s.apply_host(b, 1)
    changed = (a.reshape(-1, 128) != b.reshape(-1, 128)).any(axis=1).mean()
    assert 0.015 < changed < 0.035


def test_with_change_split():
    s = CheckpointSeries.with_change(1 << 20, 64, 0.03)
    assert abs(s.fresh_frac - 0.02) < 1e-12 and abs(s.copy_frac - 0.01) < 1e-12
