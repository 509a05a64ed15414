"""Schedule autotuner: per chunk size it picks the sweep fuse depth (and, for chunk
sizes the leaf kernel does not stage through LDS, the leaf kernel's tiles per
block) by timing incremental checkpoints of a synthetic sample buffer on the
device.  The table load factor stays at its default.  The choice never changes
results (diffs are identical for every schedule), only speed."""
from __future__ import annotations

import itertools
import time

from .tree import MURMUR3, TreeDeduplicator
from .workload import CheckpointSeries

DEFAULT = (1, 0.125, 8)
# The load factor stays at the default: a fuller table is rehashed on the next
# checkpoint, and a sparser one costs memory for little speed.
CANDIDATES = list(itertools.product((1, 2, 4, 8), (0.125,), (6, 7, 8)))
MARGIN = 0.03  # a schedule replaces the default only when it is this much faster
# Chunk sizes the leaf kernel stages through LDS.  For these, one tile per block
# selects the single-tile leaf kernel; at full size one to eight tiles per block
# run within 2% of each other (64-byte chunks, 8 GiB: 1.83-1.91 ms), below the
# margin, so a 256 MiB sample would pick among them by noise.  Only the fuse
# depth is tuned for them.
STAGED = {16, 32, 64, 128, 256, 512, 1024, 2048, 4096}


def candidates_for(chunk_size: int):
    if chunk_size in STAGED:
        return [t for t in CANDIDATES if t[0] == 1]
    return CANDIDATES


def time_schedule(series: CheckpointSeries, bufs, hash_kind: int, tuning, reps: int = 2) -> float:
    """Mean seconds of the incremental checkpoints of the pre-generated versions."""
    import torch
    d = TreeDeduplicator(series.chunk_size, hash_kind=hash_kind)
    d.set_tuning(*tuning)
    d.checkpoint_device(bufs[0], make_baseline=True)
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b in bufs[1:]:
            d.checkpoint_device(b)
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) / (len(bufs) - 1))
        d.checkpoint_device(bufs[0], make_baseline=True)
    d.close()
    return best


def autotune(chunk_sizes, hash_kind: int = MURMUR3, sample_bytes: int = 256 << 20, change: float = 0.02,
             steps: int = 3, candidates=None) -> dict:
    """{chunk size: (leaf_tiles_per_block, load_factor, fuse_levels)}."""
    import torch
    out = {}
    for C in sorted(set(int(c) for c in chunk_sizes)):
        series = CheckpointSeries.with_change(sample_bytes, C, change, seed=0xA070 + C)
        bufs = []
        for k in range(steps + 1):
            v = torch.empty(sample_bytes, dtype=torch.uint8, device="cuda:0")
            if k:
                v.copy_(bufs[-1])
            series.apply_device(v, k)
            bufs.append(v)
        timings = {t: time_schedule(series, bufs, hash_kind, t) for t in (candidates or candidates_for(C))}
        timings[DEFAULT] = min(timings.get(DEFAULT, float("inf")), time_schedule(series, bufs, hash_kind, DEFAULT))
        best = min(timings, key=timings.get)
        out[C] = best if timings[best] < (1 - MARGIN) * timings[DEFAULT] else DEFAULT
        del bufs
        torch.cuda.empty_cache()
    return out
