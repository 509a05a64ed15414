"""Schedule autotuner: per chunk size it times incremental checkpoints of a
synthetic sample buffer on the device under each schedule of a search space and
keeps the fastest, unless the library's default is within MARGIN of it.  For a
batched caller the sample is checkpointed by several deduplicators in one batched
step (`members`), so that the schedule is timed with the device as full as the
caller's batch keeps it: alone, a 256 MiB buffer favoured four stage-0 tiles per
sweep wavefront, which made the 256-buffer batch 5% slower than the default eight.  A
schedule is (leaf tiles per block, table load factor, tree levels fused per sweep
stage, sweep A tiles per wavefront, sweep B tiles per wavefront).  The choice
never changes results (diffs are identical for every schedule), only speed and,
through the load factor, the table's memory."""
from __future__ import annotations

import itertools
import time

from .tree import MURMUR3, Batch, TreeDeduplicator
from .workload import CheckpointSeries

DEFAULT = (1, 0.125, 8, 8, 8)
LOAD_FACTORS = (0.125, 0.25)   # a 0.25 table holds half the slots (half the memory)
FUSE = (6, 7, 8)
SWEEP_TILES = ((8, 8), (4, 8), (8, 16), (4, 4))
LEAF_TILES = (1, 2, 4, 8)
MARGIN = 0.03  # a schedule replaces the default only when it is this much faster
# Chunk sizes the leaf kernel stages through LDS.  For these, one tile per block
# selects the single-tile leaf kernel; at full size one to eight tiles per block
# run within 2% of each other (64-byte chunks, 8 GiB: 1.83-1.91 ms), below the
# margin, so the leaf tiles are searched only for the other chunk sizes.
STAGED = {16, 32, 64, 128, 256, 512, 1024, 2048, 4096}


def candidates_for(chunk_size: int):
    leaf = (1,) if chunk_size in STAGED else LEAF_TILES
    return [(t, lf, f, sa, sb) for t, lf, f, (sa, sb) in itertools.product(leaf, LOAD_FACTORS, FUSE, SWEEP_TILES)]


def apply(d: TreeDeduplicator, sched) -> None:
    d.set_tuning(*sched[:3])
    d.set_sweep_tiles(*sched[3:])


def time_schedule(series: CheckpointSeries, bufs, hash_kind: int, sched, reps: int = 2, members: int = 1) -> float:
    """Best mean seconds per incremental checkpoint (or batched step of `members`
    deduplicators reading the same versions) of the pre-generated versions."""
    import torch
    ds = [TreeDeduplicator(series.chunk_size, hash_kind=hash_kind) for _ in range(members)]
    for d in ds:
        apply(d, sched)
    if members == 1:
        step = lambda b, base=False: ds[0].checkpoint_device(b, make_baseline=base)
    else:
        batch = Batch()
        step = lambda b, base=False: batch.checkpoint(ds, [b] * members, make_baseline=base)
    step(bufs[0], True)
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b in bufs[1:]:
            step(b)
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) / (len(bufs) - 1))
        step(bufs[0], True)
    if members > 1:
        batch.close()
    for d in ds:
        d.close()
    return best


def autotune(chunk_sizes, hash_kind: int = MURMUR3, sample_bytes: int = 256 << 20, change: float = 0.02,
             steps: int = 3, candidates=None, report: dict | None = None, members: int = 1) -> dict:
    """{chunk size: schedule}.  `report`, when given, receives per chunk size the
    searched space, the default's and the pick's time and the gain.  members > 1:
    each schedule is timed as a batched step of that many deduplicators."""
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
        space = candidates or candidates_for(C)
        timings = {t: time_schedule(series, bufs, hash_kind, t, members=members) for t in space}
        timings[DEFAULT] = min(timings.get(DEFAULT, float("inf")),
                               time_schedule(series, bufs, hash_kind, DEFAULT, members=members))
        best = min(timings, key=timings.get)
        out[C] = best if timings[best] < (1 - MARGIN) * timings[DEFAULT] else DEFAULT
        if report is not None:
            report[C] = {"searched": len(timings), "sample_bytes": sample_bytes, "members": members,
                         "space": {"leaf_tiles": sorted({t[0] for t in space}), "load_factor": list(LOAD_FACTORS),
                                   "fuse": list(FUSE), "sweep_tiles": [list(s) for s in SWEEP_TILES]},
                         "default_ms": round(timings[DEFAULT] * 1e3, 4), "best": list(best),
                         "best_ms": round(timings[best] * 1e3, 4), "pick": list(out[C]),
                         "gain": round(timings[DEFAULT] / timings[out[C]], 4)}
        del bufs
        torch.cuda.empty_cache()
    return out
