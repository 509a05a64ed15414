"""Synthetic checkpoint series: plans drawn on the host, applied on the device.

One series is a buffer of `nbytes` split into chunks of `chunk_size`, mutated
between checkpoints the way the benchmark configurations describe:

* step 0: every byte from the pseudo-random stream of step 0, then zero-filled
  runs of U{1..zero_max} chunks at uniform positions until the run lengths add up
  to zero_frac of the chunks (overlaps allowed);
* step k >= 1: (i) runs of U{1..fresh_max} chunks re-randomised from the stream of
  step k until their lengths add up to fresh_frac of the chunks; then (ii)
  round(copy_frac * n) single-chunk copies from a uniform source chunk to a
  different uniform destination, all sources read after (i) and before any write,
  the last copy to a destination winning; (iii) everything else unchanged.

The stream byte at offset b for key K is byte (b & 7) of
splitmix64_mix(K + ((b >> 3) + 1) * 0x9E3779B97F4A7C15), so the device kernels
(libmtcheck_amd.so, include/mtcheck_workload.h) and the numpy mirror below produce
identical buffers.  Run and copy plans come from numpy's PCG64 seeded with
(seed, step).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
CONFIG0_SEED = 0x4D54434B


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def mix64(x: int) -> int:
    return int(_mix(np.array([x & 0xFFFFFFFFFFFFFFFF], np.uint64))[0])


def stream_key(seed: int, step: int) -> int:
    return mix64((seed * 0xD1B54A32D192ED03 + step + 1) & 0xFFFFFFFFFFFFFFFF)


def stream_bytes(key: int, begin: int, end: int) -> np.ndarray:
    """Bytes [begin, end) of the stream with `key` (numpy mirror of the device fill)."""
    if end <= begin:
        return np.zeros(0, np.uint8)
    w0, w1 = begin >> 3, (end + 7) >> 3
    w = np.arange(w0, w1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        words = _mix(np.uint64(key) + (w + np.uint64(1)) * GOLDEN)
    b = words.view(np.uint8)
    return b[begin - 8 * w0:end - 8 * w0]


def _draw_runs(rng: np.random.Generator, n: int, frac: float, max_len: int) -> np.ndarray:
    """(first chunk, count) runs until the counts reach round(frac * n)."""
    target = int(round(frac * n))
    if target <= 0 or n <= 0:
        return np.zeros((0, 2), np.uint32)
    max_len = max(1, min(max_len, n))
    out, total = [], 0
    while total < target:
        k = max(16, int(2 * (target - total) / ((1 + max_len) / 2)) + 16)
        lens = rng.integers(1, max_len + 1, size=k)
        starts = rng.integers(0, n - lens + 1)
        cs = total + np.cumsum(lens)
        cut = int(np.searchsorted(cs, target, side="left")) + 1
        cut = min(cut, k)
        out.append(np.stack([starts[:cut], lens[:cut]], axis=1))
        total = int(cs[cut - 1])
    return np.concatenate(out).astype(np.uint32)


@dataclass
class StepPlan:
    step: int
    key: int
    runs: np.ndarray      # (r, 2) uint32: first chunk, count
    zero_runs: bool
    pairs: np.ndarray     # (m, 2) uint32: destination chunk, source chunk (distinct destinations)


class DevicePlan:
    """A step plan's run and copy lists uploaded to the device."""

    def __init__(self, series: "CheckpointSeries", p: StepPlan, device, stream=None):
        import torch
        self.series, self.zero_runs = series, p.zero_runs
        ctx = torch.cuda.stream(stream) if stream is not None else _nullcontext()
        with ctx:
            self.runs = torch.from_numpy(np.ascontiguousarray(p.runs)).to(device) if len(p.runs) else None
            self.nruns = len(p.runs)
            self.pairs = torch.from_numpy(np.ascontiguousarray(p.pairs)).to(device) if len(p.pairs) else None
            self.npairs = len(p.pairs)
            self.tmp = (torch.empty(self.npairs * series.chunk_size, dtype=torch.uint8, device=device)
                        if self.npairs else None)

    def apply(self, buf, key: int, fill_all: bool, stream=None) -> None:
        from ._lib import check, lib
        s = self.series
        st = None if stream is None else int(getattr(stream, "cuda_stream", stream))
        ptr = buf.data_ptr()
        if fill_all:
            check(lib().mtc_wl_fill_stream(ptr, 0, s.nbytes, key, st))
        if self.nruns:
            check(lib().mtc_wl_fill_runs(ptr, s.nbytes, s.chunk_size, self.runs.data_ptr(), self.nruns, key,
                                         1 if self.zero_runs else 0, st))
        if self.npairs:
            check(lib().mtc_wl_copy_chunks(ptr, s.nbytes, s.chunk_size, self.pairs.data_ptr(), self.npairs,
                                           self.tmp.data_ptr(), st))


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


class CheckpointSeries:
    def __init__(self, nbytes: int, chunk_size: int, seed: int = CONFIG0_SEED, zero_frac: float = 0.10,
                 zero_max: int = 32, fresh_frac: float = 0.02, fresh_max: int = 64, copy_frac: float = 0.01):
        self.nbytes, self.chunk_size, self.seed = nbytes, chunk_size, seed
        self.n = (nbytes + chunk_size - 1) // chunk_size
        self.zero_frac, self.zero_max = zero_frac, zero_max
        self.fresh_frac, self.fresh_max, self.copy_frac = fresh_frac, fresh_max, copy_frac

    @classmethod
    def with_change(cls, nbytes: int, chunk_size: int, change: float, seed: int = CONFIG0_SEED):
        """Change fraction split 2:1 between fresh runs and shifted-duplicate copies."""
        return cls(nbytes, chunk_size, seed, fresh_frac=2 * change / 3, copy_frac=change / 3)

    def plan(self, step: int) -> StepPlan:
        rng = np.random.Generator(np.random.PCG64([self.seed, step]))
        key = stream_key(self.seed, step)
        if step == 0:
            return StepPlan(0, key, _draw_runs(rng, self.n, self.zero_frac, self.zero_max), True,
                            np.zeros((0, 2), np.uint32))
        runs = _draw_runs(rng, self.n, self.fresh_frac, self.fresh_max)
        m = int(round(self.copy_frac * self.n)) if self.n > 1 else 0
        if m:
            dst = rng.integers(0, self.n, size=m)
            src = rng.integers(0, self.n - 1, size=m)
            src = src + (src >= dst)
            # last copy to each destination wins
            rev_dst = dst[::-1]
            _, first = np.unique(rev_dst, return_index=True)
            keep = np.sort(m - 1 - first)
            pairs = np.stack([dst[keep], src[keep]], axis=1).astype(np.uint32)
        else:
            pairs = np.zeros((0, 2), np.uint32)
        return StepPlan(step, key, runs, False, pairs)

    # -- host mirror -----------------------------------------------------------------
    def apply_host(self, buf: np.ndarray, step: int) -> np.ndarray:
        p = self.plan(step)
        C, L = self.chunk_size, self.nbytes
        if step == 0:
            buf[:] = stream_bytes(p.key, 0, L)
        for first, cnt in p.runs:
            b, e = int(first) * C, min((int(first) + int(cnt)) * C, L)
            buf[b:e] = 0 if p.zero_runs else stream_bytes(p.key, b, e)
        if len(p.pairs):
            tmp = np.zeros((len(p.pairs), C), np.uint8)
            for i, (_, s) in enumerate(p.pairs):
                b = int(s) * C
                seg = buf[b:min(b + C, L)]
                tmp[i, :seg.size] = seg
            for i, (d, _) in enumerate(p.pairs):
                b = int(d) * C
                e = min(b + C, L)
                buf[b:e] = tmp[i, :e - b]
        return buf

    def host_buffer(self, step: int) -> np.ndarray:
        buf = np.zeros(self.nbytes, np.uint8)
        for s in range(step + 1):
            self.apply_host(buf, s)
        return buf

    # -- device ------------------------------------------------------------------------
    def apply_device(self, buf, step: int, stream=None, plan: "DevicePlan | None" = None) -> None:
        """Mutate a uint8 cuda tensor in place (kernels of libmtcheck_amd.so).  With a
        stream, the work is queued there and only that stream is waited for.  `plan`
        reuses an uploaded plan's positions with this step's stream key."""
        import torch
        dp = plan if plan is not None else DevicePlan(self, self.plan(step), buf.device, stream)
        dp.apply(buf, stream_key(self.seed, step), step == 0, stream)
        if stream is None:
            torch.cuda.synchronize(buf.device)
        else:
            stream.synchronize()

    def device_buffer(self, step: int, device="cuda:0"):
        import torch
        buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        for s in range(step + 1):
            self.apply_device(buf, s)
        return buf



# -- multi-buffer workloads ------------------------------------------------------------
CONFIG3_SEED = 0x4D544333
CONFIG4_SEED = 0x4D544334


def _log_uniform(rng: np.random.Generator, lo: int, hi: int, size: int) -> np.ndarray:
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size))).astype(np.int64)


def batch_members(count: int = 256, seed: int = CONFIG3_SEED, lo: int = 1 << 20, hi: int = 512 << 20,
                  chunks=(64, 128, 512), change=(0.005, 0.05)):
    """Independent buffers for batched deduplication: size log-uniform in [lo, hi]
    bytes, chunk size uniform from `chunks`, change fraction uniform in `change`,
    each with its own seed.  Returns a list of CheckpointSeries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = _log_uniform(rng, lo, hi, count)
    cs = rng.choice(np.array(chunks), size=count)
    ch = rng.uniform(change[0], change[1], count)
    seeds = rng.integers(1, 2**62, size=count)
    return [CheckpointSeries.with_change(int(s), int(c), float(f), seed=int(sd))
            for s, c, f, sd in zip(sizes, cs, ch, seeds)]


def serve_clients(count: int = 64, seed: int = CONFIG4_SEED, lo: int = 4 << 20, hi: int = 2 << 30, chunk: int = 128,
                  change=(0.001, 0.10)):
    """Clients of the served workload: fixed submission size log-uniform in [lo, hi],
    one chunk size, change fraction uniform in `change`, per-client seeds."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = _log_uniform(rng, lo, hi, count)
    ch = rng.uniform(change[0], change[1], count)
    seeds = rng.integers(1, 2**62, size=count)
    return [CheckpointSeries.with_change(int(s), chunk, float(f), seed=int(sd)) for s, f, sd in zip(sizes, ch, seeds)]


# -- simulation-state workload (float64 values; bytes treated opaquely) -----------------
CONFIG2_SEED = 0x4D544332


class FloatStateSeries:
    """A float64 array initialised from N(0, 1) (torch's device generator, seeded);
    each later step adds N(0, sigma) noise to `frac` of the values in contiguous runs
    of U{run_min..run_max} values at uniform positions (the last run is cut so the
    run lengths sum to frac * n; a value hit by two runs gets the noise once).  Run positions come from numpy's PCG64 seeded with (seed, step)."""

    def __init__(self, nvalues: int, seed: int = CONFIG2_SEED, frac: float = 0.01, run_min: int = 8,
                 run_max: int = 4096, sigma: float = 1e-3):
        self.n, self.seed, self.frac = nvalues, seed, frac
        self.run_min, self.run_max, self.sigma = run_min, run_max, sigma

    def initial(self, device="cuda:0"):
        import torch
        g = torch.Generator(device=device)
        g.manual_seed(self.seed)
        return torch.randn(self.n, dtype=torch.float64, device=device, generator=g)

    def perturb(self, values, step: int) -> None:
        import torch
        rng = np.random.Generator(np.random.PCG64([self.seed, step]))
        target = int(round(self.frac * self.n))
        starts, lens, total = [], [], 0
        while total < target:
            ln = int(rng.integers(self.run_min, min(self.run_max, self.n) + 1))
            ln = min(ln, target - total)  # the last run is cut so the lengths sum to the target
            st = int(rng.integers(0, self.n - ln + 1))
            starts.append(st)
            lens.append(ln)
            total += ln
        dev = values.device
        st = torch.tensor(starts, dtype=torch.int64, device=dev)
        ln = torch.tensor(lens, dtype=torch.int64, device=dev)
        base = torch.repeat_interleave(st - (torch.cumsum(ln, 0) - ln), ln)
        idx = torch.unique(base + torch.arange(int(ln.sum()), device=dev))
        g = torch.Generator(device=dev)
        g.manual_seed(self.seed * 1000003 + step)
        values[idx] += self.sigma * torch.randn(idx.numel(), dtype=torch.float64, device=dev, generator=g)
