"""Benchmark: plain single-buffer tree deduplication on one MI355X.

One step = one incremental checkpoint (hash + label + forest + diff) of a buffer
already resident in HBM.  Default workload: 8 GiB of uint8, 128 B chunks, Murmur3,
2 % of the chunks changed per checkpoint (split 2:1 between re-randomised runs of
1..64 chunks and single-chunk copies that create shifted duplicates), generated on
the device; step 0 (the baseline) is always a warm-up step.

Prints one JSON line (rank 0).  `value` is input GB/s (1e9 B) over the timed steps.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "dedup GB/s + %HBM peak at 128B chunks; batched vs plain GPU; p99 ckpt latency"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=9)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--size-gib", type=float, default=8.0)
    p.add_argument("--chunk", type=int, default=128)
    p.add_argument("--hash", choices=["murmur3", "md5"], default="murmur3")
    p.add_argument("--change", type=float, default=0.02)
    p.add_argument("--cpu-sample-mib", type=int, default=1024)
    p.add_argument("--cpu-steps", type=int, default=4)
    p.add_argument("--cpu-par-sample-mib", type=int, default=2048)
    p.add_argument("--cpu-par-steps", type=int, default=6)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-host-path", action="store_true",
                   help="skip the PCIe-inclusive runs (diffs copied to pinned host memory)")
    p.add_argument("--load-factor", type=float, default=0.125, help="digest-table entries per slot at a baseline")
    p.add_argument("--tiles-per-block", type=int, default=0, help="leaf-kernel tiles per block (0: the library default)")
    p.add_argument("--sweep", action="store_true",
                   help="only the sweep: chunk sizes {64,128,512,4096} x {murmur3,md5} x change {0.1,2,20}%%, "
                        "one JSON line each")
    p.add_argument("--no-sweep", action="store_true", help="skip the sweep in the default run")
    p.add_argument("--sweep-steps", type=int, default=3, help="timed checkpoints per sweep point")
    p.add_argument("--sweep-warmup", type=int, default=1, help="untimed checkpoints per sweep point (the baseline)")
    p.add_argument("--no-batched", action="store_true", help="skip the batched multi-buffer measurement")
    p.add_argument("--no-small-series", action="store_true", help="skip the 64 MiB ten-checkpoint series")
    p.add_argument("--only-batched", action="store_true", help="profiling aid: run the batched measurement alone")
    p.add_argument("--batch-count", type=int, default=256)
    p.add_argument("--batch-steps", type=int, default=10)
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--no-restart", action="store_true", help="skip the restart measurement")
    p.add_argument("--only-restart", action="store_true", help="profiling aid: run the restart measurement alone")
    p.add_argument("--restart-gib", type=float, default=16.0)
    p.add_argument("--restart-steps", type=int, default=20)
    p.add_argument("--no-served", action="store_true", help="skip the served (request loop) measurement")
    p.add_argument("--only-served", action="store_true", help="profiling aid: run the served measurement alone")
    p.add_argument("--serve-clients", type=int, default=64)
    p.add_argument("--serve-loads", default="50,100,200,350,500", help="offered loads, GB/s")
    p.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per load")
    p.add_argument("--serve-lanes", type=int, default=2)
    p.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                   help="only run a plain device copy of GIB GiB three times (scales the HBM byte counters)")
    return p.parse_args()


def note(msg: str) -> None:
    """Progress on stderr: the leg about to run and the device memory free."""
    try:
        import torch
        free = torch.cuda.mem_get_info()[0] / 2**30 if torch.cuda.is_available() else float("nan")
    except Exception:
        free = float("nan")
    print(f"[bench] {msg} (free {free:.1f} GiB)", file=sys.stderr, flush=True)


def cpu_baseline(args, hk):
    """The CPU oracle on bounded samples of the same workload (same generator, chunk
    size, hash and change rate; baseline untimed, incrementals timed): its OpenMP
    checkpoint on every host thread this job may use (the value), and its serial
    checkpoint on one core beside it."""
    import numpy as np
    import torch

    from mtcheck_amd.workload import CheckpointSeries
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)

    def timed(mib, steps, nthreads):
        nbytes = mib << 20
        series = CheckpointSeries.with_change(nbytes, args.chunk, args.change)
        buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
        orc = O.OracleDedup(args.chunk, hk)
        series.apply_device(buf, 0)
        orc.checkpoint(buf.cpu().numpy(), make_baseline=True, threads=nthreads)
        t = 0.0
        for k in range(1, steps + 1):
            series.apply_device(buf, k)
            host = buf.cpu().numpy()
            t0 = time.perf_counter()
            orc.checkpoint(host, threads=nthreads)
            t += time.perf_counter() - t0
        del buf, orc
        torch.cuda.empty_cache()
        return nbytes * steps / t / 1e9, t

    par, tp = timed(args.cpu_par_sample_mib, args.cpu_par_steps, threads)
    ser, ts = timed(args.cpu_sample_mib, args.cpu_steps, 0)
    return {"value": round(par, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{args.chunk} B chunks, {args.hash}, {args.change * 100:g}% change, same generator: "
                      f"OpenMP oracle on {threads} threads, {args.cpu_par_sample_mib} MiB series, "
                      f"{args.cpu_par_steps} incremental checkpoints timed ({tp:.1f} s); serial oracle on one core, "
                      f"{args.cpu_sample_mib} MiB series, {args.cpu_steps} checkpoints ({ts:.1f} s)",
            "serial_1core": round(ser, 4)}


def copy_calibration(gib: float):
    """A plain device copy of known size (read GIB GiB, write GIB GiB) for scaling
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.bitwise_not(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}), flush=True)


def leaf_traffic(nbytes: int, C: int, hash_name: str, change: float):
    """HBM bytes per k_leaf dispatch measured for this workload by the two
    rocprofv3 --pmc passes summarised under profiles/ (None when no summary
    covers this workload)."""
    path = os.path.join(ROOT, "profiles", f"k_leaf_pmc_traffic_{nbytes >> 30}gib_c{C}_{hash_name}.md")
    if not os.path.exists(path):
        return None
    fields = {}
    for line in open(path):
        if line.startswith("- ") and ": " in line:
            k, v = line[2:].split(": ", 1)
            fields[k.strip()] = v.strip().strip("`")
    if fields.get("workload") != f"size={nbytes} chunk={C} hash={hash_name} change={change:g}":
        return None
    return int(fields["traffic_bytes_per_dispatch"])


def leaf_profile(nbytes: int, C: int, hash_name: str, change: float):
    """(summary file, average leaf-kernel duration in ms) from the committed
    rocprofv3 --kernel-trace --stats summary of this workload, or None."""
    name = f"checkpoint_kernels_{nbytes >> 30}gib_c{C}_{hash_name}" + ("" if change == 0.02 else f"_{change * 100:g}pct")
    path = os.path.join(ROOT, "profiles", name + ".md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| `mtc::k_leaf1<") or line.startswith("| `mtc::k_leaf1_direct<"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            return f"profiles/{name}.md", round(float(cells[2]) / 1e3, 4)
    return None


def header_tuple(h):
    return (h.ref_id, h.chkpt_id, h.datalen, h.chunk_size, h.num_first_ocur, h.num_prior_chkpts, h.num_shift_dupl)


def batched_bench(args):
    """Many independent buffers of uneven sizes: one batched step (graph-replayed
    batched launches) against one checkpoint call per buffer, same inputs; the
    baseline step is untimed.  Every member's whole diff is compared between the
    two forms at every step, on the device (outside the timed calls)."""
    import torch

    import mtcheck_amd as M
    from mtcheck_amd.autotune import DEFAULT, apply, autotune
    from mtcheck_amd.workload import batch_members
    members = batch_members(args.batch_count)
    total = sum(s.nbytes for s in members)
    report = {}
    # each schedule timed as a batched step of 16 deduplicators over one 256 MiB
    # sample (4 GiB a step), so the device is about as full as in the 256-buffer step
    tuned = {} if args.no_autotune else autotune({s.chunk_size for s in members}, report=report, members=16)
    bufs = [torch.empty(s.nbytes, dtype=torch.uint8, device="cuda:0") for s in members]
    A = [M.TreeDeduplicator(s.chunk_size) for s in members]
    for d, s in zip(A, members):
        apply(d, tuned.get(s.chunk_size, DEFAULT))
    P = [M.TreeDeduplicator(s.chunk_size) for s in members]
    batch = M.Batch(use_graph=True)
    tb, tp, lat_b = [], [], []
    exact, compared = True, 0
    for k in range(args.batch_steps):
        for s, b in zip(members, bufs):
            s.apply_device(b, k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hdrs, sizes = batch.checkpoint(A, bufs, make_baseline=(k == 0))
        t1 = time.perf_counter()
        solo = [d.checkpoint_device(b, make_baseline=(k == 0)) for d, b in zip(P, bufs)]
        t2 = time.perf_counter()
        if k:
            tb.append(t1 - t0)
            tp.append(t2 - t1)
        for h, sz, (hs, _, szs) in zip(hdrs, sizes, solo):
            exact &= header_tuple(h) == header_tuple(hs) and sz == szs
        # every member's whole diff, batched against one call per buffer, on the device
        for a, p in zip(A, P):
            exact &= bool(torch.equal(a.copy_last_diff_device(), p.copy_last_diff_device()))
        compared += len(A)
    vb = total * len(tb) / sum(tb) / 1e9
    vp = total * len(tp) / sum(tp) / 1e9
    res = {"workload": f"{len(members)} buffers, log-uniform 1 MiB-512 MiB ({total / 2**30:.1f} GiB), chunk size "
                       f"from {{64,128,512}}, change U[0.5%,5%] split 2:1, {args.batch_steps} checkpoints",
           "value": round(vb, 2), "unit": "GB/s", "per_buffer_value": round(vp, 2), "speedup": round(vb / vp, 3),
           "ms_per_step": round(sum(tb) / len(tb) * 1e3, 3), "per_buffer_ms_per_step": round(sum(tp) / len(tp) * 1e3, 3),
           "step_ms": [round(t * 1e3, 2) for t in tb], "per_buffer_step_ms": [round(t * 1e3, 2) for t in tp],
           "bit_exact": bool(exact), "diffs_compared": compared, "graph": True,
           "tuning": {"schedule": "(leaf tiles per block, load factor, fuse levels, sweep A tiles per wave, "
                                  "sweep B tiles per wave), per chunk size; the default unless 3% faster, "
                                  "each timed as a batched step of 16 deduplicators over a 256 MiB sample",
                      "default": list(DEFAULT), **{str(c): report.get(c, list(t)) for c, t in sorted(tuned.items())}}}
    for d in A + P:
        d.close()
    batch.close()
    del bufs
    torch.cuda.empty_cache()
    return res


def state_checksum(raw, M):
    """sha256 over the Murmur3 digests of every 4 KiB of a device buffer."""
    import hashlib
    dg = M.hash_chunks(raw, 4096, M.MURMUR3)
    return hashlib.sha256(dg.cpu().numpy().tobytes()).hexdigest()


def restart_bench(args):
    """A float64 simulation state (N(0,1), seeded) checkpointed `restart_steps` times
    with 1% of the values perturbed per step; the diffs stay in HBM.  Every
    checkpoint k is then rebuilt from diffs 0..k and compared with the generated
    state (checksum of 4 KiB digests; the last one byte for byte).  Restart GB/s =
    state bytes / restart time."""
    import numpy as np
    import torch

    import mtcheck_amd as M
    from mtcheck_amd.workload import FloatStateSeries
    nvals = int(args.restart_gib * (1 << 30)) // 8
    fs = FloatStateSeries(nvals)
    vals = fs.initial()
    raw = vals.view(torch.uint8)
    d = M.TreeDeduplicator(128)
    diffs, sums, ck_ms = [], [], []
    for k in range(args.restart_steps):
        if k:
            fs.perturb(vals, k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.checkpoint_device(raw, make_baseline=(k == 0))
        ck_ms.append((time.perf_counter() - t0) * 1e3)
        diffs.append(d.copy_last_diff_device().clone())
        sums.append(state_checksum(raw, M))
    d.close()
    out = torch.empty_like(raw)
    rs_ms, ok = [], True
    for k in range(args.restart_steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        M.restart_into(out, diffs[:k + 1], k)
        torch.cuda.synchronize()
        rs_ms.append((time.perf_counter() - t0) * 1e3)
        ok &= state_checksum(out, M) == sums[k]
    ok &= bool(torch.equal(out, raw))
    nbytes = raw.numel()
    diff_bytes = sum(x.numel() for x in diffs)
    meta = sum(8 * (M.parse_header(x[:32].cpu().numpy().tobytes())["num_shift_dupl"]) for x in diffs)
    res = {"workload": f"float64 state of {nbytes / 2**30:g} GiB ({nvals} values, N(0,1)), 128 B chunks, murmur3, "
                       f"{args.restart_steps} checkpoints perturbing 1% of values in runs of 8-4096 values "
                       "by N(0,1e-3); diffs resident in HBM",
           "value": round(nbytes / (float(np.mean(rs_ms[1:])) / 1e3) / 1e9, 2), "unit": "GB/s",
           "restart_ms": [round(x, 3) for x in rs_ms], "checkpoint_ms_mean": round(float(np.mean(ck_ms[1:])), 3),
           "diff_gib_total": round(diff_bytes / 2**30, 3), "byte_identical": bool(ok),
           "hbm_frac": round((2 * nbytes + meta) / (float(np.mean(rs_ms[1:])) / 1e3) / HBM_PEAK, 4)}
    del diffs, out, vals, raw
    torch.cuda.empty_cache()
    return res


def served_bench(args):
    """Clients with fixed buffer sizes submit checkpoints with Poisson arrivals at a
    total offered load; the server batches whatever is queued.  A generator thread
    (its own stream) produces a client's next version as soon as its previous
    checkpoint completes; latency runs from the scheduled arrival to completion
    (queueing included), and is also reported from the submit call to the diff
    being ready."""
    import collections
    import heapq
    import queue
    import threading

    import numpy as np
    import torch

    from mtcheck_amd.serve import Server
    from mtcheck_amd.workload import DevicePlan, serve_clients, stream_key
    clients = serve_clients(args.serve_clients)
    sizes = [c.nbytes for c in clients]
    total = sum(sizes)
    srv = Server(lanes=args.serve_lanes, max_batch=64, memory_budget=200 << 30)
    cids = [srv.add_client(c.chunk_size, c.nbytes) for c in clients]
    bufs = [torch.empty(n, dtype=torch.uint8, device="cuda:0") for n in sizes]
    K = 6  # uploaded plans per client, reused with fresh stream keys
    plans = [[DevicePlan(c, c.plan(j + 1), "cuda:0") for j in range(K)] for c in clients]
    step = [0] * len(clients)
    for c, b in zip(clients, bufs):
        c.apply_device(b, 0)
    torch.cuda.synchronize()
    for t in [srv.submit(ci, b, make_baseline=True) for ci, b in zip(cids, bufs)]:
        srv.result(t)

    ready = [threading.Event() for _ in clients]
    todo = queue.Queue()

    def generator():
        gen = torch.cuda.Stream()
        while True:
            i = todo.get()
            if i is None:
                return
            step[i] += 1
            plans[i][(step[i] - 1) % K].apply(bufs[i], stream_key(clients[i].seed, step[i]), False, gen)
            gen.synchronize()
            ready[i].set()

    worker = threading.Thread(target=generator, daemon=True)
    worker.start()
    for i in range(len(clients)):
        todo.put(i)
    for e in ready:
        e.wait()
    levels = []
    rng = np.random.default_rng(4)
    for load in [float(x) for x in args.serve_loads.split(",")]:
        lam = load * 1e9 / total  # requests per second per client
        events = []
        for i in range(len(clients)):
            t = rng.exponential(1 / lam)
            while t < args.serve_seconds:
                events.append((t, i))
                t += rng.exponential(1 / lam)
        heapq.heapify(events)
        n_events = len(events)
        backlog = collections.defaultdict(collections.deque)
        outstanding = {}
        lat, done_bytes = [], 0
        parts = []  # (submit - arrival, start - submit, done - start) per request
        cap0 = [srv.client(c).stats().table_capacity for c in cids]
        start = time.monotonic() + 0.01
        last_done = start
        while events or outstanding or any(backlog.values()):
            now = time.monotonic()
            busy = False
            while events and start + events[0][0] <= now:
                ta, i = heapq.heappop(events)
                backlog[i].append(start + ta)
            finished = []
            for i, (tk, ta) in outstanding.items():
                r = srv.result(tk, wait=False)
                if r is not None:
                    lat.append(r[4] - ta)
                    parts.append((r[2] - ta, r[3] - r[2], r[4] - r[3]))
                    done_bytes += sizes[i]
                    last_done = max(last_done, r[4])
                    finished.append(i)
            for i in finished:
                del outstanding[i]
                ready[i].clear()
                todo.put(i)  # the client produces its next version
                busy = True
            for i, q in backlog.items():
                if q and i not in outstanding and ready[i].is_set():
                    outstanding[i] = (srv.submit(cids[i], bufs[i]), q.popleft())
                    busy = True
            if not busy:
                time.sleep(0.0001)
        lat_ms = np.array(lat) * 1e3
        pm = np.array(parts) * 1e3
        sub_ms = pm[:, 1] + pm[:, 2]  # submit -> diff ready
        grown = sum(srv.client(c).stats().table_capacity != c0 for c, c0 in zip(cids, cap0))
        levels.append({"offered_gbs": load, "achieved_gbs": round(done_bytes / (last_done - start) / 1e9, 2),
                       "requests": n_events, "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                       "p50_submit_to_ready_ms": round(float(np.percentile(sub_ms, 50)), 3),
                       "p99_submit_to_ready_ms": round(float(np.percentile(sub_ms, 99)), 3),
                       "p99_parts_ms": {k: round(float(np.percentile(pm[:, j], 99)), 3)
                                        for j, k in enumerate(("to_submit", "queued", "service"))},
                       "max_service_ms": round(float(pm[:, 2].max()), 3), "tables_grown": int(grown)})
        for e in ready:
            e.wait()
    todo.put(None)
    worker.join()
    st = srv.stats()
    srv.close()
    torch.cuda.empty_cache()
    check = served_check(clients, bufs, plans, step, K)
    del bufs, plans
    torch.cuda.empty_cache()
    return {"workload": f"{len(clients)} clients, log-uniform 4 MiB-2 GiB ({total / 2**30:.1f} GiB), 128 B chunks, "
                        f"murmur3, change U[0.1%,10%], Poisson arrivals, {args.serve_seconds:g} s per load, "
                        f"{args.serve_lanes} lanes, budget 200 GiB",
            "levels": levels, "steps": st["steps"], "served": st["served"],
            "mean_batch": round(st["served"] / max(1, st["steps"]), 2), **check}


class Versions:
    """Versions of a checkpoint series in two alternating device buffers: version k
    is built in buffer k % 2 from version k - 2 by applying steps k - 1 and k, so a
    run of any length holds two buffers.  The step plans of versions 0..upto-1 are
    drawn and uploaded up front (and shared between runs of the same series), so
    producing a version between timed checkpoints is device work of a few
    milliseconds rather than host work that leaves the GPU idle."""

    def __init__(self, series, device, plans=None, upto: int = 0):
        import torch
        self.series, self.device = series, device
        self.plans = {} if plans is None else plans
        self.buf = [torch.empty(series.nbytes, dtype=torch.uint8, device=device) for _ in range(2)]
        self.have = [-1, -1]
        for j in range(upto):
            self._plan(j)

    def _plan(self, j: int):
        from mtcheck_amd.workload import DevicePlan
        if j not in self.plans:
            self.plans[j] = DevicePlan(self.series, self.series.plan(j), self.device)
        return self.plans[j]

    def get(self, k: int):
        b = k % 2
        if self.have[b] > k:
            raise ValueError("versions are produced in increasing order")
        for j in range(self.have[b] + 1, k + 1):
            self.series.apply_device(self.buf[b], j, plan=self._plan(j))
        self.have[b] = k
        return self.buf[b]

    def release(self) -> None:
        self.buf = []


def served_check(clients, bufs, plans, step, K, rounds: int = 4, groups: int = 2) -> dict:
    """Served diffs against serial per-client dedup, untimed: the clients in
    `groups` halves, each on a fresh server (ids from 0) beside a serial
    deduplicator per client.  Every round submits all clients of the half at once
    (so their checkpoints run mixed in the server's batched steps), then checkpoints
    the same buffers serially and compares each client's whole diff on the device."""
    import torch

    import mtcheck_amd as M
    from mtcheck_amd.serve import Server
    from mtcheck_amd.workload import stream_key
    exact, checked = True, 0
    for g in range(groups):
        members = [i for i in range(len(clients)) if i % groups == g]
        srv = Server(lanes=2, max_batch=64, memory_budget=200 << 30)
        cids = {i: srv.add_client(clients[i].chunk_size, clients[i].nbytes) for i in members}
        serial = {i: M.TreeDeduplicator(clients[i].chunk_size) for i in members}
        for r in range(rounds):
            if r:
                for i in members:
                    step[i] += 1
                    plans[i][(step[i] - 1) % K].apply(bufs[i], stream_key(clients[i].seed, step[i]), False)
                torch.cuda.synchronize()
            tickets = [srv.submit(cids[i], bufs[i], make_baseline=(r == 0)) for i in members]
            for t in tickets:
                srv.result(t)
            for i in members:
                serial[i].checkpoint_device(bufs[i], make_baseline=(r == 0))
                exact &= bool(torch.equal(srv.client(cids[i]).copy_last_diff_device(),
                                          serial[i].copy_last_diff_device()))
                checked += 1
        for d in serial.values():
            d.close()
        srv.close()
        torch.cuda.empty_cache()
    return {"bit_exact_vs_serial": bool(exact), "checked_vs_serial": checked}


def host_path_runs(series, C: int, hk: int, warmup: int, steps: int, diff_max: int, tune, plans) -> dict:
    """The same checkpoints with each diff handed to pinned host memory: serially
    (checkpoint, then a synchronous copy) and overlapped (the copy of checkpoint k
    runs during checkpoint k+1; two pinned buffers, a buffer is reused once its copy
    has finished).  PCIe-inclusive rates; the device-resident rate is `value`.
    Versions are generated between the timed checkpoints; the first asynchronous
    copy (which allocates the handle's second diff arena) is made in the warm-up."""
    import torch

    import mtcheck_amd as M
    nbytes = series.nbytes
    cap = int(diff_max * 1.5) + (1 << 20)
    pinned = [torch.empty(cap, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
    out = {}
    for mode in ("serial", "overlapped"):
        vs = Versions(series, "cuda:0", plans)
        d = M.TreeDeduplicator(C, hash_kind=hk)
        if tune:
            d.set_tuning(*tune)
        for k in range(warmup):
            d.checkpoint_device(vs.get(k), make_baseline=(k == 0))
        if mode == "overlapped":
            d.copy_last_diff_async(pinned[0])
            d.wait_copies()
        torch.cuda.synchronize()
        copy_ms, elapsed = [], 0.0
        for j, k in enumerate(range(warmup, warmup + steps)):
            buf = vs.get(k)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d.checkpoint_device(buf)
            if mode == "serial":
                d.copy_last_diff_device(pinned[0])
                copy_ms.append(d.stats().ms_copy)
            else:
                if j >= 1:
                    d.wait_copies()  # the copy of checkpoint k-1 (into the other pinned buffer) is done
                d.copy_last_diff_async(pinned[(j + 1) % 2])
            elapsed += time.perf_counter() - t0
        t0 = time.perf_counter()
        d.wait_copies()
        elapsed += time.perf_counter() - t0
        out[f"{mode}_gbs"] = round(nbytes * steps / elapsed / 1e9, 2)
        if copy_ms:
            out["d2h_ms_mean"] = round(sum(copy_ms) / len(copy_ms), 3)
        d.close()
        vs.release()
        torch.cuda.empty_cache()
    return out


def plain_run(nbytes: int, C: int, hash_name: str, change: float, steps: int, warmup: int,
              load_factor: float = 0.125, tiles_per_block: int = 0, host_path: bool = False) -> dict:
    """Incremental checkpoints of one buffer resident in HBM.  Version k = version
    k-1 + step k, generated in one of two alternating buffers between the timed
    checkpoints; each timed checkpoint starts and ends on an idle device."""
    import numpy as np
    import torch

    import mtcheck_amd as M
    from mtcheck_amd.workload import CheckpointSeries
    hk = M.MD5 if hash_name == "md5" else M.MURMUR3
    n = (nbytes + C - 1) // C
    total = warmup + steps
    series = CheckpointSeries.with_change(nbytes, C, change)
    plans = {}
    vs = Versions(series, torch.device("cuda:0"), plans, upto=total)
    d = M.TreeDeduplicator(C, hash_kind=hk)
    if tiles_per_block or load_factor != 0.125:
        d.set_tuning(tiles_per_block or 1, load_factor, 8)
    for k in range(warmup):
        d.checkpoint_device(vs.get(k), make_baseline=(k == 0))
    torch.cuda.synchronize()
    lat, leaf_ms, changed, balg, dsize, leaf_bytes = [], [], [], [], [], []
    for k in range(warmup, total):
        buf = vs.get(k)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        hdr, _, size = d.checkpoint_device(buf)
        lat.append(time.perf_counter() - s0)
        st = d.stats()
        leaf_ms.append(st.ms_leaf)
        changed.append(st.changed_chunks)
        # Leaf kernel: C B read and the 16 B previous digest per chunk; per changed
        # chunk a 4 B list entry and, unless the compaction hashes it again, its
        # 16 B digest; a label byte per chunk only when it labels every chunk.
        leaf_bytes.append(nbytes + 16 * n + (n if st.leaf_labels_written else 0)
                          + (20 if st.leaf_digests_stored else 4) * st.changed_chunks)
        dsize.append(size)
        meta = 32 + 4 * hdr.num_first_ocur + 8 * hdr.num_prior_chkpts + 8 * hdr.num_shift_dupl
        balg.append(nbytes + 32 * n + 2 * (size - meta) + meta)
    elapsed = sum(lat)
    leaf_avg_ms = float(np.mean(leaf_ms))
    achieved = float(np.mean(leaf_bytes)) / (leaf_avg_ms / 1e3)
    d.close()
    vs.release()
    torch.cuda.empty_cache()
    extra = {}
    if host_path:
        tune = (tiles_per_block or 1, load_factor, 8) if (tiles_per_block or load_factor != 0.125) else None
        hp = host_path_runs(series, C, hk, warmup, steps, max(dsize), tune, plans)
        hp["d2h_gbs"] = round(float(np.mean(dsize)) / (hp["d2h_ms_mean"] / 1e3) / 1e9, 2) if hp.get("d2h_ms_mean") else None
        extra["host_path"] = {"workload": "the same checkpoints, each diff copied to pinned host memory", **hp}
    plans.clear()
    gc.collect()
    torch.cuda.empty_cache()
    prof = leaf_profile(nbytes, C, hash_name, change)
    roof = {"kernel": "k_leaf1 (hash + fixed test + label + list of chunks to insert)", "bound": "hbm",
            "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": leaf_traffic(nbytes, C, hash_name, change),
            "avg_ms": round(leaf_avg_ms, 4)}
    if prof:
        roof["profile"] = {"summary": prof[0], "avg_ms": prof[1],
                           "frac": round(float(np.mean(leaf_bytes)) / (prof[1] / 1e3) / HBM_PEAK, 4)}
    return {**extra,
        "value": round(nbytes * steps / elapsed / 1e9, 3), "ms_per_step": round(elapsed / steps * 1e3, 4),
        "config": {"workload": f"plain single-buffer dedup, {nbytes / 2**30:g} GiB uint8, {C} B chunks, "
                               f"{hash_name}, {change * 100:g}% change per checkpoint (2:1 fresh runs:copies)",
                   "size_bytes": nbytes, "chunk_size": C, "hash": hash_name, "change": change},
        "hbm_frac_checkpoint": round(float(np.mean(balg)) / (elapsed / steps) / HBM_PEAK, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "diff_mib_mean": round(float(np.mean(dsize)) / (1 << 20), 2),
        "roofline": roof,
    }


def sweep(args, emit=None):
    """The 8 GiB sweep: chunk sizes {64,128,512,4096} x {murmur3, md5} x change
    {0.1, 2, 20}%, each point a baseline, sweep_warmup - 1 untimed and sweep_steps
    timed incremental checkpoints.  Returns one record per point (and passes each to
    `emit` as it is measured)."""
    nbytes = int(args.size_gib * (1 << 30))
    out = []
    for C in (64, 128, 512, 4096):
        for h in ("murmur3", "md5"):
            for ch in (0.001, 0.02, 0.2):
                note(f"sweep point C={C} {h} change={ch}")
                try:
                    r = plain_run(nbytes, C, h, ch, args.sweep_steps, args.sweep_warmup)
                except Exception as e:  # a point that cannot run is reported, not fatal
                    rec = {"chunk": C, "hash": h, "change": ch, "error": str(e)[:200]}
                else:
                    rec = {"chunk": C, "hash": h, "change": ch, "value": r["value"], "unit": "GB/s",
                           "ms_per_step": r["ms_per_step"], "hbm_frac_checkpoint": r["hbm_frac_checkpoint"],
                           "leaf_gbs": r["roofline"]["achieved"], "leaf_frac": r["roofline"]["frac"],
                           "leaf_ms": r["roofline"]["avg_ms"], "diff_mib_mean": r["diff_mib_mean"]}
                out.append(rec)
                if emit:
                    emit(rec)
    return out


def small_series():
    """The 64 MiB, ten-checkpoint series (10% zero runs at checkpoint 0, then 2% fresh
    runs and 1% copies per checkpoint, 128 B chunks, Murmur3): device throughput over
    the nine incremental checkpoints, and the serial oracle's on the same series on
    one host core."""
    import numpy as np

    from mtcheck_amd.workload import CheckpointSeries
    from oracle import oracle as O
    nbytes, C = 64 << 20, 128
    r = plain_run(nbytes, C, "murmur3", 0.03, 9, 1)
    series = CheckpointSeries.with_change(nbytes, C, 0.03)
    orc = O.OracleDedup(C)
    host = np.zeros(nbytes, np.uint8)
    t = 0.0
    for k in range(10):
        series.apply_host(host, k)
        t0 = time.perf_counter()
        orc.checkpoint(host, make_baseline=(k == 0))
        if k:
            t += time.perf_counter() - t0
    return {"workload": "64 MiB uint8, 128 B chunks, murmur3, 10 checkpoints: 10% zero runs at checkpoint 0, "
                        "then 2% fresh runs and 1% copies per checkpoint; the 9 incremental ones timed",
            "value": r["value"], "unit": "GB/s", "ms_per_step": r["ms_per_step"],
            "leaf_frac": r["roofline"]["frac"], "cpu_serial_1core_gbs": round(nbytes * 9 / t / 1e9, 4)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(independent buffers shard one per GPU with no communication)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only_batched:
        print(json.dumps({"batched": batched_bench(args)}), flush=True)
        return
    if args.only_restart:
        print(json.dumps({"restart": restart_bench(args)}), flush=True)
        return
    if args.only_served:
        print(json.dumps({"served": served_bench(args)}), flush=True)
        return
    import mtcheck_amd as M
    if args.sweep:
        sweep(args, emit=lambda rec: print(json.dumps(rec), flush=True))
        return
    note("headline")
    r = plain_run(int(args.size_gib * (1 << 30)), args.chunk, args.hash, args.change, args.steps, args.warmup,
                  args.load_factor, args.tiles_per_block, host_path=not args.no_host_path)
    out = {
        "metric": METRIC, "value": r["value"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic", "config": r["config"],
        "hbm_frac_checkpoint": r["hbm_frac_checkpoint"], "latency_ms": r["latency_ms"],
        "diff_mib_mean": r["diff_mib_mean"], "roofline": r["roofline"],
    }
    if "host_path" in r:
        out["host_path"] = r["host_path"]
    hk = M.MD5 if args.hash == "md5" else M.MURMUR3
    if not args.no_small_series:
        note("small series")
        out["small_series"] = small_series()
    if not args.no_batched:
        note("batched")
        out["batched"] = batched_bench(args)
    if not args.no_restart:
        note("restart")
        out["restart"] = restart_bench(args)
    if not args.no_served:
        note("served")
        out["served"] = served_bench(args)
    if not args.no_sweep:
        out["sweep"] = {"workload": f"{args.size_gib:g} GiB uint8, the headline generator, chunk x hash x change; "
                                    f"per point a baseline and {args.sweep_steps} timed checkpoints",
                        "points": sweep(args)}
    if not args.no_cpu_baseline:
        note("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args, hk)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
