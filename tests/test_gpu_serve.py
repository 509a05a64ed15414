"""Served checkpoints: every client's diffs are the same alone and in the mix of
concurrent clients; admission against the memory budget."""
import time

import pytest

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.serve import Server  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_served_results_equal_serial():
    sizes = [1 << 20, 3 << 20, 700_001, 2 << 20, 65_536, 5 << 20]
    series = [CheckpointSeries.with_change(nb, 128, 0.03, seed=50 + i) for i, nb in enumerate(sizes)]
    srv = Server(lanes=2, max_batch=4)
    cids = [srv.add_client(128, nb) for nb in sizes]
    bufs = [torch.empty(nb, dtype=torch.uint8, device="cuda:0") for nb in sizes]
    serial = [M.TreeDeduplicator(128) for _ in sizes]
    steps = 4
    for k in range(steps):
        say(f"serve step {k}")
        for s, b in zip(series, bufs):
            s.apply_device(b, k)
        torch.cuda.synchronize()
        tickets = [srv.submit(c, b, make_baseline=(k == 0)) for c, b in zip(cids, bufs)]
        got = []
        for c, t in zip(cids, tickets):
            h, size, ts, t0, t1 = srv.result(t)
            assert ts <= t0 <= t1
            d = srv.client(c).copy_last_diff()
            assert len(d) == size and M.parse_header(d)["chkpt_id"] == k
            got.append(d)
        exp = [d.checkpoint(b, make_baseline=(k == 0)) for d, b in zip(serial, bufs)]
        assert got == exp
    st = srv.stats()
    assert st["served"] == steps * len(sizes) and st["queued"] == 0
    srv.close()


def test_budget_admission_waits_for_memory():
    """Client B does not fit next to client A; its checkpoint runs once A is removed."""
    srv = Server(lanes=1, max_batch=8, memory_budget=20 << 20)  # one client's state is about 13 MiB
    a = srv.add_client(64, 1 << 20)
    b = srv.add_client(64, 1 << 20)
    buf = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda:0")
    say("admission: first client")
    srv.result(srv.submit(a, buf, make_baseline=True))
    tb = srv.submit(b, buf, make_baseline=True)
    time.sleep(0.5)
    assert srv.result(tb, wait=False) is None, "second client must wait for memory"
    say("admission: remove first client")
    srv.remove_client(a)
    h, size, *_ = srv.result(tb)
    assert h.num_first_ocur >= 1
    srv.close()


def test_queue_orders_each_clients_checkpoints():
    """Several queued checkpoints of one client run in submission order."""
    series = CheckpointSeries.with_change(1 << 20, 128, 0.05, seed=3)
    bufs = [series.device_buffer(k) for k in range(4)]
    srv = Server(lanes=2, max_batch=8)
    c = srv.add_client(128, 1 << 20)
    say("queued checkpoints of one client")
    tickets = [srv.submit(c, bufs[k], make_baseline=(k == 0)) for k in range(4)]
    ids = [srv.result(t)[0].chkpt_id for t in tickets]
    assert ids == [0, 1, 2, 3]
    srv.close()


def test_tables_grow_between_checkpoints():
    """A client whose table fills is grown by its lane after its result is posted;
    its diffs stay equal to a serial handle's, which grows inside its checkpoints."""
    nb = 64 << 10  # 512 chunks: a 4096-slot table, full after about seven steps at 40% change
    series = CheckpointSeries.with_change(nb, 128, 0.4, seed=77)
    srv = Server(lanes=2, max_batch=4)
    c = srv.add_client(128, nb)
    serial = M.TreeDeduplicator(128)
    buf = torch.empty(nb, dtype=torch.uint8, device="cuda:0")
    for k in range(14):
        say(f"serve growth step {k}")
        series.apply_device(buf, k)
        torch.cuda.synchronize()
        srv.result(srv.submit(c, buf, make_baseline=(k == 0)))
        got = srv.client(c).copy_last_diff()
        assert got == serial.checkpoint(buf, make_baseline=(k == 0)), k
        if k == 1:
            resident0 = srv.stats()["resident_bytes"]
    assert srv.client(c).stats().table_capacity > 4096
    assert serial.stats().table_capacity > 4096
    # the grown table (at least 4096 more 32-byte slots) counts against the budget
    assert srv.stats()["resident_bytes"] >= resident0 + 32 * 4096
    srv.close()
