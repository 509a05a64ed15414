"""Asynchronous diff copies to pinned host memory (mtc_copy_diff_async): each
checkpoint's copy overlaps the next checkpoint, the handle alternates two device
diff arenas, and every copied diff equals the oracle's bytes and restarts."""
import numpy as np
import pytest
import torch

import mtcheck_amd as M
from mtcheck_amd.workload import CheckpointSeries
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("nbytes,C", [((1 << 22) + 333, 128), (1 << 21, 64)])
def test_async_copies_overlap_next_checkpoint(nbytes, C):
    series = CheckpointSeries(nbytes, C, seed=9, fresh_frac=0.05, copy_frac=0.03)
    d = M.TreeDeduplicator(C)
    orc = O.OracleDedup(C)
    host = np.zeros(nbytes, np.uint8)
    data = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    pinned, sizes, expect = [], [], []
    for k in range(7):
        series.apply_host(host, k)
        data.copy_(torch.from_numpy(host))
        say(f"async copy series nbytes={nbytes} C={C} step={k}")
        d.checkpoint_device(data, make_baseline=(k == 0))
        expect.append(orc.checkpoint(host, make_baseline=(k == 0)))
        # the device-side views still give the last diff while copies are in flight
        if k == 3:
            assert d.copy_last_diff() == expect[-1]
            assert bytes(d.copy_last_diff_device().cpu().numpy()) == expect[-1]
        buf = torch.empty(len(expect[-1]) + 64, dtype=torch.uint8, pin_memory=True)
        sizes.append(d.copy_last_diff_async(buf))
        pinned.append(buf)
    d.wait_copies()
    got = [bytes(b[:s].numpy()) for b, s in zip(pinned, sizes)]
    for k in range(7):
        assert got[k] == expect[k], f"diff {k}"
    out = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
    d.restart(out, got, 6)
    assert np.array_equal(out.cpu().numpy(), host)
    d.close()


def test_async_copy_rejects_small_buffer():
    d = M.TreeDeduplicator(128)
    data = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda:0")
    d.checkpoint_device(data, make_baseline=True)
    with pytest.raises(ValueError):
        d.copy_last_diff_async(torch.empty(8, dtype=torch.uint8, pin_memory=True))
    d.close()
