"""mtc_restart_noalloc against mtc_restart, through the C ABI.

The no-allocation restart keeps its scratch in one static device array per GPU,
which holds the previous call's contents; two restarts in a row, the second of a
smaller buffer, must each equal the allocating restart (and the buffer the
checkpoint was taken of) byte for byte.  Limits are refused with an error.
"""
import ctypes

import pytest

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd._lib import check, lib  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu

INVALID_ARG = 1


def say(msg):
    print(msg, flush=True)


def device_chain(nbytes, C, hk, steps, rebase_at=()):
    series = CheckpointSeries(nbytes, C, seed=nbytes % 1000 + C)
    buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    d = M.TreeDeduplicator(C, hash_kind=hk)
    chain, versions = [], []
    for k in range(steps):
        series.apply_device(buf, k)
        d.checkpoint_device(buf, make_baseline=(k == 0 or k in rebase_at))
        chain.append(d.copy_last_diff_device().clone())
        versions.append(buf.clone())
    d.close()
    return chain, versions


def restart(fn, chain, k, out):
    n = len(chain)
    ptrs = (ctypes.c_void_p * n)(*[c.data_ptr() for c in chain])
    sizes = (ctypes.c_uint64 * n)(*[c.numel() for c in chain])
    torch.cuda.synchronize()
    if fn == "noalloc":
        rc = lib().mtc_restart_noalloc(ptrs, sizes, n, k, ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_uint64(out.numel()), 0, None)
    else:
        rc = lib().mtc_restart(ptrs, sizes, n, k, 1, ctypes.c_void_p(out.data_ptr()), ctypes.c_uint64(out.numel()),
                               0, None)
    return rc


@pytest.mark.parametrize("first,second", [
    ((24 << 20, 128, M.MURMUR3, 5, ()), (3 * (1 << 20) + 77, 64, M.MD5, 4, (2,))),
])
def test_noalloc_restart_equals_restart_twice_in_a_row(first, second):
    for nbytes, C, hk, steps, rebase in (first, second):
        say(f"no-alloc restart: {nbytes} B, C={C}, hash={hk}, {steps} checkpoints, rebase at {rebase}")
        chain, versions = device_chain(nbytes, C, hk, steps, rebase)
        for k in range(steps):
            a = torch.full((nbytes,), 0xA5, dtype=torch.uint8, device="cuda:0")
            b = torch.full((nbytes,), 0x5A, dtype=torch.uint8, device="cuda:0")
            check(restart("alloc", chain, k, a))
            check(restart("noalloc", chain, k, b))
            assert torch.equal(a, b), (nbytes, k)
            assert torch.equal(b, versions[k]), (nbytes, k)


def test_noalloc_restart_refuses_what_does_not_fit():
    say("no-alloc restart: 160 MiB at 16-byte chunks (10 M chunks) needs more than 64 MiB of scratch")
    chain, versions = device_chain(160 << 20, 16, M.MURMUR3, 2)
    out = torch.empty(160 << 20, dtype=torch.uint8, device="cuda:0")
    assert restart("noalloc", chain, 1, out) == INVALID_ARG
    del chain, versions, out
    say("no-alloc restart: a chain of 257 diffs")
    chain, versions = device_chain(4096, 64, M.MURMUR3, 257)
    out = torch.empty(4096, dtype=torch.uint8, device="cuda:0")
    assert restart("noalloc", chain, 256, out) == INVALID_ARG  # ids 0..256: 257 diffs
    check(restart("noalloc", chain, 255, out))  # ids 0..255: 256 diffs
    assert torch.equal(out, versions[255])
    say("no-alloc restart: a valid call after the refusals")
    small, sv = device_chain(1 << 20, 128, M.MURMUR3, 3)
    out = torch.empty(1 << 20, dtype=torch.uint8, device="cuda:0")
    check(restart("noalloc", small, 2, out))
    assert torch.equal(out, sv[2])
