"""Tiny trees on the device, bit-exact against the oracle through the C ABI:
one to three chunks (every pair of strings over a two-letter alphabet with 1-byte
chunks), one or two 128-byte chunks with short tails and copies, a data pointer
that is not 16-byte aligned, and an empty buffer (refused on the host).
Every test prints its case first.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import first_mismatch

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def check_sequence(versions, C, data_of=None):
    """Checkpoints of `versions` (uint8 arrays of one length) on the device and in
    the oracle, then a restart of each from the device's diffs."""
    d = M.TreeDeduplicator(C)
    orc = O.OracleDedup(C)
    diffs = []
    for k, v in enumerate(versions):
        data = data_of(v) if data_of else torch.from_numpy(v.copy()).to("cuda:0")
        diff = d.checkpoint(data, make_baseline=(k == 0))
        exp = orc.checkpoint(v, make_baseline=(k == 0))
        assert diff == exp, (k, first_mismatch(diff, exp))
        diffs.append(diff)
    out = torch.zeros(len(versions[0]), dtype=torch.uint8, device="cuda:0")
    for k, v in enumerate(versions):
        out.zero_()
        d.restart(out, diffs, k)
        assert np.array_equal(out.cpu().numpy(), v), k
    d.close()


@pytest.mark.parametrize("L", [1, 2, 3])
def test_every_pair_of_short_strings(L):
    say(f"tiny: all string pairs of length {L}, 1-byte chunks")
    words = [np.frombuffer("".join(w).encode(), np.uint8) for w in itertools.product("ab", repeat=L)]
    for s0, s1 in itertools.product(words, repeat=2):
        check_sequence([s0, s1, s0], 1)


@pytest.mark.parametrize("nbytes", [1, 100, 128, 129, 200, 256])
def test_one_or_two_chunks(nbytes):
    say(f"tiny: {nbytes} bytes, 128-byte chunks")
    rng = np.random.default_rng(nbytes)
    a = rng.integers(0, 256, nbytes, dtype=np.uint8)
    b = a.copy()
    b[0] ^= 0x5A                      # first chunk rewritten
    c = b.copy()
    if nbytes > 128:
        c[:nbytes - 128] = b[128:]    # the tail copied over the head
    check_sequence([a, b, c, a, a], 128)


def test_unaligned_data_pointer():
    """A buffer starting one byte into an allocation takes the byte-granular leaf path."""
    say("tiny: unaligned data pointer, 128-byte chunks")
    rng = np.random.default_rng(9)
    n = 64 * 128 + 17
    vs = [rng.integers(0, 256, n, dtype=np.uint8)]
    for k in range(3):
        v = vs[-1].copy()
        v[(k * 1000) % n:(k * 1000) % n + 300] = rng.integers(0, 256, 300, dtype=np.uint8)
        v[5000:5128] = v[128:256]
        vs.append(v)
    backing = torch.zeros(n + 1, dtype=torch.uint8, device="cuda:0")

    def data_of(v):
        backing[1:].copy_(torch.from_numpy(v.copy()))
        return backing[1:]

    assert backing[1:].data_ptr() % 16 != 0
    check_sequence(vs, 128, data_of)


def test_empty_buffer_is_refused():
    say("tiny: empty buffer")
    d = M.TreeDeduplicator(128)
    with pytest.raises(M.MtcError):
        d.checkpoint(torch.zeros(0, dtype=torch.uint8, device="cuda:0"), make_baseline=True)
    d.close()
