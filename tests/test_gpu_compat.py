"""mtc_set_compat: the library's diffs with the reference's own bytes where it
departs from them on purpose -- the reference's leaf range of a region root
(include/kokkos_merkle_tree.hpp:102-128) and no shifted root 0
(include/reference_impl.hpp:105-106) -- against the oracle's compat_ref mode, byte
for byte, on chunk counts that are not powers of two and on a buffer equal to an
earlier checkpoint; on power-of-two counts the two modes write the same bytes."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import first_mismatch

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.workload import CheckpointSeries  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run(bufs, C, compat, hk=M.MURMUR3):
    d = M.TreeDeduplicator(C, hash_kind=hk)
    d.set_compat(compat)
    orc = O.OracleDedup(C, hk, compat_ref=compat)
    out = []
    for k, b in enumerate(bufs):
        a = np.frombuffer(b, np.uint8) if isinstance(b, (bytes, bytearray)) else b
        g = d.checkpoint(torch.from_numpy(a.copy()).to("cuda:0"), make_baseline=(k == 0))
        e = orc.checkpoint(a, make_baseline=(k == 0))
        assert g == e, (compat, k, first_mismatch(g, e))
        out.append(g)
    d.close()
    return out


@pytest.mark.parametrize("compat", [False, True])
def test_the_two_departures(compat):
    say(f"departures compat={compat}")
    mixed = run([b"ABCDE", b"VWXYZ"], 1, compat)
    assert O.parse_diff(mixed[1])["data"] == (b"YZ" if compat else b"VWXYZ")
    shifted = run([b"ABCDEFGH", b"IJKLMNOP", b"ABCDEFGH"], 1, compat)
    assert O.parse_header(shifted[2])["num_shift_dupl"] == (0 if compat else 1)


@pytest.mark.parametrize("nbytes,C,hk", [(300_007, 128, M.MURMUR3), (5 * 128 + 3, 128, M.MURMUR3),
                                         (7 * 64, 64, M.MD5), (1000 * 16 + 9, 16, M.MURMUR3),
                                         (3 << 20, 128, M.MURMUR3)])
def test_compat_series_equal_oracle_compat(nbytes, C, hk):
    say(f"compat series nbytes={nbytes} C={C} hk={hk}")
    series = CheckpointSeries(nbytes, C, seed=nbytes % 97, fresh_frac=0.05, copy_frac=0.05)
    host = np.zeros(nbytes, np.uint8)
    bufs = []
    for k in range(5):
        series.apply_host(host, k)
        bufs.append(host.copy())
    bufs.append(bufs[2].copy())  # equal to an earlier checkpoint: node 0 shifted
    run(bufs, C, True, hk)


def test_compat_identical_on_power_of_two():
    say("compat on a power-of-two chunk count")
    series = CheckpointSeries(1 << 20, 128, seed=3)
    host = np.zeros(1 << 20, np.uint8)
    bufs = []
    for k in range(4):
        series.apply_host(host, k)
        bufs.append(host.copy())
    assert run(bufs, 128, True) == run(bufs, 128, False)
