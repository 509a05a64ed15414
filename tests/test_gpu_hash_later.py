"""Steps whose leaf kernel stores no digest, with k_compact hashing the listed
chunks again from their bytes (the mode a large buffer of 128-byte chunks takes
after a checkpoint that changed few chunks; forced here for every buffer and chunk
size through MTC_HASH_LATER_MIN_CHUNKS=0), write the oracle's diffs byte for byte:
chunk sizes from 16 to 4096 bytes, chunk counts that are not powers of two and
short last chunks."""
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


@pytest.mark.parametrize("nbytes,C", [(300_007, 128), (1 << 20, 128), (1000 * 16 + 9, 16),
                                      (77 * 4096 + 100, 4096), (5 * 64 * 1024, 64), (3 << 20, 512)])
def test_hash_later_series_equal_oracle(nbytes, C, monkeypatch):
    monkeypatch.setenv("MTC_HASH_LATER_MIN_CHUNKS", "0")
    say(f"hash later nbytes={nbytes} C={C}")
    # 2% fresh runs and 1% copies: under 1/16 of the chunks change, so every step
    # after the first incremental one hashes its listed chunks again
    series = CheckpointSeries(nbytes, C, seed=nbytes % 89, fresh_frac=0.02, copy_frac=0.01)
    host = np.zeros(nbytes, np.uint8)
    d = M.TreeDeduplicator(C)
    orc = O.OracleDedup(C, M.MURMUR3)
    for k in range(6):
        series.apply_host(host, k)
        g = d.checkpoint(torch.from_numpy(host.copy()).to("cuda:0"), make_baseline=(k == 0))
        e = orc.checkpoint(host, make_baseline=(k == 0))
        assert g == e, (k, first_mismatch(g, e))
    d.close()
