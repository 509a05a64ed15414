"""Restart of every checkpoint of a float64 simulation-state series from its diff
chain (device-resident and host diffs), byte-identical to the generated state."""
import numpy as np
import pytest

from oracle import oracle as O

torch = pytest.importorskip("torch")
import mtcheck_amd as M  # noqa: E402
from mtcheck_amd.workload import FloatStateSeries  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("nvalues,C", [(1 << 20, 128), (777_777, 64)])
def test_float_state_restart_every_checkpoint(nvalues, C):
    fs = FloatStateSeries(nvalues, seed=11, frac=0.02, run_max=512)
    vals = fs.initial()
    d = M.TreeDeduplicator(C)
    orc = O.OracleDedup(C)
    dev_diffs, host_diffs, states = [], [], []
    for k in range(6):
        say(f"float state n={nvalues} C={C} checkpoint {k}")
        if k:
            fs.perturb(vals, k)
        raw = vals.view(torch.uint8)
        d.checkpoint_device(raw, make_baseline=(k == 0))
        dev_diffs.append(d.copy_last_diff_device().clone())
        host_diffs.append(d.copy_last_diff())
        states.append(raw.clone())
        assert host_diffs[-1] == orc.checkpoint(raw.cpu().numpy(), make_baseline=(k == 0)), k
    out = torch.empty_like(states[0])
    for k in range(6):
        say(f"restart {k}")
        out.fill_(0)
        M.restart_into(out, dev_diffs[:k + 1], k)
        assert torch.equal(out, states[k]), k
        out.fill_(0)
        M.restart_into(out, host_diffs[:k + 1], k)
        assert torch.equal(out, states[k]), k


def test_perturbation_touches_the_planned_fraction():
    fs = FloatStateSeries(1 << 20, seed=3)
    v = fs.initial()
    w = v.clone()
    fs.perturb(w, 1)
    changed = (v != w).double().mean().item()
    assert 0.008 < changed <= 0.01 + 1e-6
