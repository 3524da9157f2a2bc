"""What the trained closure gains over the plain viscosity.

The network committed with the package was trained on the device (by
inmsom_bt.closure_experiment) on snapshots of the 2048 x 1536, 1 km FP64
run taken before fine step meta['t_start'].  From that (held-out) time, the
4 km coarse model runs with and without the closure and both are scored
against the coarsened 1 km run.  The closure acts on momentum: it must
reduce the u and v errors while the SSH error stays close to the plain
model's (its strength was chosen on the training period as the largest
that does not raise the mean SSH error there)."""
import os

import numpy as np
import pytest

from conftest import ROOT
from inmsom_bt.closure import ClosureNet
from inmsom_bt.closure_experiment import load_params
from inmsom_bt.learned import FineTruth, rollout, snapshots
from inmsom_bt.workloads import single_basin

pytestmark = pytest.mark.gpu

PARAMS = os.path.join(ROOT, "inmsom-elfs_amd", "inmsom_bt", "data", "closure_cfg1.npz")
STEPS = 2000  # coarse 40 s steps (22 simulated hours)


@pytest.mark.skipif(not os.path.exists(PARAMS), reason="no trained closure parameters committed")
def test_trained_closure_beats_plain_viscosity():
    params, nm, meta = load_params(PARAMS)
    fine = single_basin(meta["workload"])
    truth = FineTruth(fine)
    cb = truth.coarse
    t_start = int(meta["t_start"])
    print(f"fine truth to step {t_start + truth.per_step * STEPS}", flush=True)
    start = snapshots(truth, [t_start])[0]
    truth.advance_to(t_start + truth.per_step * STEPS)
    tr = {STEPS: truth.coarse_state()}
    truth.close()
    with ClosureNet(1, cb.ny, cb.nx) as net:
        net.set_output(meta.get("output", "uv"))
        net.set_params(params)
        net.set_norm(nm.shift, nm.scale, nm.out_scale)
        plain = rollout(cb, start, STEPS, tr, None, mode="fp64")
        learned = rollout(cb, start, STEPS, tr, net, ratio=1e300, mode="fp64")
    print(f"ssh rel err vs coarsened truth after {STEPS} steps: plain {plain['ssh'][-1]:.4f}, "
          f"learned {learned['ssh'][-1]:.4f} (u {plain['u'][-1]:.4f} -> {learned['u'][-1]:.4f}, "
          f"v {plain['v'][-1]:.4f} -> {learned['v'][-1]:.4f}); fallback steps {learned['fallback_steps']}",
          flush=True)
    assert learned["bad"] == 0
    assert learned["u"][-1] < 0.98 * plain["u"][-1]
    assert learned["v"][-1] < 0.98 * plain["v"][-1]
    assert learned["ssh"][-1] < 1.15 * plain["ssh"][-1]
