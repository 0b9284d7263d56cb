"""The learned closure on its own workload (needs a GPU): a 64^3 FP32 coarse
model trained on 2x-filtered 128^3 FP64 trajectories whose activity zeta is
drawn uniformly from [0.5, 2] x the default, judged on held-out trajectories
(their own seeds and zeta) against the filtered fine run.

Asserts what the trained network gains over the coarse method without it:
the held-out one-step error and the 40-coarse-step rollout error are both
smaller with the closure than without it (the fine 128^3 FP64 run being the
judge of both).
"""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_closure_gain_on_heldout_trajectories():
    from lbman3d_amd import closure
    print("case C4 closure workflow 64^3 / 128^3", flush=True)
    rep = closure.run(n_fine=128, samples=32, train_traj=4, heldout=4, heldout_traj=2, warmup=1000, iters=1000,
                      plain_grids=(64, 128), log=lambda s: print(s, flush=True))
    summary = {k: rep[k] for k in ("onestep_relL2_no_closure", "onestep_relL2_closure", "rollout_relL2_no_closure",
                                   "rollout_relL2_closure", "fp64_plain_by_grid", "heldout_trajectories")}
    print(json.dumps(summary), flush=True)
    assert rep["onestep_relL2_closure"] < rep["onestep_relL2_no_closure"]
    assert rep["rollout_relL2_closure"][-1] < rep["rollout_relL2_no_closure"][-1]
    # every held-out trajectory gains, not only the mean
    for no, cl in zip(rep["rollout_per_trajectory_no_closure"], rep["rollout_per_trajectory_closure"]):
        assert cl[-1] < no[-1]
