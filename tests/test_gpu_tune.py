"""The autotuner: launch shapes by kernel time, sweeps over a window of
steps, all from a snapshot of the solver state, which is restored afterwards
(the simulation continues exactly where it was); the chosen parameters keep
every trial step within the residual target and are never slower over the
window than the ones before tuning."""
import numpy as np
import pytest

import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(**kw):
    lv = W.beam_hierarchy((4, 1))
    s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, pre_sweeps=5, post_sweeps=5, coarse_sweeps=2,
                 max_cycles=12, check_every=2, tol=1e-3, rho_smooth=0.9, use_graph=1, warm_start=1, **kw,
                 **W.physics())
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(W.seeded_velocity(lv[0], W.top_fixflag(lv[0])))
    return s


def test_autotune_restores_state_and_picks_valid_parameters():
    a, b = make(), make()
    a.step(30)
    b.step(30)
    x0, v0 = a.positions().copy(), a.velocity().copy()
    r = a.autotune(trial_steps=8)
    print(f"[autotune] {r}")
    assert np.array_equal(a.positions(), x0) and np.array_equal(a.velocity(), v0)
    assert r["block_tet"] in (64, 128, 256) and r["block_vert"] in (64, 128, 256)
    assert r["pre_sweeps"] in (3, 4, 5, 6) and r["pre_sweeps"] == r["post_sweeps"]
    assert r["vert_lanes"] in (1, 2, 4, 8) and r["tile_gather"] in (0, 1) and r["fuse_sweeps"] in (0, 1)
    # 35 launch shapes (one of them the fused sweeps), the window for the old
    # and the new shape, three other sweep counts
    assert r["trials"] == 40 and r["ms_per_step"] <= r["ms_per_step_default"]
    assert 0 < r["sweep_us"] <= r["sweep_us_default"]
    assert a.params.pre_sweeps == r["pre_sweeps"] and a.params.tile_gather == r["tile_gather"]
    # with the original parameters restored the run continues as an untuned copy
    b.step(1)
    c = make()
    c.step(30)
    c.autotune(trial_steps=4)
    c.set_params(pre_sweeps=5, post_sweeps=5, block_tet=0, block_vert=0, vert_lanes=0, tile_gather=1, fuse_sweeps=1)
    c.step(1)
    assert np.array_equal(c.positions(), b.positions())
    # tuned steps meet the target
    a.run_stats(reset=True)
    a.step(40)
    run = a.run_stats()
    assert run["failed_steps"] == 0 and run["max_rel"] <= 1e-3
