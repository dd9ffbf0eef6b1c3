"""Organ-like blob on nested covering voxel levels (h, 3h, 9h), pinned at
x < -5: the V-cycle meets the residual target every step; the network's
coarse correction applied to s alone, trained on recorded plain steps of the
same blob, saves half-cycles against plain MG; the learned schedule (temporal
start + network) stays within the target and within a bounded cost of the
temporal start, which its judge guarantees only step by step."""
import numpy as np
import pytest

import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu


def blob_solver(lv, schedule, **kw):
    base = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=W.BLOB["max_cycles"], check_every=2, tol=1e-3,
                rho_smooth=0.9, use_graph=1)
    base.update(kw)
    s = M.Solver(lv, schedule=schedule, if_mode=M.IF_STABLE, **base, **W.physics())
    fix = W.blob_fixflag(lv[0])
    s.set_fixed(fix)
    s.set_velocity(W.seeded_velocity(lv[0], fix))
    return s


@pytest.mark.parametrize("h", [0.45, W.BLOB_H])
def test_blob_vcycle_and_learned(h):
    """The blob at a coarser copy (h = 0.45) and at its workload size
    (h = 0.225, 194,148 tets), 120 steps per schedule.  The blob swings about
    its pinned end by up to 15 m and a few steps hold inverted tets next to
    the pinned region, where the polar decomposition's force field is not
    smooth and the Chebyshev extrapolation can make a check jump (the stall
    guard then sweeps without it).  Checked: finite positions, pinned
    vertices at rest, every step within the target, and in half-cycles per
    step:
    - the temporal start and the learned schedule below plain MG;
    - the network on s alone (warm_start = 2) at least half a half-cycle
      below plain MG: what the trained network itself gains;
    - the learned schedule at most 20 % above the temporal start.  Its judge
      keeps the network's correction only where it lowers that step's start
      residual, but which later steps are hard moves with the trajectory
      (the measured runs put it 0-15 % above or below the temporal start),
      so the bound is on the cost, not a gain."""
    lv = W.blob_hierarchy(h=h)
    assert lv[0].ntets > 20000
    if h == W.BLOB_H:
        assert lv[0].ntets == 194148
    nets = {}
    for name, warm in (("learned", 1), ("learned_s", 0)):
        rec = blob_solver(lv, M.SCHED_VCYCLE, warm_start=warm)
        rec.learn_record(150)
        rec.step(150)
        hist = rec.learn_train(epochs=15, batch=1024, lr=2e-3, seed=0)
        print(f"[blob h={h}] {name} training loss {hist[0]:.3f} -> {hist[-1]:.3f}")
        nets[name] = rec
    res = {}
    for name, sched, kw in (("plain", M.SCHED_VCYCLE, dict(check_every=1)),
                            ("temporal", M.SCHED_VCYCLE, dict(warm_start=1, coarse_first=1)),
                            ("learned_s", M.SCHED_LEARNED, dict(warm_start=2)),
                            ("learned", M.SCHED_LEARNED, dict(coarse_first=1))):
        s = blob_solver(lv, sched, **kw)
        if name in nets:
            s.learn_weights(nets[name].learn_weights())
            s.learn_norm(nets[name].learn_norm())
        s.step(120)
        run = s.run_stats()
        res[name] = run["total_halves"] / run["steps"]
        print(f"[blob h={h}] {name:9s} half-cycles/step {res[name]:.2f} fallbacks {run['fallbacks']} "
              f"steps over tol {run['failed_steps']} max rel {run['max_rel']:.2e}")
        assert run["steps"] == 120 and run["failed_steps"] == 0 and run["max_rel"] <= 1e-3
        x = s.positions()
        assert np.all(np.isfinite(x))
        pinned = W.blob_fixflag(lv[0]) == 0
        assert np.array_equal(x[pinned], lv[0].pos[pinned])
    assert res["temporal"] < res["plain"]
    assert res["learned"] < res["plain"]
    assert res["learned_s"] < res["plain"] - 0.5
    assert res["learned"] <= 1.2 * res["temporal"]
