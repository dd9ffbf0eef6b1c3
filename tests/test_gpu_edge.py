"""Edge cases of the device solver: a body at rest with no load (g(s) = 0,
so every V-cycle step converges at its start check), a fully pinned body, a
single tet against the oracle, and invalid arguments that must come back as
error codes."""
import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import MgsbError
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

VC = dict(pre_sweeps=3, post_sweeps=3, coarse_sweeps=2, max_cycles=8, tol=1e-3, rho_smooth=0.9)


@pytest.mark.parametrize("warm,check", [(0, 1), (1, 2), (1, 1)])
def test_rest_without_load_converges_at_start(warm, check):
    """No gravity, zero velocity, x at rest: s = x and g(s) = 0.  Each step
    ends at its start check with no half-cycle, the positions stay exactly
    at rest (a rejected or skipped extrapolation must leave x = s), and the
    velocities stay zero."""
    lv = W.beam_hierarchy((4, 1))
    s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, warm_start=warm, check_every=check, use_graph=1,
                 **VC, **W.physics(gravity=(0.0, 0.0, 0.0)))
    s.set_fixed(W.top_fixflag(lv[0]))
    for _ in range(5):
        st = s.step(1)
        assert st.halves == 0 and st.cycles == 0
    run = s.run_stats()
    assert run["steps"] == 5 and run["failed_steps"] == 0 and run["total_halves"] == 0
    assert np.array_equal(s.positions(), lv[0].pos)
    assert np.abs(s.velocity()).max() == 0.0


def test_fully_pinned_body_does_not_move():
    """Every vertex pinned, gravity on: no free DOF, so the residual over free
    DOFs is zero and nothing moves under any schedule."""
    lv = W.beam_hierarchy((4, 1))
    pinned = np.zeros(lv[0].nverts, np.uint8)
    for sched, kw in ((M.SCHED_SINGLE, dict(iters=10)), (M.SCHED_VCYCLE, dict(VC, check_every=2, warm_start=1))):
        levels = lv[:1] if sched == M.SCHED_SINGLE else lv
        s = M.Solver(levels, schedule=sched, if_mode=M.IF_STABLE, use_graph=1, **kw, **W.physics())
        s.set_fixed(pinned)
        s.set_velocity(np.ones((lv[0].nverts, 3), np.float32))
        s.step(4)
        assert np.array_equal(s.positions(), lv[0].pos), sched
        assert np.isfinite(s.velocity()).all()


def test_single_tet_matches_oracle():
    """One free corner (three pinned) of a single tet, falling under gravity
    through the reference's single-level schedule (Ac local step)."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.3, 1.0]], np.float32)
    tets = np.array([[0, 1, 2, 3]], np.int32)
    mesh = M.Mesh(pos, tets)
    fix = np.array([0, 0, 0, 1], np.uint8)
    phys = W.physics()
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=10, **phys)  # the reference's Ac local step
    s.set_fixed(fix)
    s.step(30)
    L = O.Level(pos, tets, mesh.tris, stiffness=phys["stiffness"], density=phys["density"])
    L.set_fix(fix.astype(np.float32))
    L.step(O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], iters=10), 30)
    g, o = s.positions(), L.positions()
    assert np.abs(g[:3] - pos[:3]).max() == 0.0
    err = float(np.abs(g - o).max() / np.abs(o).max())
    print(f"[single tet] moved {np.abs(g[3] - pos[3]).max():.3e}, rel err vs oracle {err:.2e}")
    assert np.abs(g[3] - pos[3]).max() > 1e-3 and err < 1e-4


def test_tet_soup_tiles_over_256_vertices_match_oracle():
    """600 disjoint tets (2,400 vertices): every tile of 256 consecutive tets
    holds 1,024 vertices, more than a workgroup's lanes, so the tiled local
    step's lanes also take the tile's vertices past the padded rows (the
    overflow path).  One corner of each tet pinned, the others fall and
    stretch under gravity through the reference's single-level schedule;
    positions against the oracle."""
    rng = np.random.default_rng(3)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.3, 1.0]], np.float32)
    n = 600
    off = np.stack([np.arange(n) % 25 * 2.0, np.arange(n) // 25 * 2.0, np.zeros(n)], 1).astype(np.float32)
    pos = (base[None] * rng.uniform(0.8, 1.2, (n, 1, 1)).astype(np.float32) + off[:, None]).reshape(-1, 3)
    pos = pos.astype(np.float32)
    tets = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    mesh = M.Mesh(pos, tets)
    fix = np.ones(4 * n, np.uint8)
    fix[0::4] = 0
    phys = W.physics()
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=10, **phys)
    s.set_fixed(fix)
    s.step(20)
    L = O.Level(pos, tets, mesh.tris, stiffness=phys["stiffness"], density=phys["density"])
    L.set_fix(fix.astype(np.float32))
    L.step(O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], iters=10), 20)
    g, o = s.positions(), L.positions()
    assert np.array_equal(g[0::4], pos[0::4])
    moved = float(np.abs(g - pos).max())
    err = float(np.abs(g - o).max() / np.abs(o).max())
    print(f"[tet soup] moved {moved:.3e}, rel err vs oracle {err:.2e}")
    assert moved > 1e-3 and err < 1e-4


@pytest.mark.parametrize("bad", [dict(vert_lanes=3), dict(block_vert=100), dict(warm_start=5), dict(learn_gate=2)])
def test_invalid_parameters_are_error_codes(bad):
    lv = W.beam_hierarchy((4, 1))
    with pytest.raises(MgsbError):
        M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, **VC, **bad, **W.physics())
    s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, **VC, **W.physics())
    with pytest.raises(MgsbError):
        s.set_params(**bad)
    s.step(1)  # the handle stays usable with its old parameters


def test_empty_mesh_is_an_error():
    mesh = M.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 4), np.int32), np.zeros((0, 3), np.uint32))
    with pytest.raises(MgsbError):
        M.Solver([mesh], schedule=M.SCHED_SINGLE, **W.physics())
