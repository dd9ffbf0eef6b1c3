"""Device V-cycle (MGSB_SCHED_VCYCLE) against the oracle's restatement of the
same schedule: trajectories with a fixed cycle count, adaptive cycle counts
and the per-step residual target, per-level rho, graph replay."""
import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def make_pair(ns, steps_kw, fp64=False, gravity=True, **vkw):
    lv = W.beam_hierarchy(ns)
    phys = W.physics(gravity=(0, -9.8, 0) if gravity else (0, 0, 0))
    fix = W.top_fixflag(lv[0])
    v0 = W.seeded_velocity(lv[0], fix)
    s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, **vkw, **phys)
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(v0)
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=fp64)
          for m in lv]
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(v0)
    ML = O.MultiLevel(Ls, relax_coarse=vkw.get("relax_coarse", 0.6), dt=phys["dt"])
    p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], gravity=phys["gravity"],
                 if_mode=3)
    return lv, s, ML, Ls, p


@pytest.mark.parametrize("ns,rs", [((4, 1), 0.0), ((8, 4, 1), 0.9)])
def test_vcycle_fixed_cycles_matches_oracle(ns, rs):
    kw = dict(pre_sweeps=2, post_sweeps=2, coarse_sweeps=2, max_cycles=3, check_every=0, rho_smooth=rs)
    lv, s, ML, Ls, p = make_pair(ns, None, **kw)
    for l in range(1, len(ns)):
        assert abs(s.level_rho(l) - ML.rho(l)) < 1e-4 * ML.rho(l)
    assert s.coarse_nfree() == ML.nfree()
    vp = O.vparams(pre=2, post=2, coarse_sweeps=2, max_cycles=3, check=0, rho_smooth=rs)
    s.step(60)
    ML.step(p, vp, 60)
    # FP64 yardstick of the same schedule
    lv64, _, ML64, Ls64, p64 = make_pair(ns, None, fp64=True, **kw)
    ML64.step(p64, vp, 60)
    g, o, o64 = s.positions(), Ls[0].positions(), Ls64[0].positions()
    err = rel_err(g, o)
    st = s.stats
    r32, r64 = ML.stats()["rel"], ML64.stats()["rel"]
    print(f"[vcycle fixed {ns}] rel(gpu,o32)={err:.3e} rel(gpu,o64)={rel_err(g, o64):.3e} "
          f"rel res gpu={st.rel:.3e} o32={r32:.3e} o64={r64:.3e}")
    assert st.cycles == 3
    assert err < REL_TOL and rel_err(g, o64) < REL_TOL
    # the device iterates on x = s + u (u the in-step correction), so its
    # residual tracks the FP64 schedule, below the absolute-FP32 floor
    assert st.rel <= max(2 * r64, 1e-5), (st.rel, r64)


@pytest.mark.parametrize("lanes", [1, 2, 8])
def test_vcycle_gather_lanes_match_oracle(lanes):
    """Rows of corner forces gathered by 1, 2 or 8 lanes per vertex (the
    default, 4, runs in every other test): the same schedule, the same
    oracle trajectory, and the same residual-norm partials consumed by the
    adaptive checks (a wider launch gives more partials)."""
    kw = dict(pre_sweeps=2, post_sweeps=2, coarse_sweeps=2, max_cycles=8, check_every=2, rho_smooth=0.9,
              tol=1e-3, vert_lanes=lanes)
    lv, s, ML, Ls, p = make_pair((8, 4, 1), None, **kw)
    vp = O.vparams(pre=2, post=2, coarse_sweeps=2, max_cycles=8, check=2, rho_smooth=0.9, tol=1e-3)
    halves_g, halves_o = [], []
    for _ in range(30):
        st = s.step(1)
        ML.step(p, vp, 1)
        halves_g.append(st.halves)
        halves_o.append(ML.stats()["halves"])
        assert st.rel <= 1e-3
    err = rel_err(s.positions(), Ls[0].positions())
    print(f"[lanes {lanes}] rel(gpu,oracle)={err:.2e} halves gpu {sum(halves_g)} oracle {sum(halves_o)}")
    assert err < REL_TOL
    assert sum(a == b for a, b in zip(halves_g, halves_o)) >= 27


@pytest.mark.parametrize("check,cfirst", [(1, 0), (2, 0), (1, 1), (2, 1)])
def test_vcycle_adaptive_meets_target_each_step(check, cfirst):
    """Adaptive cycle counts against the oracle's, in the textbook order and
    with coarse_first (the first half-cycle the coarse correction of g(s))."""
    kw = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=check, tol=1e-3,
              rho_smooth=0.9, coarse_first=cfirst)
    lv, s, ML, Ls, p = make_pair((4, 1), None, **kw)
    vp = O.vparams(pre=5, post=5, coarse_sweeps=2, max_cycles=12, check=check, tol=1e-3, rho_smooth=0.9,
                   coarse_first=cfirst)
    gc, oc = [], []
    for _ in range(80):
        st = s.step(1)
        ML.step(p, vp, 1)
        o = ML.stats()
        gc.append((st.cycles, st.halves))
        oc.append((o["cycles"], o["halves"]))
        assert st.rel <= 1e-3, st.rel
        if check == 1:
            assert st.halves == 2 * st.cycles - cfirst
    agree = np.mean([a == b for a, b in zip(gc, oc)])
    print(f"[vcycle adaptive check={check} coarse_first={cfirst}] halves gpu mean {np.mean([h for _, h in gc]):.2f} "
          f"oracle mean {np.mean([h for _, h in oc]):.2f} agree {agree:.2f}")
    assert agree >= 0.9
    assert rel_err(s.positions(), Ls[0].positions()) < REL_TOL
    run = s.run_stats()
    assert run["steps"] == 80 and run["failed_steps"] == 0
    assert run["total_halves"] == sum(h for _, h in gc)
    assert run["total_cycles"] == sum(c for c, _ in gc)


@pytest.mark.parametrize("ns", [(8, 4, 1), (12, 4, 1)])
def test_vcycle_dense_tail_matches_sequential_tail(ns):
    """coarse_dense = 1 applies the coarse levels below level 1 (sweeps,
    restriction, the direct coarsest solve, prolongation, sweeps) as one
    dense matrix formed at set-up; coarse_dense = 0 runs them in sequence.
    The map is linear, so the two agree but for FP32 rounding: the same
    half-cycle counts, positions within 1e-5 of each other, and the same
    trajectory as the oracle; bitwise equality would mean the dense path
    never ran."""
    kw = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=2, tol=1e-3, rho_smooth=0.9)
    runs = []
    for dense in (1, 0):
        lv, s, ML, Ls, p = make_pair(ns, None, coarse_dense=dense, **kw)
        hs = [s.step(1).halves for _ in range(40)]
        runs.append((s.positions(), hs, s.run_stats()))
    (g1, h1, r1), (g0, h0, r0) = runs
    vp = O.vparams(pre=5, post=5, coarse_sweeps=2, max_cycles=12, check=2, tol=1e-3, rho_smooth=0.9)
    ML.step(p, vp, 40)
    err, erro = rel_err(g1, g0), rel_err(g1, Ls[0].positions())
    agree = sum(a == b for a, b in zip(h1, h0))
    print(f"[dense tail {ns}] rel(dense, sequential)={err:.2e} rel(dense, oracle)={erro:.2e} halves agree {agree}/40")
    assert not np.array_equal(g1, g0)
    assert err < 1e-5 and erro < REL_TOL
    assert agree >= 38
    assert r1["failed_steps"] == 0 and r0["failed_steps"] == 0


@pytest.mark.parametrize("check", [1, 2])
def test_vcycle_graph_replay_matches_eager(check):
    """Captured pieces (start, half-cycles, finish) replayed by the host's
    cycle loop give the eager trajectory bit for bit."""
    lv = W.beam_hierarchy((4, 1))
    fix = W.top_fixflag(lv[0])
    v0 = W.seeded_velocity(lv[0], fix)
    res = []
    for g in (0, 1):
        s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, max_cycles=8, check_every=check, tol=1e-3,
                     use_graph=g, **W.physics())
        s.fix_plane(1, 10.0, 1e-5)
        s.set_velocity(v0)
        for _ in range(4):
            st = s.step(5)
            s.sync()
            res.append((s.positions(), st.cycles, st.rel, st.halves))
    n = len(res) // 2
    for a, b in zip(res[:n], res[n:]):
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


@pytest.mark.parametrize("warm", [0, 1])
def test_vcycle_collision_matches_oracle(warm):
    """The large-box collider set on the small box: no gravity, a sphere
    (r 1.5, moving -x at 2 m/s through y = 0) and the ground plane y = -12,
    penalty 1e6; adaptive V-cycles, with and without the temporal start."""
    lv = W.beam_hierarchy((4, 1))
    phys = W.physics(gravity=(0, 0, 0))
    c = W.large_colliders()
    sp = c["sphere"]
    fix = W.top_fixflag(lv[0])
    v0 = W.seeded_velocity(lv[0], fix)
    kw = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=2, tol=1e-3, rho_smooth=0.9)
    s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, warm_start=warm, kcol=c["k"], **kw, **phys)
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(v0)
    s.set_sphere(0, sp["center"], sp["radius"], sp["vel"])
    s.set_plane(c["plane"]["normal"], c["plane"]["d"], c["k"])
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"]) for m in lv]
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(v0)
    ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
    p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], gravity=(0, 0, 0), if_mode=3,
                 kcol=c["k"], spheres=[(*sp["center"], sp["radius"], 1, *sp["vel"])],
                 plane=(*c["plane"]["normal"], c["plane"]["d"], c["k"]))
    vp = O.vparams(pre=5, post=5, coarse_sweeps=2, max_cycles=12, check=2, tol=1e-3, rho_smooth=0.9)
    u1 = u2 = np.zeros((lv[0].nverts, 3), np.float32)
    agree = 0
    for k in range(90):
        st = s.step(1)
        if warm:
            ML.set_warm(None, (fix[:, None] * (2 * u1 - u2)).astype(np.float32))
        ML.step(p, vp, 1)
        if warm:
            u2, u1 = u1, (Ls[0].positions() - Ls[0].get(O.S).reshape(-1, 3)).astype(np.float32)
        o = ML.stats()
        agree += (st.halves, st.fallback) == (o["halves"], o["fallback"])
        assert st.rel <= 1e-3, (k, st.rel)
    g, x = s.positions(), Ls[0].positions()
    moved = np.abs(x - lv[0].pos).max()
    print(f"[vcycle collision warm={warm}] rel={rel_err(g, x):.2e} agree {agree / 90:.2f} max displacement {moved:.2f}")
    assert moved > 1.0, "the sphere never pushed the beam"
    assert rel_err(g, x) < REL_TOL and agree >= 81


def test_large_box_contact_rim_converges():
    """The large box (24x240x24, 829,440 tets, 4 levels) with its sphere and
    plane at penalty 1e6, temporal start, 30-cycle cap.  V-cycle sweeps take
    an exact 3x3 block Jacobi step on contact vertices.  With the Chebyshev
    step there the contact rim stalled (steps ended at the cap near 0.2); with
    the per-axis diagonal k n_c^2 the rim could not slide along the sphere
    (most steps ended at the cap near 3e-3).  Every step must meet the
    target in a few cycles."""
    lv = W.beam_hierarchy((24, 12, 4, 1))
    phys = W.physics(gravity=(0, 0, 0))
    c = W.large_colliders()
    sp = c["sphere"]
    fix = W.top_fixflag(lv[0])
    kw = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=30, check_every=2, tol=1e-3, rho_smooth=0.9,
              use_graph=1)
    s = M.Solver(lv, schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, warm_start=1, kcol=c["k"], **kw, **phys)
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(W.seeded_velocity(lv[0], fix))
    s.set_sphere(0, sp["center"], sp["radius"], sp["vel"])
    s.set_plane(c["plane"]["normal"], c["plane"]["d"], c["k"])
    s.step(10)
    s.run_stats(reset=True)
    s.step(40)
    r = s.run_stats()
    print(f"[large box contact] steps {r['steps']} over tol {r['failed_steps']} max rel {r['max_rel']:.2e} "
          f"half-cycles/step {r['total_halves'] / r['steps']:.1f}")
    assert r["failed_steps"] == 0 and r["max_rel"] <= 1e-3
    assert r["total_halves"] / r["steps"] < 12


def _fuse_pair_solver(case, fuse):
    if case == "bodies":
        b = W.batched(nbodies=5, seed=7, sizes=(2, 4))
        lv = b["levels"]
    else:
        lv = W.beam_hierarchy((8, 4, 1))
    grav = (0, 0, 0) if case == "collision" else (0, -9.8, 0)
    phys = W.physics(gravity=grav)
    kw = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=2, tol=1e-3, rho_smooth=0.9,
              fuse_sweeps=fuse, use_graph=1)
    if case in ("temporal_cf", "collision", "bodies"):
        kw.update(warm_start=1, coarse_first=1)
    if case == "learned_rejected":
        kw.update(coarse_first=1, learn_gate=0)
    if case == "plain":
        kw.update(check_every=1)
    if case == "collision":
        kw.update(kcol=W.large_colliders()["k"])
    sched = M.SCHED_LEARNED if case == "learned_rejected" else M.SCHED_VCYCLE
    s = M.Solver(lv, schedule=sched, if_mode=M.IF_STABLE, **kw, **phys)
    if case == "learned_rejected":  # weights far too large: the judge rejects every warm start
        n = s.learn_nparams()
        s.learn_weights((50.0 * np.random.default_rng(5).normal(size=n)).astype(np.float32))
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(W.seeded_velocity(lv[0], W.top_fixflag(lv[0])))
    if case == "bodies":
        s.set_body_stiffness(b["E"])
    if case == "collision":
        c = W.large_colliders()
        s.set_sphere(0, c["sphere"]["center"], c["sphere"]["radius"], c["sphere"]["vel"])
        s.set_plane(c["plane"]["normal"], c["plane"]["d"], c["k"])
    return s


@pytest.mark.parametrize("case", ["plain", "temporal_cf", "collision", "bodies", "learned_rejected"])
def test_fused_sweeps_bit_identical(case):
    """fuse_sweeps = 1 runs each fine-level sweep's vertex update (and the
    coarse correction's prolongation) inside the next local step's launch,
    every tile updating its own vertices.  Every copy of a vertex's update
    reads the same inputs in the same order, so the trajectory, the collision
    output and the run statistics equal those of separate launches bit for
    bit: the plain V-cycle, the temporal coarse-first start, a sphere and a
    plane at penalty 1e6, five bodies with per-body checks, and a learned
    start the judge rejects on every step.  With fused sweeps the warm start
    is one launch and its judge runs inside the first coarse half's
    restriction (a rejection restricts g(s) and prolongs from u = 0); with
    separate launches the judge is its own kernel that returns x to s."""
    a, b = _fuse_pair_solver(case, 0), _fuse_pair_solver(case, 1)
    us = b.time_kernel(20, kernel=1)
    halves = []
    for k in range(40):
        sa, sb = a.step(1), b.step(1)
        assert (sa.halves, sa.cycles, sa.fallback) == (sb.halves, sb.cycles, sb.fallback), k
        assert sa.rel == sb.rel, (k, sa.rel, sb.rel)
        halves.append(sb.halves)
    xa, xb = a.positions(), b.positions()
    print(f"[fused {case}] 40 steps, half-cycles {sum(halves)}, max |dx| {np.abs(xa - xb).max():.3e}, "
          f"fused sweep {us:.2f} us")
    assert np.array_equal(xa, xb)
    assert np.array_equal(a.velocity(), b.velocity())
    assert a.run_stats() == b.run_stats()
    if case == "learned_rejected":
        assert a.run_stats()["fallbacks"] >= 30
    if case == "collision":
        ca, cb = a.collision(), b.collision()
        for q, r in zip(ca, cb):
            assert np.array_equal(q, r)
    # the unfused solver has no fused sweep to time
    with pytest.raises(M.MgsbError):
        a.time_kernel(5, kernel=1)


@pytest.mark.parametrize("graph", [0, 1])
def test_stall_guard_matches_oracle_on_a_diverging_smoother(graph):
    """rho_smooth 1.1 over-estimates the fine smoother's spectral radius: the
    Chebyshev weights pass 2 and the first sweeps diverge, the first check's
    residual jumps past 3 times the start's, and the guard runs the rest of
    the step without extrapolation.  The device takes the same decisions as
    the oracle (half-cycle counts), meets the target every step and stays on
    its trajectory; the step's check history shows the jump."""
    kw = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=2, tol=1e-3,
              rho_smooth=1.1, stall_guard=1, use_graph=graph)
    lv, s, ML, Ls, p = make_pair((4, 1), None, **kw)
    vp = O.vparams(pre=5, post=5, coarse_sweeps=2, max_cycles=12, check=2, tol=1e-3, rho_smooth=1.1, stall_guard=1)
    hg, ho = [], []
    for k in range(20):
        st = s.step(1)
        ML.step(p, vp, 1)
        hg.append(st.halves)
        ho.append(ML.stats()["halves"])
        assert st.rel <= 1e-3, (k, st.rel)
        h = s.step_history()
        assert h[0] == 1.0 and abs(h[-1] - st.rel) <= 1e-12 and len(h) == 1 + st.halves
        assert np.any(h[1:] > 3 * h[:-1]), h
    err = rel_err(s.positions(), Ls[0].positions())
    print(f"[stall guard graph={graph}] halves gpu {hg} oracle {ho} rel(gpu,oracle)={err:.2e}")
    assert sum(a == b for a, b in zip(hg, ho)) >= 18
    assert err < REL_TOL
