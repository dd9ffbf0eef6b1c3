"""Learned coarse correction on the device: the BF16-MFMA MLP against the
oracle's restatement (forward and gradients), the fused inference kernel
against the training GEMM path, and the learned schedule end to end (record
plain-solver trajectories, train on the device, step with the warm start)
with an A/B comparison of V-cycles per step against the plain schedule."""
import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_solver(schedule, **kw):
    lv = W.beam_hierarchy((4, 1))
    base = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=1, tol=1e-3,
                rho_smooth=0.9)
    base.update(kw)
    s = M.Solver(lv, schedule=schedule, if_mode=M.IF_STABLE, **base, **W.physics())
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(W.seeded_velocity(lv[0], W.top_fixflag(lv[0])))
    return s


def rand_weights(n, seed=0, scale=0.15):
    return (scale * np.random.default_rng(seed).normal(size=n)).astype(np.float32)


def test_mlp_forward_paths_and_oracle():
    s = small_solver(M.SCHED_VCYCLE)
    n = s.learn_nparams()
    assert n == O.mlp_nparams(12, 3)
    w = rand_weights(n)
    s.learn_weights(w)
    assert np.array_equal(s.learn_weights(), w)
    x = np.random.default_rng(1).normal(size=(203, 12)).astype(np.float32)
    y_fused = s.mlp_forward(x, path=0)
    y_gemm = s.mlp_forward(x, path=1)
    y_orc = O.mlp_forward(w, x, bf16=True)
    y_64 = O.mlp_forward(w, x, bf16=False)
    print(f"[mlp fwd] |fused-gemm|={np.abs(y_fused - y_gemm).max():.2e} |fused-oracle_bf16|="
          f"{np.abs(y_fused - y_orc).max():.2e} |fused-fp64|={np.abs(y_fused - y_64).max():.2e} "
          f"scale={np.abs(y_64).max():.2e}")
    assert np.abs(y_fused - y_gemm).max() <= 1e-5 * np.abs(y_gemm).max()
    assert np.abs(y_fused - y_orc).max() <= 2e-2 * np.abs(y_orc).max()


@pytest.mark.parametrize("rows", [96, 1100])
def test_mlp_gradients_match_oracle(rows):
    """Gradients of the training GEMM path against the oracle; with 1100
    rows the weight gradients (K = rows) take the split-K form (5 chunks of
    224 rows, the last one ragged)."""
    s = small_solver(M.SCHED_VCYCLE)
    w = rand_weights(s.learn_nparams(), seed=2)
    s.learn_weights(w)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(rows, 12)).astype(np.float32)
    t = rng.normal(size=(rows, 3)).astype(np.float32)
    g, loss = s.mlp_grad(x, t)
    go, loss_o = O.mlp_grad(w, x, t, bf16=True)
    assert abs(loss - loss_o) <= 1e-2 * loss_o
    H = O.MLP_H
    sizes = [12 * H, H, H * H, H, H * H, H, H * 3, 3]
    o = 0
    for k, n in enumerate(sizes):
        a, b = g[o:o + n], go[o:o + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"[mlp grad rows={rows}] block {k} rel err {err:.2e}")
        assert err < 3e-2, (k, err)
        o += n


def oracle_pair(ns, fp64=False):
    lv = W.beam_hierarchy(ns)
    phys = W.physics()
    fix = W.top_fixflag(lv[0])
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=fp64)
          for m in lv]
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(W.seeded_velocity(lv[0], fix))
    ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
    p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], gravity=phys["gravity"],
                 if_mode=3)
    return lv, Ls, ML, p, fix


VP = dict(pre=5, post=5, coarse_sweeps=2, max_cycles=12, tol=1e-3, rho_smooth=0.9)


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


class Temporal:
    """Temporal extrapolation u_t = 2 u_{n-1} - u_{n-2} of an oracle run."""

    def __init__(self, L0, fix):
        self.L0, self.fix = L0, fix[:, None]
        self.u1 = self.u2 = np.zeros((L0.nv, 3), np.float32)

    def warm(self, pe=None):
        """fix (u_t + pe): the extrapolation plus a prolongated coarse
        correction pe (the learned schedule's warm start)."""
        u = 2 * self.u1 - self.u2
        if pe is not None:
            u = u + pe
        return (self.fix * u).astype(np.float32)

    def push(self):
        u = self.L0.positions() - self.L0.get(O.S).reshape(-1, 3)
        self.u2, self.u1 = self.u1, u.astype(np.float32)


@pytest.mark.parametrize("check,cfirst", [(1, 0), (2, 0), (1, 1), (2, 1)])
def test_temporal_warm_start_matches_oracle(check, cfirst):
    lv, Ls, ML, p, fix = oracle_pair((4, 1))
    g = small_solver(M.SCHED_VCYCLE, warm_start=1, check_every=check, coarse_first=cfirst)
    vp = O.vparams(check=check, coarse_first=cfirst, **VP)
    hist = Temporal(Ls[0], fix)
    agree = 0
    for k in range(60):
        st = g.step(1)
        ML.set_warm(None, hist.warm())
        ML.step(p, vp, 1)
        hist.push()
        o = ML.stats()
        agree += (st.halves, st.fallback) == (o["halves"], o["fallback"])
        assert st.rel <= 1e-3
    err = rel_err(g.positions(), Ls[0].positions())
    print(f"[temporal check={check} coarse_first={cfirst}] rel(gpu,oracle)={err:.2e} agree {agree / 60:.2f} last halves {st.halves}")
    assert err < 1e-4 and agree >= 54


def record_and_train(ns, steps, epochs, seed=0, warm=1):
    """Record `steps` steps of the V-cycle (warm = 1: temporal start, targets
    u - u_t; warm = 0: start from s, targets u) and train the MLP on them."""
    rec = small_solver(M.SCHED_VCYCLE, warm_start=warm, check_every=2) if ns == (4, 1) else None
    if rec is None:
        rec = medium_solver(M.SCHED_VCYCLE, warm_start=warm)
    rec.learn_record(steps)
    rec.step(steps)
    assert rec.learn_samples() == steps * rec.levels[1].nverts
    hist = rec.learn_train(epochs=epochs, batch=512, lr=2e-3, seed=seed)
    return rec, hist


def medium_solver(schedule, **kw):
    lv = W.beam_hierarchy((12, 4, 1))
    base = dict(pre_sweeps=5, post_sweeps=5, coarse_sweeps=2, max_cycles=12, check_every=2, tol=1e-3,
                rho_smooth=0.9, use_graph=1)
    base.update(kw)
    s = M.Solver(lv, schedule=schedule, if_mode=M.IF_STABLE, **base, **W.physics())
    s.fix_plane(1, 10.0, 1e-5)
    s.set_velocity(W.seeded_velocity(lv[0], W.top_fixflag(lv[0])))
    return s


def prolong(lv, e1):
    """P e1: a coarse (level-1) vector interpolated at the fine vertices
    through their binding in coarse tets (PDSolver_MG.cpp:5-58), summed in
    the device's order."""
    ids, w, _ = lv[1].bind(lv[0].pos, 0)
    e = np.asarray(e1, np.float32)
    w = w.astype(np.float32)
    out = w[:, 0:1] * e[ids[:, 0]]
    for q in (1, 2, 3):
        out = out + w[:, q:q + 1] * e[ids[:, q]]
    return out


@pytest.mark.parametrize("cfirst", [0, 1])
def test_learned_schedule_matches_oracle_given_prediction(cfirst):
    """The learned schedule's machinery (the MLP's coarse prediction
    prolongated into the temporal warm start, one judge of the warm start,
    cycles, in either half-cycle order) against the oracle fed the same warm
    start; the MLP itself is checked against the oracle above."""
    rec, hist = record_and_train((4, 1), 100, 20)
    w, norm = rec.learn_weights(), rec.learn_norm()
    lv, Ls, ML, p, fix = oracle_pair((4, 1))
    g = small_solver(M.SCHED_LEARNED, check_every=2, learn_gate=0, coarse_first=cfirst)
    g.learn_weights(w)
    g.learn_norm(norm)
    vp = O.vparams(check=2, coarse_first=cfirst, **VP)
    tmp = Temporal(Ls[0], fix)
    agree = used = 0
    for k in range(60):
        st = g.step(1)
        ML.set_warm(None, tmp.warm(prolong(lv, g.learn_last())))
        ML.step(p, vp, 1)
        tmp.push()
        o = ML.stats()
        agree += (st.halves, bool(st.fallback)) == (o["halves"], bool(o["fallback"]))
        used += (st.fallback & 2) == 0
        assert st.rel <= 1e-3
    err = rel_err(g.positions(), Ls[0].positions())
    print(f"[learned vs oracle] rel={err:.2e} agree {agree / 60:.2f} prediction kept in {used}/60 steps")
    assert err < 1e-4 and agree >= 54
    assert used > 0


def test_learned_only_matches_oracle_given_prediction():
    """warm_start = 2: the MLP's coarse correction applied to s alone (no
    temporal extrapolation), judged against g(s), against the oracle fed the
    same warm start."""
    rec, hist = record_and_train((4, 1), 100, 20, warm=0)
    w, norm = rec.learn_weights(), rec.learn_norm()
    lv, Ls, ML, p, fix = oracle_pair((4, 1))
    g = small_solver(M.SCHED_LEARNED, check_every=2, learn_gate=0, warm_start=2)
    g.learn_weights(w)
    g.learn_norm(norm)
    vp = O.vparams(check=2, **VP)
    agree = used = 0
    for k in range(60):
        st = g.step(1)
        ML.set_warm(None, (fix[:, None] * prolong(lv, g.learn_last())).astype(np.float32))
        ML.step(p, vp, 1)
        o = ML.stats()
        agree += (st.halves, bool(st.fallback)) == (o["halves"], bool(o["fallback"]))
        used += (st.fallback & 2) == 0
        assert st.rel <= 1e-3
    err = rel_err(g.positions(), Ls[0].positions())
    print(f"[learned on s vs oracle] rel={err:.2e} agree {agree / 60:.2f} prediction kept in {used}/60 steps")
    assert err < 1e-4 and agree >= 54
    assert used > 0


def test_learned_ab_medium_box():
    """What the trained network gains on the 103,680-tet box: half-cycles
    per step to rel. residual 1e-3 of the plain V-cycle (start from s), the
    network's coarse correction on s alone (trained on plain-schedule
    recordings), the V-cycle with the temporal warm start alone, and the
    learned schedule (temporal warm start + MLP coarse delta), after the same
    200 steps."""
    rec, hist = record_and_train((12, 4, 1), 200, 20)
    print(f"[learned] training loss first {hist[0]:.4f} last {hist[-1]:.4f}")
    assert hist[-1] < 0.7 * hist[0]
    w, norm = rec.learn_weights(), rec.learn_norm()
    rec_s, hist_s = record_and_train((12, 4, 1), 200, 20, warm=0)
    print(f"[learned on s] training loss first {hist_s[0]:.4f} last {hist_s[-1]:.4f}")
    sols = {"plain": medium_solver(M.SCHED_VCYCLE, check_every=1),
            "learned_s": medium_solver(M.SCHED_LEARNED, warm_start=2, check_every=2),
            "temporal": medium_solver(M.SCHED_VCYCLE, warm_start=1),
            "learned": medium_solver(M.SCHED_LEARNED)}
    sols["learned"].learn_weights(w)
    sols["learned"].learn_norm(norm)
    sols["learned_s"].learn_weights(rec_s.learn_weights())
    sols["learned_s"].learn_norm(rec_s.learn_norm())
    res = {}
    for name, s in sols.items():
        s.step(200)
        s.run_stats(reset=True)
        s.step(100)
        r = s.run_stats()
        res[name] = r["total_halves"] / r["steps"]
        assert r["failed_steps"] == 0 and r["max_rel"] <= 1e-3, (name, r)
        print(f"[A/B medium] {name:9s} half-cycles/step {res[name]:.2f} fallbacks {r['fallbacks']} "
              f"max rel {r['max_rel']:.2e}")
    # the temporal stage carries the schedule on this smooth workload; the
    # network's correction must not cost half-cycles (the judge undoes it)
    assert res["temporal"] < 0.5 * res["plain"]
    assert res["learned"] <= res["temporal"] + 0.05
    # on its own (start from s) the network's coarse correction saves
    # half-cycles: steps stop after the pre-smoothing half of the second cycle
    assert res["learned_s"] < res["plain"] - 0.5


def test_learn_gate_bounds_rejected_stages():
    """A network whose every correction the judge rejects (weights far too
    large: the prediction overflows): each step still meets the target from
    s, nothing non-finite is kept, and the gate runs the MLP stage in
    shrinking windows (8 rejections, then shut for 8, 16, 32, ... steps),
    against every step with the gate off."""
    res = {}
    for gate in (1, 0):
        s = small_solver(M.SCHED_LEARNED, check_every=2, learn_gate=gate)
        s.learn_weights(rand_weights(s.learn_nparams(), seed=5, scale=50.0))
        s.step(120)
        run = s.run_stats()
        x = s.positions()
        print(f"[gate {gate}] steps {run['steps']} rejected stages {run['fallbacks']} "
              f"half-cycles/step {run['total_halves'] / run['steps']:.2f} max rel {run['max_rel']:.2e}")
        assert np.all(np.isfinite(x))
        assert run["failed_steps"] == 0 and run["max_rel"] <= 1e-3
        res[gate] = run["fallbacks"]
    # 8 rejections, shut 8, 8, shut 16, 8, shut 32, 8, shut 64: 32 of 120 steps
    assert res[1] <= 40 and res[0] >= 100 and res[0] >= 2.5 * res[1]


def _hand_net(rows, H=128):
    """A network whose output component c is sum_k rows[c][k] * z_k (z the
    standardised features), exact in BF16 for small integer weights: ReLU
    pairs carry +-value through the hidden layers."""
    W1 = np.zeros((12, H)); b1 = np.zeros(H); W2 = np.zeros((H, H)); b2 = np.zeros(H)
    W3 = np.zeros((H, H)); b3 = np.zeros(H); W4 = np.zeros((H, 3)); b4 = np.zeros(3)
    for c in range(3):
        for k, a in rows[c].items():
            W1[k, 2 * c] += a
            W1[k, 2 * c + 1] -= a
        for u in (2 * c, 2 * c + 1):
            W2[u, u] = W3[u, u] = 1
        W4[2 * c, c], W4[2 * c + 1, c] = 1, -1
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3, W4.ravel(), b4]).astype(np.float32)


def test_fused_coarse_stage_features():
    """The one-launch learned stage (features, network, prior, mask): with a
    network that outputs 2 z_v - D1 for standardised velocity features
    z_v = (v - 0.5) / 2 (feature order: injected velocity, rest position,
    last coarse delta D1, the one before), and output mean 0.5, the stage's
    correction vfix (D1 + y + 0.5) is the injected fine velocity of the
    coarse vertices, masked by their Dirichlet flags."""
    lv = W.beam_hierarchy((4, 1))
    g = small_solver(M.SCHED_LEARNED, check_every=2, learn_gate=0)
    g.learn_weights(_hand_net([{0: 2.0, 6: -1.0}, {1: 2.0, 7: -1.0}, {2: 2.0, 8: -1.0}]))
    norm = np.concatenate([[0.5, 0.5, 0.5], np.zeros(9), [2.0, 2.0, 2.0], np.ones(9), [0.5, 0.5, 0.5], np.ones(3)])
    g.learn_norm(norm.astype(np.float32))
    g.step(5)
    v = g.velocity()
    g.step(1)
    e = g.learn_last()
    ids, w, _ = lv[0].bind(lv[1].pos, 1)
    want = W.top_fixflag(lv[1])[:, None] * (w[:, :, None] * v[ids]).sum(axis=1)
    err = np.abs(e - want).max() / np.abs(want).max()
    print(f"[fused stage] |v_inj| max {np.abs(want).max():.3e}, rel err {err:.2e}")
    assert np.abs(want).max() > 0.05 and err < 1e-2
    assert np.all(e[W.top_fixflag(lv[1]) == 0] == 0)
