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


def test_mlp_gradients_match_oracle():
    s = small_solver(M.SCHED_VCYCLE)
    w = rand_weights(s.learn_nparams(), seed=2)
    s.learn_weights(w)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(96, 12)).astype(np.float32)
    t = rng.normal(size=(96, 3)).astype(np.float32)
    g, loss = s.mlp_grad(x, t)
    go, loss_o = O.mlp_grad(w, x, t, bf16=True)
    assert abs(loss - loss_o) <= 1e-2 * loss_o
    H = O.MLP_H
    sizes = [12 * H, H, H * H, H, H * H, H, H * 3, 3]
    o = 0
    for k, n in enumerate(sizes):
        a, b = g[o:o + n], go[o:o + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"[mlp grad] block {k} rel err {err:.2e}")
        assert err < 3e-2, (k, err)
        o += n


def test_learned_schedule_record_train_ab():
    plain = small_solver(M.SCHED_VCYCLE)
    plain.learn_record(150)
    plain.step(150)
    assert plain.learn_samples() == 150 * 44
    hist = plain.learn_train(epochs=30, batch=512, lr=2e-3, seed=0)
    print(f"[learned] loss first {hist[0]:.4f} last {hist[-1]:.4f}")
    assert hist[-1] < 0.5 * hist[0]
    w, norm = plain.learn_weights(), plain.learn_norm()
    # A/B from the same state: continue the plain run vs a learned copy
    learned = small_solver(M.SCHED_LEARNED)
    learned.learn_weights(w)
    learned.learn_norm(norm)
    learned.learn_record(150)  # to carry the previous-correction feature
    learned.step(150)
    learned.learn_record(0)
    for sol in (plain, learned):
        sol.run_stats(reset=True)
    plain.step(60)
    learned.step(60)
    a, b = plain.run_stats(), learned.run_stats()
    print(f"[learned A/B] plain cycles/step {a['total_cycles'] / a['steps']:.2f} learned "
          f"{b['total_cycles'] / b['steps']:.2f} fallbacks {b['fallbacks']} max rel {b['max_rel']:.2e}")
    assert b["failed_steps"] == 0 and b["max_rel"] <= 1e-3
    assert a["failed_steps"] == 0
