"""Pins the oracle's V-cycle parts (the build's own multigrid schedule) on
the CPU: the Chebyshev spectral radius against a dense eigen-solve, the
coarsest Cholesky solve against an independently assembled system, the
deviatoric-invariant polar against SVD at small strains, and convergence of
the whole cycle to the residual target."""
import numpy as np
import pytest

import oracle as O
from multigridsoftbody_amd import Mesh
from multigridsoftbody_amd import workloads as W


def dense_system(mesh, fix, k, density, dt):
    """A = M/h^2 + sum_t k vol G_t^T G_t on free vertices, from the mesh alone."""
    X = mesh.pos.astype(np.float64)
    V = mesh.nverts
    A = np.zeros((V, V))
    mass = np.zeros(V)
    for t in mesh.tets:
        Dm = np.stack([X[t[1]] - X[t[0]], X[t[2]] - X[t[0]], X[t[3]] - X[t[0]]], axis=1)
        vol = abs(np.linalg.det(Dm)) / 6
        G = np.zeros((3, 4))  # gradient of the 4 hat functions
        Binv = np.linalg.inv(Dm)
        G[:, 1:] = Binv.T
        G[:, 0] = -Binv.T.sum(axis=1)
        A[np.ix_(t, t)] += k * vol * (G.T @ G)
        mass[t] += density * vol / 4
    A[np.diag_indices(V)] += mass / dt ** 2
    free = np.flatnonzero(fix != 0)
    return A, free, mass


def test_coarsest_cholesky_solve_matches_numpy():
    fine, coarse = W.beam(2), W.beam(1)
    phys = W.physics()
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=True)
          for m in (fine, coarse)]
    Ls[0].set_fix(W.top_fixflag(fine))
    ML = O.MultiLevel(Ls, dt=phys["dt"])
    vfix = ML.vfix(1)
    assert np.array_equal(vfix, W.top_fixflag(coarse))  # nested: coincident vertices
    A, free, _ = dense_system(coarse, vfix, phys["stiffness"], phys["density"], phys["dt"])
    assert ML.nfree() == len(free)
    r = np.random.default_rng(0).normal(size=(coarse.nverts, 3)).astype(np.float32)
    r[vfix == 0] = 0
    e = ML.coarse_solve(r)
    ref = np.zeros_like(e, dtype=np.float64)
    ref[free] = np.linalg.solve(A[np.ix_(free, free)], r[free].astype(np.float64))
    assert np.abs(e - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.all(e[vfix == 0] == 0)


def test_power_iteration_rho_matches_dense_eigenvalues():
    fine, coarse = W.beam(2), W.beam(1)
    phys = W.physics()
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=True)
          for m in (fine, coarse)]
    Ls[0].set_fix(W.top_fixflag(fine))
    ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
    A, free, _ = dense_system(coarse, ML.vfix(1), phys["stiffness"], phys["density"], phys["dt"])
    Af = A[np.ix_(free, free)]
    J = np.eye(len(free)) - 0.6 * Af / np.diag(Af)[:, None]
    rho = np.abs(np.linalg.eigvals(J)).max()
    assert abs(ML.rho(1) - rho) < 1e-3 * rho, (ML.rho(1), rho)


def test_coarse_smoother_against_the_assembled_system():
    """ml_smooth (damped Jacobi with Chebyshev on A_l e = r, the V-cycle's
    coarse-level smoother) on level 1 of a 3-level beam, against the system
    assembled from the mesh alone (dense_system): k sweeps from e = 0 follow
        y = e + relax D^-1 (r - A e),  e_new = w_k (y - e_prev) + e_prev,
        w_k = 4 / (4 - rho^2 w_{k-1}),  w_0 = 1,
    on the free DOFs (pinned rows stay 0), and many sweeps converge to the
    solution of A_ff e = r_f.  A wrong assembly, sign or omega seed fails one
    of the two."""
    lv = W.beam_hierarchy((4, 2, 1))
    phys = W.physics()
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=True)
          for m in lv]
    Ls[0].set_fix(W.top_fixflag(lv[0]))
    ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
    vfix = ML.vfix(1)
    A, free, _ = dense_system(lv[1], vfix, phys["stiffness"], phys["density"], phys["dt"])
    Af = A[np.ix_(free, free)]
    d = np.diag(Af)
    rho, relax = ML.rho(1), float(np.float32(0.6))
    r = np.random.default_rng(4).normal(size=(lv[1].nverts, 3)).astype(np.float32)
    r[vfix == 0] = 0
    rf = r[free].astype(np.float64)
    for k in (1, 2, 5):
        e, ep, w = np.zeros_like(rf), np.zeros_like(rf), 1.0
        for _ in range(k):
            w = 4 / (4 - rho * rho * w)
            y = e + relax * (rf - Af @ e) / d[:, None]
            e, ep = w * (y - ep) + ep, e
        got = ML.lin_smooth(1, k, r)
        # the entry point returns FP32: agreement to its rounding, 6e-8 of the max
        assert np.abs(got[free] - e).max() <= 3e-7 * np.abs(e).max(), k
        assert np.all(got[vfix == 0] == 0)
    got = ML.lin_smooth(1, 400, r)
    exact = np.linalg.solve(Af, rf)
    assert np.abs(got[free] - exact).max() <= 1e-5 * np.abs(exact).max()


def test_stable_polar_keeps_small_anisotropic_strain():
    """At strains ~1e-4 the reference's FP32 branch (k < 1e-6) returns the
    isotropic factor; the deviatoric-invariant form (variant 2) still matches
    the SVD polar factor."""
    rng = np.random.default_rng(2)
    th = 0.4
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    for _ in range(20):
        S = np.eye(3) + 1e-4 * rng.normal(size=(3, 3))
        S = (S + S.T) / 2
        F = (Q @ S).astype(np.float32).astype(np.float64)
        U, _, Vt = np.linalg.svd(F)
        Rs = U @ Vt
        R2, _ = O.polar(F, 2, False)
        R0, _ = O.polar(F, 0, False)
        e2 = np.abs((R2 - F) - (Rs - F)).max()
        e0 = np.abs((R0 - F) - (Rs - F)).max()
        assert e2 < 2e-6 * 1 + 0.05 * np.abs(Rs - F).max(), e2
        assert e0 > 10 * e2  # the reference branch loses the deviatoric part


@pytest.mark.parametrize("fp64", [False, True])
def test_vcycle_reaches_tolerance_every_step(fp64):
    lv = W.beam_hierarchy((4, 1))
    phys = W.physics()
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=fp64) for m in lv]
    fix = W.top_fixflag(lv[0])
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(W.seeded_velocity(lv[0], fix))
    ML = O.MultiLevel(Ls, dt=phys["dt"])
    p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], if_mode=3)
    vp = O.vparams(max_cycles=12, tol=1e-3)
    worst = 0
    for _ in range(60):
        ML.step(p, vp, 1)
        st = ML.stats()
        worst = max(worst, st["rel"])
        assert st["cycles"] <= 6
    assert worst <= 1e-3


def _small_ml():
    lv = W.beam_hierarchy((4, 1))
    phys = W.physics()
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"]) for m in lv]
    fix = W.top_fixflag(lv[0])
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(W.seeded_velocity(lv[0], fix))
    ML = O.MultiLevel(Ls, dt=phys["dt"])
    p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], gravity=phys["gravity"],
                 if_mode=3)
    return lv, Ls, ML, p


def test_warm_start_judge_rejects_and_restores():
    """A warm-start stage that raises the residual is undone exactly: the step
    then equals the step without that stage, bit for bit; a stage that helps
    is kept."""
    lv, Ls, ML, p = _small_ml()
    vp = O.vparams(pre=3, post=3, max_cycles=8, check=2, tol=1e-3, rho_smooth=0.9)
    ML.step(p, vp, 10)
    x0, v0 = Ls[0].positions().copy(), Ls[0].get(O.V).copy()
    V1 = lv[1].nverts

    def run(e1=None, u0=None):
        Ls[0].set_positions(x0)
        Ls[0].set_velocity(v0)
        ML.set_warm(e1, u0)
        ML.step(p, vp, 1)
        return Ls[0].positions().copy(), ML.stats()

    xp, sp = run()
    assert sp["fallback"] == 0
    # a wild fine correction: rejected, identical to the plain step
    bad_u = np.full((lv[0].nverts, 3), 0.5, np.float32)
    xb, sb = run(None, bad_u)
    assert sb["fallback"] == 1 and sb["res_t"] > sb["res0"]
    assert np.array_equal(xb, xp) and sb["halves"] == sp["halves"]
    # the exact step correction as fine warm start: converged with no cycles
    s = Ls[0].get(O.S).reshape(-1, 3)
    u_exact = xp - s
    xe, se = run(None, u_exact)
    assert se["fallback"] == 0 and se["halves"] == 0 and se["res_t"] <= 1e-3 * se["res0"]
    # good temporal stage, then a wild coarse stage: the coarse stage is undone
    half = 0.5 * u_exact
    xh, sh = run(None, half)
    bad_e = np.full((V1, 3), 0.5, np.float32)
    xhb, shb = run(bad_e, half)
    assert shb["fallback"] == 2 and shb["res_ws"] > shb["res_t"]
    assert np.array_equal(xhb, xh) and shb["halves"] == sh["halves"]


def arap_forces(mesh, x, k):
    """Per-vertex ARAP forces sum_t k vol_t (R_t - F_t) AC_t in FP64, R the
    SVD polar factor (proper rotation), AC[:, 0] = -sum of B's rows,
    AC[:, 1:] = B^T with B = Dm^-1."""
    X = mesh.pos.astype(np.float64)
    x = np.asarray(x, np.float64)
    f = np.zeros_like(X)
    for tet in mesh.tets:
        Dm = np.stack([X[tet[q]] - X[tet[0]] for q in (1, 2, 3)], 1)
        B = np.linalg.inv(Dm)
        F = np.stack([x[tet[q]] - x[tet[0]] for q in (1, 2, 3)], 1) @ B
        U, _, Vt = np.linalg.svd(F)
        if np.linalg.det(U @ Vt) < 0:
            U[:, -1] *= -1
        AC = np.concatenate([-B.sum(axis=0)[:, None], B.T], axis=1)
        f[tet] += (k * abs(np.linalg.det(Dm)) / 6 * (U @ Vt - F) @ AC).T
    return f


def test_coarse_first_half_is_the_coarse_correction_of_g_s():
    """coarse_first: a step's first half-cycle is the coarse correction of
    the start's residual.  Two levels, no warm start, no post-smoothing and
    no further cycles: the step's correction u = x - s must equal
    fix * P A_1^-1 (vfix_1 * P^T g(s)), assembled here from the oracle's
    corner forces at s (summed per vertex in numpy), the binding weights and
    the coarsest solve (pinned above against numpy).  A pre-smoothing-first
    order, a residual taken at x instead of s, or a transfer without the
    Dirichlet masks fails it."""
    fine, coarse = W.beam(2), W.beam(1)
    phys = W.physics()
    fix = W.top_fixflag(fine)
    v0 = W.seeded_velocity(fine, fix)
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=True)
          for m in (fine, coarse)]
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(v0)
    ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
    p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], gravity=phys["gravity"],
                 if_mode=3)
    vp = O.vparams(pre=3, post=0, coarse_sweeps=2, max_cycles=0, check=2, tol=0.0, rho_smooth=0.9, coarse_first=1)
    ML.step(p, vp, 1)
    st = ML.stats()
    assert st["halves"] == 1 and st["cycles"] == 1
    s = Ls[0].get(O.S).reshape(-1, 3)
    u = Ls[0].positions() - s
    # g(s) = fix * f_el(s) (x = s: the inertial term vanishes; no colliders)
    g = fix[:, None] * arap_forces(fine, s, phys["stiffness"])
    ids, w, _ = coarse.bind(fine.pos, 0)
    r1 = np.zeros((coarse.nverts, 3))
    np.add.at(r1, ids.reshape(-1), (w[:, :, None] * g[:, None, :]).reshape(-1, 3))
    r1 *= ML.vfix(1)[:, None]
    e1 = ML.coarse_solve(r1.astype(np.float32)).astype(np.float64)
    ref = fix[:, None] * np.einsum("nk,nkc->nc", w.astype(np.float64), e1[ids])
    err = np.abs(u - ref).max() / np.abs(ref).max()
    assert err < 1e-6, err
    # the textbook order smooths first: a different correction
    Ls2 = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=True)
           for m in (fine, coarse)]
    Ls2[0].set_fix(fix)
    Ls2[0].set_velocity(v0)
    ML2 = O.MultiLevel(Ls2, relax_coarse=0.6, dt=phys["dt"])
    ML2.step(p, O.vparams(pre=3, post=0, coarse_sweeps=2, max_cycles=1, check=2, tol=0.0, rho_smooth=0.9), 1)
    u2 = Ls2[0].positions() - Ls2[0].get(O.S).reshape(-1, 3)
    assert np.abs(u2 - ref).max() > 100 * err * np.abs(ref).max()


def test_stall_guard_inert_while_converging_and_rescues_a_diverging_smoother():
    """stall_guard: a check whose residual exceeds 3 times the previous
    check's turns the fine Chebyshev extrapolation off for the rest of the
    step.  On steps whose checks decrease (rho 0.9, FP64, tolerance 1e-3) the
    trajectory is the same with and without it, bit for bit.  With rho 1.1
    (an over-estimate of the smoother's spectral radius) the Chebyshev
    weights grow past 2 (1.43, 1.77, 2.15, 2.85, 7.3) and the sweeps
    diverge: unguarded, no step reaches the target in 12 cycles and the
    residual ends above the start's; guarded, every step meets it within
    5 half-cycles, in FP32 as in FP64, on the same trajectory."""
    lv = W.beam_hierarchy((4, 1))
    phys = W.physics()
    fix = W.top_fixflag(lv[0])
    v0 = W.seeded_velocity(lv[0], fix)

    def run(guard, fp64, rho, steps):
        Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=fp64)
              for m in lv]
        Ls[0].set_fix(fix)
        Ls[0].set_velocity(v0)
        ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
        p = O.params(dt=phys["dt"], stiffness=phys["stiffness"], density=phys["density"], gravity=phys["gravity"],
                     if_mode=3)
        vp = O.vparams(pre=5, post=5, coarse_sweeps=2, max_cycles=12, check=2, tol=1e-3, rho_smooth=rho,
                       stall_guard=guard)
        st = []
        for _ in range(steps):
            ML.step(p, vp, 1)
            st.append(ML.stats())
        return Ls[0].positions(), st

    assert np.array_equal(run(0, True, 0.9, 20)[0], run(1, True, 0.9, 20)[0])
    _, wild = run(0, False, 1.1, 3)
    assert all(s["halves"] == 24 and s["rel"] > 1.0 for s in wild)
    x32, calm32 = run(1, False, 1.1, 20)
    x64, calm64 = run(1, True, 1.1, 20)
    for st in (calm32, calm64):
        assert all(s["rel"] <= 1e-3 and s["halves"] <= 5 for s in st)
    assert float(np.abs(x32 - x64).max() / np.abs(x64).max()) < 1e-5
