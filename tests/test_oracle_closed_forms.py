"""Closed-form pins of the oracle's step pieces that the geometry tests leave
open: calculateST + calculateV (oracle/mgsb_oracle.c calc_st, calc_v;
PDSolver.cu:72-116, :245-252), and calculatePOS's under-relaxation and
Chebyshev line together with the omega recurrence of PDSolver::Step
(calc_pos, step_single; PDSolver.cu:192-235, PDSolver.cpp:261,270).

Each expected trajectory is derived by hand from the physics of a problem
whose elastic force is zero or exactly linear, so a dropped term, a swapped
order (damping after gravity) or a wrong seed of omega (4/(4 - rho^2) in
place of 1) changes the numbers.  CPU only."""
import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M

F32 = lambda v: float(np.float32(v))  # noqa: E731  (the oracle's float parameters)

# unit right-angle tet: Dm = I, vol = 1/6
_POS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
_TET = np.array([[0, 1, 2, 3]], np.int32)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("if_mode", [0, 1])
def test_free_fall_closed_form(fp64, if_mode):
    """A free tet translating rigidly feels no elastic force (F = I), so each
    step is calculateST then calculateV alone:
        v_n = damp * v_{n-1} + g h      (damping first, then gravity)
        x_n = x_{n-1} + h v_n
    i.e. v_n = damp^n v0 + g h sum_{k<n} damp^k.  Damping applied after
    gravity would give v_n = damp (v_{n-1} + g h) and fail."""
    h, damp = F32(1 / 60), F32(0.95)
    g = np.array([F32(0.7), F32(-9.8), F32(0.3)])
    v0 = np.array([0.4, -0.25, 1.5], np.float32)
    L = O.Level(_POS, _TET, None, stiffness=1e5, density=1000.0, fp64=fp64)
    L.set_fix(np.ones(4, np.float32))
    L.set_velocity(np.tile(v0, (4, 1)))
    p = O.params(dt=h, damping=damp, stiffness=1e5, density=1000.0, gravity=g, iters=10, if_mode=if_mode)
    n = 25
    L.step(p, n)
    v = v0.astype(np.float64)
    x = _POS.astype(np.float64)
    for _ in range(n):
        v = damp * v + g * h
        x = x + h * v
    closed_v = damp ** n * v0 + g * h * sum(damp ** k for k in range(n))
    assert np.allclose(v, closed_v, rtol=1e-12)
    tol = dict(rtol=1e-10, atol=1e-12) if fp64 else dict(rtol=2e-5, atol=2e-5)
    assert np.allclose(L.positions(), x, **tol)
    assert np.allclose(L.get(O.V).reshape(-1, 3), np.tile(v, (4, 1)), **tol)
    # the wrong order differs by about (1 - damp) g h per step: far outside tol
    vw = v0.astype(np.float64)
    for _ in range(n):
        vw = damp * (vw + g * h)
    assert np.abs(vw - v).max() > 100 * (tol["atol"] + tol["rtol"] * np.abs(v).max())


@pytest.mark.parametrize("fp64", [True, False])
def test_pinned_vertices_keep_rest_position_and_zero_velocity(fp64):
    """fixflag 0 multiplies every update of calculateST and calculatePOS, so
    a pinned vertex keeps its position bit for bit and its velocity is 0,
    while the free vertex falls."""
    L = O.Level(_POS, _TET, None, stiffness=1e5, density=1000.0, fp64=fp64)
    fix = np.array([0, 1, 0, 0], np.float32)
    L.set_fix(fix)
    L.set_velocity(np.full((4, 3), 0.3, np.float32))
    L.step(O.params(stiffness=1e5, density=1000.0, iters=10), 30)
    x = L.positions()
    assert np.array_equal(x[fix == 0], _POS[fix == 0].astype(np.float64))
    assert np.array_equal(L.get(O.V).reshape(-1, 3)[fix == 0], np.zeros((3, 3)))
    assert np.abs(x[1] - _POS[1]).max() > 0.01


def _chebyshev_closed_form(n_steps, iters, rho, h, damp, g, v0, K, m):
    """One free vertex (vertex 1 of the unit tet, moving along x only) with the
    other three pinned.  F = diag(1 + a, 1, 1) with a = x - 1 is symmetric
    positive, so R = I and the elastic force on the vertex is exactly -K a
    (K = k vol; AC column 1 is e_x), and its Jacobi diagonal is K.  The
    Jacobi step of calculatePOS therefore lands on the exact minimiser
        x* = (dc s + K) / (dc + K),   dc = m / h^2,
    and with errors e = x - x*:
        relaxed = 0.4 e_k                 (x + 0.6 (x_J - x))
        e_{k+1} = w_k (0.4 e_k - e_{k-1}) + e_{k-1}
        w_k = 4 / (4 - rho^2 w_{k-1}),  w_0 = 1,
    starting from e_0 = e_{-1} = s - x* (calculateST sets x = x_prev = s)."""
    x, v = 1.0, v0
    dc = m / (h * h)
    for _ in range(n_steps):
        xlast = x
        v = damp * v + g * h
        x = x + h * v
        s = x
        xs = (dc * s + K * 1.0) / (dc + K)
        e, ep = s - xs, s - xs
        w = 1.0
        for _ in range(iters):
            w = 4 / (4 - rho * rho * w)
            e, ep = w * (0.4 * e - ep) + ep, e
        x = xs + e
        v = (x - xlast) / h
    return x, v


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("if_mode", [0, 1])
def test_chebyshev_jacobi_one_free_vertex_closed_form(fp64, if_mode):
    h, damp, rho = F32(1 / 60), F32(0.95), F32(0.9992)
    k, dens = 1e5, 1000.0
    gx, v0 = F32(-2.0), F32(3.0)
    iters, n = 10, 4
    L = O.Level(_POS, _TET, None, stiffness=k, density=dens, fp64=fp64)
    L.set_fix(np.array([0, 1, 0, 0], np.float32))
    vel = np.zeros((4, 3), np.float32)
    vel[1, 0] = v0
    L.set_velocity(vel)
    p = O.params(dt=h, damping=damp, stiffness=k, density=dens, rho=rho, gravity=(gx, 0, 0), iters=iters,
                 if_mode=if_mode)
    L.step(p, n)
    K, m = k / 6, dens / 24
    xw, vw = _chebyshev_closed_form(n, iters, rho, h, damp, gx, v0, K, m)
    x = L.positions()
    tol = dict(rtol=1e-10, atol=1e-12) if fp64 else dict(rtol=0, atol=3e-6)
    assert np.allclose(x[1], [xw, 0, 0], **tol), (x[1], xw)
    assert np.isclose(L.get(O.V).reshape(-1, 3)[1, 0], vw, rtol=1e-8 if fp64 else 1e-3)
    assert np.array_equal(x[[0, 2, 3]], _POS[[0, 2, 3]].astype(np.float64))
    # the step has not converged (iteration error well above tol), so the
    # recurrence's details decide x: a seed of omega = 4/(4 - rho^2) instead
    # of 1, or relaxation 0.5 instead of 0.6, move it far outside tol
    dc = m / (h * h)
    xstar = (dc * (1 + h * (damp * v0 + gx * h)) + K) / (dc + K)
    assert abs(_chebyshev_closed_form(1, iters, rho, h, damp, gx, v0, K, m)[0] - xstar) > 1e-4

    def variant(seed_bad=False, relax=0.6):
        x, v = 1.0, v0
        for _ in range(n):
            xlast = x
            v = damp * v + gx * h
            x = x + h * v
            xs = (dc * x + K) / (dc + K)
            e = ep = x - xs
            w = 4 / (4 - rho * rho) if seed_bad else 1.0
            for _ in range(iters):
                w = 4 / (4 - rho * rho * w)
                e, ep = w * ((1 - relax) * e - ep) + ep, e
            x = xs + e
            v = (x - xlast) / h
        return x

    assert abs(variant() - xw) < 1e-12
    assert abs(variant(seed_bad=True) - xw) > 1e-4
    assert abs(variant(relax=0.5) - xw) > 1e-4


def test_default_params_are_the_references():
    """mgsb_default_params: the constants of PDSolver::Init
    (PDSolver.cpp:166-176: 20 iterations, dt 1/60, damping 0.95, stiffness
    8000, collision stiffness 1000, rho 0.9992, gravity (0, -9.8, 0)), the
    0.6 under-relaxation of calculatePOS (PDSolver.cu:219) and the 20 coarse
    iterations of PDSolver_MG::Step (PDSolver_MG.cpp:202); unit density (the
    reference's lumped mass is vol/4, PDSolver.cpp:46-49)."""
    p = M.default_params()
    assert p.iters == 20 and p.iters_coarse == 20
    assert p.dt == F32(1.0 / 60.0)
    assert p.damping == F32(0.95) and p.stiffness == 8000.0 and p.kcol == 1000.0 and p.rho == F32(0.9992)
    assert list(p.gravity) == [0.0, F32(-9.8), 0.0]
    assert p.relax == F32(0.6) and p.density == 1.0
    assert p.schedule == M.SCHED_SINGLE and p.if_mode == M.IF_AC  # PDSolver::Step runs runCalculateIFAc
    assert p.tri_dcd_coarse == 1  # PDSolver_MG's coarse level uses DCDByTriangle (PDSolver_MG.cpp:204)


def test_two_level_schedule_free_fall_closed_form():
    """PDSolver_MG::Step (PDSolver_MG.cpp:183-253) on an unpinned nested pair
    translating rigidly: no elastic force on either level, and the barycentric
    transfers reproduce affine fields, so both levels follow the free-fall
    closed form of calculateST + calculateV (a level missing its ST or V
    update, or a transfer that is not affine-exact, breaks it)."""
    from multigridsoftbody_amd import workloads as W

    h, damp = F32(1 / 60), F32(0.95)
    g = np.array([0.0, F32(-9.8), F32(1.1)])
    v0 = np.array([0.2, 0.5, -0.3], np.float32)
    fine, coarse = W.beam(2), W.beam(1)
    Lf = O.Level(fine.pos, fine.tets, fine.tris, stiffness=1e5, density=1000.0, fp64=True)
    Lc = O.Level(coarse.pos, coarse.tets, coarse.tris, stiffness=1e5, density=1000.0, fp64=True)
    for L, m in ((Lf, fine), (Lc, coarse)):
        L.set_fix(np.ones(m.nverts, np.float32))
        L.set_velocity(np.tile(v0, (m.nverts, 1)))
    mg = O.TwoLevel(Lc, Lf)
    n = 12
    mg.step(O.params(dt=h, damping=damp, stiffness=1e5, density=1000.0, gravity=g, iters=5, iters_coarse=5,
                     if_mode=0), n)
    v = v0.astype(np.float64)
    d = np.zeros(3)
    for _ in range(n):
        v = damp * v + g * h
        d = d + h * v
    for L, m in ((Lf, fine), (Lc, coarse)):
        assert np.allclose(L.positions(), m.pos.astype(np.float64) + d, rtol=0, atol=1e-9)
        assert np.allclose(L.get(O.V).reshape(-1, 3), v, rtol=0, atol=1e-7)


@pytest.mark.parametrize("sig", [(1.1, 0.9, 1.25), (0.9, 0.95, 1.05)])
def test_ep_density_uniform_stretch_closed_form(sig):
    """calEveryVertEp (PDSolver.cu:893-924) under x = Q diag(sig) X: every
    tet has |F - R|^2 = sum (sig - 1)^2 and current volume det(F) vol, so
    every vertex's density is +-(k/2) sum (sig - 1)^2 / prod(sig), negative
    when det F < 1."""
    from multigridsoftbody_amd import Mesh

    m = Mesh.box(2, 4, 2, (-1.0, -2.0, -1.0), (1.0, 2.0, 1.0))
    k = 5e3
    sig = np.array(sig)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    x = (m.pos.astype(np.float64) @ (Q @ np.diag(sig)).T).astype(np.float32)
    want = 0.5 * k * ((sig - 1) ** 2).sum() / sig.prod() * (1 if sig.prod() >= 1 else -1)
    # the reference sums the densities in float vectors (Simulator.h:32-34),
    # so the FP64 oracle carries their rounding too
    for fp64, rtol in ((True, 3e-6), (False, 3e-5)):
        L = O.Level(m.pos, m.tets, m.tris, stiffness=k, fp64=fp64)
        L.set_positions(x)
        d = L.ep_density()
        assert np.allclose(d, want, rtol=rtol), (fp64, d[:4], want)


@pytest.mark.parametrize("fp64", [True, False])
def test_convergence_reference_step_one_free_vertex(fp64):
    """PDSolver::StepForConvergence (PDSolver.cpp:202-221) on the one-free-vertex
    tet: the energy after each iteration is E(x) = m (x - s)^2 / (2 h^2) +
    K (x - 1)^2 / 2 (Ek + Ep; no collider) at the iterates of the
    hand-derived Chebyshev-Jacobi sequence, E[0] = E(s), the last tends to
    the minimum E(x*), and positions and velocity come back unchanged."""
    h, damp, rho = F32(1 / 60), F32(0.95), F32(0.9992)
    k, dens = 1e5, 1000.0
    L = O.Level(_POS, _TET, None, stiffness=k, density=dens, fp64=fp64)
    L.set_fix(np.array([0, 1, 0, 0], np.float32))
    vel = np.zeros((4, 3), np.float32)
    vel[1, 0] = 3.0
    L.set_velocity(vel)
    p = O.params(dt=h, damping=damp, stiffness=k, density=dens, rho=rho, gravity=(0, 0, 0), iters=10, if_mode=1)
    x_before, v_before = L.positions().copy(), L.get(O.V).copy()
    iters = 128
    E = L.step_convergence(p, iters)
    assert np.array_equal(L.positions(), x_before) and np.array_equal(L.get(O.V), v_before)
    K, m = k / 6, dens / 24
    dc = m / (h * h)
    s = 1.0 + h * (damp * 3.0)
    xs = (dc * s + K) / (dc + K)
    Ef = lambda x: m * (x - s) ** 2 / (2 * h * h) + K * (x - 1) ** 2 / 2  # noqa: E731
    e = ep = s - xs
    w = 1.0
    want = [Ef(s)]
    for _ in range(iters):
        w = 4 / (4 - rho * rho * w)
        e, ep = w * (0.4 * e - ep) + ep, e
        want.append(Ef(xs + e))
    want = np.array(want)
    rtol = 1e-9 if fp64 else 2e-4
    assert np.allclose(E, want, rtol=rtol, atol=rtol * want[0])
    assert abs(E[-1] - Ef(xs)) <= (1e-6 if fp64 else 1e-4) * (want[0] - Ef(xs))  # converged to the minimum
    assert want[0] > want[1] and E[-1] < E[1]
