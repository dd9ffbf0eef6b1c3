"""Device checks of single pieces against the oracle and against closed
forms: the per-tet local step in every mode, the SphereFixer rule on every
level, all five energy terms, and the free-fall and one-free-vertex
Chebyshev trajectories of the reference schedule."""
import os

import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

# device local-step mode -> oracle tet_forces_on mode
ORACLE_MODE = {M.IF_PLAIN: 0, M.IF_AC: 1, M.IF_STABLE: 3, M.IF_LINEAR: 2}


def _rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def _deformed(mesh, strain, rng, invert=False):
    """x = A X + noise: A a rotation times a stretch of size `strain`
    (reflected along one axis when invert), plus per-vertex noise of the same
    relative size, so every tet sees a different F of that strain level."""
    X = mesh.pos.astype(np.float64)
    d = 1.0 + strain * rng.uniform(-1, 1, size=3)
    if invert:
        d[2] = -d[2]
    A = _rot(rng) @ np.diag(d) @ _rot(rng)
    h = np.ptp(X[:, 0]) / 4
    x = X @ A.T + strain * h * rng.normal(size=X.shape)
    return x.astype(np.float32)


def _F_and_scale(mesh, x, pre, k):
    """Per-tet F = Ds Dm^-1 and the force scale k vol max|F AC| (the size of
    the terms that cancel in (R - F) AC)."""
    t = mesh.tets
    x = x.astype(np.float64)
    D = np.stack([x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]], x[t[:, 3]] - x[t[:, 0]]], axis=2)
    B = pre["B"].reshape(-1, 3, 3).astype(np.float64)
    F = D @ B
    AC = pre["AC"].reshape(-1, 3, 4).astype(np.float64)
    scale = k * pre["vol"].astype(np.float64) * np.abs(F @ AC).max(axis=(1, 2))
    return F, scale


@pytest.mark.parametrize("if_mode", [M.IF_PLAIN, M.IF_AC, M.IF_STABLE, M.IF_LINEAR])
def test_tet_forces_every_mode_match_oracle(if_mode):
    """Per-tet corner forces (calculateIF, PDSolver.cu:128-183; calTetIFV1,
    PDSolverAc.cu:630-662; the stable and linear variants) on the device
    against the FP32 oracle, at strains 1e-4 .. 1e-1 and on inverted tets.
    Error per tet: max |cf_gpu - cf_oracle| / (k vol max|F AC|).

    Both sides run the same FP32 operations; what remains are the ulp
    differences of acosf / cosf between the device and host libraries.  With
    det F > 0 they move the polar's eigenvalue by an ulp: error < 1e-5
    (median 0, the forces are bit-identical for most tets).  On inverted tets
    the closed form is ill-conditioned (sqrt(-lambda^2 + I_c + 2 III_u /
    lambda) of a small difference), so there an ulp changes R visibly; the
    device must then be as accurate as the FP32 oracle itself against the
    FP64 oracle: max error vs FP64 within 1.5x the FP32 oracle's (+1e-5).

    The stable variant has no counterpart in the reference, and the device
    evaluates it with fused multiply-adds and the hardware reciprocal, square
    root and cosine, the oracle with IEEE operations: there the FP64 oracle is
    the yardstick everywhere, the device within 1.5x the FP32 oracle's error
    (+1e-6; +1e-5 inverted), and with det F > 0 the median difference of the
    two FP32 evaluations below 3e-6 (a few ulps of the forces' scale)."""
    rng = np.random.default_rng(5)
    mesh = W.beam(4)
    k = 1e5
    pre = mesh.precompute(k, 1000.0)
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=1, stiffness=k, density=1000.0)
    cases = [(e, False) for e in (1e-4, 3e-4, 1e-3, 1e-2, 1e-1)] + [(0.3, True), (0.02, True)]
    for strain, inv in cases:
        x = _deformed(mesh, strain, rng, inv)
        gpu = s.tet_forces(x, if_mode).astype(np.float64)
        orc = []
        for fp64 in (False, True):
            L = O.Level(mesh.pos, mesh.tets, mesh.tris, stiffness=k, density=1000.0, fp64=fp64)
            L.set_positions(x)
            orc.append(L.corner_forces(ORACLE_MODE[if_mode]))
        F, scale = _F_and_scale(mesh, x, pre, k)
        e32 = np.abs(gpu - orc[0]).max(axis=(1, 2)) / scale
        e64 = np.abs(gpu - orc[1]).max(axis=(1, 2)) / scale
        o64 = np.abs(orc[0] - orc[1]).max(axis=(1, 2)) / scale
        ninv = int((np.linalg.det(F) < 0).sum())
        print(f"[tet forces mode {if_mode}] strain {strain:.0e}{' inverted' if inv else ''}: "
              f"{ninv} tets with det F < 0; gpu vs o32 max {e32.max():.2e} p99.9 {np.quantile(e32, 0.999):.2e} "
              f"median {np.median(e32):.2e}; gpu vs o64 max {e64.max():.2e}; o32 vs o64 max {o64.max():.2e}",
              flush=True)
        assert np.isfinite(gpu).all()
        if if_mode == M.IF_STABLE:
            assert inv or np.median(e32) < 3e-6
            assert e64.max() <= 1.5 * o64.max() + (1e-5 if inv else 1e-6), (strain, e64.max(), o64.max())
            assert (ninv > 0) == inv
            continue
        assert np.median(e32) < 1e-6
        if inv:
            assert ninv > 0
            assert e64.max() <= 1.5 * o64.max() + 1e-5
        else:
            assert ninv == 0
            assert e32.max() < 1e-5, (strain, e32.max())


def test_plain_polar_is_get_rotation_d_not_the_float_form():
    """IF_PLAIN evaluates GetRotation_D's l to double accuracy (utils.cu:161):
    at strain 1e-3 it agrees with oracle mode 0 to FP32 rounding, while the
    FP32 form of getTetFRV2 (the device's IF_AC without its 1e-5 drop, which
    never fires here) is off by about sqrt(k) of |F|."""
    rng = np.random.default_rng(9)
    mesh = W.beam(4)
    k = 1e5
    pre = mesh.precompute(k, 1000.0)
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=1, stiffness=k, density=1000.0)
    x = _deformed(mesh, 1e-3, rng)
    L = O.Level(mesh.pos, mesh.tets, mesh.tris, stiffness=k, density=1000.0)
    L.set_positions(x)
    ref = L.corner_forces(0)
    _, scale = _F_and_scale(mesh, x, pre, k)
    e_plain = (np.abs(s.tet_forces(x, M.IF_PLAIN) - ref).max(axis=(1, 2)) / scale).max()
    e_float = (np.abs(s.tet_forces(x, M.IF_AC) - ref).max(axis=(1, 2)) / scale).max()
    print(f"[polar variants] strain 1e-3: IF_PLAIN vs mode 0 {e_plain:.2e}, FP32 l vs mode 0 {e_float:.2e}", flush=True)
    assert e_plain < 1e-5
    assert e_float > 100 * e_plain


def _sphere_inside(X, spheres):
    X = X.astype(np.float32)
    inside = np.zeros(len(X), bool)
    for c in spheres:
        d = X - np.asarray(c[:3], np.float32)
        dist = np.sqrt((d * d).sum(axis=1, dtype=np.float32), dtype=np.float32)
        inside |= dist < np.float32(c[3])
    return inside


def test_fix_spheres_pins_strictly_inside_on_every_level():
    """SetFixedVert's SphereFixer rule (PDSolver.cpp:145-161) on a 3-level
    beam hierarchy: a vertex is pinned iff its rest position lies strictly
    inside a sphere (|X - c| < r).  The first sphere is centred on a vertex
    with r = the cell size, so its three axis neighbours at distance exactly r
    stay free; the second holds vertices of every level."""
    lv = W.beam_hierarchy((4, 2, 1))
    spheres = [(-1.0, 10.0, -1.0, 0.5), (0.0, -9.0, 0.0, 1.9)]
    s = M.Solver(lv, schedule=M.SCHED_SINGLE)
    s.fix_spheres(spheres)
    for l, m in enumerate(lv):
        want = np.where(_sphere_inside(m.pos, spheres), 0.0, 1.0).astype(np.float32)
        got = s.fixed(l)
        assert np.array_equal(got, want), l
        assert (want == 0).sum() > 0, l
    X = lv[0].pos
    on_r = np.isclose(np.linalg.norm(X - np.array([-1.0, 10.0, -1.0]), axis=1), 0.5)
    assert on_r.sum() == 3 and np.all(s.fixed(0)[on_r] == 1)
    # pinned vertices do not move under the reference schedule; the others fall
    s.step(10)
    x = s.positions(0)
    pinned = s.fixed(0) == 0
    assert np.array_equal(x[pinned], X[pinned])
    assert np.abs(x[~pinned] - X[~pinned]).max(axis=1).min() > 0


def test_reference_sphere_fixers_on_the_liver(golden):
    """The reference's live configuration (Simulator.cpp:52-54): three
    SphereFixers on liver700_3500.msh, reference constants and local step;
    30 steps on the device against the oracle with the same pinned set."""
    mesh = M.Mesh.load_msh(os.path.join(golden, "liver700_3500.msh"))
    fixers = [(-4.0, 4.7, -3.0, 1.0), (-4.0, 4.7, 3.0, 1.0), (7.5, 4.2, 2.0, 1.0)]
    s = M.PDSolver(mesh)
    s.fix_spheres(fixers)
    fix = s.fixed()
    assert np.array_equal(fix == 0, _sphere_inside(mesh.pos, fixers))
    assert (fix == 0).sum() > 0
    s.step(30)
    gpu = s.positions()
    L = O.Level(mesh.pos, mesh.tets, mesh.tris)
    L.set_fix(fix)
    L.step(O.params(), 30)
    ref = L.positions()
    err = float(np.abs(gpu - ref).max() / np.abs(ref).max())
    print(f"[liver + sphere fixers] {int((fix == 0).sum())} pinned, rel err {err:.2e}", flush=True)
    assert err < 1e-4
    assert np.array_equal(gpu[fix == 0], mesh.pos[fix == 0])


def test_energy_all_terms_match_oracle():
    """runCalEnergy (PDSolver.cu:816-913): Ek, Ep, Ec, Nc and dX after one
    step of one PD iteration with a sphere in contact, from the same deformed
    state and velocity on the device and in the FP32 oracle, so s, x, x_prev
    and the collision arrays agree to FP32 rounding on both sides."""
    mesh = W.beam(4)
    phys = W.physics()
    rng = np.random.default_rng(3)
    x0 = _deformed(mesh, 0.05, rng)
    fix = W.top_fixflag(mesh)
    v0 = W.seeded_velocity(mesh, fix)
    sph = (0.8, 0.0, 0.0, 0.6)
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=1, kcol=1e4, **phys)
    s.set_fixed(fix)
    s.set_positions(x0)
    s.set_velocity(v0)
    s.set_sphere(0, sph[:3], sph[3])
    s.step(1)
    e = s.energy()
    L = O.Level(mesh.pos, mesh.tets, mesh.tris, stiffness=phys["stiffness"], density=phys["density"])
    L.set_fix(fix)
    L.set_positions(x0)
    L.set_velocity(v0)
    p = O.params(dt=phys["dt"], damping=phys["damping"], stiffness=phys["stiffness"], density=phys["density"],
                 rho=phys["rho"], gravity=phys["gravity"], iters=1, kcol=1e4, spheres=[(*sph, 1)])
    L.step(p, 1)
    eo = L.energy(p)
    print(f"[energy] gpu {e}\n[energy] oracle {eo}", flush=True)
    assert eo["Nc"] > 0 and eo["Ec"] > 0 and eo["Ek"] > 0 and eo["dX"] > 0
    assert e["Nc"] == eo["Nc"]
    for key in ("Ek", "Ep", "Ec", "dX"):
        assert abs(e[key] - eo[key]) <= 1e-4 * abs(eo[key]), key


_POS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
_TET = np.array([[0, 1, 2, 3]], np.int32)


def test_free_fall_and_chebyshev_closed_forms_on_device():
    """The reference schedule on the unit tet against the hand-derived
    trajectories of the oracle's closed-form tests: a free tet in free fall
    (v_n = damp v_{n-1} + g h) and one free vertex whose Jacobi step is exact,
    so its error follows e_{k+1} = w_k (0.4 e_k - e_{k-1}) + e_{k-1},
    w_k = 4/(4 - rho^2 w_{k-1}), w_0 = 1."""
    from test_oracle_closed_forms import F32, _chebyshev_closed_form

    mesh = M.Mesh(_POS, _TET)
    h, damp, rho, k, dens = F32(1 / 60), F32(0.95), F32(0.9992), 1e5, 1000.0
    g = (F32(0.7), F32(-9.8), F32(0.3))
    for if_mode in (M.IF_AC, M.IF_PLAIN):
        s = M.Solver([mesh], schedule=M.SCHED_SINGLE, if_mode=if_mode, iters=10, dt=h, damping=damp, rho=rho,
                     stiffness=k, density=dens, gravity=g)
        v0 = np.array([0.4, -0.25, 1.5], np.float32)
        s.set_velocity(np.tile(v0, (4, 1)))
        s.step(25)
        v = v0.astype(np.float64)
        x = _POS.astype(np.float64)
        for _ in range(25):
            v = damp * v + np.array(g) * h
            x = x + h * v
        assert np.allclose(s.positions(), x, rtol=2e-5, atol=2e-5)
        assert np.allclose(s.velocity(), np.tile(v, (4, 1)), rtol=2e-5, atol=2e-5)

        s = M.Solver([mesh], schedule=M.SCHED_SINGLE, if_mode=if_mode, iters=10, dt=h, damping=damp, rho=rho,
                     stiffness=k, density=dens, gravity=(F32(-2.0), 0.0, 0.0))
        s.set_fixed(np.array([0, 1, 0, 0], np.float32))
        vel = np.zeros((4, 3), np.float32)
        vel[1, 0] = 3.0
        s.set_velocity(vel)
        s.step(4)
        xw, _ = _chebyshev_closed_form(4, 10, rho, h, damp, F32(-2.0), 3.0, k / 6, dens / 24)
        got = s.positions()
        assert np.allclose(got[1], [xw, 0, 0], rtol=0, atol=3e-6), (got[1], xw)
        assert np.array_equal(got[[0, 2, 3]], _POS[[0, 2, 3]])
