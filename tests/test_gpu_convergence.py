"""The reference's convergence diagnostics on the device: StepForConvergence
(PDSolver.cpp:202-221) with its per-iteration energies, the energy error
(E - E_conv) / (E0 - E_conv) of a step, and the per-vertex Ep density of
runCalEnergy (PDSolver.cu:893-924), against the oracle."""
import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

SPHERE = ((0.8, -4.0, 0.0), 0.6, (-0.5, 0.0, 0.0))


def _oracle_params(phys, **kw):
    c, r, v = SPHERE
    return O.params(dt=phys["dt"], damping=phys["damping"], stiffness=phys["stiffness"], density=phys["density"],
                    rho=phys["rho"], gravity=phys["gravity"], kcol=1e4, spheres=[(*c, r, 1, *v)], **kw)


def _pair(n=4, steps=5, iters=10):
    """Device solver and FP32 oracle level, both stepped `steps` reference
    steps from the same seeded state with a moving sphere."""
    mesh = W.beam(n)
    fix = W.top_fixflag(mesh)
    v0 = W.seeded_velocity(mesh, fix)
    phys = W.physics()
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=iters, kcol=1e4, **phys)
    s.set_fixed(fix)
    s.set_velocity(v0)
    s.set_sphere(0, SPHERE[0], SPHERE[1], vel=SPHERE[2])
    s.step(steps)
    L = O.Level(mesh.pos, mesh.tets, mesh.tris, stiffness=phys["stiffness"], density=phys["density"])
    L.set_fix(fix)
    L.set_velocity(v0)
    p = _oracle_params(phys, iters=iters)
    L.step(p, steps)
    return mesh, s, L, p


def test_step_convergence_matches_oracle():
    mesh, s, L, p = _pair()
    E, last = s.step_convergence(128)
    Eo = L.step_convergence(p, 128)
    rel = np.abs(E - Eo) / np.abs(Eo)
    print(f"[convergence] E0 {E[0]:.6e} E_conv {E[-1]:.6e} (oracle {Eo[0]:.6e} {Eo[-1]:.6e}), "
          f"max rel diff {rel.max():.2e}, Nc {last['Nc']}", flush=True)
    assert rel.max() < 1e-4
    assert E[-1] < E[0] and last["Nc"] > 0


@pytest.mark.parametrize("sched", ["single", "vcycle"])
def test_step_convergence_leaves_the_trajectory_unchanged(sched):
    """The convergence step restores positions, velocity and colliders: the
    steps after it are bit-identical to those of a twin solver without it."""
    lv = W.beam_hierarchy((4, 1))
    fix = W.top_fixflag(lv[0])
    v0 = W.seeded_velocity(lv[0], fix)
    phys = W.physics()
    kw = dict(schedule=M.SCHED_SINGLE, iters=10) if sched == "single" else \
        dict(schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, warm_start=1, check_every=2, max_cycles=12)
    out = []
    for with_conv in (False, True):
        s = M.Solver(lv if sched == "vcycle" else [lv[0]], kcol=1e4, **kw, **phys)
        s.set_fixed(fix)
        s.set_velocity(v0)
        s.set_sphere(0, SPHERE[0], SPHERE[1], vel=SPHERE[2])
        s.step(4)
        if with_conv:
            x = s.positions()
            s.step_convergence(64)
            assert np.array_equal(s.positions(), x)
        s.step(6)
        out.append((s.positions(), s.velocity()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_step_error_matches_oracle():
    """(E - E_conv) / (E0 - E_conv) of one reference step (10 iterations)
    against the oracle's StepForConvergence, step and energy at the end with
    a DCD at the final positions."""
    mesh, s, L, p = _pair()
    err, E = s.step_error(128)
    Eo = L.step_convergence(p, 128)
    L.step(p, 1)
    L.collide(p)
    eo = L.energy(p)
    Eend = eo["Ek"] + eo["Ep"] + eo["Ec"]
    erro = (Eend - Eo[-1]) / (Eo[0] - Eo[-1])
    print(f"[energy error] gpu {err:.4e} (E0 {E['E0']:.6e} E_conv {E['E_conv']:.6e} E {E['E']:.6e}), "
          f"oracle {erro:.4e}", flush=True)
    assert 0 < err < 1
    # the metric divides E - E_conv (here ~0.02 of E0 - E_conv, itself ~1 %
    # of E) by E0 - E_conv: device and oracle states that agree to ~1e-5
    # (FP32 trajectories) bound the difference by 2e-5 E / (E0 - E_conv)
    assert abs(err - erro) <= 2e-5 * abs(Eend) / abs(Eo[0] - Eo[-1])
    assert abs(E["E"] - Eend) <= 1e-4 * abs(Eend)


def test_ep_density_matches_oracle():
    """The density at the device's positions after 30 steps, on the device
    and in the FP32 oracle set to the same positions."""
    mesh, s, L, p = _pair(steps=30)
    d = s.ep_density()
    L.set_positions(s.positions())
    do = L.ep_density()
    err = np.abs(d - do).max() / np.abs(do).max()
    print(f"[ep density] range [{do.min():.4g}, {do.max():.4g}], rel err {err:.2e}", flush=True)
    assert (do < 0).any() and (do > 0).any()
    assert err < 1e-5


def test_ep_density_uniform_stretch_on_device():
    """Closed form: under x = Q diag(sig) X every vertex's density is
    (k/2) sum (sig - 1)^2 / prod(sig) (negated when prod(sig) < 1)."""
    m = M.Mesh.box(2, 4, 2, (-1.0, -2.0, -1.0), (1.0, 2.0, 1.0))
    k = 5e3
    for sig in ((1.1, 0.9, 1.25), (0.9, 0.95, 1.05)):
        sig = np.array(sig)
        th = 0.7
        Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        s = M.Solver([m], schedule=M.SCHED_SINGLE, stiffness=k)
        s.set_positions((m.pos.astype(np.float64) @ (Q @ np.diag(sig)).T).astype(np.float32))
        want = 0.5 * k * ((sig - 1) ** 2).sum() / sig.prod() * (1 if sig.prod() >= 1 else -1)
        assert np.allclose(s.ep_density(), want, rtol=3e-5)
