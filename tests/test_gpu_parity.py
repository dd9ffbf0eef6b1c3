"""Device parity: libmgsb (HIP, through the C ABI) against the CPU oracle on
the same seeded inputs.  Positions must agree within 1e-4 relative
(||x_gpu - x_oracle||_inf / ||x_oracle||_inf, FP32 arithmetic); max and RMS
errors against the FP64 oracle are printed."""
import os

import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def rms(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))


def oracle_params(phys, **kw):
    return O.params(dt=phys["dt"], damping=phys["damping"], stiffness=phys["stiffness"], density=phys["density"],
                    rho=phys["rho"], relax=phys["relax"], gravity=phys["gravity"], **kw)


def report(tag, gpu, o32, o64):
    print(f"[{tag}] rel(gpu,o32)={rel_err(gpu, o32):.3e} rel(gpu,o64)={rel_err(gpu, o64):.3e} "
          f"max|gpu-o64|={np.abs(gpu - o64).max():.3e} rms(gpu,o64)={rms(gpu, o64):.3e} "
          f"rel(o32,o64)={rel_err(o32, o64):.3e}", flush=True)


@pytest.mark.parametrize("if_mode", [M.IF_AC, M.IF_PLAIN])
def test_single_level_small_box(if_mode):
    """PDSolver::Step on the 4x40x4 beam (3,840 tets), 10 iterations/step,
    E = 1e5, density 1000, seeded v0 ~ N(0, 0.1), 200 steps."""
    mesh = W.beam(4)
    fix = W.top_fixflag(mesh)
    v0 = W.seeded_velocity(mesh, fix)
    phys = W.physics()
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, if_mode=if_mode, iters=10, **phys)
    s.set_fixed(fix)
    s.set_velocity(v0)
    s.step(200)
    gpu = s.positions()
    outs = []
    for fp64 in (False, True):
        L = O.Level(mesh.pos, mesh.tets, mesh.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=fp64)
        L.set_fix(fix)
        L.set_velocity(v0)
        L.step(oracle_params(phys, iters=10, if_mode=if_mode), 200)
        outs.append(L.positions())
    report(f"single if_mode={if_mode}", gpu, *outs)
    assert np.isfinite(gpu).all()
    assert rel_err(gpu, outs[0]) < REL_TOL
    assert rel_err(gpu, outs[1]) < REL_TOL


def test_single_level_reference_constants_msh_with_sphere(golden):
    """Reference defaults (k 8000, unit density, 20 iterations, Ac local step)
    on the reference's own Y_4_40_4.msh with a moving sphere collider."""
    mesh = M.Mesh.load_msh(os.path.join(golden, "Y_4_40_4.msh"))
    fix = W.top_fixflag(mesh)
    s = M.PDSolver(mesh)
    s.SetFixedVert(fix)
    s.set_sphere(0, (3.0, 0.0, 0.0), 1.0, vel=(-3.0, 0.0, 0.0))
    s.step(60)  # sphere centre reaches x = 0, inside the beam
    gpu = s.positions()
    flags, depth = s.collision()
    outs = []
    for fp64 in (False, True):
        L = O.Level(mesh.pos, mesh.tets, mesh.tris, fp64=fp64)
        L.set_fix(fix)
        L.step(O.params(spheres=[(3.0, 0, 0, 1.0, 1, -3.0, 0, 0)]), 60)
        outs.append(L.positions())
        if not fp64:
            oflags = L.get(O.CFLAG)
            odepth = L.get(O.CDEPTH)
    report("single Y_4_40_4 + sphere", gpu, *outs)
    # the collision changed the trajectory (compare with no sphere)
    L = O.Level(mesh.pos, mesh.tets, mesh.tris)
    L.set_fix(fix)
    L.step(O.params(), 60)
    assert np.abs(outs[0] - L.positions()).max() > 0.05, "sphere never touched the beam"
    # flags / depths of the last iteration (t > r/2 criterion, PDSolver.cu:301-304)
    assert np.array_equal(flags, oflags.astype(np.int32))
    assert np.abs(depth - odepth).max() < 1e-4
    assert rel_err(gpu, outs[0]) < REL_TOL


def test_two_level_reference_schedule():
    """PDSolver_MG::Step: coarse 1x10x1 (tri DCD) -> prolong -> fine 4x40x4 ->
    restrict -> updateMapping, 20 + 20 iterations, sphere active."""
    fine, coarse = W.beam(4), W.beam(1)
    phys = W.physics()
    s = M.Solver([fine, coarse], schedule=M.SCHED_TWOLEVEL, if_mode=M.IF_PLAIN, iters=20, iters_coarse=20,
                 kcol=1e4, **phys)
    s.fix_plane(1, 10.0, 1e-5)
    s.set_sphere(0, (2.2, 0.0, 0.0), 1.0, vel=(-2.0, 0.0, 0.0))
    s.step(80)
    gpu_f, gpu_c = s.positions(0), s.positions(1)
    outs = []
    for fp64 in (False, True):
        Lf = O.Level(fine.pos, fine.tets, fine.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=fp64)
        Lc = O.Level(coarse.pos, coarse.tets, coarse.tris, stiffness=phys["stiffness"], density=phys["density"],
                     fp64=fp64)
        Lf.set_fix(W.top_fixflag(fine))
        Lc.set_fix(W.top_fixflag(coarse))
        mg = O.TwoLevel(Lc, Lf)
        mg.step(oracle_params(phys, iters=20, iters_coarse=20, if_mode=0, kcol=1e4,
                              spheres=[(2.2, 0, 0, 1.0, 1, -2.0, 0, 0)]), 80)
        outs.append((Lf.positions(), Lc.positions()))
    report("two-level fine", gpu_f, outs[0][0], outs[1][0])
    report("two-level coarse", gpu_c, outs[0][1], outs[1][1])
    assert rel_err(gpu_f, outs[0][0]) < REL_TOL
    assert rel_err(gpu_c, outs[0][1]) < REL_TOL


def test_normals_energy_and_render_layout():
    mesh = W.beam(4)
    fix = W.top_fixflag(mesh)
    phys = W.physics()
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=10, **phys)
    s.set_fixed(fix)
    s.set_velocity(W.seeded_velocity(mesh, fix))
    s.step(30)
    x = s.positions()
    L = O.Level(mesh.pos, mesh.tets, mesh.tris, stiffness=phys["stiffness"], density=phys["density"], fp64=True)
    L.set_positions(x)
    n_o = L.vertex_normals()
    n_g = s.normals()
    assert np.abs(n_g - n_o).max() < 1e-5
    r9 = s.render9()
    assert r9.shape == (mesh.nverts, 9)
    assert np.array_equal(r9[:, :3], x) and np.array_equal(r9[:, 3:6], n_g) and not r9[:, 6:].any()
    e = s.energy()
    eo = L.energy(oracle_params(phys))
    assert abs(e["Ep"] - eo["Ep"]) <= 1e-4 * abs(eo["Ep"]) + 1e-9
    assert e["Ep"] > 0 and e["Ek"] >= 0


def test_graph_replay_matches_eager():
    mesh = W.beam(4)
    fix = W.top_fixflag(mesh)
    v0 = W.seeded_velocity(mesh, fix)
    res = []
    for g in (0, 1):
        s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=10, use_graph=g, **W.physics())
        s.set_fixed(fix)
        s.set_velocity(v0)
        for _ in range(4):  # several replays with eager host work between them
            s.step(5)
            s.sync()
            res.append(s.positions())
    n = len(res) // 2
    for a, b in zip(res[:n], res[n:]):
        assert np.array_equal(a, b)


def test_kernel_timing_counts_launches():
    mesh = W.beam(4)
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=10, **W.physics())
    s.fix_plane(1, 10.0)
    s.set_timing(True)
    s.step(3)
    ms, n = s.kernel_time(0)
    assert n == 30 and ms > 0
