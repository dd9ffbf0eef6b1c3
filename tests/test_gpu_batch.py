"""Several bodies in one solver (the batched workload): a disjoint union of
beams, each with its own stiffness, advanced by the same launches.  Parity
of the fixed-cycle V-cycle against the oracle stepping the same union with
per-tet stiffness; per-body residual checks (every body meets the target);
the 64-body workload end to end."""
import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

KW = dict(pre_sweeps=3, post_sweeps=3, coarse_sweeps=2, rho_smooth=0.9, tol=1e-3)


def small_batch():
    return W.batched(nbodies=5, seed=7, sizes=(2, 4))


def body_tet_k(levels, ns, E):
    """Per-tet stiffness of each union level (bodies in order)."""
    out = []
    for l, m in enumerate(levels):
        k = np.concatenate([np.full(60 * W.BATCH_LEVELS[int(n)][l] ** 3, e, np.float32) for n, e in zip(ns, E)])
        assert k.size == m.ntets
        out.append(k)
    return out


def gpu_solver(b, **kw):
    s = M.Solver(b["levels"], schedule=M.SCHED_VCYCLE, if_mode=M.IF_STABLE, **kw, **W.physics())
    s.fix_plane(1, 10.0, 1e-5)
    lv0 = b["levels"][0]
    s.set_velocity(W.seeded_velocity(lv0, W.top_fixflag(lv0)))
    s.set_body_stiffness(b["E"])
    return s


def test_batch_fixed_cycles_matches_oracle_union():
    b = small_batch()
    lv = b["levels"]
    s = gpu_solver(b, max_cycles=2, check_every=0, **KW)
    assert s.body_count() == (5, True)
    phys = W.physics()
    Ls = [O.Level(m.pos, m.tets, m.tris, stiffness=1.0, density=phys["density"]) for m in lv]
    for L, k in zip(Ls, body_tet_k(lv, b["ns"], b["E"])):
        L.set_tet_stiffness(k)
    fix = W.top_fixflag(lv[0])
    Ls[0].set_fix(fix)
    Ls[0].set_velocity(W.seeded_velocity(lv[0], fix))
    ML = O.MultiLevel(Ls, relax_coarse=0.6, dt=phys["dt"])
    p = O.params(dt=phys["dt"], density=phys["density"], gravity=phys["gravity"], if_mode=3)
    vp = O.vparams(pre=3, post=3, coarse_sweeps=2, max_cycles=2, check=0, rho_smooth=0.9)
    for l in range(1, 3):
        assert abs(s.level_rho(l) - ML.rho(l)) < 1e-4 * ML.rho(l)
    assert s.coarse_nfree() == ML.nfree()
    s.step(40)
    ML.step(p, vp, 40)
    g, o = s.positions(), Ls[0].positions()
    err = float(np.abs(g - o).max() / np.abs(o).max())
    print(f"[batch fixed] rel(gpu, oracle union)={err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("cfirst", [0, 1])
def test_batch_every_body_meets_target(cfirst):
    b = small_batch()
    s = gpu_solver(b, max_cycles=12, check_every=2, warm_start=1, coarse_first=cfirst, **KW)
    for _ in range(60):
        st = s.step(1)
        r0, r = s.body_residuals()
        rel = r / r0
        assert rel.max() <= 1e-3, rel
        assert abs(st.rel - rel.max()) <= 1e-12 * max(1.0, rel.max())
    run = s.run_stats()
    assert run["failed_steps"] == 0 and run["max_rel"] <= 1e-3
    # the bodies sag by their own stiffness: stiffer bodies of the same size move less
    x = s.positions()
    lv0 = b["levels"][0]
    o, sag = 0, []
    for n in b["ns"]:
        nv = (n + 1) ** 2 * (10 * n + 1)
        sag.append(np.abs(x[o:o + nv] - lv0.pos[o:o + nv]).max())
        o += nv
    print(f"[batch adaptive] E {np.round(b['E'] / 1e3)} kPa sag {np.round(sag, 3)}")


def test_batch_64_bodies_runs():
    b = W.batched()
    s = gpu_solver(b, max_cycles=12, check_every=2, warm_start=1, use_graph=1, **KW)
    n, per_body = s.body_count()
    assert n == 64 and per_body
    s.step(20)
    r0, r = s.body_residuals()
    assert np.all(np.isfinite(s.positions()))
    assert (r / r0).max() <= 1e-3
    run = s.run_stats()
    print(f"[batch 64] half-cycles/step {run['total_halves'] / run['steps']:.2f} max rel {run['max_rel']:.2e}")
    assert run["failed_steps"] == 0
