"""The reference's own gmsh hierarchies through PDSolver_MG::Step
(PDSolver_MG.cpp:60-253) on the device against the oracle: Y_4_40_4 on
Y_1_10_1, cube40_4_4 on cube10_1_1, and the pair of the reference's slides,
cube120_12_12 (48,087 tets) on cube40_4_4 (analysis/display.pptx slide 2).
The levels are not nested: fine vertices bind to coarse tets and coarse
vertices to fine tets by point location (PDSolver_MG.cpp:5-58, 133-168)."""
import os

import numpy as np
import pytest

import oracle as O
import multigridsoftbody_amd as M

pytestmark = pytest.mark.gpu

# fine, coarse, long axis (its +10 end pinned), sphere centre on the way in
PAIRS = [
    ("Y_4_40_4.msh", "Y_1_10_1.msh", 1, (2.2, 0.0, 0.0)),
    ("cube40_4_4.msh", "cube10_1_1.msh", 2, (2.2, 1.0, 0.0)),
    ("cube120_12_12.msh", "cube40_4_4.msh", 2, (2.2, 1.0, 0.0)),
]


@pytest.mark.parametrize("fine,coarse,axis,center", PAIRS)
def test_two_level_reference_schedule_on_gmsh_pairs(golden, fine, coarse, axis, center):
    mf = M.Mesh.load_msh(os.path.join(golden, fine))
    mc = M.Mesh.load_msh(os.path.join(golden, coarse))
    s = M.PDSolver_MG(mc, mf)  # reference constants: calculateIF, 20 + 20 iterations, coarse triangle DCD
    s.fix_plane(axis, 10.0)
    vel = (-2.0, 0.0, 0.0)
    s.set_sphere(0, center, 1.0, vel=vel)
    n = 40
    s.step(n)
    gf, gc = s.positions(0), s.positions(1)
    Lf = O.Level(mf.pos, mf.tets, mf.tris)
    Lc = O.Level(mc.pos, mc.tets, mc.tris)
    Lf.set_fix(s.fixed(0))
    Lc.set_fix(s.fixed(1))
    assert (s.fixed(0) == 0).sum() > 0 and (s.fixed(1) == 0).sum() > 0
    mg = O.TwoLevel(Lc, Lf)
    assert mg.unfound() == (0, 0)
    mg.step(O.params(if_mode=0, iters=20, iters_coarse=20, spheres=[(*center, 1.0, 1, *vel)]), n)
    of, oc = Lf.positions(), Lc.positions()
    ef = float(np.abs(gf - of).max() / np.abs(of).max())
    ec = float(np.abs(gc - oc).max() / np.abs(oc).max())
    moved = float(np.abs(of - mf.pos).max())
    print(f"[gmsh {fine} on {coarse}] {mf.ntets}/{mc.ntets} tets, max displacement {moved:.3f}, "
          f"rel err fine {ef:.2e} coarse {ec:.2e}", flush=True)
    assert np.isfinite(gf).all() and moved > 0.05
    assert ef < 1e-4 and ec < 1e-4
