"""Pins the oracle's per-tet stiffness, which the batched workload (several
bodies with their own E in one mesh) relies on: a disjoint union of bodies
stepped as one mesh equals the bodies stepped one by one, bit for bit (each
vertex gathers its own tets in the same order and nothing couples bodies)."""
import numpy as np

import oracle as O
from multigridsoftbody_amd import workloads as W


def parts_and_union():
    offs = [(0.0, 0.0), (4.0, 0.0), (0.0, 4.0)]
    ns = [2, 1, 2]
    E = [3e4, 2e5, 8e5]
    parts = []
    for (ox, oz), n in zip(offs, ns):
        lo = (W.BEAM_LO[0] + ox, W.BEAM_LO[1], W.BEAM_LO[2] + oz)
        hi = (W.BEAM_HI[0] + ox, W.BEAM_HI[1], W.BEAM_HI[2] + oz)
        parts.append(W.Mesh.box(n, 10 * n, n, lo, hi))
    return parts, W.mesh_union(parts), E


def test_union_with_per_tet_stiffness_equals_separate_bodies():
    parts, U, E = parts_and_union()
    p = O.params(dt=1 / 60, density=1000.0, iters=6, if_mode=1)
    outs = []
    for m, e in zip(parts, E):
        L = O.Level(m.pos, m.tets, m.tris, stiffness=e, density=1000.0)
        fix = W.top_fixflag(m)
        L.set_fix(fix)
        L.set_velocity(W.seeded_velocity(m, fix, seed=3))
        L.step(p, 12)
        outs.append(L.positions())
    LU = O.Level(U.pos, U.tets, U.tris, stiffness=1.0, density=1000.0)
    LU.set_tet_stiffness(np.concatenate([np.full(m.ntets, e, np.float32) for m, e in zip(parts, E)]))
    fixu = W.top_fixflag(U)
    LU.set_fix(fixu)
    LU.set_velocity(np.concatenate([W.seeded_velocity(m, W.top_fixflag(m), seed=3) for m in parts]))
    LU.step(p, 12)
    xu = LU.positions()
    o = 0
    for m, x in zip(parts, outs):
        assert np.array_equal(xu[o:o + m.nverts], x)
        o += m.nverts
    # the stiffer body deforms less under the same gravity load
    sag = [np.abs(x - m.pos).max() for m, x in zip(parts, outs)]
    assert sag[2] < sag[0]
