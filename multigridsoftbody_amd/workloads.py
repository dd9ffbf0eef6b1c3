"""Workload definitions: the box, collision, batched and organ-blob cases.

Each builder returns plain data (meshes, parameter dicts, pinned sets,
initial velocities) so that the device solver and the CPU oracle can be fed
the same inputs.  Geometry: the beam x[-1,1] y[-10,10] z[-1,1] with the end
y = 10 pinned and gravity along -y (README of the reference: fixed points at
[0,10,0]; PDSolver.cpp:137-143 holds the y = 10 rule).
"""
import numpy as np

from . import Mesh

BEAM_LO = (-1.0, -10.0, -1.0)
BEAM_HI = (1.0, 10.0, 1.0)


def beam(n):
    """n x 10n x n hex cells, 6 tets each, on the standard beam."""
    return Mesh.box(n, 10 * n, n, BEAM_LO, BEAM_HI)


def beam_hierarchy(ns):
    """Nested beam levels, finest first, e.g. (12, 4, 1)."""
    return [beam(n) for n in ns]


def top_fixflag(mesh, y=10.0, tol=1e-5):
    """fixflag (1 free, 0 pinned) with |y - 10| < tol pinned."""
    return (np.abs(mesh.pos[:, 1] - y) >= tol).astype(np.float32)


def seeded_velocity(mesh, fix, sigma=0.1, seed=0):
    """v0 ~ N(0, sigma^2) per component from numpy's PCG64 stream; pinned 0."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, sigma, size=(mesh.nverts, 3)).astype(np.float32)
    v[fix == 0] = 0.0
    return v


def physics(density=1000.0, E=1e5, dt=1.0 / 60.0, gravity=(0.0, -9.8, 0.0)):
    """Material constants of the box workloads: ARAP weight = E * rest volume
    (stiffness E per unit volume), lumped mass density * volume / 4.
    Damping 0.95, under-relaxation 0.6 and rho 0.9992 are the reference's
    (PDSolver.cpp:169-172, PDSolver.cu:219)."""
    return dict(dt=dt, damping=0.95, stiffness=E, density=density, rho=0.9992, relax=0.6, gravity=gravity)


# Named configurations ---------------------------------------------------------
SMALL = dict(name="small_box_4x40x4", ns=(4, 1), steps=200, iters=10, sweeps=2, gravity=True)
MEDIUM = dict(name="medium_box_12x120x12", ns=(12, 4, 1), steps=500, iters=10, sweeps=2, gravity=True)
# max_cycles 120: the cap is a safety net, not a budget.  The sphere starts
# with the beam inside its contact radius (1.5 r), so step 1 needs about 240
# half-cycles; when the beam meets the ground plane (steps 300-330 and
# 450-480) single steps needed more than 60 cycles
LARGE = dict(name="large_box_24x240x24_collision", ns=(24, 12, 4, 1), steps=300, iters=10, sweeps=2, gravity=False,
             max_cycles=120)


def large_colliders():
    """Sphere r = 1.5 at x0 = +2.5 moving -x at 2 m/s through y = 0, ground
    plane y = -12, penalty 1e6."""
    return dict(sphere=dict(center=(2.5, 0.0, 0.0), radius=1.5, vel=(-2.0, 0.0, 0.0)),
                plane=dict(normal=(0.0, 1.0, 0.0), d=-12.0), k=1e6)


# Batched bodies -----------------------------------------------------------------
BATCH_SIZES = (2, 4, 6, 8, 12, 16)
# level cell counts per body (coarsest 1x10x1 for all; n = 2 repeats it so that
# every body has three levels and the union stays nested level by level)
BATCH_LEVELS = {2: (2, 1, 1), 4: (4, 2, 1), 6: (6, 3, 1), 8: (8, 4, 1), 12: (12, 4, 1), 16: (16, 4, 1)}


def mesh_union(parts):
    """Disjoint union of meshes: vertices concatenated in part order, so each
    part is a contiguous vertex range (a body of the solver)."""
    pos, tets, tris, opp, off = [], [], [], [], 0
    for m in parts:
        pos.append(m.pos)
        tets.append(m.tets + off)
        tris.append(m.tris.astype(np.int64) + off)
        opp.append((m.tri_opp.astype(np.int64) if m.tri_opp is not None else np.zeros(m.ntris)) + off)
        off += m.nverts
    return Mesh(np.concatenate(pos), np.concatenate(tets), np.concatenate(tris).astype(np.uint32),
                np.concatenate(opp).astype(np.uint32))


def batched(nbodies=64, seed=0, spacing=4.0, sizes=BATCH_SIZES):
    """nbodies beams of n x 10n x n hexes, n drawn uniformly from `sizes`,
    with their own E ~ U[1e4, 1e6] and nu ~ U[0.25, 0.45] (nu is recorded;
    the ARAP energy has no Poisson term), laid out on an 8-wide grid in x-z
    `spacing` apart.  Returns the three union levels (finest first) and the
    per-body data."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.array(sizes), size=nbodies)
    E = rng.uniform(1e4, 1e6, size=nbodies).astype(np.float32)
    nu = rng.uniform(0.25, 0.45, size=nbodies).astype(np.float32)
    levels = []
    for l in range(3):
        parts = []
        for b in range(nbodies):
            n = BATCH_LEVELS[int(ns[b])][l]
            ox, oz = spacing * (b % 8), spacing * (b // 8)
            lo = (BEAM_LO[0] + ox, BEAM_LO[1], BEAM_LO[2] + oz)
            hi = (BEAM_HI[0] + ox, BEAM_HI[1], BEAM_HI[2] + oz)
            parts.append(Mesh.box(n, 10 * n, n, lo, hi))
        levels.append(mesh_union(parts))
    return dict(levels=levels, ns=ns, E=E, nu=nu)


BATCHED = dict(name="batched_64_boxes", nbodies=64, steps=200, gravity=True)


# Organ-like blob ------------------------------------------------------------------
BLOB_H = 0.225  # ~194k tets at the finest level (6 per voxel)


def blob_hierarchy(h=BLOB_H, seed=0):
    """Superquadric |x/6|^2.5 + |y/3|^2.5 + |z/4|^2.5 <= r^2.5 with a seeded
    8-sinusoid radial perturbation of amplitude 0.15, voxelised at h, 3h, 9h
    (each coarser level covers the finer one), finest first."""
    return [Mesh.blob(h=h, seed=seed, coarsen=c) for c in (1, 3, 9)]


def blob_fixflag(mesh, xmax=-5.0):
    """fixflag (1 free, 0 pinned) with the region x < -5 pinned."""
    return (mesh.pos[:, 0] >= xmax).astype(np.float32)


# max_cycles 30: a few steps hold inverted tets next to the pinned region and
# plain MG from s then needs more than 12 cycles; the cap is a safety net
BLOB = dict(name="organ_blob_h0.225", steps=300, gravity=True, max_cycles=30)
