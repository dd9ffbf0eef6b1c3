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
LARGE = dict(name="large_box_24x240x24_collision", ns=(24, 12, 4, 1), steps=300, iters=10, sweeps=2, gravity=False)


def large_colliders():
    """Sphere r = 1.5 at x0 = +2.5 moving -x at 2 m/s through y = 0, ground
    plane y = -12, penalty 1e6."""
    return dict(sphere=dict(center=(2.5, 0.0, 0.0), radius=1.5, vel=(-2.0, 0.0, 0.0)),
                plane=dict(normal=(0.0, 1.0, 0.0), d=-12.0), k=1e6)
