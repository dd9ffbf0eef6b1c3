"""mgsb_energy_noalloc (energy diagnostics with no allocation: FP64 partials
in a fixed device array) against mgsb_energy through the C ABI: two calls in
a row, the second on a smaller mesh, then the larger one again, each
bit-identical to mgsb_energy (same kernels, same reduction order)."""
import ctypes as C

import numpy as np
import pytest

import multigridsoftbody_amd as M
from multigridsoftbody_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("Ek", "Ep", "Ec", "Nc", "dX")


def _stepped(n):
    mesh = W.beam(n)
    fix = W.top_fixflag(mesh)
    s = M.Solver([mesh], schedule=M.SCHED_SINGLE, iters=5, kcol=1e4, **W.physics())
    s.set_fixed(fix)
    s.set_velocity(W.seeded_velocity(mesh, fix))
    s.set_sphere(0, (0.8, 0.0, 0.0), 0.6)
    s.step(5)
    return s


def _noalloc(s):
    e = M.mgsb_energy_t()
    rc = M.lib.mgsb_energy_noalloc(s._h, C.byref(e))
    assert rc == 0, M.lib.mgsb_last_error()
    return {k: getattr(e, k) for k in KEYS}


def test_energy_noalloc_matches_energy_twice_then_smaller():
    big, small = _stepped(8), _stepped(2)  # 30,720 and 480 tets on the same GPU
    print(f"[no-alloc] big {big.levels[0].ntets} tets, small {small.levels[0].ntets} tets", flush=True)
    for s in (big, small, big):
        got = _noalloc(s)
        ref = s.energy()
        print(f"[no-alloc] {s.levels[0].ntets} tets: {got}", flush=True)
        assert got["Nc"] > 0 and got["Ec"] > 0 and got["Ep"] > 0
        for k in KEYS:
            assert got[k] == ref[k], (k, got[k], ref[k])
    # the array keeps the last call's contents: a repeat on the same state
    # gives the same numbers
    assert _noalloc(small) == small.energy()


def test_energy_noalloc_bad_arguments():
    assert M.lib.mgsb_energy_noalloc(None, None) == -1
    s = _stepped(2)
    assert M.lib.mgsb_energy_noalloc(s._h, None) == -1
    assert b"bad arguments" in M.lib.mgsb_last_error()
