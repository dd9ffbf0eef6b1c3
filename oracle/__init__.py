"""CPU oracle of the reference PD solver -- TEST INFRASTRUCTURE ONLY.

ctypes loader for ``oracle/liboracle.so`` (built from ``mgsb_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it, as the checker.  The product never
imports it.

Two precisions share one source: ``Level(..., fp64=False)`` follows the
reference's FP32 operation order; ``fp64=True`` is the FP64 yardstick.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

MAX_SPHERES = 8


class orc_params(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("damping", C.c_float),
        ("stiffness", C.c_float),
        ("density", C.c_float),
        ("rho", C.c_float),
        ("relax", C.c_double),
        ("gravity", C.c_float * 3),
        ("kcol", C.c_float),
        ("iters", C.c_int),
        ("iters_coarse", C.c_int),
        ("if_mode", C.c_int),
        ("tri_dcd_coarse", C.c_int),
        ("nspheres", C.c_int),
        ("sphere", (C.c_float * 8) * MAX_SPHERES),
        ("plane", C.c_float * 5),
    ]


class orc_vparams(C.Structure):
    _fields_ = [
        ("pre", C.c_int),
        ("post", C.c_int),
        ("coarse_sweeps", C.c_int),
        ("max_cycles", C.c_int),
        ("coarse_direct", C.c_int),
        ("check", C.c_int),
        ("tol", C.c_float),
        ("rho_smooth", C.c_float),
        ("relax_coarse", C.c_float),
        ("rho_coarse", C.c_float),
        ("coarse_first", C.c_int),
        ("stall_guard", C.c_int),
    ]


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

_SIGS = {
    "level_create": (_vp, [C.c_int, C.c_int, C.c_int, _fp, _ip, C.POINTER(C.c_uint32), C.c_float, C.c_float]),
    "level_free": (None, [_vp]),
    "level_bad_tet": (C.c_int, [_vp]),
    "set_threads": (None, [C.c_int]),
    "level_set_fix": (None, [_vp, _fp]),
    "level_set_velocity": (None, [_vp, _fp]),
    "level_set_positions": (None, [_vp, _fp]),
    "level_set_s": (None, [_vp, _fp]),
    "sweep": (None, [_vp, C.POINTER(orc_params), C.c_double, C.c_int]),
    "level_set_tet_k": (None, [_vp, _fp]),
    "level_get": (C.c_int, [_vp, C.c_int, _dp]),
    "step_single": (None, [_vp, C.POINTER(orc_params), C.c_int]),
    "mg_create": (_vp, [_vp, _vp]),
    "mg_free": (None, [_vp]),
    "mg_unfound": (C.c_int, [_vp, C.c_int]),
    "mg_get_maps": (None, [_vp, _ip, _fp, _ip, _fp]),
    "step_mg": (None, [_vp, C.POINTER(orc_params), C.c_int]),
    "energy": (None, [_vp, C.POINTER(orc_params), _dp]),
    "vertex_normals": (None, [_vp]),
    "step_convergence": (None, [_vp, C.POINTER(orc_params), C.c_int, _dp]),
    "ep_density": (None, [_vp, _fp]),
    "residual_norm": (C.c_double, [_vp, C.POINTER(orc_params)]),
    "polar_f": (C.c_float, [_fp, _fp, C.c_int]),
    "inverse3_f": (C.c_float, [_fp, _fp]),
    "barycentric4": (None, [_fp, _fp, _fp]),
    "tet_corner_forces": (None, [_vp, C.c_int, _dp]),
    "collide": (None, [_vp, C.POINTER(orc_params), C.c_int]),
    "ml_create": (_vp, [C.POINTER(_vp), C.c_int, C.c_float, C.c_float, C.c_float]),
    "ml_free": (None, [_vp]),
    "ml_rho": (C.c_double, [_vp, C.c_int]),
    "ml_nfree": (C.c_int, [_vp]),
    "ml_get_vfix": (None, [_vp, C.c_int, _fp]),
    "ml_stats": (None, [_vp, _dp]),
    "ml_get_mask": (None, [_vp, _fp]),
    "ml_set_warm": (None, [_vp, _fp, _fp]),
    "step_vcycle": (None, [_vp, C.POINTER(orc_params), C.POINTER(orc_vparams), C.c_int]),
    "ml_coarse_solve": (None, [_vp, _fp, _fp]),
    "ml_lin_smooth": (None, [_vp, C.c_int, C.c_int, _fp, _fp]),
    "mlp_forward": (None, [_fp, C.c_int, C.c_int, _fp, C.c_int, _dp, C.c_int]),
    "mlp_grad": (C.c_double, [_fp, C.c_int, C.c_int, _fp, _fp, C.c_int, _dp, C.c_int]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built; run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    for pfx in ("orc32_", "orc64_"):
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, pfx + name)
            fn.restype = res
            fn.argtypes = args
    return lib


lib = _load()

# level_get selectors
X, V, S, XPREV, XLAST, MASS, DIAG, VOL, B, AC, CDIAG, CFORCE, CFLAG, CDEPTH, CENERGY, VNORMAL, TRIN, FIX, FORCE, \
    CENTER, FACEN, FACEA = range(22)


def params(dt=1 / 60, damping=0.95, stiffness=8000.0, density=1.0, rho=0.9992, relax=0.6, gravity=(0, -9.8, 0),
           kcol=1000.0, iters=20, iters_coarse=20, if_mode=1, tri_dcd_coarse=1, spheres=(), plane=None):
    """orc_params with the reference defaults (PDSolver.cpp:166-176).
    spheres: iterable of (cx, cy, cz, r, active, vx, vy, vz)."""
    p = orc_params()
    p.dt, p.damping, p.stiffness, p.density, p.rho, p.relax = dt, damping, stiffness, density, rho, relax
    p.gravity[:] = [float(g) for g in gravity]
    p.kcol, p.iters, p.iters_coarse, p.if_mode, p.tri_dcd_coarse = kcol, iters, iters_coarse, if_mode, tri_dcd_coarse
    spheres = list(spheres)
    assert len(spheres) <= MAX_SPHERES
    p.nspheres = len(spheres)
    for i, s in enumerate(spheres):
        vals = list(s) + [0.0] * (8 - len(s))
        p.sphere[i][:] = [float(v) for v in vals]
    if plane is not None:
        p.plane[:] = [float(v) for v in plane]
    return p


def set_threads(n):
    lib.orc32_set_threads(n)
    lib.orc64_set_threads(n)


class Level:
    """One oracle mesh level (the state of a reference PDSolver)."""

    def __init__(self, pos, tets, tris=None, stiffness=8000.0, density=1.0, fp64=False):
        self.pfx = "orc64_" if fp64 else "orc32_"
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        self.tris = np.ascontiguousarray(tris if tris is not None else np.zeros((0, 3)), np.uint32).reshape(-1, 3)
        self.nv, self.nt, self.ns = self.pos.shape[0], self.tets.shape[0], self.tris.shape[0]
        self._h = self._f("level_create")(self.nv, self.nt, self.ns, self.pos.ctypes.data_as(_fp),
                                          self.tets.ctypes.data_as(_ip),
                                          self.tris.ctypes.data_as(C.POINTER(C.c_uint32)), stiffness, density)

    def _f(self, name):
        return getattr(lib, self.pfx + name)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._f("level_free")(h)
            self._h = None

    @property
    def bad_tet(self):
        return self._f("level_bad_tet")(self._h)

    def set_fix(self, fix):
        a = np.ascontiguousarray(fix, np.float32)
        self._f("level_set_fix")(self._h, a.ctypes.data_as(_fp))

    def set_velocity(self, v):
        a = np.ascontiguousarray(v, np.float32).reshape(-1)
        self._f("level_set_velocity")(self._h, a.ctypes.data_as(_fp))

    def set_positions(self, x):
        a = np.ascontiguousarray(x, np.float32).reshape(-1)
        self._f("level_set_positions")(self._h, a.ctypes.data_as(_fp))

    def set_s(self, s):
        a = np.ascontiguousarray(s, np.float32).reshape(-1)
        self._f("level_set_s")(self._h, a.ctypes.data_as(_fp))

    def sweep(self, p, omega=1.0, contact_block=False):
        """One PD sweep (point DCD, local step, gather, Jacobi step); with
        contact_block the V-cycle's exact 3x3 step on contact vertices."""
        self._f("sweep")(self._h, C.byref(p), float(omega), int(contact_block))

    def set_tet_stiffness(self, k):
        """Per-tet ARAP weight per unit volume (bodies with their own E)."""
        a = np.ascontiguousarray(k, np.float32).reshape(-1)
        assert a.size == self.nt
        self._f("level_set_tet_k")(self._h, a.ctypes.data_as(_fp))

    def get(self, which):
        sizes = {X: 3 * self.nv, V: 3 * self.nv, S: 3 * self.nv, XPREV: 3 * self.nv, XLAST: 3 * self.nv,
                 MASS: self.nv, DIAG: self.nv, VOL: self.nt, B: 9 * self.nt, AC: 12 * self.nt, CDIAG: 3 * self.nv,
                 CFORCE: 3 * self.nv, CFLAG: self.nv, CDEPTH: self.nv, CENERGY: self.nv, VNORMAL: 3 * self.nv,
                 TRIN: 3 * self.ns, FIX: self.nv, FORCE: 3 * self.nv, CENTER: 3 * self.nt, FACEN: 12 * self.nt,
                 FACEA: 4 * self.nt}
        out = np.zeros(max(1, sizes[which]), np.float64)
        n = self._f("level_get")(self._h, which, out.ctypes.data_as(_dp))
        return out[:n]

    def positions(self):
        return self.get(X).reshape(-1, 3)

    def step(self, p, nsteps=1):
        self._f("step_single")(self._h, C.byref(p), nsteps)

    def energy(self, p):
        out = np.zeros(5, np.float64)
        self._f("energy")(self._h, C.byref(p), out.ctypes.data_as(_dp))
        return dict(zip(("Ek", "Ep", "Ec", "Nc", "dX"), out.tolist()))

    def step_convergence(self, p, iters=128):
        """PDSolver::StepForConvergence: Ek + Ep + Ec at s and after each of
        `iters` iterations; positions and velocity restored."""
        E = np.zeros(iters + 1)
        self._f("step_convergence")(self._h, C.byref(p), iters, E.ctypes.data_as(_dp))
        return E

    def ep_density(self):
        out = np.zeros(self.nv, np.float32)
        self._f("ep_density")(self._h, out.ctypes.data_as(_fp))
        return out

    def vertex_normals(self):
        self._f("vertex_normals")(self._h)
        return self.get(VNORMAL).reshape(-1, 3)

    def residual(self, p):
        return self._f("residual_norm")(self._h, C.byref(p))

    def corner_forces(self, mode=0):
        out = np.zeros(self.nt * 12, np.float64)
        self._f("tet_corner_forces")(self._h, mode, out.ctypes.data_as(_dp))
        return out.reshape(self.nt, 4, 3)

    def collide(self, p, tri=False):
        self._f("collide")(self._h, C.byref(p), int(tri))


class TwoLevel:
    """Oracle of PDSolver_MG (coarse + fine)."""

    def __init__(self, coarse, fine):
        assert coarse.pfx == fine.pfx
        self.coarse, self.fine, self.pfx = coarse, fine, coarse.pfx
        self._h = getattr(lib, self.pfx + "mg_create")(coarse._h, fine._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(lib, self.pfx + "mg_free")(h)
            self._h = None

    def unfound(self):
        f = getattr(lib, self.pfx + "mg_unfound")
        return f(self._h, 0), f(self._h, 1)

    def maps(self):
        nf, nc = self.fine.nv, self.coarse.nv
        iids = np.zeros(nf * 4, np.int32)
        iw = np.zeros(nf * 4, np.float32)
        aids = np.zeros(nc * 4, np.int32)
        aw = np.zeros(nc * 4, np.float32)
        getattr(lib, self.pfx + "mg_get_maps")(self._h, iids.ctypes.data_as(_ip), iw.ctypes.data_as(_fp),
                                               aids.ctypes.data_as(_ip), aw.ctypes.data_as(_fp))
        return iids.reshape(nf, 4), iw.reshape(nf, 4), aids.reshape(nc, 4), aw.reshape(nc, 4)

    def step(self, p, nsteps=1):
        getattr(lib, self.pfx + "step_mg")(self._h, C.byref(p), nsteps)


def polar(F, variant=0, fp64=False):
    F = np.ascontiguousarray(F, np.float32).reshape(9)
    R = np.zeros(9, np.float32)
    det = getattr(lib, ("orc64_" if fp64 else "orc32_") + "polar_f")(F.ctypes.data_as(_fp), R.ctypes.data_as(_fp),
                                                                    variant)
    return R.reshape(3, 3), det


def barycentric4(P, tet4x3, fp64=False):
    P = np.ascontiguousarray(P, np.float32).reshape(3)
    T = np.ascontiguousarray(tet4x3, np.float32).reshape(12)
    w = np.zeros(4, np.float32)
    getattr(lib, ("orc64_" if fp64 else "orc32_") + "barycentric4")(P.ctypes.data_as(_fp), T.ctypes.data_as(_fp),
                                                                   w.ctypes.data_as(_fp))
    return w


def vparams(pre=2, post=2, coarse_sweeps=2, max_cycles=8, coarse_direct=1, check=1, tol=1e-3, rho_smooth=0.0,
            relax_coarse=0.6, rho_coarse=0.0, coarse_first=0, stall_guard=1):
    vp = orc_vparams()
    vp.pre, vp.post, vp.coarse_sweeps, vp.max_cycles = pre, post, coarse_sweeps, max_cycles
    vp.coarse_direct, vp.check, vp.tol, vp.rho_smooth = coarse_direct, check, tol, rho_smooth
    vp.relax_coarse, vp.rho_coarse, vp.coarse_first = relax_coarse, rho_coarse, coarse_first
    vp.stall_guard = stall_guard
    return vp


class MultiLevel:
    """Oracle of the residual-correction V-cycle over nested levels
    (levels[0] finest; its fixflag must be set before construction)."""

    def __init__(self, levels, relax_coarse=0.6, dt=1 / 60, rho_coarse=0.0):
        self.levels = list(levels)
        self.pfx = levels[0].pfx
        arr = (_vp * len(levels))(*[lv._h for lv in levels])
        self._h = getattr(lib, self.pfx + "ml_create")(arr, len(levels), relax_coarse, dt, rho_coarse)
        if not self._h:
            raise ValueError("ml_create failed")

    def _f(self, name):
        return getattr(lib, self.pfx + name)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._f("ml_free")(h)
            self._h = None

    def rho(self, level):
        return self._f("ml_rho")(self._h, level)

    def nfree(self):
        return self._f("ml_nfree")(self._h)

    def vfix(self, level):
        out = np.zeros(self.levels[level].nv, np.float32)
        self._f("ml_get_vfix")(self._h, level, out.ctypes.data_as(_fp))
        return out

    def active_mask(self):
        """Active-set mask of the last fine residual evaluation (0 where a
        collider acts)."""
        out = np.zeros(self.levels[0].nv, np.float32)
        self._f("ml_get_mask")(self._h, out.ctypes.data_as(_fp))
        return out

    def step(self, p, vp, nsteps=1):
        self._f("step_vcycle")(self._h, C.byref(p), C.byref(vp), nsteps)

    def coarse_solve(self, r):
        r = np.ascontiguousarray(r, np.float32).reshape(-1)
        e = np.zeros_like(r)
        self._f("ml_coarse_solve")(self._h, r.ctypes.data_as(_fp), e.ctypes.data_as(_fp))
        return e.reshape(-1, 3)

    def lin_smooth(self, level, sweeps, r):
        """`sweeps` coarse-level smoothing sweeps on A_l e = r from e = 0."""
        r = np.ascontiguousarray(r, np.float32).reshape(-1)
        e = np.zeros_like(r)
        self._f("ml_lin_smooth")(self._h, level, sweeps, r.ctypes.data_as(_fp), e.ctypes.data_as(_fp))
        return e.reshape(-1, 3)

    def stats(self):
        out = np.zeros(7, np.float64)
        self._f("ml_stats")(self._h, out.ctypes.data_as(_dp))
        return {"res0": out[0], "res": out[1], "cycles": int(out[2]), "halves": int(out[3]),
                "fallback": int(out[4]), "res_ws": out[5], "res_t": out[6],
                "rel": out[1] / out[0] if out[0] else 0.0}

    def set_warm(self, e1=None, u0=None):
        """Warm start of the next step: x = s + fix u0 + P e1 (level-1
        correction e1 (V1, 3), fine correction u0 (V0, 3)); None omits a part."""
        a = None if e1 is None else np.ascontiguousarray(e1, np.float32).reshape(-1)
        b = None if u0 is None else np.ascontiguousarray(u0, np.float32).reshape(-1)
        self._f("ml_set_warm")(self._h, None if a is None else a.ctypes.data_as(_fp),
                               None if b is None else b.ctypes.data_as(_fp))


MLP_H = 128


def mlp_nparams(inp=12, out=3):
    H = MLP_H
    return inp * H + H + 2 * (H * H + H) + H * out + out


def mlp_forward(w, x, inp=12, out=3, bf16=True):
    w = np.ascontiguousarray(w, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, inp)
    y = np.zeros((x.shape[0], out), np.float64)
    lib.orc64_mlp_forward(w.ctypes.data_as(_fp), inp, out, x.ctypes.data_as(_fp), x.shape[0], y.ctypes.data_as(_dp),
                          int(bf16))
    return y


def mlp_grad(w, x, t, inp=12, out=3, bf16=True):
    w = np.ascontiguousarray(w, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, inp)
    t = np.ascontiguousarray(t, np.float32).reshape(-1, out)
    g = np.zeros(w.size, np.float64)
    loss = lib.orc64_mlp_grad(w.ctypes.data_as(_fp), inp, out, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp),
                              x.shape[0], g.ctypes.data_as(_dp), int(bf16))
    return g, loss
