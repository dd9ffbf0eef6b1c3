"""multigridsoftbody_amd -- MI355X-native multigrid projective-dynamics solver.

Python host layer over the C ABI of ``libmgsb.so`` (declared in
``include/mgsb.h``).  It mirrors the reference solver interface
(``PDSolver`` / ``PDSolver_MG`` of MultiGridSoftBody/src/PDSolver.h:70-110 and
PDSolver_MG.h:4-25): build from tet-mesh arrays, set the fixed-vertex set,
``Step()``, read positions back.  All arithmetic runs in the HIP kernels of
the library; this module only marshals arrays.

The library must have been built;
importing raises if it is missing -- there is no CPU fallback.
Build with ``python multigridsoftbody_amd/build.py``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmgsb.so")

SCHED_SINGLE, SCHED_TWOLEVEL, SCHED_VCYCLE, SCHED_LEARNED = 0, 1, 2, 3
IF_PLAIN, IF_AC, IF_STABLE, IF_LINEAR = 0, 1, 2, 3
MAX_SPHERES = 8


class MgsbError(RuntimeError):
    """A libmgsb call returned a negative status."""


class mgsb_mesh(C.Structure):
    _fields_ = [
        ("nverts", C.c_int32),
        ("ntets", C.c_int32),
        ("ntris", C.c_int32),
        ("pos", C.POINTER(C.c_float)),
        ("tets", C.POINTER(C.c_int32)),
        ("tris", C.POINTER(C.c_uint32)),
        ("tri_opp", C.POINTER(C.c_uint32)),
    ]


class mgsb_params(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("damping", C.c_float),
        ("stiffness", C.c_float),
        ("density", C.c_float),
        ("rho", C.c_float),
        ("relax", C.c_float),
        ("gravity", C.c_float * 3),
        ("kcol", C.c_float),
        ("iters", C.c_int32),
        ("iters_coarse", C.c_int32),
        ("schedule", C.c_int32),
        ("if_mode", C.c_int32),
        ("tri_dcd_coarse", C.c_int32),
        ("pre_sweeps", C.c_int32),
        ("post_sweeps", C.c_int32),
        ("coarse_sweeps", C.c_int32),
        ("max_cycles", C.c_int32),
        ("tol", C.c_float),
        ("check_every", C.c_int32),
        ("use_graph", C.c_int32),
        ("block_tet", C.c_int32),
        ("block_vert", C.c_int32),
        ("device", C.c_int32),
        ("rho_smooth", C.c_float),
        ("relax_coarse", C.c_float),
        ("rho_coarse", C.c_float),
        ("coarse_direct", C.c_int32),
        ("warm_start", C.c_int32),
        ("learn_gate", C.c_int32),
        ("vert_lanes", C.c_int32),
        ("learn_slack", C.c_float),
        ("coarse_dense", C.c_int32),
        ("tile_gather", C.c_int32),
        ("coarse_first", C.c_int32),
        ("fuse_sweeps", C.c_int32),
        ("stall_guard", C.c_int32),
    ]


class mgsb_step_stats(C.Structure):
    _fields_ = [
        ("res0", C.c_double),
        ("res", C.c_double),
        ("rel", C.c_double),
        ("cycles", C.c_int32),
        ("fallback", C.c_int32),
        ("steps", C.c_int32),
        ("halves", C.c_int32),
    ]


class mgsb_run_stats_t(C.Structure):
    _fields_ = [("max_rel", C.c_double), ("total_cycles", C.c_int64), ("steps", C.c_int32),
                ("failed_steps", C.c_int32), ("fallbacks", C.c_int32), ("total_halves", C.c_int32)]


class mgsb_tune_result(C.Structure):
    _fields_ = [("block_tet", C.c_int32), ("block_vert", C.c_int32), ("pre_sweeps", C.c_int32),
                ("post_sweeps", C.c_int32), ("ms_per_step", C.c_double), ("ms_per_step_default", C.c_double),
                ("trials", C.c_int32), ("vert_lanes", C.c_int32), ("tile_gather", C.c_int32),
                ("sweep_us", C.c_double), ("sweep_us_default", C.c_double), ("fuse_sweeps", C.c_int32)]


class mgsb_energy_t(C.Structure):
    _fields_ = [("Ek", C.c_double), ("Ep", C.c_double), ("Ec", C.c_double), ("Nc", C.c_double), ("dX", C.c_double)]


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/mgsb.h
SIGNATURES = {
    "mgsb_load_msh": (C.c_int, [C.c_char_p, C.POINTER(mgsb_mesh)]),
    "mgsb_extract_surface": (C.c_int, [C.POINTER(mgsb_mesh)]),
    "mgsb_make_box": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _fp, _fp, C.POINTER(mgsb_mesh)]),
    "mgsb_make_blob": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_float,
                                 C.POINTER(mgsb_mesh)]),
    "mgsb_make_blob_level": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_float,
                                       C.c_int32, C.POINTER(mgsb_mesh)]),
    "mgsb_mesh_free": (None, [C.POINTER(mgsb_mesh)]),
    "mgsb_precompute": (C.c_int, [C.POINTER(mgsb_mesh), C.c_float, C.c_float, _fp, _fp, _fp, _fp, _fp]),
    "mgsb_bind_points": (C.c_int, [C.POINTER(mgsb_mesh), _fp, C.c_int32, C.c_int32, _ip, _fp]),
    "mgsb_mesh_tiles": (C.c_int, [C.POINTER(mgsb_mesh), _ip, _ip, _ip, _ip, _ip, C.POINTER(C.c_uint32)]),
    "mgsb_default_params": (None, [C.POINTER(mgsb_params)]),
    "mgsb_create": (C.c_int, [C.POINTER(mgsb_mesh), C.c_int32, C.POINTER(mgsb_params), C.POINTER(_vp)]),
    "mgsb_destroy": (None, [_vp]),
    "mgsb_set_params": (C.c_int, [_vp, C.POINTER(mgsb_params)]),
    "mgsb_set_fixed": (C.c_int, [_vp, C.c_int32, _fp]),
    "mgsb_fix_spheres": (C.c_int, [_vp, _fp, C.c_int32]),
    "mgsb_fix_plane": (C.c_int, [_vp, C.c_int32, C.c_float, C.c_float]),
    "mgsb_read_fixed": (C.c_int, [_vp, C.c_int32, _fp]),
    "mgsb_set_sphere": (C.c_int, [_vp, C.c_int32, _fp, C.c_float, _fp, C.c_int32]),
    "mgsb_set_plane": (C.c_int, [_vp, _fp, C.c_float, C.c_float, C.c_int32]),
    "mgsb_set_velocity": (C.c_int, [_vp, _fp]),
    "mgsb_set_positions": (C.c_int, [_vp, C.c_int32, _fp]),
    "mgsb_step": (C.c_int, [_vp, C.c_int32, C.POINTER(mgsb_step_stats)]),
    "mgsb_sync": (C.c_int, [_vp]),
    "mgsb_read_positions": (C.c_int, [_vp, C.c_int32, _fp]),
    "mgsb_read_velocity": (C.c_int, [_vp, C.c_int32, _fp]),
    "mgsb_read_normals": (C.c_int, [_vp, _fp]),
    "mgsb_read_render9": (C.c_int, [_vp, _fp]),
    "mgsb_read_collision": (C.c_int, [_vp, _ip, _fp]),
    "mgsb_energy": (C.c_int, [_vp, C.POINTER(mgsb_energy_t)]),
    "mgsb_energy_noalloc": (C.c_int, [_vp, C.POINTER(mgsb_energy_t)]),
    "mgsb_step_convergence": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(mgsb_energy_t)]),
    "mgsb_step_error": (C.c_int, [_vp, C.c_int32, _dp, _dp]),
    "mgsb_read_ep_density": (C.c_int, [_vp, _fp]),
    "mgsb_residual": (C.c_int, [_vp, _dp]),
    "mgsb_level_info": (C.c_int, [_vp, C.c_int32, _ip, _ip, _ip]),
    "mgsb_level_rho": (C.c_int, [_vp, C.c_int32, _dp]),
    "mgsb_learn_record": (C.c_int, [_vp, C.c_int32]),
    "mgsb_learn_samples": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "mgsb_learn_clear": (C.c_int, [_vp]),
    "mgsb_learn_train": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_float, C.c_uint32, _fp]),
    "mgsb_learn_nparams": (C.c_int, [_vp, _ip]),
    "mgsb_learn_weights": (C.c_int, [_vp, _fp, C.c_int32, C.c_int32]),
    "mgsb_learn_norm": (C.c_int, [_vp, _fp, C.c_int32]),
    "mgsb_learn_last": (C.c_int, [_vp, _fp]),
    "mgsb_mlp_forward": (C.c_int, [_vp, _fp, C.c_int32, _fp, C.c_int32]),
    "mgsb_mlp_grad": (C.c_int, [_vp, _fp, _fp, C.c_int32, _fp, _dp]),
    "mgsb_run_stats": (C.c_int, [_vp, C.POINTER(mgsb_run_stats_t), C.c_int32]),
    "mgsb_step_history": (C.c_int, [_vp, _dp, C.c_int32, _ip]),
    "mgsb_coarse_nfree": (C.c_int, [_vp, _ip]),
    "mgsb_set_timing": (C.c_int, [_vp, C.c_int32]),
    "mgsb_kernel_time": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "mgsb_autotune": (C.c_int, [_vp, C.c_int32, C.POINTER(mgsb_tune_result)]),
    "mgsb_body_count": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mgsb_set_body_stiffness": (C.c_int, [_vp, _fp, C.c_int32]),
    "mgsb_body_residuals": (C.c_int, [_vp, _dp, _dp, C.c_int32]),
    "mgsb_time_kernel": (C.c_int, [_vp, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "mgsb_tet_forces": (C.c_int, [_vp, C.c_int32, C.c_int32, _fp, _fp]),
    "mgsb_last_error": (C.c_char_p, []),
    "mgsb_version": (C.c_char_p, []),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python multigridsoftbody_amd/build.py` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc, what):
    if rc < 0:
        raise MgsbError(f"{what} failed ({rc}): {lib.mgsb_last_error().decode(errors='replace')}")
    return rc


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


def _fptr(a):
    return a.ctypes.data_as(_fp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


class Mesh:
    """A tet mesh: positions (V,3) float32, tets (T,4) int32, surface tris
    (S,3) uint32 and the opposite tet vertex of each (S,) -- the arrays of
    SoftObject (SoftObject.h:22-27)."""

    def __init__(self, pos, tets, tris=None, tri_opp=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        if tris is None:
            self.tris = None
            m = self._as_c()
            _check(lib.mgsb_extract_surface(C.byref(m)), "mgsb_extract_surface")
            S = m.ntris
            tris = np.ctypeslib.as_array(m.tris, shape=(max(S, 1) * 3,))[: S * 3].copy()
            tri_opp = np.ctypeslib.as_array(m.tri_opp, shape=(max(S, 1),))[:S].copy()
            m.pos = None  # numpy-owned: only the surface arrays belong to the library
            m.tets = None
            lib.mgsb_mesh_free(C.byref(m))
        self.tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        self.tri_opp = None if tri_opp is None else np.ascontiguousarray(tri_opp, dtype=np.uint32)

    @property
    def nverts(self):
        return self.pos.shape[0]

    @property
    def ntets(self):
        return self.tets.shape[0]

    @property
    def ntris(self):
        return self.tris.shape[0]

    def _as_c(self):
        m = mgsb_mesh()
        m.nverts, m.ntets = self.nverts, self.ntets
        m.pos = _fptr(self.pos)
        m.tets = _iptr(self.tets)
        if getattr(self, "tris", None) is not None:
            m.ntris = self.tris.shape[0]
            m.tris = self.tris.ctypes.data_as(C.POINTER(C.c_uint32))
        else:
            m.ntris = 0
        return m

    @staticmethod
    def _take(m):
        """Copy arrays out of a library-owned mgsb_mesh and free it."""
        try:
            V, T, S = m.nverts, m.ntets, m.ntris
            pos = np.ctypeslib.as_array(m.pos, shape=(V * 3,)).copy().reshape(V, 3)
            tets = np.ctypeslib.as_array(m.tets, shape=(T * 4,)).copy().reshape(T, 4)
            tris = np.ctypeslib.as_array(m.tris, shape=(max(S, 1) * 3,))[: S * 3].copy().reshape(S, 3)
            opp = np.ctypeslib.as_array(m.tri_opp, shape=(max(S, 1),))[:S].copy()
        finally:
            if m.pos or m.tets:
                lib.mgsb_mesh_free(C.byref(m))
        return pos, tets, tris, opp

    @classmethod
    def load_msh(cls, path):
        """SoftObject::ReadFromFile + TetFaceExtraction (SoftObject.cpp:85-251)."""
        m = mgsb_mesh()
        _check(lib.mgsb_load_msh(os.fsencode(path), C.byref(m)), f"mgsb_load_msh({path})")
        return cls(*cls._take(m))

    @classmethod
    def box(cls, nx, ny, nz, lo=(-1.0, -10.0, -1.0), hi=(1.0, 10.0, 1.0)):
        """Structured box of nx*ny*nz hexes, 6 tets each."""
        m = mgsb_mesh()
        lo_ = _f32(lo, 3)
        hi_ = _f32(hi, 3)
        _check(lib.mgsb_make_box(nx, ny, nz, _fptr(lo_), _fptr(hi_), C.byref(m)), "mgsb_make_box")
        return cls(*cls._take(m))

    @classmethod
    def blob(cls, a=6.0, b=3.0, c=4.0, p=2.5, amp=0.15, seed=0, h=0.25, coarsen=1):
        """Voxelised superquadric blob, 6 tets per voxel; coarsen > 1 gives the
        covering level of cell size coarsen * h of a nested hierarchy."""
        m = mgsb_mesh()
        _check(lib.mgsb_make_blob_level(a, b, c, p, amp, seed, h, coarsen, C.byref(m)), "mgsb_make_blob_level")
        return cls(*cls._take(m))

    def precompute(self, stiffness, density=1.0):
        """InitVolumeConstraint (PDSolver.cpp:12-132) on the host."""
        T, V = self.ntets, self.nverts
        B = np.zeros(T * 9, np.float32)
        AC = np.zeros(T * 12, np.float32)
        vol = np.zeros(T, np.float32)
        mass = np.zeros(V, np.float32)
        diag = np.zeros(V, np.float32)
        m = self._as_c()
        _check(lib.mgsb_precompute(C.byref(m), stiffness, density, _fptr(B), _fptr(AC), _fptr(vol), _fptr(mass),
                                   _fptr(diag)), "mgsb_precompute")
        return {"B": B.reshape(T, 9), "AC": AC.reshape(T, 12), "vol": vol, "mass": mass, "diag": diag}

    def tiles(self):
        """The tiled local step's gather layout (tiles of 256 consecutive
        tets): prp, tvo, tvs, tlr and the four corner positions per tet."""
        m = self._as_c()
        sz = np.zeros(2, np.int32)
        _check(lib.mgsb_mesh_tiles(C.byref(m), _iptr(sz), None, None, None, None, None), "mgsb_mesh_tiles")
        nt, npart = int(sz[0]), int(sz[1])
        prp = np.zeros(self.nverts + 1, np.int32)
        tvo = np.zeros(nt + 1, np.int32)
        tvs = np.zeros(max(npart, 1), np.int32)
        tlr = np.zeros(npart + nt, np.int32)
        tls = np.zeros(2 * self.ntets, np.uint32)
        _check(lib.mgsb_mesh_tiles(C.byref(m), _iptr(sz), _iptr(prp), _iptr(tvo), _iptr(tvs), _iptr(tlr),
                                   tls.ctypes.data_as(C.POINTER(C.c_uint32))), "mgsb_mesh_tiles")
        pos = np.stack([tls[0::2] & 0xFFFF, tls[0::2] >> 16, tls[1::2] & 0xFFFF, tls[1::2] >> 16], axis=1)
        return {"prp": prp, "tvo": tvo, "tvs": tvs[:npart], "tlr": tlr, "pos": pos.astype(np.int64)}

    def bind(self, points, mode=0):
        """Barycentric binding of points into this mesh's tets
        (PDSolver_MG.cpp:5-58 for mode 0, :133-168 for mode 1)."""
        pts = _f32(points).reshape(-1, 3)
        n = pts.shape[0]
        ids = np.zeros(n * 4, np.int32)
        w = np.zeros(n * 4, np.float32)
        m = self._as_c()
        unf = _check(lib.mgsb_bind_points(C.byref(m), _fptr(pts), n, mode, _iptr(ids), _fptr(w)), "mgsb_bind_points")
        return ids.reshape(n, 4), w.reshape(n, 4), unf


def default_params(**kw):
    """PDSolver::Init constants (PDSolver.cpp:166-176), overridden by kw."""
    p = mgsb_params()
    lib.mgsb_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity[:] = [float(x) for x in v]
        elif not hasattr(p, k):
            raise KeyError(k)
        else:
            setattr(p, k, v)
    return p


class Solver:
    """One solver handle over 1..n nested levels (levels[0] finest).

    nlevels == 1 mirrors PDSolver (PDSolver.h:70-110); nlevels == 2 with the
    two-level schedule mirrors PDSolver_MG (PDSolver_MG.h:4-25)."""

    def __init__(self, levels, params=None, **kw):
        if isinstance(levels, Mesh):
            levels = [levels]
        self.levels = list(levels)
        self.params = params if params is not None else default_params(**kw)
        arr = (mgsb_mesh * len(self.levels))(*[m._as_c() for m in self.levels])
        h = _vp()
        _check(lib.mgsb_create(arr, len(self.levels), C.byref(self.params), C.byref(h)), "mgsb_create")
        self._h = h
        self.stats = mgsb_step_stats()

    def close(self):
        if getattr(self, "_h", None):
            lib.mgsb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- configuration ----------------------------------------------------
    def set_params(self, **kw):
        for k, v in kw.items():
            if k == "gravity":
                self.params.gravity[:] = [float(x) for x in v]
            else:
                setattr(self.params, k, v)
        _check(lib.mgsb_set_params(self._h, C.byref(self.params)), "mgsb_set_params")

    def set_fixed(self, fixflag, level=0):
        """fixflag per vertex: 1 free, 0 pinned (PDSolver.cpp:134-162)."""
        f = _f32(fixflag, self.levels[level].nverts)
        _check(lib.mgsb_set_fixed(self._h, level, _fptr(f)), "mgsb_set_fixed")

    def fix_spheres(self, spheres):
        s = _f32(spheres).reshape(-1, 4)
        _check(lib.mgsb_fix_spheres(self._h, _fptr(s), s.shape[0]), "mgsb_fix_spheres")

    def fixed(self, level=0):
        """fixflag per vertex of a level (1 free, 0 pinned)."""
        out = np.zeros(self.levels[level].nverts, np.float32)
        _check(lib.mgsb_read_fixed(self._h, level, _fptr(out)), "mgsb_read_fixed")
        return out

    def fix_plane(self, axis, value, tol=1e-5):
        _check(lib.mgsb_fix_plane(self._h, axis, value, tol), "mgsb_fix_plane")

    def set_sphere(self, sid, center, radius, vel=(0.0, 0.0, 0.0), active=True):
        c = _f32(center, 3)
        v = _f32(vel, 3)
        _check(lib.mgsb_set_sphere(self._h, sid, _fptr(c), radius, _fptr(v), int(active)), "mgsb_set_sphere")

    def set_plane(self, normal, d, k, active=True):
        n = _f32(normal, 3)
        _check(lib.mgsb_set_plane(self._h, _fptr(n), d, k, int(active)), "mgsb_set_plane")

    def set_velocity(self, v):
        a = _f32(v, self.levels[0].nverts * 3)
        _check(lib.mgsb_set_velocity(self._h, _fptr(a)), "mgsb_set_velocity")

    def set_positions(self, x, level=0):
        a = _f32(x, self.levels[level].nverts * 3)
        _check(lib.mgsb_set_positions(self._h, level, _fptr(a)), "mgsb_set_positions")

    # --- stepping -----------------------------------------------------------
    def step(self, nsteps=1):
        _check(lib.mgsb_step(self._h, nsteps, C.byref(self.stats)), "mgsb_step")
        return self.stats

    def sync(self):
        _check(lib.mgsb_sync(self._h), "mgsb_sync")

    # --- readback -----------------------------------------------------------
    def positions(self, level=0):
        out = np.zeros(self.levels[level].nverts * 3, np.float32)
        _check(lib.mgsb_read_positions(self._h, level, _fptr(out)), "mgsb_read_positions")
        return out.reshape(-1, 3)

    def velocity(self, level=0):
        out = np.zeros(self.levels[level].nverts * 3, np.float32)
        _check(lib.mgsb_read_velocity(self._h, level, _fptr(out)), "mgsb_read_velocity")
        return out.reshape(-1, 3)

    def normals(self):
        out = np.zeros(self.levels[0].nverts * 3, np.float32)
        _check(lib.mgsb_read_normals(self._h, _fptr(out)), "mgsb_read_normals")
        return out.reshape(-1, 3)

    def render9(self):
        out = np.zeros(self.levels[0].nverts * 9, np.float32)
        _check(lib.mgsb_read_render9(self._h, _fptr(out)), "mgsb_read_render9")
        return out.reshape(-1, 9)

    def collision(self):
        V = self.levels[0].nverts
        flags = np.zeros(V, np.int32)
        depth = np.zeros(V, np.float32)
        _check(lib.mgsb_read_collision(self._h, _iptr(flags), _fptr(depth)), "mgsb_read_collision")
        return flags, depth

    def energy(self):
        e = mgsb_energy_t()
        _check(lib.mgsb_energy(self._h, C.byref(e)), "mgsb_energy")
        return {"Ek": e.Ek, "Ep": e.Ep, "Ec": e.Ec, "Nc": e.Nc, "dX": e.dX}

    def step_convergence(self, iters=128):
        """PDSolver::StepForConvergence: energies Ek + Ep + Ec after each of
        `iters` reference iterations (index 0: at s) and the final terms; the
        state is restored."""
        E = np.zeros(iters + 1)
        t = mgsb_energy_t()
        _check(lib.mgsb_step_convergence(self._h, iters, E.ctypes.data_as(_dp), C.byref(t)),
               "mgsb_step_convergence")
        return E, {"Ek": t.Ek, "Ep": t.Ep, "Ec": t.Ec, "Nc": t.Nc, "dX": t.dX}

    def step_error(self, iters=128):
        """One step with the reference's energy error (E - E_conv) / (E0 - E_conv)."""
        err = C.c_double()
        E = np.zeros(3)
        _check(lib.mgsb_step_error(self._h, iters, C.byref(err), E.ctypes.data_as(_dp)), "mgsb_step_error")
        return err.value, {"E0": E[0], "E_conv": E[1], "E": E[2]}

    def ep_density(self):
        """Per-vertex Ep density of the finest level (calEveryVertEp)."""
        out = np.zeros(self.levels[0].nverts, np.float32)
        _check(lib.mgsb_read_ep_density(self._h, _fptr(out)), "mgsb_read_ep_density")
        return out

    def residual(self):
        r = C.c_double()
        _check(lib.mgsb_residual(self._h, C.byref(r)), "mgsb_residual")
        return r.value

    def step_history(self):
        """Relative residuals at the last step's checks (start, judge, then
        each checked half-cycle)."""
        out = np.zeros(64, np.float64)
        n = C.c_int32()
        _check(lib.mgsb_step_history(self._h, out.ctypes.data_as(_dp), 64, C.byref(n)), "mgsb_step_history")
        return out[:min(64, n.value)]

    def run_stats(self, reset=False):
        st = mgsb_run_stats_t()
        _check(lib.mgsb_run_stats(self._h, C.byref(st), int(reset)), "mgsb_run_stats")
        return {"max_rel": st.max_rel, "total_cycles": st.total_cycles, "steps": st.steps,
                "failed_steps": st.failed_steps, "fallbacks": st.fallbacks, "total_halves": st.total_halves}

    # --- learned coarse correction -------------------------------------------
    def learn_record(self, max_steps):
        _check(lib.mgsb_learn_record(self._h, max_steps), "mgsb_learn_record")

    def learn_samples(self):
        n = C.c_int64()
        _check(lib.mgsb_learn_samples(self._h, C.byref(n)), "mgsb_learn_samples")
        return n.value

    def learn_clear(self):
        _check(lib.mgsb_learn_clear(self._h), "mgsb_learn_clear")

    def learn_train(self, epochs=5, batch=4096, lr=1e-3, seed=0):
        hist = np.zeros(epochs, np.float32)
        _check(lib.mgsb_learn_train(self._h, epochs, batch, lr, seed, _fptr(hist)), "mgsb_learn_train")
        return hist

    def learn_nparams(self):
        n = C.c_int32()
        _check(lib.mgsb_learn_nparams(self._h, C.byref(n)), "mgsb_learn_nparams")
        return n.value

    def learn_weights(self, w=None):
        n = self.learn_nparams()
        if w is None:
            out = np.zeros(n, np.float32)
            _check(lib.mgsb_learn_weights(self._h, _fptr(out), n, 0), "mgsb_learn_weights")
            return out
        a = _f32(w, n)
        _check(lib.mgsb_learn_weights(self._h, _fptr(a), n, 1), "mgsb_learn_weights")

    def learn_norm(self, norm=None):
        if norm is None:
            out = np.zeros(30, np.float32)
            _check(lib.mgsb_learn_norm(self._h, _fptr(out), 0), "mgsb_learn_norm")
            return out
        a = _f32(norm, 30)
        _check(lib.mgsb_learn_norm(self._h, _fptr(a), 1), "mgsb_learn_norm")

    def learn_last(self):
        """Coarse correction (V1, 3) the last learned step prolongated into x."""
        out = np.zeros(self.levels[1].nverts * 3, np.float32)
        _check(lib.mgsb_learn_last(self._h, _fptr(out)), "mgsb_learn_last")
        return out.reshape(-1, 3)
    def mlp_forward(self, x, path=0):
        x = _f32(x).reshape(-1, 12)
        y = np.zeros((x.shape[0], 3), np.float32)
        _check(lib.mgsb_mlp_forward(self._h, _fptr(x), x.shape[0], _fptr(y), path), "mgsb_mlp_forward")
        return y

    def mlp_grad(self, x, t):
        x = _f32(x).reshape(-1, 12)
        t = _f32(t).reshape(-1, 3)
        g = np.zeros(self.learn_nparams(), np.float32)
        loss = C.c_double()
        _check(lib.mgsb_mlp_grad(self._h, _fptr(x), _fptr(t), x.shape[0], _fptr(g), C.byref(loss)), "mgsb_mlp_grad")
        return g, loss.value

    def level_rho(self, level):
        r = C.c_double()
        _check(lib.mgsb_level_rho(self._h, level, C.byref(r)), "mgsb_level_rho")
        return r.value

    def coarse_nfree(self):
        n = C.c_int32()
        _check(lib.mgsb_coarse_nfree(self._h, C.byref(n)), "mgsb_coarse_nfree")
        return n.value

    def set_timing(self, on):
        _check(lib.mgsb_set_timing(self._h, int(on)), "mgsb_set_timing")

    def kernel_time(self, kernel=0):
        ms = C.c_double()
        n = C.c_int64()
        _check(lib.mgsb_kernel_time(self._h, kernel, C.byref(ms), C.byref(n)), "mgsb_kernel_time")
        return ms.value, n.value

    def autotune(self, trial_steps=20):
        """Launch shapes by kernel time, then sweeps over a window of steps
        (mgsb_autotune; the solver's state is restored)."""
        r = mgsb_tune_result()
        _check(lib.mgsb_autotune(self._h, trial_steps, C.byref(r)), "mgsb_autotune")
        for k in ("block_tet", "block_vert", "pre_sweeps", "post_sweeps", "vert_lanes",
                  "tile_gather"):  # keep the host copy in step
            setattr(self.params, k, getattr(r, k))
        self.params.fuse_sweeps = 2 if r.fuse_sweeps else 0
        return {k: getattr(r, k) for k, _ in r._fields_}

    def body_count(self):
        """(number of bodies, per-body residual checks on)."""
        n, c = C.c_int32(), C.c_int32()
        _check(lib.mgsb_body_count(self._h, C.byref(n), C.byref(c)), "mgsb_body_count")
        return n.value, bool(c.value)

    def set_body_stiffness(self, E):
        a = _f32(E)
        _check(lib.mgsb_set_body_stiffness(self._h, _fptr(a), a.size), "mgsb_set_body_stiffness")

    def body_residuals(self):
        n = self.body_count()[0]
        r0, r = np.zeros(n), np.zeros(n)
        _check(lib.mgsb_body_residuals(self._h, r0.ctypes.data_as(_dp), r.ctypes.data_as(_dp), n),
               "mgsb_body_residuals")
        return r0, r

    def tet_forces(self, x, if_mode, level=0):
        """Per-tet corner forces (T, 4, 3) of the local step at positions x."""
        m = self.levels[level]
        a = _f32(x, m.nverts * 3)
        out = np.zeros(m.ntets * 12, np.float32)
        _check(lib.mgsb_tet_forces(self._h, level, if_mode, _fptr(a), _fptr(out)), "mgsb_tet_forces")
        return out.reshape(-1, 4, 3)

    def time_kernel(self, reps=200, kernel=0):
        """Average microseconds of `reps` back-to-back launches of kernel 0 (the
        finest local step) or 1 (the fused fine-level sweep)."""
        us = C.c_double()
        _check(lib.mgsb_time_kernel(self._h, kernel, reps, C.byref(us)), "mgsb_time_kernel")
        return us.value


class PDSolver(Solver):
    """Single-level solver with the reference's method names
    (PDSolver.h:70-110)."""

    def __init__(self, mesh, params=None, **kw):
        kw.setdefault("schedule", SCHED_SINGLE)
        super().__init__([mesh], params, **kw)

    def SetFixedVert(self, fixflag):  # noqa: N802  (reference name)
        self.set_fixed(fixflag)

    def Step(self):  # noqa: N802
        self.step(1)


class PDSolver_MG(Solver):  # noqa: N801  (reference name)
    """Two-level solver, coarse + fine (PDSolver_MG.h:4-25); levels are
    passed fine first."""

    def __init__(self, coarse, fine, params=None, **kw):
        kw.setdefault("schedule", SCHED_TWOLEVEL)
        kw.setdefault("if_mode", IF_PLAIN)
        super().__init__([fine, coarse], params, **kw)

    def Step(self):  # noqa: N802
        self.step(1)


__all__ = ["lib", "Mesh", "Solver", "PDSolver", "PDSolver_MG", "default_params", "MgsbError", "SCHED_SINGLE",
           "SCHED_TWOLEVEL", "SCHED_VCYCLE", "SCHED_LEARNED", "IF_PLAIN", "IF_AC", "IF_STABLE", "IF_LINEAR"]
