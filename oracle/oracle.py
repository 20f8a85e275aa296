"""ctypes binding of the CPU oracle (liblbm_oracle.so) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module, and only as the checker.  The oracle restates the reference's
collide/stream/force shaders and host setup (MarcosAsh/Lattice_Fluid_Dynamics
simulation/shaders/lbm_*.comp, simulation/src/lbm.c) in their own AoS
two-buffer order; see oracle_dyn.c / oracle_geom.c for line citations.

How each oracle function is pinned (tests/test_oracle.py):
  feq, weights           reference KATs (simulation/test/test_lbm.c:191-227)
  forward/inverse MRT     reference KATs (test_lbm.c:300-366) + M M^T = diag(norms)
  collide (BGK/MRT/reg)   exact conservation of rho, rho*u; equilibrium fixed point;
                          Taylor-Green shear-wave decay rate nu = (tau-1/2)/3
  Smagorinsky             closed-form tau_eff for a single-mode stress
  stream (periodic)       pure permutation (bitwise shift)
  bounce-back/Bouzidi     Poiseuille channel profile; q=1/2 reduces to halfway BB
  Zou-He                  imposed inlet velocity / outlet density after the BC
  force                   brute-force momentum balance on tiny grids; symmetry
  geometry                closed-form AABB cells/area; analytic sphere q; mesh of
                          a box equals the AABB voxelisation
  samplers                exact on linear fields; nearest index
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LBM_ORACLE_LIBRARY names another build of the oracle (the sanitizer build)
_LIB_PATH = os.environ.get("LBM_ORACLE_LIBRARY") or os.path.join(_HERE, "liblbm_oracle.so")
_lib: Optional[ctypes.CDLL] = None

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)


def build() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for pfx, real, rp in (("or32_", ctypes.c_float, c_float_p), ("or64_", ctypes.c_double, c_double_p)):
            getattr(L, pfx + "create").restype = ctypes.c_void_p
            getattr(L, pfx + "create").argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float]
            getattr(L, pfx + "free").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "set_params").argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                        ctypes.c_int, ctypes.c_double, ctypes.c_int]
            for name in ("f", "f_new", "vel"):
                getattr(L, pfx + name).restype = rp
                getattr(L, pfx + name).argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "solid").restype = c_int_p
            getattr(L, pfx + "solid").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "q").restype = c_float_p
            getattr(L, pfx + "q").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "tau").restype = ctypes.c_double
            getattr(L, pfx + "tau").argtypes = [ctypes.c_void_p]
            getattr(L, pfx + "initialize_flow").restype = ctypes.c_int
            getattr(L, pfx + "initialize_flow").argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 3
            getattr(L, pfx + "step").argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 3
            getattr(L, pfx + "run").argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_float] * 3
            getattr(L, pfx + "force").argtypes = [ctypes.c_void_p, c_double_p, c_ll_p]
            getattr(L, pfx + "feq").restype = real
            getattr(L, pfx + "feq").argtypes = [ctypes.c_int, real, real, real, real]
            getattr(L, pfx + "forward_mrt").argtypes = [rp, rp]
            getattr(L, pfx + "inverse_mrt").argtypes = [rp, rp]
            getattr(L, pfx + "meq").argtypes = [real, real, real, real, rp]
            getattr(L, pfx + "collide_cell").argtypes = [rp, rp, rp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                          real, ctypes.c_int, ctypes.c_int, ctypes.c_int, real,
                                                          real, real, real]
        L.or_feq_f32.restype = ctypes.c_float
        L.or_feq_f32.argtypes = [ctypes.c_int] + [ctypes.c_float] * 4
        L.or_tau_from_viscosity.restype = ctypes.c_float
        L.or_tau_from_viscosity.argtypes = [ctypes.c_float]
        geo = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, c_float_p]
        L.or_set_solid_aabb.argtypes = geo + [ctypes.c_float] * 6
        L.or_set_solid_sphere.argtypes = geo + [ctypes.c_float] * 4
        L.or_add_ground_plane.argtypes = geo + [ctypes.c_float]
        L.or_set_solid_mesh.argtypes = geo + [c_float_p, ctypes.c_int] + [ctypes.c_float] * 6
        L.or_projected_area.restype = ctypes.c_float
        L.or_projected_area.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_int]
        L.or_sample_trilinear.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_float] * 3 + [c_float_p]
        L.or_sample_nearest.argtypes = L.or_sample_trilinear.argtypes
        L.or_particles_seed.argtypes = [c_float_p, ctypes.c_int, ctypes.c_uint32]
        L.or_trace_streamlines.argtypes = [c_float_p] + [ctypes.c_int] * 5 + [c_float_p]
        L.or_particles_advect.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_float_p,
                                          ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                          ctypes.c_uint32]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


D3Q19_E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                    [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], dtype=np.int64)
D3Q19_OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15], dtype=np.int64)
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


class OracleGrid:
    """The reference LBMGrid restated on the CPU (AoS f / f_new, vec4 vel)."""

    def __init__(self, nx: int, ny: int, nz: int, viscosity: float, precision: int = 32):
        assert precision in (32, 64)
        self.L = lib()
        self.pfx = "or32_" if precision == 32 else "or64_"
        self.dtype = np.float32 if precision == 32 else np.float64
        self.real = ctypes.c_float if precision == 32 else ctypes.c_double
        self.nx, self.ny, self.nz = nx, ny, nz
        self.ncell = nx * ny * nz
        self.h = self._fn("create")(nx, ny, nz, viscosity)
        if not self.h:
            raise MemoryError("oracle grid allocation failed")
        self.tau = float(self._fn("tau")(self.h))
        self.useRegularized = 0
        self.useMRT = 0
        self.useSmagorinsky = 0
        self.smagorinskyCs = 0.1
        self.periodicYZ = 0

    def _fn(self, name):
        return getattr(self.L, self.pfx + name)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self._fn("free")(self.h)
                self.h = None
        except Exception:
            pass

    def _sync_params(self):
        self._fn("set_params")(self.h, float(self.tau), int(self.useRegularized), int(self.useMRT),
                               int(self.useSmagorinsky), float(self.smagorinskyCs), int(self.periodicYZ))

    # views onto oracle memory
    def _view(self, name, n, dtype):
        p = self._fn(name)(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).view(dtype)

    @property
    def f(self):
        return self._view("f", 19 * self.ncell, self.dtype).reshape(self.ncell, 19)

    @property
    def f_new(self):
        return self._view("f_new", 19 * self.ncell, self.dtype).reshape(self.ncell, 19)

    @property
    def vel(self):
        return self._view("vel", 4 * self.ncell, self.dtype).reshape(self.ncell, 4)

    @property
    def solid(self):
        return self._view("solid", self.ncell, np.int32)

    @property
    def q(self):
        return self._view("q", 19 * self.ncell, np.float32).reshape(self.ncell, 19)

    def set_solid_q(self, solid: np.ndarray, q: np.ndarray):
        self.solid[:] = solid.reshape(-1)
        self.q[:] = q.reshape(self.ncell, 19)

    def set_solid_aabb(self, *b):
        self.L.or_set_solid_aabb(self.nx, self.ny, self.nz, self._fn("solid")(self.h), self._fn("q")(self.h), *b)

    def set_solid_sphere(self, cx, cy, cz, r):
        self.L.or_set_solid_sphere(self.nx, self.ny, self.nz, self._fn("solid")(self.h), self._fn("q")(self.h),
                                   cx, cy, cz, r)

    def set_solid_mesh(self, tris: np.ndarray, bbox):
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
        self.L.or_set_solid_mesh(self.nx, self.ny, self.nz, self._fn("solid")(self.h), self._fn("q")(self.h),
                                 _ptr(t, ctypes.c_float), len(t) // 12, *bbox)

    def add_ground_plane(self, z):
        self.L.or_add_ground_plane(self.nx, self.ny, self.nz, self._fn("solid")(self.h), self._fn("q")(self.h), z)

    def projected_area(self, axis=0):
        return float(self.L.or_projected_area(self.nx, self.ny, self.nz, self._fn("solid")(self.h), axis))

    def initialize_flow(self, ux, uy, uz):
        return int(self._fn("initialize_flow")(self.h, ux, uy, uz))

    def step(self, ux, uy=0.0, uz=0.0, n: int = 1):
        self._sync_params()
        self._fn("run")(self.h, int(n), ux, uy, uz)

    def force(self):
        """Returns (exact[6] as FP64 sums of per-cell values, fixed[7] reference int accumulators)."""
        ex = np.zeros(6, np.float64)
        fx = np.zeros(7, np.int64)
        self._fn("force")(self.h, _ptr(ex, ctypes.c_double), _ptr(fx, ctypes.c_longlong))
        return ex, fx

    def post_stream_soa(self):
        """The reference's post-stream f as [19, cell]."""
        return np.ascontiguousarray(self.f.T)


def feq(i, rho, ux, uy, uz, precision=32):
    L = lib()
    return float((L.or32_feq if precision == 32 else L.or64_feq)(i, rho, ux, uy, uz))


def feq_f32_host(i, rho, ux, uy, uz):
    return float(lib().or_feq_f32(i, rho, ux, uy, uz))


def forward_mrt(fi, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    ct = ctypes.c_float if precision == 32 else ctypes.c_double
    fi = np.ascontiguousarray(fi, dtype=dt)
    m = np.zeros(19, dt)
    (lib().or32_forward_mrt if precision == 32 else lib().or64_forward_mrt)(_ptr(fi, ct), _ptr(m, ct))
    return m


def inverse_mrt(m, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    ct = ctypes.c_float if precision == 32 else ctypes.c_double
    m = np.ascontiguousarray(m, dtype=dt)
    fi = np.zeros(19, dt)
    (lib().or32_inverse_mrt if precision == 32 else lib().or64_inverse_mrt)(_ptr(m, ct), _ptr(fi, ct))
    return fi


def meq(rho, ux, uy, uz, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    ct = ctypes.c_float if precision == 32 else ctypes.c_double
    out = np.zeros(19, dt)
    (lib().or32_meq if precision == 32 else lib().or64_meq)(rho, ux, uy, uz, _ptr(out, ct))
    return out


def collide_cell(fin, solid=0, x=5, nx=10, tau=0.6, reg=0, mrt=0, smag=0, cs=0.1, inlet=(0.05, 0.0, 0.0),
                 precision=32):
    """One cell of the collide pass; returns (f after Zou-He rewrite, f_new, vel)."""
    dt = np.float32 if precision == 32 else np.float64
    ct = ctypes.c_float if precision == 32 else ctypes.c_double
    fi = np.array(fin, dtype=dt)
    fo = np.zeros(19, dt)
    v = np.zeros(4, dt)
    fn = lib().or32_collide_cell if precision == 32 else lib().or64_collide_cell
    fn(_ptr(fi, ct), _ptr(fo, ct), _ptr(v, ct), solid, x, nx, tau, reg, mrt, smag, cs, *inlet)
    return fi, fo, v


def sample_trilinear(vel: np.ndarray, nx, ny, nz, p):
    v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
    out = np.zeros(3, np.float32)
    lib().or_sample_trilinear(_ptr(v, ctypes.c_float), nx, ny, nz, float(p[0]), float(p[1]), float(p[2]),
                              _ptr(out, ctypes.c_float))
    return out


def sample_nearest(vel: np.ndarray, nx, ny, nz, p):
    v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
    out = np.zeros(3, np.float32)
    lib().or_sample_nearest(_ptr(v, ctypes.c_float), nx, ny, nz, float(p[0]), float(p[1]), float(p[2]),
                            _ptr(out, ctypes.c_float))
    return out


def trace_streamlines(vel: np.ndarray, nx, ny, nz, num_seeds: int, trace_len: int) -> np.ndarray:
    """RK4 streamlines of simulation/shaders/streamline_trace.comp through the
    (ux, uy, uz, rho) field: (num_seeds, trace_len, 4) vertices (x, y, z, speed)."""
    v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
    assert v.size == 4 * nx * ny * nz and min(nx, ny, nz) >= 2
    out = np.zeros((num_seeds, trace_len, 4), np.float32)
    lib().or_trace_streamlines(_ptr(v, ctypes.c_float), nx, ny, nz, num_seeds, trace_len, _ptr(out, ctypes.c_float))
    return out


def particles_seed(n: int, seed: int) -> np.ndarray:
    p = np.zeros((n, 8), np.float32)
    lib().or_particles_seed(_ptr(p, ctypes.c_float), n, seed)
    return p


class OrParticleFlow(ctypes.Structure):
    """oracle_particles.c OrParticleFlow (host pointers)."""
    _fields_ = [("wind", ctypes.c_float * 3), ("carMin", ctypes.c_float * 3), ("carMax", ctypes.c_float * 3),
                ("carCenter", ctypes.c_float * 3), ("collisionMode", ctypes.c_int), ("useLBM", ctypes.c_int),
                ("vizMode", ctypes.c_int), ("triangles", ctypes.c_void_p), ("numTriangles", ctypes.c_int),
                ("cellStart", ctypes.c_void_p), ("cellCount", ctypes.c_void_p), ("triIndices", ctypes.c_void_p),
                ("numIndices", ctypes.c_int), ("gridMin", ctypes.c_float * 3), ("gridCellSize", ctypes.c_float * 3),
                ("gridRes", ctypes.c_int)]


def particles_live(vel: np.ndarray, solid: np.ndarray, nx, ny, nz, parts: np.ndarray, dt: float, wind, car_min,
                   car_max, car_center, collision_mode=0, use_lbm=1, viz_mode=0, triangles=None, grid=None,
                   grid_res=8, seed=0, step=0) -> np.ndarray:
    """One live update (simulation/shaders/particle.comp:218-407) of (n, 8) particles."""
    v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
    s = np.ascontiguousarray(solid, dtype=np.int32).reshape(-1)
    p = np.ascontiguousarray(parts, dtype=np.float32).copy()
    f = OrParticleFlow()
    f.wind[:], f.carMin[:], f.carMax[:], f.carCenter[:] = wind, car_min, car_max, car_center
    f.collisionMode, f.useLBM, f.vizMode = collision_mode, use_lbm, viz_mode
    keep = []
    if triangles is not None:
        t = np.ascontiguousarray(triangles, np.float32)
        keep.append(t)
        f.triangles, f.numTriangles = t.ctypes.data, t.shape[0]
    if grid is not None:
        arrs = [np.ascontiguousarray(grid[k], np.int32) for k in ("cellStart", "cellCount", "triIndices")]
        keep += arrs
        f.cellStart, f.cellCount, f.triIndices = [a.ctypes.data for a in arrs]
        f.numIndices = arrs[2].size
        f.gridMin[:], f.gridCellSize[:] = [float(x) for x in grid["origin"]], [float(x) for x in grid["cell"]]
        f.gridRes = grid_res
    lib().or_particles_live(_ptr(v, ctypes.c_float), _ptr(s, ctypes.c_int), nx, ny, nz, _ptr(p, ctypes.c_float),
                            p.shape[0], ctypes.c_float(dt), ctypes.byref(f), ctypes.c_uint32(seed),
                            ctypes.c_uint32(step))
    return p


def point_triangle(p, v0, v1, v2):
    """(distance, closest point, unit normal) of oracle_particles.c."""
    a = [np.ascontiguousarray(x, np.float32) for x in (p, v0, v1, v2)]
    c = np.zeros(3, np.float32)
    n = np.zeros(3, np.float32)
    lib().or_point_triangle.restype = ctypes.c_float
    d = lib().or_point_triangle(*[_ptr(x, ctypes.c_float) for x in a], _ptr(c, ctypes.c_float),
                                _ptr(n, ctypes.c_float))
    return float(d), c, n


def particles_advect(vel: np.ndarray, nx, ny, nz, parts: np.ndarray, dt: float, vel_scale: float, seed: int,
                     step: int) -> np.ndarray:
    v = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
    p = np.ascontiguousarray(parts, dtype=np.float32).copy()
    lib().or_particles_advect(_ptr(v, ctypes.c_float), nx, ny, nz, _ptr(p, ctypes.c_float), p.shape[0], dt,
                              vel_scale, seed, step)
    return p
