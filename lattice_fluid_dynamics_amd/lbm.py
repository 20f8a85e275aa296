"""Python binding of the C ABI in include/lbm.h (liblbm_amd.so).

The functions keep the names and argument order of the reference solver API
(MarcosAsh/Lattice_Fluid_Dynamics simulation/lib/lbm.h:58-129) so host code
and tests read like the reference's own: ``g = LBM_Create(128, 64, 64, nu)``,
``g.tau = 0.6``, ``LBM_Step(g, 0.05, 0, 0)``, ``LBM_ComputeDragForce(g)``.
Out-pointer arguments of the C API become return values.

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing, or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LBM_AMD_LIBRARY names another build of the same library (the sanitizer
# build of csrc/sanitize.mk)
LIB_PATH = os.environ.get("LBM_AMD_LIBRARY") or os.path.join(_HERE, "liblbm_amd.so")

LBM_STORAGE_FP32 = 0
LBM_STORAGE_FP16 = 1
LBM_STORAGE_BF16 = 2
LBM_STORAGE_FP64 = 3

c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)


class LBMGridStruct(ctypes.Structure):
    """Field-for-field mirror of LBMGrid in include/lbm.h."""

    _fields_ = [
        ("sizeX", c_int), ("sizeY", c_int), ("sizeZ", c_int), ("totalCells", c_int),
        ("tau", c_float),
        ("velocityBuffer", c_void_p), ("solidBuffer", c_void_p),
        ("useRegularized", c_int), ("useMRT", c_int), ("useSmagorinsky", c_int),
        ("smagorinskyCs", c_float), ("periodicYZ", c_int),
        ("numChunks", c_int), ("slabZ", c_int), ("haloZ", c_int),
        ("forceFixedPoint", c_int), ("populationStorage", c_int),
        ("stepCount", ctypes.c_longlong), ("impl", c_void_p),
    ]


GridP = ctypes.POINTER(LBMGridStruct)

# Every symbol include/lbm.h declares, with its ctypes signature.
SIGNATURES = {
    "LBM_Create": (GridP, [c_int, c_int, c_int, c_float]),
    "LBM_Free": (None, [GridP]),
    "LBM_SetSolidAABB": (None, [GridP] + [c_float] * 6),
    "LBM_InitializeFlow": (c_int, [GridP, c_float, c_float, c_float]),
    "LBM_InitializeFlowFromField": (c_int, [GridP, c_float_p, c_int, c_int, c_int]),
    "LBM_InitializeFlowFromFieldNoAlloc": (c_int, [GridP, c_float_p, c_int, c_int, c_int]),
    "LBM_Step": (None, [GridP, c_float, c_float, c_float]),
    "LBM_GetVelocityBuffer": (c_void_p, [GridP]),
    "LBM_ComputeDragForce": (None, [GridP, c_float_p, c_float_p, c_float_p]),
    "LBM_ComputeDragForceDecomposed": (None, [GridP] + [c_float_p] * 6),
    "LBM_ComputeDragCoefficient": (c_float, [GridP, c_float, c_float]),
    "LBM_ComputeLiftCoefficient": (c_float, [GridP, c_float, c_float]),
    "LBM_SetSolidSphere": (None, [GridP, c_float, c_float, c_float, c_float]),
    "LBM_AddGroundPlane": (None, [GridP, c_float]),
    "LBM_ComputeProjectedArea": (c_float, [GridP, c_int]),
    "LBM_SetSolidMesh": (None, [GridP, c_float_p, c_int] + [c_float] * 6),
    "LBM_StepN": (None, [GridP, c_int, c_float, c_float, c_float]),
    "LBM_CopyVelocityToHost": (c_int, [GridP, c_float_p]),
    "LBM_CopySolidToHost": (c_int, [GridP, c_int_p]),
    "LBM_CopyQToHost": (c_int, [GridP, c_float_p]),
    "LBM_CopyPopulationsToHost": (c_int, [GridP, c_float_p, c_float_p]),
    "LBM_CopyPopulationsToHostF64": (c_int, [GridP, c_double_p, c_double_p]),
    "LBM_ComputeTotalMass": (c_double, [GridP]),
    "LBM_ComputeDragForceF64": (None, [GridP, c_double_p]),
    "LBM_SeedParticles": (None, [GridP, c_void_p, c_int, ctypes.c_uint]),
    "LBM_AdvectParticles": (None, [GridP, c_void_p, c_int, c_float, c_float, ctypes.c_uint, ctypes.c_uint]),
    "LBM_DeviceAlloc": (c_void_p, [ctypes.c_ulonglong]),
    "LBM_DeviceFree": (None, [c_void_p]),
    "LBM_CopyToDevice": (c_int, [c_void_p, c_void_p, ctypes.c_ulonglong]),
    "LBM_CopyToHost": (c_int, [c_void_p, c_void_p, ctypes.c_ulonglong]),
    "LBM_CopyDeviceToDevice": (c_int, [c_void_p, c_void_p, ctypes.c_ulonglong]),
    "LBM_Synchronize": (None, [GridP]),
    "LBM_ProfileSteps": (c_int, [GridP, c_int, c_float, c_float, c_float, c_float_p, c_float_p, c_float_p]),
    "LBM_StepNTimed": (c_int, [GridP, c_int, c_float, c_float, c_float]),
    "LBM_GetStateSize": (ctypes.c_longlong, [GridP]),
    "LBM_SaveState": (c_int, [GridP, c_void_p, ctypes.c_longlong]),
    "LBM_LoadState": (c_int, [GridP, c_void_p, ctypes.c_longlong]),
    "LBM_CollectStepTimings": (c_int, [GridP, c_float_p, c_float_p, ctypes.POINTER(c_int)]),
    "LBM_GetBoundaryCellCount": (c_int, [GridP]),
    "LBM_AdvectParticlesLive": (None, [GridP, c_void_p, c_int, c_float, c_void_p, ctypes.c_uint, ctypes.c_uint]),
    "LBM_TraceStreamlines": (c_int, [GridP, c_void_p, c_int, c_int]),
    "LBM_GetGuardResetCount": (ctypes.c_longlong, [GridP]),
    "LBM_GetLastError": (ctypes.c_char_p, []),
    "LBM_CdMonitorCreate": (c_void_p, [GridP, c_float, c_float, c_int]),
    "LBM_CdMonitorFree": (None, [c_void_p]),
    "LBM_CdMonitorSample": (c_int, [c_void_p, GridP, c_float_p]),
    "LBM_CdMonitorSetBody": (None, [c_void_p, GridP] + [c_float] * 7),
    "LBM_EffectiveReynolds": (c_float, [GridP, c_float, c_float, c_float, c_int]),
    "LBM_CdMonitorConverged": (c_int, [c_void_p, c_float_p, c_float_p]),
    "LBM_CdMonitorStrouhal": (c_float, [c_void_p, c_int, c_float]),
    "LBM_ComputeStrouhal": (c_float, [c_float_p, c_int, c_int, c_float, c_float]),
    "LBM_WriteVTI": (c_int, [GridP, ctypes.c_char_p, c_int]),
    "LBM_Autotune": (c_int, [GridP, c_int, c_float, c_float, c_float, c_float_p]),
    "LBM_SetBlockSize": (c_int, [GridP, c_int]),
    "LBM_SetStreamingAccess": (c_int, [GridP, c_int]),
    "LBM_SetBlockOrder": (c_int, [GridP, c_int]),
    "LBM_GetKernelConfig": (c_int, [GridP, c_int_p, c_int_p, c_int_p]),
    "LBM_MeasureCopyBandwidth": (c_float, [ctypes.c_ulonglong, c_int]),
    "LBM_BatchCreate": (c_void_p, [c_int, c_int, c_int, c_int, c_float, c_int]),
    "LBM_BatchFree": (None, [c_void_p]),
    "LBM_BatchAddTunnel": (c_int, [c_void_p, GridP, c_float, c_float, c_float, c_float]),
    "LBM_BatchInitialize": (c_int, [c_void_p]),
    "LBM_BatchStep": (None, [c_void_p, c_int]),
    "LBM_BatchProfile": (c_int, [c_void_p, c_int, c_int, c_float_p]),
    "LBM_BatchAutotune": (c_int, [c_void_p, c_int, c_float_p]),
    "LBM_BatchSetBlockSize": (c_int, [c_void_p, c_int]),
    "LBM_BatchSetStreamingAccess": (c_int, [c_void_p, c_int]),
    "LBM_BatchSetBlockOrder": (c_int, [c_void_p, c_int]),
    "LBM_BatchGetKernelConfig": (c_int, [c_void_p, c_int_p, c_int_p, c_int_p]),
    "LBM_BatchTunnelCount": (c_int, [c_void_p]),
    "LBM_BatchGuardResetCount": (ctypes.c_longlong, [c_void_p]),
    "LBM_BatchTotalCells": (ctypes.c_longlong, [c_void_p]),
    "LBM_BatchComputeForcesF64": (c_int, [c_void_p, c_double_p, ctypes.POINTER(ctypes.c_longlong)]),
    "LBM_BatchComputeCoefficients": (c_int, [c_void_p, c_float_p]),
    "LBM_BatchCopyPopulationsToHostF64": (c_int, [c_void_p, c_int, c_double_p, c_double_p]),
    "LBM_BatchCopyVelocityToHost": (c_int, [c_void_p, c_int, c_float_p]),
    "LBM_BatchInitializeTunnelFromField": (c_int, [c_void_p, c_int, c_float_p, c_int, c_int, c_int]),
    "LBM_BatchCopyCoarseVelocityToHost": (c_int, [c_void_p, c_int, c_int, c_float_p]),
    "LBM_CopyCoarseVelocityToHost": (c_int, [GridP, c_int, c_float_p]),
}

_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads liblbm_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `make -C lattice_fluid_dynamics_amd/csrc`")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return load_library().LBM_GetLastError().decode()


def _fp(a: np.ndarray):
    return a.ctypes.data_as(c_float_p)


class Grid:
    """Owning handle of an LBMGrid*; struct fields read and write through."""

    def __init__(self, ptr):
        object.__setattr__(self, "_p", ptr)

    @property
    def ptr(self):
        return self._p

    def __getattr__(self, name):
        return getattr(self._p.contents, name)

    def __setattr__(self, name, value):
        if name in dict(LBMGridStruct._fields_):
            setattr(self._p.contents, name, value)
        else:
            object.__setattr__(self, name, value)

    def free(self):
        if self._p:
            load_library().LBM_Free(self._p)
            object.__setattr__(self, "_p", None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def shape(self) -> Tuple[int, int, int]:
        return (self.sizeX, self.sizeY, self.sizeZ)


# ---- the reference API --------------------------------------------------------

def LBM_Create(sizeX: int, sizeY: int, sizeZ: int, viscosity: float) -> Grid:
    p = load_library().LBM_Create(sizeX, sizeY, sizeZ, viscosity)
    if not p:
        raise RuntimeError(f"LBM_Create failed: {last_error()}")
    return Grid(p)


def LBM_Free(g: Grid) -> None:
    g.free()


def LBM_SetSolidAABB(g: Grid, minX, minY, minZ, maxX, maxY, maxZ) -> None:
    load_library().LBM_SetSolidAABB(g.ptr, minX, minY, minZ, maxX, maxY, maxZ)


def LBM_SetSolidSphere(g: Grid, cx, cy, cz, radius) -> None:
    load_library().LBM_SetSolidSphere(g.ptr, cx, cy, cz, radius)


def LBM_SetSolidMesh(g: Grid, triangles: np.ndarray, bbox) -> None:
    t = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1)
    load_library().LBM_SetSolidMesh(g.ptr, _fp(t), len(t) // 12, *[float(v) for v in bbox])


def LBM_AddGroundPlane(g: Grid, worldZ: float) -> None:
    load_library().LBM_AddGroundPlane(g.ptr, worldZ)


def LBM_ComputeProjectedArea(g: Grid, axis: int) -> float:
    return float(load_library().LBM_ComputeProjectedArea(g.ptr, axis))


def LBM_InitializeFlow(g: Grid, ux: float, uy: float, uz: float) -> int:
    return int(load_library().LBM_InitializeFlow(g.ptr, ux, uy, uz))


def LBM_InitializeFlowFromField(g: Grid, field: np.ndarray) -> int:
    """Equilibrium initial state from a coarse (cnz, cny, cnx, 4) field of
    (ux, uy, uz, rho), trilinearly upsampled on the device."""
    f = np.ascontiguousarray(field, dtype=np.float32)
    assert f.ndim == 4 and f.shape[3] == 4
    cnz, cny, cnx = f.shape[:3]
    return load_library().LBM_InitializeFlowFromField(g.ptr, _fp(f), cnx, cny, cnz)


def LBM_InitializeFlowFromFieldNoAlloc(g: Grid, field: np.ndarray) -> int:
    """As LBM_InitializeFlowFromField, without device allocation (the field
    is staged in a static 4 MiB per-GPU array; larger fields are refused)."""
    f = np.ascontiguousarray(field, dtype=np.float32)
    assert f.ndim == 4 and f.shape[3] == 4
    cnz, cny, cnx = f.shape[:3]
    return load_library().LBM_InitializeFlowFromFieldNoAlloc(g.ptr, _fp(f), cnx, cny, cnz)


def LBM_Step(g: Grid, ux: float, uy: float, uz: float) -> None:
    load_library().LBM_Step(g.ptr, ux, uy, uz)


def LBM_StepN(g: Grid, n: int, ux: float, uy: float, uz: float) -> None:
    load_library().LBM_StepN(g.ptr, n, ux, uy, uz)


def LBM_GetVelocityBuffer(g: Grid) -> int:
    return int(load_library().LBM_GetVelocityBuffer(g.ptr) or 0)


def LBM_ComputeDragForce(g: Grid):
    f = [c_float(), c_float(), c_float()]
    load_library().LBM_ComputeDragForce(g.ptr, *[ctypes.byref(v) for v in f])
    return tuple(v.value for v in f)


def LBM_ComputeDragForceDecomposed(g: Grid):
    f = [c_float() for _ in range(6)]
    load_library().LBM_ComputeDragForceDecomposed(g.ptr, *[ctypes.byref(v) for v in f])
    return tuple(v.value for v in f)


def LBM_ComputeDragCoefficient(g: Grid, U: float, A: float) -> float:
    return float(load_library().LBM_ComputeDragCoefficient(g.ptr, U, A))


def LBM_ComputeLiftCoefficient(g: Grid, U: float, A: float) -> float:
    return float(load_library().LBM_ComputeLiftCoefficient(g.ptr, U, A))


# ---- additions ---------------------------------------------------------------

def LBM_CopyVelocityToHost(g: Grid) -> np.ndarray:
    out = np.empty((g.totalCells, 4), np.float32)
    if not load_library().LBM_CopyVelocityToHost(g.ptr, _fp(out)):
        raise RuntimeError(last_error())
    return out


def LBM_CopyCoarseVelocityToHost(g: Grid, factor: int) -> np.ndarray:
    """Block means of (ux, uy, uz, rho) over factor^3 cells, shape (nz/f, ny/f, nx/f, 4)."""
    nx, ny, nz = g.shape
    out = np.empty((nz // factor, ny // factor, nx // factor, 4), np.float32)
    if not load_library().LBM_CopyCoarseVelocityToHost(g.ptr, factor, _fp(out)):
        raise RuntimeError(last_error())
    return out


def LBM_CopySolidToHost(g: Grid) -> np.ndarray:
    out = np.empty(g.totalCells, np.int32)
    if not load_library().LBM_CopySolidToHost(g.ptr, out.ctypes.data_as(c_int_p)):
        raise RuntimeError(last_error())
    return out


def LBM_CopyQToHost(g: Grid) -> np.ndarray:
    out = np.empty((g.totalCells, 19), np.float32)
    if not load_library().LBM_CopyQToHost(g.ptr, _fp(out)):
        raise RuntimeError(last_error())
    return out


def LBM_CopyPopulationsToHost(g: Grid, f64: bool = False):
    """(f post-stream, f_new post-collision) in the reference layout [cell, 19]."""
    dt = np.float64 if f64 else np.float32
    f = np.empty((g.totalCells, 19), dt)
    fn = np.empty((g.totalCells, 19), dt)
    L = load_library()
    if f64:
        ok = L.LBM_CopyPopulationsToHostF64(g.ptr, f.ctypes.data_as(c_double_p), fn.ctypes.data_as(c_double_p))
    else:
        ok = L.LBM_CopyPopulationsToHost(g.ptr, _fp(f), _fp(fn))
    if not ok:
        raise RuntimeError(last_error())
    return f, fn


def LBM_ComputeTotalMass(g: Grid) -> float:
    return float(load_library().LBM_ComputeTotalMass(g.ptr))


def LBM_ComputeDragForceF64(g: Grid) -> np.ndarray:
    out = np.zeros(6, np.float64)
    load_library().LBM_ComputeDragForceF64(g.ptr, out.ctypes.data_as(c_double_p))
    return out


def LBM_Synchronize(g: Grid) -> None:
    load_library().LBM_Synchronize(g.ptr)


def LBM_ProfileSteps(g: Grid, n: int, ux: float, uy: float, uz: float, per_launch: bool = True):
    tot, mn, bd = c_float(), c_float(), c_float()
    L = load_library()
    ok = L.LBM_ProfileSteps(g.ptr, n, ux, uy, uz, ctypes.byref(tot),
                            ctypes.byref(mn) if per_launch else None, ctypes.byref(bd) if per_launch else None)
    if not ok:
        raise RuntimeError(last_error())
    return tot.value, mn.value, bd.value


def LBM_SaveState(g: Grid) -> np.ndarray:
    """The grid's checkpoint (header, stored populations, velocity) as bytes."""
    L = load_library()
    n = int(L.LBM_GetStateSize(g.ptr))
    if n <= 0:
        raise RuntimeError(last_error())
    buf = np.empty(n, np.uint8)
    if not L.LBM_SaveState(g.ptr, buf.ctypes.data, n):
        raise RuntimeError(last_error())
    return buf


def LBM_LoadState(g: Grid, buf: np.ndarray) -> int:
    b = np.ascontiguousarray(buf, np.uint8)
    return int(load_library().LBM_LoadState(g.ptr, b.ctypes.data, b.size))


def LBM_StepNTimed(g: Grid, n: int, ux: float, uy: float, uz: float) -> None:
    if not load_library().LBM_StepNTimed(g.ptr, n, ux, uy, uz):
        raise RuntimeError(last_error())


def LBM_CollectStepTimings(g: Grid):
    """(main_ms, boundary_ms, launches) over the steps timed since the last call."""
    mn, bd, n = c_float(), c_float(), c_int()
    if not load_library().LBM_CollectStepTimings(g.ptr, ctypes.byref(mn), ctypes.byref(bd), ctypes.byref(n)):
        raise RuntimeError(last_error())
    return mn.value, bd.value, n.value


def LBM_GetBoundaryCellCount(g: Grid) -> int:
    return int(load_library().LBM_GetBoundaryCellCount(g.ptr))


def LBM_GetGuardResetCount(g: Grid) -> int:
    n = int(load_library().LBM_GetGuardResetCount(g.ptr))
    if n < 0:
        raise RuntimeError("LBM_GetGuardResetCount failed: " + last_error())
    return n


class LBMParticleFlow(ctypes.Structure):
    """include/lbm.h LBMParticleFlow: the uniforms of the live particle update."""
    _fields_ = ([(n, c_float) for n in ("windX", "windY", "windZ", "carMinX", "carMinY", "carMinZ", "carMaxX",
                                        "carMaxY", "carMaxZ", "carCenterX", "carCenterY", "carCenterZ")]
                + [("collisionMode", c_int), ("useLBM", c_int), ("vizMode", c_int),
                   ("triangles", c_void_p), ("numTriangles", c_int), ("cellStart", c_void_p),
                   ("cellCount", c_void_p), ("triIndices", c_void_p), ("numIndices", c_int)]
                + [(n, c_float) for n in ("gridMinX", "gridMinY", "gridMinZ", "gridCellSizeX", "gridCellSizeY",
                                          "gridCellSizeZ")]
                + [("gridRes", c_int)])


def particle_flow(wind, car_min, car_max, car_center, collision_mode=0, use_lbm=1, viz_mode=0, triangles=None,
                  grid=None, grid_res=8):
    """An LBMParticleFlow; for collision mode 2, `triangles` and the arrays of
    `grid` (model.collision_grid) are device buffers: (ptr, count) pairs as
    DeviceBuffer gives them, and grid["origin"], grid["cell"] the geometry."""
    f = LBMParticleFlow()
    f.windX, f.windY, f.windZ = wind
    f.carMinX, f.carMinY, f.carMinZ = car_min
    f.carMaxX, f.carMaxY, f.carMaxZ = car_max
    f.carCenterX, f.carCenterY, f.carCenterZ = car_center
    f.collisionMode, f.useLBM, f.vizMode = collision_mode, use_lbm, viz_mode
    if triangles is not None:
        f.triangles, f.numTriangles = triangles
    if grid is not None:
        f.cellStart = grid["cellStart_dev"]
        f.cellCount = grid["cellCount_dev"]
        f.triIndices, f.numIndices = grid["triIndices_dev"]
        f.gridMinX, f.gridMinY, f.gridMinZ = [float(v) for v in grid["origin"]]
        f.gridCellSizeX, f.gridCellSizeY, f.gridCellSizeZ = [float(v) for v in grid["cell"]]
        f.gridRes = grid_res
    return f


class DeviceBuffer:
    """Device copy of a host array (LBM_DeviceAlloc / LBM_CopyToDevice)."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.n = a.shape[0] if a.ndim else 1
        self.ptr = load_library().LBM_DeviceAlloc(max(a.nbytes, 4))
        if not self.ptr or (a.nbytes and not load_library().LBM_CopyToDevice(self.ptr, a.ctypes.data, a.nbytes)):
            raise MemoryError("device buffer")

    def __del__(self):
        try:
            if self.ptr:
                load_library().LBM_DeviceFree(self.ptr)
                self.ptr = None
        except Exception:
            pass


def LBM_AdvectParticlesLive(g: Grid, parts: "DeviceParticles", dt: float, flow: LBMParticleFlow, seed: int,
                            step: int) -> None:
    load_library().LBM_AdvectParticlesLive(g.ptr, parts.ptr, parts.n, dt, ctypes.byref(flow), seed, step)


def LBM_TraceStreamlines(g: Grid, num_seeds: int = 256, trace_len: int = 64) -> np.ndarray:
    """RK4 streamlines of the grid's velocity buffer (the application's
    visualisation mode 9): (num_seeds, trace_len, 4) vertices (x, y, z,
    speed), traced into a device buffer and copied back."""
    L = load_library()
    nbytes = 16 * max(num_seeds, 1) * max(trace_len, 1)
    dev = L.LBM_DeviceAlloc(nbytes)
    if not dev:
        raise MemoryError("device allocation for streamlines failed")
    try:
        if not L.LBM_TraceStreamlines(g.ptr, dev, num_seeds, trace_len):
            raise RuntimeError(last_error())
        L.LBM_Synchronize(g.ptr)
        out = np.empty((num_seeds, trace_len, 4), np.float32)
        if not L.LBM_CopyToHost(out.ctypes.data, dev, out.nbytes):
            raise RuntimeError(last_error())
        return out
    finally:
        L.LBM_DeviceFree(dev)


class DeviceParticles:
    """Device array of n 32-byte Particle records."""

    def __init__(self, n: int):
        self.n = n
        self.ptr = load_library().LBM_DeviceAlloc(32 * max(n, 1))
        if not self.ptr:
            raise MemoryError("device allocation for particles failed")

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=np.float32).reshape(self.n, 8)
        if not load_library().LBM_CopyToDevice(self.ptr, h.ctypes.data, 32 * self.n):
            raise RuntimeError(last_error())

    def download(self) -> np.ndarray:
        h = np.empty((self.n, 8), np.float32)
        if not load_library().LBM_CopyToHost(h.ctypes.data, self.ptr, 32 * self.n):
            raise RuntimeError(last_error())
        return h

    def free(self):
        if self.ptr:
            load_library().LBM_DeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def LBM_SeedParticles(g: Grid, parts: DeviceParticles, seed: int) -> None:
    load_library().LBM_SeedParticles(g.ptr, parts.ptr, parts.n, seed)


def LBM_AdvectParticles(g: Grid, parts: DeviceParticles, dt: float, vel_scale: float, seed: int, step: int) -> None:
    load_library().LBM_AdvectParticles(g.ptr, parts.ptr, parts.n, dt, vel_scale, seed, step)


class CdMonitor:
    """The caller-side Cd/Cl bookkeeping of the reference main loop (C, lbm_monitor.c)."""

    def __init__(self, g: Grid, U: float, ref_area: float, verbose: bool = False):
        self.g = g
        self.ptr = load_library().LBM_CdMonitorCreate(g.ptr, U, ref_area, int(verbose))
        if not self.ptr:
            raise MemoryError("LBM_CdMonitorCreate failed")

    def sample(self):
        """Returns (newly_converged, dict of Cd, Cd_corr, Cl, Cd_pressure, Cd_friction, ema)."""
        out = np.zeros(6, np.float32)
        newly = load_library().LBM_CdMonitorSample(self.ptr, self.g.ptr, _fp(out))
        keys = ("Cd", "Cd_corr", "Cl", "Cd_pressure", "Cd_friction", "ema")
        return bool(newly), dict(zip(keys, out.tolist()))

    def set_body(self, reynolds: float, bbox):
        """World bounding box (minX, minY, minZ, maxX, maxY, maxZ) and Re of
        the model: the next verbose sample prints the Cd breakdown."""
        load_library().LBM_CdMonitorSetBody(self.ptr, self.g.ptr, reynolds, *bbox)

    def converged(self):
        m, r = c_float(), c_float()
        ok = load_library().LBM_CdMonitorConverged(self.ptr, ctypes.byref(m), ctypes.byref(r))
        return bool(ok), m.value, r.value

    def strouhal(self, sample_interval: int, char_length: float) -> float:
        return float(load_library().LBM_CdMonitorStrouhal(self.ptr, sample_interval, char_length))

    def __del__(self):
        try:
            if self.ptr:
                load_library().LBM_CdMonitorFree(self.ptr)
                self.ptr = None
        except Exception:
            pass


def LBM_ComputeStrouhal(cl: np.ndarray, sample_interval: int, char_length: float, U: float) -> float:
    c = np.ascontiguousarray(cl, dtype=np.float32)
    return float(load_library().LBM_ComputeStrouhal(_fp(c), len(c), sample_interval, char_length, U))


def LBM_EffectiveReynolds(g: Grid, U: float, minX: float, maxX: float, print_line: bool = True) -> float:
    return float(load_library().LBM_EffectiveReynolds(g.ptr, U, minX, maxX, int(print_line)))


def LBM_WriteVTI(g: Grid, directory: str, step: int) -> int:
    return int(load_library().LBM_WriteVTI(g.ptr, directory.encode(), step))


def LBM_SetBlockSize(g: Grid, bs: int) -> None:
    if not load_library().LBM_SetBlockSize(g.ptr, bs):
        raise ValueError(last_error())


STREAM_PHASE_A, STREAM_PHASE_B = 1, 2


def LBM_SetStreamingAccess(g: Grid, mode: int) -> None:
    """mode: STREAM_PHASE_A | STREAM_PHASE_B bits, 0 = default cache policy."""
    if not load_library().LBM_SetStreamingAccess(g.ptr, int(mode)):
        raise ValueError(last_error())


def LBM_SetBlockOrder(g: Grid, xcd_contiguous: bool) -> None:
    if not load_library().LBM_SetBlockOrder(g.ptr, int(xcd_contiguous)):
        raise ValueError(last_error())


def LBM_GetKernelConfig(g: Grid):
    """(workgroup size, streaming mode, XCD-contiguous order) of the step kernel."""
    bs, nt, xo = c_int(), c_int(), c_int()
    if not load_library().LBM_GetKernelConfig(g.ptr, ctypes.byref(bs), ctypes.byref(nt), ctypes.byref(xo)):
        raise RuntimeError(last_error())
    return bs.value, nt.value, bool(xo.value)


def LBM_Autotune(g: Grid, steps: int, ux: float, uy: float = 0.0, uz: float = 0.0):
    """Returns (chosen workgroup size, best ms per step for 128/256/512 lanes);
    the chosen streaming mode is read back with LBM_GetKernelConfig."""
    ms = np.zeros(3, np.float32)
    bs = load_library().LBM_Autotune(g.ptr, steps, ux, uy, uz, _fp(ms))
    if not bs:
        raise RuntimeError(last_error())
    return int(bs), ms


class Batch:
    """Many tunnels advanced together (include/lbm.h LBM_Batch*)."""

    def __init__(self, storage=LBM_STORAGE_FP32, useRegularized=0, useMRT=1, useSmagorinsky=0,
                 smagorinskyCs=0.1, periodicYZ=1):
        self.ptr = load_library().LBM_BatchCreate(storage, useRegularized, useMRT, useSmagorinsky,
                                                   smagorinskyCs, periodicYZ)
        if not self.ptr:
            raise MemoryError("LBM_BatchCreate failed")

    def add(self, geometry: Grid, tau: float, ux: float, uy: float = 0.0, uz: float = 0.0) -> int:
        k = load_library().LBM_BatchAddTunnel(self.ptr, geometry.ptr, tau, ux, uy, uz)
        if k < 0:
            raise RuntimeError("LBM_BatchAddTunnel failed")
        return k

    def initialize(self):
        if not load_library().LBM_BatchInitialize(self.ptr):
            raise RuntimeError("LBM_BatchInitialize failed: " + last_error())

    def step(self, n: int):
        load_library().LBM_BatchStep(self.ptr, n)

    def profile(self, n: int, use_graph: bool = True) -> float:
        ms = c_float()
        if not load_library().LBM_BatchProfile(self.ptr, n, int(use_graph), ctypes.byref(ms)):
            raise RuntimeError(last_error())
        return ms.value

    def autotune(self, steps: int):
        ms = np.zeros(3, np.float32)
        bs = load_library().LBM_BatchAutotune(self.ptr, steps, _fp(ms))
        if not bs:
            raise RuntimeError(last_error())
        return int(bs), ms

    def initialize_from_field(self, tunnel: int, field: np.ndarray):
        """Replaces the tunnel's uniform start with the equilibrium of a coarse
        (cnz, cny, cnx, 4) (ux, uy, uz, rho) field (before the first step)."""
        f = np.ascontiguousarray(field, dtype=np.float32)
        cnz, cny, cnx = f.shape[:3]
        if not load_library().LBM_BatchInitializeTunnelFromField(self.ptr, tunnel, _fp(f), cnx, cny, cnz):
            raise RuntimeError("LBM_BatchInitializeTunnelFromField failed")

    def set_block_size(self, bs: int):
        if not load_library().LBM_BatchSetBlockSize(self.ptr, bs):
            raise RuntimeError("invalid block size")

    def set_streaming(self, mode: int):
        if not load_library().LBM_BatchSetStreamingAccess(self.ptr, int(mode)):
            raise ValueError("invalid streaming mode (0-3) or batch not initialised")

    def set_block_order(self, xcd_contiguous: bool):
        if not load_library().LBM_BatchSetBlockOrder(self.ptr, int(xcd_contiguous)):
            raise ValueError("batch not initialised")

    def kernel_config(self):
        """(workgroup size, streaming mode, XCD-contiguous order) of the batch step kernel."""
        bs, nt, xo = c_int(), c_int(), c_int()
        if not load_library().LBM_BatchGetKernelConfig(self.ptr, ctypes.byref(bs), ctypes.byref(nt),
                                                       ctypes.byref(xo)):
            raise RuntimeError("batch not initialised")
        return bs.value, nt.value, bool(xo.value)

    @property
    def ntunnels(self) -> int:
        return int(load_library().LBM_BatchTunnelCount(self.ptr))

    @property
    def guard_resets(self) -> int:
        n = int(load_library().LBM_BatchGuardResetCount(self.ptr))
        if n < 0:
            raise RuntimeError("LBM_BatchGuardResetCount failed: " + last_error())
        return n

    @property
    def total_cells(self) -> int:
        return int(load_library().LBM_BatchTotalCells(self.ptr))

    def forces(self):
        n = self.ntunnels
        out = np.zeros((n, 6), np.float64)
        fixed = np.zeros((n, 7), np.int64)
        if not load_library().LBM_BatchComputeForcesF64(self.ptr, out.ctypes.data_as(c_double_p),
                                                         fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))):
            raise RuntimeError(last_error())
        return out, fixed

    def coefficients(self) -> np.ndarray:
        out = np.zeros((self.ntunnels, 4), np.float32)
        if not load_library().LBM_BatchComputeCoefficients(self.ptr, _fp(out)):
            raise RuntimeError(last_error())
        return out

    def populations(self, tunnel: int, cells: int):
        f = np.empty((cells, 19), np.float64)
        fn = np.empty((cells, 19), np.float64)
        if not load_library().LBM_BatchCopyPopulationsToHostF64(self.ptr, tunnel, f.ctypes.data_as(c_double_p),
                                                                 fn.ctypes.data_as(c_double_p)):
            raise RuntimeError(last_error())
        return f, fn

    def velocity(self, tunnel: int, cells: int) -> np.ndarray:
        out = np.empty((cells, 4), np.float32)
        if not load_library().LBM_BatchCopyVelocityToHost(self.ptr, tunnel, _fp(out)):
            raise RuntimeError(last_error())
        return out

    def coarse_velocity(self, tunnel: int, dims, factor: int) -> np.ndarray:
        """Block means of (ux, uy, uz, rho) over factor^3 cells of one tunnel
        (dims = its (nx, ny, nz)), shape (nz/f, ny/f, nx/f, 4)."""
        nx, ny, nz = dims
        out = np.empty((nz // factor, ny // factor, nx // factor, 4), np.float32)
        if not load_library().LBM_BatchCopyCoarseVelocityToHost(self.ptr, tunnel, factor, _fp(out)):
            raise RuntimeError("LBM_BatchCopyCoarseVelocityToHost failed")
        return out

    def __del__(self):
        try:
            if self.ptr:
                load_library().LBM_BatchFree(self.ptr)
                self.ptr = None
        except Exception:
            pass


def LBM_MeasureCopyBandwidth(nbytes: int, iters: int = 20) -> float:
    return float(load_library().LBM_MeasureCopyBandwidth(nbytes, iters))
