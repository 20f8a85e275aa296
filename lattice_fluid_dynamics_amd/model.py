"""Python binding of include/lbm_model.h: real meshes from OBJ files placed
as the reference's caller places them, and the occupancy voxeliser.

``mesh_from_obj(path)`` runs the reference caller's sequence (simulation/src/
main.c:239-370): loadOBJ, automatic placement, computeModelBounds and
createTriangleBuffer, and returns the (n, 12) float32 triangle array and
bounding box that ``lbm.LBM_SetSolidMesh`` takes.  ``voxelize(path, R)`` is
Voxelize_FromModel (simulation/lib/voxelize.h:33-37).

Every function takes an optional ``lib``: any shared library exporting the
same C functions, so the same wrapper drives another build of them.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import lbm

VOXELIZE_MODE_SOLID = 0
VOXELIZE_MODE_SURFACE = 1
VOXELIZE_MODE_AUTO = 2


class Vertex(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class Face(ctypes.Structure):
    _fields_ = [("v1", ctypes.c_int), ("v2", ctypes.c_int), ("v3", ctypes.c_int)]


class Model(ctypes.Structure):
    _fields_ = [("vertices", ctypes.POINTER(Vertex)), ("faces", ctypes.POINTER(Face)),
                ("vertexCount", ctypes.c_int), ("faceCount", ctypes.c_int)]


class CarBounds(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("minX", "minY", "minZ", "maxX", "maxY", "maxZ", "centerX", "centerY", "centerZ")]


class CollisionGrid(ctypes.Structure):
    _fields_ = [("cellStart", ctypes.c_void_p), ("cellCount", ctypes.c_void_p), ("triIndices", ctypes.c_void_p),
                ("totalIndices", ctypes.c_int), ("totalCells", ctypes.c_int),
                ("minX", ctypes.c_float), ("minY", ctypes.c_float), ("minZ", ctypes.c_float),
                ("cellSizeX", ctypes.c_float), ("cellSizeY", ctypes.c_float), ("cellSizeZ", ctypes.c_float)]


class ModelPlacement(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("scale", "offsetX", "offsetY", "offsetZ", "rotationY")]


class VoxelGrid(ctypes.Structure):
    _fields_ = [("resolution", ctypes.c_int), ("data", ctypes.POINTER(ctypes.c_uint8))]


_F = ctypes.c_float
SIGNATURES = {
    "loadOBJ": (Model, [ctypes.c_char_p]),
    "freeModel": (None, [ctypes.POINTER(Model)]),
    "computeModelBounds": (CarBounds, [ctypes.POINTER(Model), _F, _F, _F, _F, _F]),
    "createTriangleBuffer": (ctypes.c_void_p, [ctypes.POINTER(Model), _F, _F, _F, _F, _F,
                                               ctypes.POINTER(ctypes.c_int)]),
    "Voxelize_FromModel": (ctypes.POINTER(VoxelGrid), [ctypes.POINTER(Model), ctypes.c_int, _F, ctypes.c_int,
                                                      ctypes.c_int]),
    "Voxelize_WriteBinary": (ctypes.c_int, [ctypes.POINTER(VoxelGrid), ctypes.c_char_p]),
    "Voxelize_SolidFraction": (_F, [ctypes.POINTER(VoxelGrid)]),
    "Voxelize_Free": (None, [ctypes.POINTER(VoxelGrid)]),
    "buildCollisionGrid": (CollisionGrid, [ctypes.c_void_p, ctypes.c_int, _F, _F, _F, _F, _F, _F]),
}
PLACEMENT_SIGNATURE = (ModelPlacement, [ctypes.POINTER(Model), _F, ctypes.c_int])

_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]
_bound = set()


def bind(lib: ctypes.CDLL, placement: bool = True) -> ctypes.CDLL:
    """Sets the argument and result types of the model functions on ``lib``."""
    if id(lib) not in _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        if placement:
            fn = lib.LBM_AutoPlaceModel
            fn.restype, fn.argtypes = PLACEMENT_SIGNATURE
        _bound.add(id(lib))
    return lib


def _lib(lib: Optional[ctypes.CDLL]) -> ctypes.CDLL:
    return bind(lib if lib is not None else lbm.load_library(), placement=lib is None)


class LoadedModel:
    """A Model owned by Python (freed with freeModel)."""

    def __init__(self, path: str, lib: Optional[ctypes.CDLL] = None):
        self.lib = _lib(lib)
        self.m = self.lib.loadOBJ(path.encode())

    def vertices(self) -> np.ndarray:
        n = self.m.vertexCount
        if n == 0:
            return np.zeros((0, 3), np.float32)
        return np.ctypeslib.as_array(ctypes.cast(self.m.vertices, ctypes.POINTER(ctypes.c_float)), (n, 3)).copy()

    def faces(self) -> np.ndarray:
        n = self.m.faceCount
        if n == 0:
            return np.zeros((0, 3), np.int32)
        return np.ctypeslib.as_array(ctypes.cast(self.m.faces, ctypes.POINTER(ctypes.c_int)), (n, 3)).copy()

    def bounds(self, scale, ox, oy, oz, rot) -> Tuple[float, ...]:
        b = self.lib.computeModelBounds(ctypes.byref(self.m), scale, ox, oy, oz, rot)
        return tuple(getattr(b, n) for n, _ in CarBounds._fields_)

    def triangles(self, scale, ox, oy, oz, rot) -> np.ndarray:
        n = ctypes.c_int(0)
        p = self.lib.createTriangleBuffer(ctypes.byref(self.m), scale, ox, oy, oz, rot, ctypes.byref(n))
        if not p:
            return np.zeros((0, 12), np.float32)
        out = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), (n.value, 12)).copy()
        _libc.free(p)
        return out

    def voxelize(self, resolution: int, padding: float = 0.1, align_longest_x: int = 1,
                 mode: int = VOXELIZE_MODE_SOLID) -> Optional[np.ndarray]:
        """(R, R, R) uint8 array indexed [z, y, x], or None where the C call returns NULL."""
        g = self.lib.Voxelize_FromModel(ctypes.byref(self.m), resolution, padding, align_longest_x, mode)
        if not g:
            return None
        R = g.contents.resolution
        out = np.ctypeslib.as_array(g.contents.data, (R, R, R)).copy()
        self.lib.Voxelize_Free(g)
        return out

    def write_voxbin(self, path: str, resolution: int, padding: float = 0.1, align_longest_x: int = 1,
                     mode: int = VOXELIZE_MODE_SOLID) -> int:
        g = self.lib.Voxelize_FromModel(ctypes.byref(self.m), resolution, padding, align_longest_x, mode)
        if not g:
            return -1
        rc = self.lib.Voxelize_WriteBinary(g, path.encode())
        self.lib.Voxelize_Free(g)
        return rc

    def close(self):
        if self.lib is not None:
            self.lib.freeModel(ctypes.byref(self.m))
            self.lib = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


COLL_GRID_RES = 8


def collision_grid(triangles: np.ndarray, bbox, lib: Optional[ctypes.CDLL] = None):
    """buildCollisionGrid over (n, 12) placed triangles and the model box:
    a dict of cellStart, cellCount, triIndices (int32 arrays, R^3 cells of
    R = 8), origin and cell size (float32 triples)."""
    L = _lib(lib)
    t = np.ascontiguousarray(triangles, np.float32)
    g = L.buildCollisionGrid(t.ctypes.data, t.shape[0], *[float(v) for v in bbox])
    if g.totalCells and not (g.cellStart and g.cellCount):
        raise MemoryError("buildCollisionGrid failed")

    def take(ptr, n):
        if not ptr or n == 0:
            return np.zeros(0, np.int32)
        out = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int)), (n,)).copy()
        _libc.free(ptr)
        return out

    out = {"cellStart": take(g.cellStart, g.totalCells), "cellCount": take(g.cellCount, g.totalCells)}
    out["triIndices"] = take(g.triIndices, g.totalIndices) if g.triIndices else np.zeros(0, np.int32)
    if g.triIndices and g.totalIndices == 0:
        _libc.free(g.triIndices)
    out["origin"] = np.array([g.minX, g.minY, g.minZ], np.float32)
    out["cell"] = np.array([g.cellSizeX, g.cellSizeY, g.cellSizeZ], np.float32)
    return out


def auto_placement(model: LoadedModel, scale: float = 0.0) -> Tuple[float, float, float, float, float]:
    """(scale, offsetX, offsetY, offsetZ, rotationY) of the reference caller's
    automatic placement; ``scale`` > 0 keeps that scale (its --scale option)."""
    p = _lib(None).LBM_AutoPlaceModel(ctypes.byref(model.m), scale, int(scale > 0))
    return (p.scale, p.offsetX, p.offsetY, p.offsetZ, p.rotationY)


def mesh_from_obj(path: str, scale: float = 0.0):
    """Triangles (n, 12) float32 and the bounding box (minX, minY, minZ, maxX,
    maxY, maxZ) of the OBJ model placed in the world box as the reference's
    caller places it; the arguments LBM_SetSolidMesh takes."""
    m = LoadedModel(path)
    try:
        place = auto_placement(m, scale)
        tris = m.triangles(*place)
        b = m.bounds(*place)
    finally:
        m.close()
    return tris, b[:6], place


def voxelize(path: str, resolution: int, padding: float = 0.1, align_longest_x: int = 1,
             mode: int = VOXELIZE_MODE_SOLID) -> Optional[np.ndarray]:
    m = LoadedModel(path)
    try:
        return m.voxelize(resolution, padding, align_longest_x, mode)
    finally:
        m.close()
