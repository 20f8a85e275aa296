"""Python binding of include/lbm_superres.h (the super-resolution upscaler)
and the LTWS / normaliser file writers it reads.

``python -m lattice_fluid_dynamics_amd.superres [cx cy cz]`` times the
upscale of a coarse field of that size with random weights and prints one
JSON line (ms per call, TFLOP/s of the MLP).
"""
from __future__ import annotations

import ctypes
import os
import struct
import tempfile
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import lbm

H, IN, OUT = 128, 108, 32
FLOP_PER_CELL = 2 * (IN * H + 2 * H * H + H * OUT)
# file order of the tensors (simulation/src/superres.c:124-154)
SHAPES = [(H,), (H, IN), (H,), (H, H), (H,), (H, H), (OUT,), (OUT, H)]

SR_SIGNATURES = {
    "SR_Create": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p]),
    "SR_Free": (None, [ctypes.c_void_p]),
    "SR_Upscale": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "SR_GetFineVelocityBuffer": (ctypes.c_void_p, [ctypes.c_void_p]),
    "SR_GetFineSize": (None, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int)]),
    "SR_Profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
}
_bound = False


def lib() -> ctypes.CDLL:
    global _bound
    L = lbm.load_library()
    if not _bound:
        for name, (res, args) in SR_SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _bound = True
    return L


def write_ltws(path: str, tensors: Sequence[np.ndarray], version: int = 1, magic: int = 0x4C545753) -> None:
    """u32 magic, u32 version, u32 count, then per tensor u32 ndim, the u32
    sizes and the float32 data (simulation/src/superres.c:43-101)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<III", magic, version, len(tensors)))
        for t in tensors:
            a = np.ascontiguousarray(t, dtype="<f4")
            f.write(struct.pack("<I", a.ndim))
            f.write(struct.pack("<%dI" % a.ndim, *a.shape))
            f.write(a.tobytes())


def write_normalizer(path: str, mean: Sequence[float], std: Sequence[float]) -> None:
    np.concatenate([np.asarray(mean, "<f4"), np.asarray(std, "<f4")]).tofile(path)


def random_weights(seed: int = 0, scale: float = 1.0) -> List[np.ndarray]:
    """Tensors in file order, He-style scaled (fan-in) times `scale`."""
    rng = np.random.default_rng(seed)
    out = []
    for s in SHAPES:
        if len(s) == 1:
            out.append((0.1 * rng.standard_normal(s)).astype(np.float32))
        else:
            out.append((scale * rng.standard_normal(s) * np.sqrt(2.0 / s[1])).astype(np.float32))
    return out


class SuperRes:
    """SR_Create / SR_Upscale / SR_GetFineVelocityBuffer on a device field."""

    def __init__(self, cx: int, cy: int, cz: int, weights_path: str, norm_path: Optional[str]):
        self.ptr = lib().SR_Create(cx, cy, cz, weights_path.encode(), norm_path.encode() if norm_path else None)
        if not self.ptr:
            raise RuntimeError("SR_Create failed")
        self.coarse = (cx, cy, cz)

    def fine_size(self) -> Tuple[int, int, int]:
        v = [ctypes.c_int() for _ in range(3)]
        lib().SR_GetFineSize(self.ptr, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def upscale(self, coarse_dev: int) -> None:
        lib().SR_Upscale(self.ptr, ctypes.c_void_p(coarse_dev))

    def fine_to_host(self) -> np.ndarray:
        fx, fy, fz = self.fine_size()
        out = np.empty((fz, fy, fx, 4), np.float32)
        dev = lib().SR_GetFineVelocityBuffer(self.ptr)
        if not lbm.load_library().LBM_CopyToHost(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dev),
                                                 out.nbytes):
            raise RuntimeError("copy of the fine field failed")
        return out

    def profile(self, coarse_dev: int, n: int) -> float:
        ms = ctypes.c_float()
        if not lib().SR_Profile(self.ptr, ctypes.c_void_p(coarse_dev), n, ctypes.byref(ms)):
            raise RuntimeError("SR_Profile failed")
        return ms.value

    def close(self):
        if self.ptr:
            lib().SR_Free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceField:
    """A device float4 field (coarse input for SuperRes)."""

    def __init__(self, host: np.ndarray):
        a = np.ascontiguousarray(host, np.float32)
        L = lbm.load_library()
        self.nbytes = a.nbytes
        self.ptr = L.LBM_DeviceAlloc(a.nbytes)
        if not self.ptr or not L.LBM_CopyToDevice(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                                  a.nbytes):
            raise RuntimeError("device field allocation failed")

    def close(self):
        if self.ptr:
            lbm.load_library().LBM_DeviceFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def benchmark(cx: int = 256, cy: int = 128, cz: int = 128, iters: int = 10, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    with tempfile.TemporaryDirectory() as d:
        w, nrm = os.path.join(d, "w.ltws"), os.path.join(d, "n.bin")
        write_ltws(w, random_weights(seed))
        write_normalizer(nrm, [0.05, 0.0, 0.0, 1.0], [0.02, 0.01, 0.01, 0.01])
        sr = SuperRes(cx, cy, cz, w, nrm)
    field = np.empty((cz, cy, cx, 4), np.float32)
    field[..., 0] = 0.05 + 0.02 * rng.standard_normal((cz, cy, cx))
    field[..., 1:3] = 0.01 * rng.standard_normal((cz, cy, cx, 2))
    field[..., 3] = 1.0 + 0.01 * rng.standard_normal((cz, cy, cx))
    dev = DeviceField(field)
    sr.profile(dev.ptr, 2)
    ms = sr.profile(dev.ptr, iters)
    cells = cx * cy * cz
    res = {"coarse": [cx, cy, cz], "ms_per_call": ms, "tflops": cells * FLOP_PER_CELL / (ms * 1e-3) / 1e12,
           "gflop_per_call": cells * FLOP_PER_CELL / 1e9}
    sr.close()
    dev.close()
    return res


if __name__ == "__main__":
    import json
    import sys

    dims = [int(v) for v in sys.argv[1:4]] if len(sys.argv) >= 4 else [256, 128, 128]
    print(json.dumps(benchmark(*dims)), flush=True)
