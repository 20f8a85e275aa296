"""Learned warm start: Python side of include/lbm_warmstart.h.

The network, its training step and its inference run in liblbm_amd.so (C host
code + HIP kernels on the matrix cores).  This module binds them and adds
what surrounds them:
  * training data from the batched solver: seeded random bodies
    (geometry.random_body) in tunnels of one size, run from a cold start,
    then the block-averaged (ux, uy, uz, rho) field averaged over the last
    steps (which removes most of the acoustic oscillation);
  * the training loop: the packed dataset stays on the device (four copies:
    mirrored in y and/or z, with uy/uz negated), batches are gathered by
    device copies, AdamW with a cosine learning-rate schedule;
  * the time-to-converged-Cd measurement for cold and warm starts.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import convergence, geometry, lbm
from .lbm import _fp, c_float, c_float_p, c_int, c_void_p

c_u16_p = ctypes.POINTER(ctypes.c_uint16)

# Every symbol include/lbm_warmstart.h declares, with its ctypes signature.
WS_SIGNATURES = {
    "LBM_WarmStartCreate": (c_void_p, [c_int, c_int, c_int, c_int, ctypes.c_uint]),
    "LBM_WarmStartFree": (None, [c_void_p]),
    "LBM_WarmStartParamTensors": (c_int, [c_void_p]),
    "LBM_WarmStartParamCount": (ctypes.c_longlong, [c_void_p]),
    "LBM_WarmStartSave": (c_int, [c_void_p, ctypes.c_char_p]),
    "LBM_WarmStartLoad": (c_int, [c_void_p, ctypes.c_char_p]),
    "LBM_WarmStartGetParams": (c_int, [c_void_p, c_int, c_float_p]),
    "LBM_WarmStartSetParams": (c_int, [c_void_p, c_float_p]),
    "LBM_WarmStartPackInput": (c_int, [c_float_p, c_float, c_float, c_int, c_int, c_int, c_u16_p]),
    "LBM_WarmStartPackTarget": (c_int, [c_float_p, c_float, c_int, c_int, c_int, c_float_p]),
    "LBM_WarmStartUnpackOutput": (c_int, [c_float_p, c_float, c_int, c_int, c_int, c_float_p]),
    "LBM_WarmStartGradients": (c_float, [c_void_p, c_void_p, c_void_p, c_int]),
    "LBM_WarmStartTrainStep": (c_float, [c_void_p, c_void_p, c_void_p, c_int, c_float, c_float]),
    "LBM_WarmStartForward": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "LBM_WarmStartPredict": (c_int, [c_void_p, c_float_p, c_float, c_float, c_float_p]),
    "LBM_CoarseSolidFraction": (c_int, [lbm.GridP, c_int, c_float_p]),
    "LBM_WarmStartInitialize": (c_int, [lbm.GridP, c_void_p, c_float]),
    "lbmnn_conv_fwd": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p] + [c_int] * 7),
    "lbmnn_conv_fwd_packed": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p] + [c_int] * 7),
    "lbmnn_conv_mirror": (c_int, [c_void_p, c_void_p, c_int, c_int]),
    "lbmnn_conv_pack": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "lbmnn_conv_wgrad": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p] + [c_int] * 6),
    "lbmnn_relu_bwd": (c_int, [c_void_p, c_void_p, ctypes.c_longlong]),
    "lbmnn_bias_grad": (c_int, [c_void_p, c_void_p, ctypes.c_longlong, c_int]),
    "lbmnn_pool_fwd": (c_int, [c_void_p, c_void_p] + [c_int] * 5),
    "lbmnn_pool_bwd": (c_int, [c_void_p, c_void_p] + [c_int] * 5),
    "lbmnn_up_fwd": (c_int, [c_void_p, c_void_p, c_void_p] + [c_int] * 5),
    "lbmnn_up_bwd": (c_int, [c_void_p, c_void_p] + [c_int] * 5),
    "lbmnn_add_inplace": (c_int, [c_void_p, c_void_p, ctypes.c_longlong]),
    "lbmnn_mse": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_longlong]),
    "lbmnn_adamw": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_longlong] + [c_float] * 5
                    + [c_int]),
    "lbmnn_f32_to_bf16": (c_int, [c_void_p, c_void_p, ctypes.c_longlong]),
}

_bound = False


def lib() -> ctypes.CDLL:
    """liblbm_amd.so with the warm-start signatures set (raises if missing)."""
    global _bound
    L = lbm.load_library()
    if not _bound:
        for name, (res, args) in WS_SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


class DeviceArray:
    """A device buffer (LBM_DeviceAlloc) with host copies in and out."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lbm.load_library().LBM_DeviceAlloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if not lbm.load_library().LBM_CopyToDevice(self.ptr + offset, a.ctypes.data, a.nbytes):
            raise RuntimeError(lbm.last_error())

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        if not lbm.load_library().LBM_CopyToHost(out.ctypes.data, self.ptr, out.nbytes):
            raise RuntimeError(lbm.last_error())
        return out

    def free(self):
        if self.ptr:
            lbm.load_library().LBM_DeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------ packing
def pack_input(mask: np.ndarray, u: float, nu: float) -> np.ndarray:
    """(cnz, cny, cnx) solid fractions -> (cnz/2, cny/2, cnx/2, 32) BF16 bits."""
    m = np.ascontiguousarray(mask, np.float32)
    cz, cy, cx = m.shape
    out = np.empty((cz // 2, cy // 2, cx // 2, 32), np.uint16)
    if not lib().LBM_WarmStartPackInput(_fp(m), u, nu, cx, cy, cz, out.ctypes.data_as(c_u16_p)):
        raise ValueError("LBM_WarmStartPackInput failed")
    return out


def pack_target(field: np.ndarray, u: float) -> np.ndarray:
    f = np.ascontiguousarray(field, np.float32)
    cz, cy, cx, _ = f.shape
    out = np.empty((cz // 2, cy // 2, cx // 2, 32), np.float32)
    if not lib().LBM_WarmStartPackTarget(_fp(f), u, cx, cy, cz, _fp(out)):
        raise ValueError("LBM_WarmStartPackTarget failed")
    return out


def unpack_output(packed: np.ndarray, u: float) -> np.ndarray:
    p = np.ascontiguousarray(packed, np.float32)
    hz, hy, hx, _ = p.shape
    out = np.empty((2 * hz, 2 * hy, 2 * hx, 4), np.float32)
    if not lib().LBM_WarmStartUnpackOutput(_fp(p), u, 2 * hx, 2 * hy, 2 * hz, _fp(out)):
        raise ValueError("LBM_WarmStartUnpackOutput failed")
    return out


def coarse_solid_fraction(g: lbm.Grid, factor: int) -> np.ndarray:
    nx, ny, nz = g.shape
    out = np.empty((nz // factor, ny // factor, nx // factor), np.float32)
    if not lib().LBM_CoarseSolidFraction(g.ptr, factor, _fp(out)):
        raise RuntimeError(lbm.last_error())
    return out


class WarmStart:
    """Owning handle of an LBMWarmStart* (coarse lattice cnx x cny x cnz)."""

    def __init__(self, cnx: int = 64, cny: int = 32, cnz: int = 32, max_batch: int = 16, seed: int = 0):
        self.cshape = (cnz, cny, cnx)
        self.max_batch = max_batch
        self.ptr = lib().LBM_WarmStartCreate(cnx, cny, cnz, max_batch, seed)
        if not self.ptr:
            raise RuntimeError("LBM_WarmStartCreate failed: " + lbm.last_error())

    @property
    def packed_positions(self) -> int:
        cz, cy, cx = self.cshape
        return (cz // 2) * (cy // 2) * (cx // 2)

    @property
    def param_count(self) -> int:
        return int(lib().LBM_WarmStartParamCount(self.ptr))

    def get_params(self, which: int = 0) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        if not lib().LBM_WarmStartGetParams(self.ptr, which, _fp(out)):
            raise RuntimeError("LBM_WarmStartGetParams failed")
        return out

    def set_params(self, vec: np.ndarray):
        v = np.ascontiguousarray(vec, np.float32)
        assert v.size == self.param_count
        if not lib().LBM_WarmStartSetParams(self.ptr, _fp(v)):
            raise RuntimeError("LBM_WarmStartSetParams failed")

    def save(self, path: str):
        if not lib().LBM_WarmStartSave(self.ptr, path.encode()):
            raise RuntimeError(lbm.last_error())

    def load(self, path: str):
        if not lib().LBM_WarmStartLoad(self.ptr, path.encode()):
            raise RuntimeError(lbm.last_error())

    def gradients(self, x: DeviceArray, t: DeviceArray, batch: int) -> float:
        loss = lib().LBM_WarmStartGradients(self.ptr, x.ptr, t.ptr, batch)
        if loss < 0:
            raise RuntimeError(lbm.last_error())
        return loss

    def train_step(self, x: DeviceArray, t: DeviceArray, batch: int, lr: float, wd: float) -> float:
        loss = lib().LBM_WarmStartTrainStep(self.ptr, x.ptr, t.ptr, batch, lr, wd)
        if loss < 0:
            raise RuntimeError(lbm.last_error())
        return loss

    def forward(self, x: DeviceArray, batch: int, out: DeviceArray):
        if not lib().LBM_WarmStartForward(self.ptr, x.ptr, batch, out.ptr):
            raise RuntimeError(lbm.last_error())

    def predict(self, mask: np.ndarray, u: float, nu: float) -> np.ndarray:
        """(cnz, cny, cnx) solid fractions -> (cnz, cny, cnx, 4) (ux, uy, uz, rho)."""
        m = np.ascontiguousarray(mask, np.float32)
        assert m.shape == self.cshape
        out = np.empty(self.cshape + (4,), np.float32)
        if not lib().LBM_WarmStartPredict(self.ptr, _fp(m), u, nu, _fp(out)):
            raise RuntimeError(lbm.last_error())
        return out

    def predict_many(self, masks, us, nus) -> list:
        """predict for many tunnels, max_batch of them per pass of the network."""
        P = self.packed_positions
        xin = DeviceArray(2 * P * 32 * self.max_batch)
        xout = DeviceArray(4 * P * 32 * self.max_batch)
        out = []
        for s in range(0, len(masks), self.max_batch):
            idx = list(range(s, min(s + self.max_batch, len(masks))))
            packed = np.stack([pack_input(masks[i], us[i], nus[i]) for i in idx])
            xin.upload(packed)
            self.forward(xin, len(idx), xout)
            o = xout.download(np.float32, (len(idx),) + packed.shape[1:4] + (32,))
            out += [unpack_output(o[k], us[i]) for k, i in enumerate(idx)]
        return out

    def initialize(self, g: lbm.Grid, u: float):
        """LBM_InitializeFlowFromField with this network's prediction for g."""
        if not lib().LBM_WarmStartInitialize(g.ptr, self.ptr, u):
            raise RuntimeError(lbm.last_error())

    def free(self):
        if self.ptr:
            lib().LBM_WarmStartFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# network layout of include/lbm_warmstart.h (convolution widths and levels)
CONV_CIN = [32, 32, 32, 64, 64, 128, 128, 256, 256, 128, 128, 64, 64, 32, 32]
CONV_COUT = [32, 32, 64, 64, 128, 128, 256, 256, 128, 128, 64, 64, 32, 32, 32]
CONV_LEVEL = [0, 0, 1, 1, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0]


def flops_per_sample(cshape, training: bool = True) -> float:
    """Matrix-core FLOP of one sample: 2 * positions * 27 * cin * cout per
    convolution; training adds the input gradient (all but the first
    convolution) and the weight gradient."""
    cz, cy, cx = cshape
    total = 0.0
    for k, (ci, co, lv) in enumerate(zip(CONV_CIN, CONV_COUT, CONV_LEVEL)):
        pos = (cz >> (lv + 1)) * (cy >> (lv + 1)) * (cx >> (lv + 1))
        f = 2.0 * pos * 27 * ci * co
        total += f
        if training:
            total += f + (f if k > 0 else 0.0)
    return total


def train_benchmark(cshape=(32, 32, 64), batch: int = 16, iters: int = 50, seed: int = 0):
    """Training steps on random packed data: ms per step and matrix-core TFLOP/s."""
    cz, cy, cx = cshape
    ws = WarmStart(cx, cy, cz, max_batch=batch, seed=seed)
    rng = np.random.default_rng(seed)
    P = ws.packed_positions
    x = DeviceArray.from_host(np.where(rng.uniform(size=(batch, P, 32)) > 0.5, 0x3F80, 0).astype(np.uint16))
    t = DeviceArray.from_host(rng.normal(size=(batch, P, 32)).astype(np.float32))
    for _ in range(3):
        ws.train_step(x, t, batch, 1e-4, 0.0)
    t0 = time.perf_counter()
    for _ in range(iters):
        ws.train_step(x, t, batch, 1e-4, 0.0)
    dt = (time.perf_counter() - t0) / iters
    fl = flops_per_sample(cshape) * batch
    return {"ms_per_step": dt * 1e3, "tflops": fl / dt / 1e12, "gflop_per_step": fl / 1e9}


# ------------------------------------------------------------ data
@dataclass
class TunnelSpec:
    seed: int   # body seed (geometry.random_body)
    u: float    # inlet speed (lattice units)
    tau: float  # relaxation time; nu = (tau - 1/2) / 3

    @property
    def nu(self) -> float:
        return (self.tau - 0.5) / 3.0


def random_specs(seeds: Sequence[int], u_range=(0.03, 0.07), tau_range=(0.6, 0.9)) -> List[TunnelSpec]:
    """Inlet speed and tau drawn from the body seed, so a seed names a tunnel."""
    out = []
    for s in seeds:
        r = np.random.default_rng(10_000_019 + int(s))
        out.append(TunnelSpec(int(s), float(r.uniform(*u_range)), float(r.uniform(*tau_range))))
    return out


def build_batch(specs: Sequence[TunnelSpec], dims, factor: int, storage: int):
    """A batch of tunnels (MRT, periodic y/z) with the specs' bodies, and the
    coarse solid fraction of each."""
    nx, ny, nz = dims
    b = lbm.Batch(storage=storage, useRegularized=0, useMRT=1, useSmagorinsky=0, periodicYZ=1)
    masks, areas = [], []
    for sp in specs:
        g = lbm.LBM_Create(nx, ny, nz, sp.nu)
        g.tau = sp.tau
        tris, bbox, _ = geometry.random_body(sp.seed)
        lbm.LBM_SetSolidMesh(g, tris, bbox)
        masks.append(coarse_solid_fraction(g, factor))
        areas.append(lbm.LBM_ComputeProjectedArea(g, 0))
        b.add(g, sp.tau, sp.u)
        g.free()
    return b, np.stack(masks), np.asarray(areas)


def solve_fields(specs: Sequence[TunnelSpec], dims=(256, 128, 128), factor: int = 4, steps: int = 8000,
                 avg_steps: int = 2000, avg_every: int = 100, storage: int = 0, chunk: int = 128, log=None,
                 ws: Optional["WarmStart"] = None):
    """Solutions: masks (N, cz, cy, cx) and fields (N, cz, cy, cx, 4), the
    block-averaged field averaged over `avg_steps` after `steps`, from a cold
    start, or from the prediction of `ws` when it is given."""
    masks, fields = [], []
    for c0 in range(0, len(specs), chunk):
        part = specs[c0:c0 + chunk]
        t0 = time.time()
        b, m, _ = build_batch(part, dims, factor, storage)
        b.initialize()
        if ws is not None:
            preds = ws.predict_many(m, [sp.u for sp in part], [sp.nu for sp in part])
            for t in range(len(part)):
                b.initialize_from_field(t, preds[t])
        t1 = time.time()
        if log:
            log(f"tunnels {c0}..{c0 + len(part) - 1}: set up in {t1 - t0:.1f}s")
        for k in range(0, steps, 2000):
            b.step(min(2000, steps - k))
            if log:
                b.coefficients()  # synchronises
                log(f"  step {k + min(2000, steps - k)} ({time.time() - t1:.1f}s)")
        acc = np.zeros((len(part),) + m.shape[1:] + (4,), np.float64)
        nav = 0
        for _ in range(avg_steps // avg_every):
            b.step(avg_every)
            for t in range(len(part)):
                acc[t] += b.coarse_velocity(t, dims, factor)
            nav += 1
        masks.append(m)
        fields.append((acc / nav).astype(np.float32))
        if log:
            log(f"solved tunnels {c0}..{c0 + len(part) - 1}: setup {t1 - t0:.1f}s, "
                f"solve {time.time() - t1:.1f}s")
        del b
    return np.concatenate(masks), np.concatenate(fields)


def mirrored(mask, field, fy: bool, fz: bool):
    """The same tunnel mirrored in y and/or z (uy, uz change sign)."""
    m, f = mask, field.copy()
    if fy:
        m, f = m[:, ::-1, :], f[:, ::-1, :, :]
        f[..., 1] = -f[..., 1]
    if fz:
        m, f = m[::-1, :, :], f[::-1, :, :, :]
        f[..., 2] = -f[..., 2]
    return np.ascontiguousarray(m), np.ascontiguousarray(f)


class DeviceDataset:
    """Packed inputs and targets of every sample (and its mirror images) on the device."""

    def __init__(self, masks, fields, specs: Sequence[TunnelSpec], augment: bool = True):
        variants = [(False, False), (True, False), (False, True), (True, True)] if augment else [(False, False)]
        xs, ts = [], []
        for m, f, sp in zip(masks, fields, specs):
            for fy, fz in variants:
                mm, ff = mirrored(m, f, fy, fz)
                xs.append(pack_input(mm, sp.u, sp.nu))
                ts.append(pack_target(ff, sp.u))
        self.n = len(xs)
        self.xbytes = xs[0].nbytes
        self.tbytes = ts[0].nbytes
        self.x = DeviceArray.from_host(np.stack(xs))
        self.t = DeviceArray.from_host(np.stack(ts))

    def gather(self, idx, xb: DeviceArray, tb: DeviceArray):
        L = lib()
        for k, i in enumerate(idx):
            if not (L.LBM_CopyDeviceToDevice(xb.ptr + k * self.xbytes, self.x.ptr + int(i) * self.xbytes,
                                             self.xbytes)
                    and L.LBM_CopyDeviceToDevice(tb.ptr + k * self.tbytes, self.t.ptr + int(i) * self.tbytes,
                                                 self.tbytes)):
                raise RuntimeError("device copy failed")


def train(ws: WarmStart, data: DeviceDataset, iters: int, batch: int = 16, lr: float = 1e-3, wd: float = 1e-4,
          warmup: int = 50, seed: int = 0, log=None, log_every: int = 100):
    """AdamW with linear warm-up and cosine decay to lr/20.  Returns the losses."""
    rng = np.random.default_rng(seed)
    xb = DeviceArray(batch * data.xbytes)
    tb = DeviceArray(batch * data.tbytes)
    losses = []
    order = rng.permutation(data.n)
    pos = 0
    t0 = time.time()
    for it in range(iters):
        if pos + batch > data.n:
            order = rng.permutation(data.n)
            pos = 0
        idx = order[pos:pos + batch]
        pos += batch
        data.gather(idx, xb, tb)
        if it < warmup:
            a = lr * (it + 1) / warmup
        else:
            p = (it - warmup) / max(1, iters - warmup)
            a = lr / 20 + (lr - lr / 20) * 0.5 * (1 + np.cos(np.pi * p))
        losses.append(ws.train_step(xb, tb, batch, float(a), wd))
        if log and (it % log_every == 0 or it == iters - 1):
            log(f"iter {it}: loss {np.mean(losses[-log_every:]):.5f} lr {a:.2e} ({time.time() - t0:.1f}s)")
    return losses


# ------------------------------------------------------------ convergence
def cd_histories(specs: Sequence[TunnelSpec], dims, steps: int, storage: int, ws: Optional[WarmStart] = None,
                 factor: int = 4, interval: int = 100, avg_from: Optional[int] = None, log=None):
    """Runs the tunnels together and samples Cd every `interval` steps.
    Returns a dict: histories (convergence.CdHistory per tunnel), ms_per_step
    of the whole batch, init_s (warm-start inference + initialisation per
    tunnel), predicted (the warm-start fields) and mean_field (block-averaged
    fields averaged over the samples from step `avg_from` on)."""
    t0 = time.time()
    b, masks, _ = build_batch(specs, dims, factor, storage)
    b.initialize()
    init_s = 0.0
    predicted = None
    if ws is not None:
        t1 = time.perf_counter()
        predicted = ws.predict_many(masks, [sp.u for sp in specs], [sp.nu for sp in specs])
        for t in range(len(specs)):
            b.initialize_from_field(t, predicted[t])
        init_s = (time.perf_counter() - t1) / len(specs)
    if log:
        log(f"batch of {len(specs)} tunnels ready in {time.time() - t0:.1f}s")
    hs = [convergence.CdHistory() for _ in specs]
    acc, nav = None, 0
    ms = 0.0
    done = 0
    while done < steps:
        n = min(interval, steps - done)
        ms += b.profile(n, use_graph=True)
        done += n
        c = b.coefficients()
        for t, h in enumerate(hs):
            h.steps.append(done)
            h.cd.append(float(c[t, 0]))
            h.cl.append(float(c[t, 1]))
            h.cd_pressure.append(float(c[t, 2]))
        if avg_from is not None and done > avg_from:
            v = np.stack([b.coarse_velocity(t, dims, factor) for t in range(len(specs))])
            acc = v.astype(np.float64) if acc is None else acc + v
            nav += 1
        if log and done % 2000 == 0:
            log(f"  step {done}: {ms / done:.3f} ms/step")
    for h in hs:
        h.ms_per_step = ms / steps
    del b
    return {"histories": hs, "ms_per_step": ms / steps, "init_s": init_s, "predicted": predicted,
            "mean_field": None if acc is None else (acc / nav).astype(np.float32), "masks": masks}


if __name__ == "__main__":
    # python -m lattice_fluid_dynamics_amd.warmstart [iters]: training-step benchmark
    import json
    import sys

    print(json.dumps(train_benchmark(iters=int(sys.argv[1]) if len(sys.argv) > 1 else 50)), flush=True)
