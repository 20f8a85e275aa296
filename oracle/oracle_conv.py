"""CPU FP64 oracle of the warm-start network (test infrastructure only).

Only the tests, the smoke check and the bench's CPU leg use this module, as
the checker; the product runs the HIP kernels of
lattice_fluid_dynamics_amd/csrc/warmstart_kernels.hip.

The reference solver has no warm start, so there is no reference algorithm to
follow; this restates the network that include/lbm_warmstart.h documents:
channels-last tensors (n, d, h, w, c); 3x3x3 convolutions, stride 1, zero
padding 1, weights (cout, 27, cin) with tap t = 9 (dz+1) + 3 (dy+1) + (dx+1);
2x2x2 average pooling; nearest x2 upsampling plus skip; ReLU; mean squared
error; AdamW.

How it is pinned (tests/test_conv_oracle.py):
  * conv_fwd against a pure-Python loop over positions, taps and channels
    with explicit bounds checks (tiny shapes);
  * conv_dgrad and conv_wgrad by adjointness <conv(x, w), y> = <x, dgrad(y, w)>
    = <w, wgrad(y, x)> and by central differences;
  * pool/upsample backward by adjointness with their forward;
  * Net.backward by central differences of the loss on random parameters of
    every layer (FP64, no rounding);
  * adamw against the published update (Loshchilov & Hutter, decoupled
    decay) written out for one scalar over several steps.

`bf16=True` in Net rounds to bfloat16 (round to nearest even) at the points
where the GPU stores BF16: the input, weights, every activation, the loss
gradient, every activation gradient.  Weight gradients stay unrounded.
"""
from __future__ import annotations

import numpy as np

# network layout (include/lbm_warmstart.h)
CONVS = ["c0a", "c0b", "c1a", "c1b", "c2a", "c2b", "c3a", "c3b", "r3", "d2", "r2", "d1", "r1", "d0", "head"]
CIN = [32, 32, 32, 64, 64, 128, 128, 256, 256, 128, 128, 64, 64, 32, 32]
COUT = [32, 32, 64, 64, 128, 128, 256, 256, 128, 128, 64, 64, 32, 32, 32]
LEVEL = [0, 0, 1, 1, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0]


def tap_offsets(t):
    return t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def to_bf16(a):
    """float64/float32 -> nearest-even bfloat16 value, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """bfloat16 bit patterns (uint16) of the nearest-even rounding of a."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_from_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def conv_fwd(x, w, b=None):
    """y[p, co] = b[co] + sum_{t, ci} w[co, t, ci] x[p + off(t), ci]."""
    n, d, h, ww, cin = x.shape
    cout = w.shape[0]
    xp = _pad(np.asarray(x, np.float64))
    y = np.zeros((n, d, h, ww, cout))
    for t in range(27):
        dz, dy, dx = tap_offsets(t)
        xs = xp[:, 1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + ww, :]
        y += xs @ np.asarray(w[:, t, :], np.float64).T
    if b is not None:
        y += b
    return y


def conv_dgrad(gy, w):
    """Input gradient of conv_fwd: gx[q, ci] = sum_{t, co} w[co, t, ci] gy[q - off(t), co]."""
    n, d, h, ww, cout = gy.shape
    cin = w.shape[2]
    gp = _pad(np.asarray(gy, np.float64))
    gx = np.zeros((n, d, h, ww, cin))
    for t in range(27):
        dz, dy, dx = tap_offsets(t)
        gs = gp[:, 1 - dz:1 - dz + d, 1 - dy:1 - dy + h, 1 - dx:1 - dx + ww, :]
        gx += gs @ np.asarray(w[:, t, :], np.float64)
    return gx


def conv_wgrad(gy, x):
    """Weight gradient of conv_fwd: gw[co, t, ci] = sum_p gy[p, co] x[p + off(t), ci]."""
    n, d, h, ww, cin = x.shape
    cout = gy.shape[-1]
    xp = _pad(np.asarray(x, np.float64))
    g2 = np.asarray(gy, np.float64).reshape(-1, cout)
    gw = np.zeros((cout, 27, cin))
    for t in range(27):
        dz, dy, dx = tap_offsets(t)
        xs = xp[:, 1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + ww, :].reshape(-1, cin)
        gw[:, t, :] = g2.T @ xs
    return gw


def pool_fwd(x):
    n, d, h, w, c = x.shape
    return np.asarray(x, np.float64).reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4, 6))


def pool_bwd(gy):
    g = np.asarray(gy, np.float64) / 8.0
    return g.repeat(2, 1).repeat(2, 2).repeat(2, 3)


def up_fwd(x, skip=None):
    y = np.asarray(x, np.float64).repeat(2, 1).repeat(2, 2).repeat(2, 3)
    return y if skip is None else y + skip


def up_bwd(gy):
    n, d, h, w, c = gy.shape
    return np.asarray(gy, np.float64).reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4, 6))


def mse(y, t):
    d = np.asarray(y, np.float64) - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def adamw(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW step (bias-corrected moments, decoupled decay); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w * (1 - lr * wd) - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


class Net:
    """The encoder-decoder of include/lbm_warmstart.h.  params: list of
    (w (cout, 27, cin), b (cout,)) per convolution, in CONVS order."""

    def __init__(self, params, bf16=False):
        self.bf16 = bf16
        r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
        self.r = r
        self.w = [r(w) for w, _ in params]
        self.b = [np.asarray(b, np.float64) for _, b in params]

    def _conv(self, k, x, relu=True):
        y = conv_fwd(x, self.w[k], self.b[k])
        if relu:
            y = np.maximum(y, 0.0)
        return y

    def forward(self, x0):
        r = self.r
        A = {"x0": r(x0)}
        A["a0"] = r(self._conv(0, A["x0"]))
        A["s0"] = r(self._conv(1, A["a0"]))
        A["p1"] = r(pool_fwd(A["s0"]))
        A["a1"] = r(self._conv(2, A["p1"]))
        A["s1"] = r(self._conv(3, A["a1"]))
        A["p2"] = r(pool_fwd(A["s1"]))
        A["a2"] = r(self._conv(4, A["p2"]))
        A["s2"] = r(self._conv(5, A["a2"]))
        A["p3"] = r(pool_fwd(A["s2"]))
        A["a3"] = r(self._conv(6, A["p3"]))
        A["b3"] = r(self._conv(7, A["a3"]))
        A["r3"] = r(self._conv(8, A["b3"]))
        A["u2"] = r(up_fwd(A["r3"], A["s2"]))
        A["d2"] = r(self._conv(9, A["u2"]))
        A["r2"] = r(self._conv(10, A["d2"]))
        A["u1"] = r(up_fwd(A["r2"], A["s1"]))
        A["d1"] = r(self._conv(11, A["u1"]))
        A["r1"] = r(self._conv(12, A["d1"]))
        A["u0"] = r(up_fwd(A["r1"], A["s0"]))
        A["d0"] = r(self._conv(13, A["u0"]))
        A["yf"] = self._conv(14, A["d0"], relu=False)
        self.A = A
        return A["yf"]

    def _back(self, k, g, xin, out, relu, need_input=True):
        """Returns (gw, gb, g_in) of convolution k given the gradient of its output."""
        if relu:
            g = g * (out > 0)
        gw = conv_wgrad(g, xin)
        gb = g.reshape(-1, g.shape[-1]).sum(0)
        gin = self.r(conv_dgrad(g, self.w[k])) if need_input else None
        return gw, gb, gin

    def backward(self, target):
        """Loss and parameter gradients [(gw, gb)] for the last forward."""
        A, r = self.A, self.r
        loss, gy = mse(A["yf"], target)
        gy = r(gy)
        grads = [None] * len(CONVS)

        def back(k, g, xin, out, relu=True, need=True):
            gw, gb, gin = self._back(k, g, xin, out, relu, need)
            grads[k] = (gw, gb)
            return gin

        g_d0 = back(14, gy, A["d0"], None, relu=False)
        g_u0 = back(13, g_d0, A["u0"], A["d0"])
        g_r1 = r(up_bwd(g_u0))
        g_d1 = back(12, g_r1, A["d1"], A["r1"])
        g_u1 = back(11, g_d1, A["u1"], A["d1"])
        g_r2 = r(up_bwd(g_u1))
        g_d2 = back(10, g_r2, A["d2"], A["r2"])
        g_u2 = back(9, g_d2, A["u2"], A["d2"])
        g_r3 = r(up_bwd(g_u2))
        g_b3 = back(8, g_r3, A["b3"], A["r3"])
        g_a3 = back(7, g_b3, A["a3"], A["b3"])
        g_p3 = back(6, g_a3, A["p3"], A["a3"])
        g_s2 = r(g_u2 + r(pool_bwd(g_p3)))
        g_a2 = back(5, g_s2, A["a2"], A["s2"])
        g_p2 = back(4, g_a2, A["p2"], A["a2"])
        g_s1 = r(g_u1 + r(pool_bwd(g_p2)))
        g_a1 = back(3, g_s1, A["a1"], A["s1"])
        g_p1 = back(2, g_a1, A["p1"], A["a1"])
        g_s0 = r(g_u0 + r(pool_bwd(g_p1)))
        g_a0 = back(1, g_s0, A["a0"], A["s0"])
        back(0, g_a0, A["x0"], A["a0"], need=False)
        return loss, grads


def flatten(params):
    """Parameters in LTWS file order as one vector."""
    return np.concatenate([np.concatenate([np.ravel(w), np.ravel(b)]) for w, b in params])


def unflatten(vec):
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        n = co * 27 * ci
        w = np.asarray(vec[o:o + n], np.float64).reshape(co, 27, ci)
        o += n
        b = np.asarray(vec[o:o + co], np.float64)
        o += co
        out.append((w, b))
    return out


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return [(rng.normal(0, scale * np.sqrt(2.0 / (27 * ci)), (co, 27, ci)), rng.normal(0, 0.05, co))
            for ci, co in zip(CIN, COUT)]


# ------------------------------------------------------------ packing
def pack_input(mask, u, nu):
    """(cnz, cny, cnx) solid fractions -> (cnz/2, cny/2, cnx/2, 32) input."""
    cz, cy, cx = mask.shape
    hz, hy, hx = cz // 2, cy // 2, cx // 2
    out = np.zeros((hz, hy, hx, 32))
    for s in range(8):
        dz, dy, dx = s >> 2, (s >> 1) & 1, s & 1
        out[..., s] = mask[dz::2, dy::2, dx::2]
    out[..., 8] = 20.0 * np.float32(u)
    out[..., 9] = 10.0 * np.float32(nu)
    out[..., 10] = (2.0 * (np.arange(hx) + 0.5) / hx - 1.0)[None, None, :]
    out[..., 11] = (2.0 * (np.arange(hy) + 0.5) / hy - 1.0)[None, :, None]
    out[..., 12] = (2.0 * (np.arange(hz) + 0.5) / hz - 1.0)[:, None, None]
    return out


def pack_target(field, u):
    """(cnz, cny, cnx, 4) (ux, uy, uz, rho) -> (cnz/2, cny/2, cnx/2, 32)."""
    cz, cy, cx, _ = field.shape
    out = np.zeros((cz // 2, cy // 2, cx // 2, 32))
    for s in range(8):
        dz, dy, dx = s >> 2, (s >> 1) & 1, s & 1
        c = field[dz::2, dy::2, dx::2]
        out[..., 4 * s + 0] = (c[..., 3] - 1.0) / (u * u)
        out[..., 4 * s + 1] = c[..., 0] / u - 1.0
        out[..., 4 * s + 2] = c[..., 1] / u
        out[..., 4 * s + 3] = c[..., 2] / u
    return out


def trilinear_upsample(coarse, shape):
    """Cell-centred trilinear interpolation of (cz, cy, cx, k) to shape
    (nz, ny, nx): fine cell x at coarse coordinate (x + 1/2) cx/nx - 1/2,
    clamped to the coarse centres."""
    cz, cy, cx, _ = coarse.shape
    nz, ny, nx = shape

    def axis(n, c):
        g = np.clip((np.arange(n) + 0.5) * c / n - 0.5, 0, c - 1)
        i0 = np.floor(g).astype(int)
        i1 = np.minimum(i0 + 1, c - 1)
        return i0, i1, g - i0

    z0, z1, fz = axis(nz, cz)
    y0, y1, fy = axis(ny, cy)
    x0, x1, fx = axis(nx, cx)
    out = np.zeros((nz, ny, nx, coarse.shape[3]))
    for k in range(8):
        zi, wz = (z1, fz) if k & 4 else (z0, 1 - fz)
        yi, wy = (y1, fy) if k & 2 else (y0, 1 - fy)
        xi, wx = (x1, fx) if k & 1 else (x0, 1 - fx)
        w = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        out += w[..., None] * coarse[zi][:, yi][:, :, xi]
    return out
