"""Oracle for the super-resolution upscaler (test infrastructure only).

CPU restatement of simulation/shaders/superres_upscale.comp in FP64:
  * gather (:69-83): the 3x3x3 neighbourhood of every coarse cell, indices
    clamped to the grid (sampleCoarse :50-55), normalised (v - mean) / std
    (:57-59), flattened [dz][dy][dx][channel] with x fastest (108 values);
  * trilinear skip (:86-122): sub-cell (dz, dy, dx) at fractional offsets
    f = 0.25 + 0.5 d interpolates between the centre and its +x/+y/+z
    neighbours only (patch cells 1 + c for c in {0, 1}), weights
    wz wy wx with w = f for c = 1 and 1 - f for c = 0;
  * MLP (:124-147, :175-199): 108 -> 128 -> 128 -> 128 -> 32, ReLU after the
    first three layers, W row-major [out][in], bias first;
  * output (:201-220): skip + correction, de-normalised v * std + mean,
    sub-cell k = 4 dz + 2 dy + dx to fine cell (2x + dx, 2y + dy, 2z + dz).
Weights arrive in the LTWS file order b0, W0, b1, W1, b2, W2, b3, W3
(simulation/src/superres.c:124-154).

`upscale` is vectorised with numpy; `upscale_loops` follows the shader's
per-invocation loops literally and is for tiny grids (it cross-checks the
vectorised form).
"""
import numpy as np

H, IN, OUT = 128, 108, 32


def split_weights(flat):
    """b0, W0, b1, W1, b2, W2, b3, W3 from the concatenated LTWS floats."""
    flat = np.asarray(flat, np.float64).ravel()
    sizes = [H, H * IN, H, H * H, H, H * H, OUT, OUT * H]
    parts, o = [], 0
    for s in sizes:
        parts.append(flat[o:o + s])
        o += s
    b0, W0, b1, W1, b2, W2, b3, W3 = parts
    return b0, W0.reshape(H, IN), b1, W1.reshape(H, H), b2, W2.reshape(H, H), b3, W3.reshape(OUT, H)


def _patches(n):
    """(Z, Y, X, 108) clamped neighbourhoods of the normalised field n (Z, Y, X, 4)."""
    p = np.pad(n, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="edge")
    Z, Y, X = n.shape[:3]
    cols = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                cols.append(p[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X, :])
    return np.concatenate(cols, axis=-1)


def _skip(patch):
    """(..., 32): sub-cell k = 4 dz + 2 dy + dx, 4 channels each."""
    out = np.zeros(patch.shape[:-1] + (OUT,))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                f = (0.25 + 0.5 * dz, 0.25 + 0.5 * dy, 0.25 + 0.5 * dx)
                k = 4 * dz + 2 * dy + dx
                val = np.zeros(patch.shape[:-1] + (4,))
                for cz in (0, 1):
                    for cy in (0, 1):
                        for cx in (0, 1):
                            w = ((f[0] if cz else 1 - f[0]) * (f[1] if cy else 1 - f[1]) *
                                 (f[2] if cx else 1 - f[2]))
                            idx = ((1 + cz) * 9 + (1 + cy) * 3 + (1 + cx)) * 4
                            val += w * patch[..., idx:idx + 4]
                out[..., 4 * k:4 * k + 4] = val
    return out


def upscale(coarse, flat_weights, mean=(0, 0, 0, 0), std=(1, 1, 1, 1)):
    """coarse (Z, Y, X, 4) -> fine (2Z, 2Y, 2X, 4), FP64."""
    c = np.asarray(coarse, np.float64)
    mean, std = np.asarray(mean, np.float64), np.asarray(std, np.float64)
    b0, W0, b1, W1, b2, W2, b3, W3 = split_weights(flat_weights)
    patch = _patches((c - mean) / std)
    skip = _skip(patch)
    h = np.maximum(patch @ W0.T + b0, 0)
    h = np.maximum(h @ W1.T + b1, 0)
    h = np.maximum(h @ W2.T + b2, 0)
    v = skip + h @ W3.T + b3
    Z, Y, X = c.shape[:3]
    v = v.reshape(Z, Y, X, 2, 2, 2, 4) * std + mean  # sub-cell axes (dz, dy, dx)
    return v.transpose(0, 3, 1, 4, 2, 5, 6).reshape(2 * Z, 2 * Y, 2 * X, 4)


def upscale_loops(coarse, flat_weights, mean=(0, 0, 0, 0), std=(1, 1, 1, 1)):
    """The shader's invocation loops written out (tiny grids only)."""
    c = np.asarray(coarse, np.float64)
    Z, Y, X = c.shape[:3]
    b0, W0, b1, W1, b2, W2, b3, W3 = split_weights(flat_weights)
    fine = np.zeros((2 * Z, 2 * Y, 2 * X, 4))
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                patch = []
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            v = c[min(max(z + dz, 0), Z - 1), min(max(y + dy, 0), Y - 1), min(max(x + dx, 0), X - 1)]
                            for ch in range(4):
                                patch.append((v[ch] - mean[ch]) / std[ch])
                skip = []
                for dz in (0, 1):
                    for dy in (0, 1):
                        for dx in (0, 1):
                            fx, fy, fz = 0.25 + 0.5 * dx, 0.25 + 0.5 * dy, 0.25 + 0.5 * dz
                            val = [0.0] * 4
                            for cz in (0, 1):
                                for cy in (0, 1):
                                    for cx in (0, 1):
                                        w = (fz if cz else 1 - fz) * (fy if cy else 1 - fy) * (fx if cx else 1 - fx)
                                        p = ((1 + cz) * 9 + (1 + cy) * 3 + (1 + cx)) * 4
                                        for ch in range(4):
                                            val[ch] += w * patch[p + ch]
                            skip += val
                h0 = [max(b0[j] + sum(W0[j, i] * patch[i] for i in range(IN)), 0.0) for j in range(H)]
                h1 = [max(b1[j] + sum(W1[j, i] * h0[i] for i in range(H)), 0.0) for j in range(H)]
                h2 = [max(b2[j] + sum(W2[j, i] * h1[i] for i in range(H)), 0.0) for j in range(H)]
                corr = [b3[j] + sum(W3[j, i] * h2[i] for i in range(H)) for j in range(OUT)]
                k = 0
                for dz in (0, 1):
                    for dy in (0, 1):
                        for dx in (0, 1):
                            for ch in range(4):
                                fine[2 * z + dz, 2 * y + dy, 2 * x + dx, ch] = \
                                    (skip[4 * k + ch] + corr[4 * k + ch]) * std[ch] + mean[ch]
                            k += 1
    return fine
