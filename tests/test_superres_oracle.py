"""The super-resolution oracle (oracle/oracle_sr.py) pinned by closed forms:
the vectorised form against the shader's loops written out, a linear field
through the skip path alone, the output ordering through a bias pattern,
the MLP plumbing through identity layers, and the LTWS writer's bytes."""
import numpy as np

from lattice_fluid_dynamics_amd import superres as S
from oracle import oracle_sr as R


def flat(tensors):
    return np.concatenate([np.asarray(t, np.float64).ravel() for t in tensors])


def zero_tensors():
    return [np.zeros(s) for s in S.SHAPES]


def test_vectorised_equals_loops():
    rng = np.random.default_rng(3)
    c = rng.standard_normal((3, 4, 5, 4))
    w = flat(S.random_weights(1))
    mean, std = (0.1, -0.2, 0.05, 1.0), (0.5, 2.0, 0.25, 0.1)
    a = R.upscale(c, w, mean, std)
    b = R.upscale_loops(c, w, mean, std)
    assert a.shape == (6, 8, 10, 4)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


def test_skip_of_a_linear_field():
    """Zero network: fine cell (2x + dx) holds the field at x + 0.25 + 0.5 dx
    (interpolation towards +x/+y/+z only), per axis, away from the upper
    edges where the +1 neighbour is clamped."""
    Z, Y, X = 4, 5, 6
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    g = np.array([0.3, -0.7, 1.1, 0.2])
    c = np.stack([1 + g[k] * x + 0.5 * g[k] * y - 0.25 * g[k] * z + k for k in range(4)], axis=-1)
    out = R.upscale(c, flat(zero_tensors()), (0.5, 0, 0, 1), (2, 1, 1, 3))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                X_, Y_, Z_ = x + 0.25 + 0.5 * dx, y + 0.25 + 0.5 * dy, z + 0.25 + 0.5 * dz
                want = np.stack([1 + g[k] * X_ + 0.5 * g[k] * Y_ - 0.25 * g[k] * Z_ + k for k in range(4)], -1)
                got = out[2 * z + dz, 2 * y + dy, 2 * x + dx]
                inner = (x < X - 1) & (y < Y - 1) & (z < Z - 1)
                assert np.allclose(got[inner], want[inner], rtol=0, atol=1e-12)
    # on the upper faces the +1 neighbour is the cell itself
    assert np.allclose(out[-1, -1, -1], c[-1, -1, -1])


def test_bias_pattern_orders_sub_cells_and_channels():
    """b3[n] = n, all else zero, constant field: fine cell (dz, dy, dx) of a
    coarse cell gets channel ch = n % 4 of n = 4 (4 dz + 2 dy + dx) + ch, de-
    normalised."""
    t = zero_tensors()
    t[6] = np.arange(32.0)
    c0 = np.array([0.04, 0.01, -0.02, 1.01])
    mean, std = np.array([0.05, 0.0, 0.0, 1.0]), np.array([0.02, 0.01, 0.01, 0.01])
    c = np.broadcast_to(c0, (2, 3, 4, 4))
    out = R.upscale(c, flat(t), mean, std)
    nrm = (c0 - mean) / std
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                k = 4 * dz + 2 * dy + dx
                want = (nrm + 4 * k + np.arange(4)) * std + mean
                assert np.allclose(out[1 * 2 + dz, 2 * 2 + dy, 3 * 2 + dx], want, atol=1e-12)


def test_identity_layers_pass_relu_of_the_patch():
    """W0 = [I 0], W1 = W2 = I, W3 = [I 0]: the correction is ReLU of the
    first 32 patch values (neighbours 0-7 of the 27, x fastest), i.e. the
    MLP is row-major [out][in] and ReLU sits between layers."""
    t = zero_tensors()
    t[1] = np.eye(128, 108)
    t[3] = np.eye(128)
    t[5] = np.eye(128)
    t[7] = np.eye(32, 128)
    rng = np.random.default_rng(5)
    c = rng.standard_normal((3, 3, 3, 4))
    base = R.upscale(c, flat(zero_tensors()))
    out = R.upscale(c, flat(t))
    # centre cell (1, 1, 1): its neighbours (dz, dy, dx) in the first 8 patch slots
    nb = [(-1, -1, -1), (-1, -1, 0), (-1, -1, 1), (-1, 0, -1), (-1, 0, 0), (-1, 0, 1), (-1, 1, -1), (-1, 1, 0)]
    corr = np.maximum(np.concatenate([c[1 + d[0], 1 + d[1], 1 + d[2]] for d in nb]), 0)
    for k, (dz, dy, dx) in enumerate([(a, b, e) for a in (0, 1) for b in (0, 1) for e in (0, 1)]):
        got = out[2 + dz, 2 + dy, 2 + dx] - base[2 + dz, 2 + dy, 2 + dx]
        assert np.allclose(got, corr[4 * k:4 * k + 4], atol=1e-12)


def test_ltws_writer_layout(tmp_path):
    p = str(tmp_path / "w.ltws")
    S.write_ltws(p, [np.arange(3, dtype=np.float32), np.ones((2, 2), np.float32)])
    raw = open(p, "rb").read()
    hdr = np.frombuffer(raw[:12], "<u4")
    assert hdr.tolist() == [0x4C545753, 1, 2]
    assert raw[:4] == b"SWTL"  # the magic as little-endian bytes
    assert np.frombuffer(raw[12:16], "<u4")[0] == 1 and np.frombuffer(raw[16:20], "<u4")[0] == 3
    assert np.array_equal(np.frombuffer(raw[20:32], "<f4"), [0, 1, 2])
    assert np.frombuffer(raw[32:44], "<u4").tolist() == [2, 2, 2]
    assert len(raw) == 44 + 16
