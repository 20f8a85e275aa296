"""Pins the FP64 oracle of the warm-start network (oracle/oracle_conv.py)
against definitions that do not share its code: a pure-Python loop
convolution, adjointness, central differences and the AdamW formula."""
import numpy as np
import pytest

from oracle import oracle_conv as C


def loop_conv(x, w, b):
    n, d, h, ww, cin = x.shape
    cout = w.shape[0]
    y = np.zeros((n, d, h, ww, cout))
    for s in range(n):
        for z in range(d):
            for yy in range(h):
                for xx in range(ww):
                    for co in range(cout):
                        acc = b[co]
                        for dz in (-1, 0, 1):
                            for dy in (-1, 0, 1):
                                for dx in (-1, 0, 1):
                                    Z, Y, X = z + dz, yy + dy, xx + dx
                                    if not (0 <= Z < d and 0 <= Y < h and 0 <= X < ww):
                                        continue
                                    t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                                    for ci in range(cin):
                                        acc += w[co, t, ci] * x[s, Z, Y, X, ci]
                        y[s, z, yy, xx, co] = acc
    return y


def test_conv_matches_loop_definition():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 3, 4, 5, 2))
    w = rng.normal(size=(3, 27, 2))
    b = rng.normal(size=3)
    np.testing.assert_allclose(C.conv_fwd(x, w, b), loop_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 6, 5, 3))
    w = rng.normal(size=(4, 27, 3))
    g = rng.normal(size=(2, 4, 6, 5, 4))
    lhs = np.sum(C.conv_fwd(x, w) * g)
    assert abs(lhs - np.sum(x * C.conv_dgrad(g, w))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(w * C.conv_wgrad(g, x))) < 1e-9 * abs(lhs)


def test_conv_central_differences():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 3, 3, 4, 2))
    w = rng.normal(size=(2, 27, 2))
    g = rng.normal(size=(1, 3, 3, 4, 2))
    gw, gx = C.conv_wgrad(g, x), C.conv_dgrad(g, w)
    eps = 1e-6
    for idx in [(0, 0, 0), (1, 13, 1), (0, 26, 1), (1, 5, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = (np.sum(C.conv_fwd(x, wp) * g) - np.sum(C.conv_fwd(x, wm) * g)) / (2 * eps)
        assert abs(fd - gw[idx]) < 1e-6 * max(1, abs(fd))
    for idx in [(0, 0, 0, 0, 0), (0, 2, 1, 3, 1), (0, 1, 2, 0, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        fd = (np.sum(C.conv_fwd(xp, w) * g) - np.sum(C.conv_fwd(xm, w) * g)) / (2 * eps)
        assert abs(fd - gx[idx]) < 1e-6 * max(1, abs(fd))


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 4, 6, 8, 3))
    gy = rng.normal(size=(2, 2, 3, 4, 3))
    assert abs(np.sum(C.pool_fwd(x) * gy) - np.sum(x * C.pool_bwd(gy))) < 1e-10
    xc = rng.normal(size=(2, 2, 3, 4, 3))
    gf = rng.normal(size=(2, 4, 6, 8, 3))
    assert abs(np.sum(C.up_fwd(xc) * gf) - np.sum(xc * C.up_bwd(gf))) < 1e-10
    # pooling averages exactly the 2x2x2 block; upsampling copies it
    assert np.isclose(C.pool_fwd(x)[1, 1, 2, 3, 2], x[1, 2:4, 4:6, 6:8, 2].mean())
    assert np.all(C.up_fwd(xc)[0, 2:4, 4:6, 6:8, 1] == xc[0, 1, 2, 3, 1])


def test_network_gradients_central_differences():
    """Net.backward against central differences of the loss (FP64)."""
    params = C.random_params(4, scale=1.0)
    rng = np.random.default_rng(5)
    x0 = rng.normal(size=(1, 8, 8, 8, 32))
    target = rng.normal(size=(1, 8, 8, 8, 32))
    net = C.Net(params)
    net.forward(x0)
    loss, grads = net.backward(target)
    eps = 1e-5
    for k in range(len(C.CONVS)):
        w, b = params[k]
        # the weight with the largest gradient (away from ReLU kinks in practice), and the bias
        gw = grads[k][0]
        idx = np.unravel_index(np.argmax(np.abs(gw)), gw.shape)
        for which, pidx, analytic in (("w", idx, gw[idx]), ("b", (0,), grads[k][1][0])):
            arr = w if which == "w" else b
            old = arr[pidx]
            arr[pidx] = old + eps
            lp, _ = C.mse(C.Net(params).forward(x0), target)
            arr[pidx] = old - eps
            lm, _ = C.mse(C.Net(params).forward(x0), target)
            arr[pidx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - analytic) <= 1e-5 * max(abs(fd), 1e-6) + 1e-10, (C.CONVS[k], which, fd, analytic)


def test_adamw_written_out():
    """Three steps of AdamW on one scalar against the formula written longhand."""
    w, m, v = 1.0, 0.0, 0.0
    lr, b1, b2, eps, wd = 0.01, 0.9, 0.999, 1e-8, 0.1
    ws, ms, vs = np.array([w]), np.array([m]), np.array([v])
    for t, g in enumerate([0.5, -0.2, 0.3], start=1):
        ws, ms, vs = C.adamw(ws, np.array([g]), ms, vs, t, lr, b1, b2, eps, wd)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mhat, vhat = m / (1 - b1 ** t), v / (1 - b2 ** t)
        w = w - lr * wd * w - lr * mhat / (np.sqrt(vhat) + eps)
        assert abs(ws[0] - w) < 1e-15
    # first step moves by ~lr regardless of the gradient's scale (sign step)
    w1, _, _ = C.adamw(np.array([0.0]), np.array([1e-6]), np.zeros(1), np.zeros(1), 1, 0.01)
    assert abs(w1[0] + 0.01) < 1e-4


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 1e-30])
    r = C.to_bf16(vals)
    # ties to even: 1 + 2^-8 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5
    assert np.all(C.bf16_from_bits(C.bf16_bits(vals)) == r)


def test_packing_round_trip_and_upsampling():
    rng = np.random.default_rng(6)
    field = rng.normal(size=(4, 4, 6, 4)) * 0.01 + np.array([0.05, 0, 0, 1.0])
    packed = C.pack_target(field, 0.05)
    assert packed.shape == (2, 2, 3, 32)
    # sub-cell s = 4 dz + 2 dy + dx, target k: channel 4 s + k
    assert np.isclose(packed[1, 0, 2, 4 * 5 + 1], field[3, 0, 5, 0] / 0.05 - 1)
    assert np.isclose(packed[0, 1, 1, 4 * 2 + 0], (field[0, 3, 2, 3] - 1) / 0.05 ** 2)
    mask = rng.uniform(size=(4, 4, 6))
    pin = C.pack_input(mask, 0.05, 0.1)
    assert pin[1, 1, 2, 7] == mask[3, 3, 5] and pin[0, 0, 0, 0] == mask[0, 0, 0]
    # trilinear upsampling reproduces linear fields away from the clamped edges
    z, y, x = np.meshgrid(np.arange(4), np.arange(4), np.arange(6), indexing="ij")
    lin = (0.3 * x + 0.2 * y - 0.1 * z + 1.0)[..., None]
    up = C.trilinear_upsample(lin, (16, 16, 24))
    zf, yf, xf = np.meshgrid((np.arange(16) + 0.5) / 4 - 0.5, (np.arange(16) + 0.5) / 4 - 0.5,
                             (np.arange(24) + 0.5) / 4 - 0.5, indexing="ij")
    exact = 0.3 * xf + 0.2 * yf - 0.1 * zf + 1.0
    inner = (slice(2, 14), slice(2, 14), slice(2, 22))
    np.testing.assert_allclose(up[inner][..., 0], exact[inner], atol=1e-12)
