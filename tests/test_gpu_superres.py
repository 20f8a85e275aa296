"""The super-resolution upscaler on the GPU (include/lbm_superres.h) against
the FP64 oracle of the shader (oracle/oracle_sr.py).

Bar: FP32 products accumulated in FP32 in matrix-core order, against FP64:
relative L2 <= 1e-5 and max |diff| <= 1e-4 of the field's scale."""
import os

import numpy as np
import pytest

from lattice_fluid_dynamics_amd import lbm
from lattice_fluid_dynamics_amd import superres as S
from oracle import oracle_sr as R

pytestmark = pytest.mark.gpu

MEAN, STD = (0.05, 0.0, 0.0, 1.0), (0.02, 0.01, 0.01, 0.01)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def flow_like(rng, shape):
    c = np.empty(shape + (4,), np.float32)
    c[..., 0] = 0.05 + 0.02 * rng.standard_normal(shape)
    c[..., 1:3] = 0.01 * rng.standard_normal(shape + (2,))
    c[..., 3] = 1.0 + 0.01 * rng.standard_normal(shape)
    return c


def make(tmp_path, dims, tensors, mean=MEAN, std=STD, norm=True):
    w = str(tmp_path / "w.ltws")
    S.write_ltws(w, tensors)
    n = None
    if norm:
        n = str(tmp_path / "n.bin")
        S.write_normalizer(n, mean, std)
    return S.SuperRes(*dims, w, n)


@pytest.mark.parametrize("dims", [(8, 8, 8), (20, 12, 9), (5, 3, 2), (64, 32, 16)])
def test_upscale_matches_oracle(gpu, tmp_path, dims):
    """Random weights (He scale), a flow-like field; brick-edge tails included."""
    cx, cy, cz = dims
    rng = np.random.default_rng(sum(dims))
    tensors = S.random_weights(sum(dims))
    sr = make(tmp_path, dims, tensors)
    assert sr.fine_size() == (2 * cx, 2 * cy, 2 * cz)
    c = flow_like(rng, (cz, cy, cx))
    dev = S.DeviceField(c)
    sr.upscale(dev.ptr)
    got = sr.fine_to_host()
    want = R.upscale(c, np.concatenate([t.ravel() for t in tensors]), MEAN, STD)
    e = rel_l2(got, want)
    m = float(np.max(np.abs(got - want)))
    print(dims, "rel L2 %.2e max %.2e" % (e, m))
    assert e < 1e-5 and m < 1e-4


def test_output_ordering_and_missing_normalizer(gpu, tmp_path):
    """b3[n] = n with a zero network, constant field, no normaliser file
    (mean 0, std 1): every fine cell holds c0 + n for its sub-cell and
    channel, exactly."""
    t = [np.zeros(s, np.float32) for s in S.SHAPES]
    t[6] = np.arange(32, dtype=np.float32)
    sr = make(tmp_path, (6, 5, 4), t, norm=False)
    c0 = np.array([0.25, -0.5, 0.125, 1.0], np.float32)
    dev = S.DeviceField(np.broadcast_to(c0, (4, 5, 6, 4)).copy())
    sr.upscale(dev.ptr)
    got = sr.fine_to_host()
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                k = 4 * dz + 2 * dy + dx
                want = c0 + 4 * k + np.arange(4, dtype=np.float32)
                assert np.array_equal(got[dz::2, dy::2, dx::2], np.broadcast_to(want, (4, 5, 6, 4)))


def test_upscale_of_a_solver_velocity_buffer(gpu, tmp_path):
    """The caller's use (simulation/src/main.c:955-956): the grid's velocity
    buffer after some steps goes straight in as the coarse field."""
    g = lbm.LBM_Create(32, 16, 16, 0.02)
    g.periodicYZ = 1
    lbm.LBM_SetSolidSphere(g, -1.0, 0.0, 0.0, 0.6)
    assert lbm.LBM_InitializeFlow(g, 0.05, 0, 0) == 1
    lbm.LBM_StepN(g, 40, 0.05, 0, 0)
    lbm.LBM_Synchronize(g)
    tensors = S.random_weights(7)
    sr = make(tmp_path, (32, 16, 16), tensors)
    sr.upscale(lbm.LBM_GetVelocityBuffer(g))
    got = sr.fine_to_host()
    vel = lbm.LBM_CopyVelocityToHost(g).reshape(16, 16, 32, 4)
    want = R.upscale(vel, np.concatenate([t.ravel() for t in tensors]), MEAN, STD)
    e = rel_l2(got, want)
    print("solver field: rel L2 %.2e" % e)
    assert e < 1e-5


def test_rejected_files(gpu, tmp_path):
    good = S.random_weights(0)
    bad = str(tmp_path / "bad.ltws")
    S.write_ltws(bad, good, magic=0x12345678)
    assert not S.lib().SR_Create(8, 8, 8, bad.encode(), None)
    short = str(tmp_path / "short.ltws")
    S.write_ltws(short, good[:-1])  # W3 missing: fewer floats than the network reads
    assert not S.lib().SR_Create(8, 8, 8, short.encode(), None)
    assert not S.lib().SR_Create(8, 8, 8, str(tmp_path / "missing.ltws").encode(), None)


def test_profile_reports_time(gpu, tmp_path):
    sr = make(tmp_path, (32, 32, 32), S.random_weights(1))
    dev = S.DeviceField(flow_like(np.random.default_rng(0), (32, 32, 32)))
    ms = sr.profile(dev.ptr, 3)
    print("32^3 upscale %.3f ms" % ms)
    assert ms > 0
