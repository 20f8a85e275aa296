"""LBM_TraceStreamlines (the application's visualisation mode 9,
simulation/shaders/streamline_trace.comp dispatched by main.c:1120-1140)
against the oracle's C restatement: bit for bit on a developed flow, on a
rotation written into the velocity buffer, on a resting field (every line
stops at its seed), on ragged seed counts and lengths and on ragged walled
lattices; errors for what the kernel cannot trace."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import _rotation_field

pytestmark = pytest.mark.gpu


def _flow(gpu, dims=(64, 32, 32), steps=300, periodic=1):
    g = gpu.LBM_Create(*dims, 0.02)
    g.useMRT = 1
    g.useSmagorinsky = 1
    g.periodicYZ = periodic
    gpu.LBM_SetSolidSphere(g, -1.5, 0.1, -0.05, 0.45)
    assert gpu.LBM_InitializeFlow(g, 0.06, 0.0, 0.0) == 1
    gpu.LBM_StepN(g, steps, 0.06, 0.0, 0.0)
    return g


def _write_velocity(gpu, g, vel):
    v = np.ascontiguousarray(vel, np.float32)
    gpu.LBM_Synchronize(g)
    assert gpu.load_library().LBM_CopyToDevice(ctypes.c_void_p(g.velocityBuffer), v.ctypes.data, v.nbytes) == 1


@pytest.mark.parametrize("num_seeds,trace_len", [(256, 64), (200, 1), (1, 100), (37, 33)])
def test_streamlines_match_oracle_on_a_developed_flow(gpu, num_seeds, trace_len):
    g = _flow(gpu)
    vel = gpu.LBM_CopyVelocityToHost(g)
    got = gpu.LBM_TraceStreamlines(g, num_seeds, trace_len)
    ref = O.trace_streamlines(vel, g.sizeX, g.sizeY, g.sizeZ, num_seeds, trace_len)
    assert got.shape == ref.shape
    assert np.all(np.isfinite(got))
    bad = np.argwhere(got.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, (bad[:5], got[tuple(bad[0])], ref[tuple(bad[0])])
    # the lines move: downstream of their seeds, at the flow's speed
    if trace_len > 1:
        assert np.all(got[:, -1, 0] > got[:, 0, 0])
        assert 0.3 < np.median(got[:, 0, 3]) < 0.7  # 8 |u| with u ~ 0.06


@pytest.mark.parametrize("dims", [(50, 36, 20), (33, 2, 70)])
def test_streamlines_match_oracle_on_ragged_walled_lattices(gpu, dims):
    """Dimensions that are no multiple of the workgroup or of each other, with
    clamped (walled) y/z faces, and the thinnest lattice the tracer accepts
    (ny = 2): still bit for bit."""
    g = _flow(gpu, dims=dims, steps=200, periodic=0)
    vel = gpu.LBM_CopyVelocityToHost(g)
    got = gpu.LBM_TraceStreamlines(g, 45, 27)
    ref = O.trace_streamlines(vel, *dims, 45, 27)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_streamlines_match_oracle_on_a_rotation_and_at_rest(gpu):
    g = _flow(gpu, dims=(32, 64, 64), steps=2)
    vel = _rotation_field(32, 64, 64, 1.0)
    _write_velocity(gpu, g, vel)
    got = gpu.LBM_TraceStreamlines(g, 64, 32)
    ref = O.trace_streamlines(vel, 32, 64, 64, 64, 32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    r = np.hypot(got[:, :, 1], got[:, :, 2])
    inner = r[:, 0] < 1.3
    assert np.allclose(r[inner], r[inner, :1], rtol=2e-3)  # circles
    _write_velocity(gpu, g, np.zeros((32 * 64 * 64, 4), np.float32))
    rest = gpu.LBM_TraceStreamlines(g, 10, 5)
    assert np.array_equal(rest.view(np.uint32), O.trace_streamlines(np.zeros((32 * 64 * 64, 4), np.float32), 32, 64,
                                                                    64, 10, 5).view(np.uint32))
    assert np.all(rest[:, :, 3] == 0.0) and np.all(rest[:, :, 0] == -2.5)


def test_streamlines_refuse_what_they_cannot_trace(gpu):
    L = gpu.load_library()
    g = _flow(gpu, dims=(16, 8, 8), steps=1)
    buf = L.LBM_DeviceAlloc(16 * 64)
    try:
        assert L.LBM_TraceStreamlines(g.ptr, buf, 0, 4) == 0
        assert "numSeeds" in gpu.last_error()
        assert L.LBM_TraceStreamlines(g.ptr, buf, 4, 0) == 0
        assert L.LBM_TraceStreamlines(g.ptr, None, 4, 4) == 0
        assert L.LBM_TraceStreamlines(None, buf, 4, 4) == 0
        assert L.LBM_TraceStreamlines(g.ptr, buf, 4, 4) == 1
        flat = gpu.LBM_Create(16, 1, 8, 0.02)
        assert gpu.LBM_InitializeFlow(flat, 0.05, 0.0, 0.0) == 1
        assert L.LBM_TraceStreamlines(flat.ptr, buf, 4, 4) == 0
        assert "at least 2" in gpu.last_error()
    finally:
        L.LBM_Synchronize(g.ptr)
        L.LBM_DeviceFree(buf)
