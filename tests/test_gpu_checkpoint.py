"""Checkpoint and resume (LBM_SaveState / LBM_LoadState): a grid restored from
a saved state continues bit for bit as the saved grid does, in every storage
format and from either population layout, into a fresh grid or rewinding the
same one; states from other dimensions, formats, y/z modes or geometries, and
short or foreign buffers, are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (48, 24, 24)


def make(gpu, storage, radius=0.6, periodic=1):
    g = gpu.LBM_Create(*DIMS, 0.02)
    g.useMRT = 1
    g.useSmagorinsky = 1
    g.periodicYZ = periodic
    g.populationStorage = storage
    gpu.LBM_SetSolidSphere(g, -1.0, 0.05, -0.1, radius)
    return g


def state(gpu, g):
    f, fn = gpu.LBM_CopyPopulationsToHost(g, f64=g.populationStorage == gpu.LBM_STORAGE_FP64)
    return f, fn, gpu.LBM_CopyVelocityToHost(g), gpu.LBM_ComputeDragForceF64(g)


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("storage", [0, 1, 2, 3])
@pytest.mark.parametrize("first", [101, 100])
def test_resume_continues_bit_for_bit(gpu, storage, first):
    a = make(gpu, storage)
    assert gpu.LBM_InitializeFlow(a, 0.05, 0.0, 0.0) == 1
    gpu.LBM_StepN(a, first, 0.05, 0.0, 0.0)
    saved = gpu.LBM_SaveState(a)
    assert saved.size == gpu.load_library().LBM_GetStateSize(a.ptr)
    assert bytes(saved[:8]) == b"LBMSTAT1"
    gpu.LBM_StepN(a, 40, 0.05, 0.0, 0.0)
    want = state(gpu, a)
    b = make(gpu, storage)
    assert gpu.LBM_LoadState(b, saved) == 1, gpu.last_error()
    assert b.stepCount == first
    gpu.LBM_StepN(b, 40, 0.05, 0.0, 0.0)
    assert same(state(gpu, b), want)
    assert b.stepCount == a.stepCount == first + 40
    # rewinding the saved grid itself
    assert gpu.LBM_LoadState(a, saved) == 1
    gpu.LBM_StepN(a, 40, 0.05, 0.0, 0.0)
    assert same(state(gpu, a), want)


def test_resume_right_after_initialisation(gpu):
    a = make(gpu, 0)
    assert gpu.LBM_InitializeFlow(a, 0.05, 0.0, 0.0) == 1
    saved = gpu.LBM_SaveState(a)
    gpu.LBM_StepN(a, 7, 0.05, 0.0, 0.0)
    want = state(gpu, a)
    b = make(gpu, 0)
    assert gpu.LBM_LoadState(b, saved) == 1
    gpu.LBM_StepN(b, 7, 0.05, 0.0, 0.0)
    assert same(state(gpu, b), want)


def test_mismatched_states_are_refused(gpu):
    a = make(gpu, 1)
    assert gpu.LBM_InitializeFlow(a, 0.05, 0.0, 0.0) == 1
    gpu.LBM_StepN(a, 3, 0.05, 0.0, 0.0)
    saved = gpu.LBM_SaveState(a)
    cases = [
        (make(gpu, 0), "populationStorage"),
        (make(gpu, 1, radius=0.5), "geometry"),
        (make(gpu, 1, periodic=0), "periodicYZ"),
    ]
    for g, word in cases:
        assert gpu.LBM_LoadState(g, saved) == 0
        assert word in gpu.last_error(), (word, gpu.last_error())
    other = gpu.LBM_Create(32, 24, 24, 0.02)
    other.populationStorage = 1
    assert gpu.LBM_LoadState(other, saved) == 0 and "dimensions" in gpu.last_error()
    g = make(gpu, 1)
    assert gpu.LBM_LoadState(g, saved[:-1]) == 0 and "truncated" in gpu.last_error()
    assert gpu.LBM_LoadState(g, saved[:40]) == 0 and "truncated" in gpu.last_error()
    bad = saved.copy()
    bad[0] ^= 1
    assert gpu.LBM_LoadState(g, bad) == 0 and "not a saved" in gpu.last_error()
    L = gpu.load_library()
    fresh = make(gpu, 1)
    buf = np.empty(16, np.uint8)
    assert L.LBM_SaveState(fresh.ptr, buf.ctypes.data, buf.size) == 0 and "no flow" in gpu.last_error()
    assert L.LBM_SaveState(a.ptr, buf.ctypes.data, buf.size) == 0 and "smaller" in gpu.last_error()
    # a refused load leaves the grid as it was
    assert gpu.LBM_LoadState(g, saved) == 1
