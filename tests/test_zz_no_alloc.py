"""LBM_InitializeFlowFromFieldNoAlloc, the field initialisation that calls no
device allocator (its coarse field is staged in a static 4 MiB per-GPU
array), against LBM_InitializeFlowFromField on the same grids and fields:
the initial populations and velocity and the states 20 steps later are
bit-identical, over two calls in a row of which the second passes a smaller
field (the array keeps the first call's contents), in FP32 and FP16
storage.  Fields above the array's size, a storage switch and a changed
geometry are refused with an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (128, 64, 64)


def make(gpu, storage):
    g = gpu.LBM_Create(*DIMS, 0.05)
    g.tau = 0.56
    g.useMRT = 1
    g.periodicYZ = 1
    g.populationStorage = storage
    gpu.LBM_SetSolidSphere(g, -1.0, 0.1, 0.0, 0.6)
    assert gpu.LBM_InitializeFlow(g, 0.05, 0, 0) == 1  # the allocating set-up
    return g


def field(cnx, cny, cnz, seed):
    rng = np.random.default_rng(seed)
    f = np.empty((cnz, cny, cnx, 4), np.float32)
    f[..., 0] = 0.05 + 0.01 * rng.standard_normal((cnz, cny, cnx))
    f[..., 1:3] = 0.005 * rng.standard_normal((cnz, cny, cnx, 2))
    f[..., 3] = 1.0 + 0.002 * rng.standard_normal((cnz, cny, cnx))
    return f


def state(gpu, g):
    f, fn = gpu.LBM_CopyPopulationsToHost(g, f64=True)
    return f, fn, gpu.LBM_CopyVelocityToHost(g)


@pytest.mark.parametrize("storage", [0, 1])
def test_no_alloc_field_init_equals_the_allocating_call(gpu, storage):
    a, b = make(gpu, storage), make(gpu, storage)
    for dims, seed in (((64, 32, 32), 1), ((16, 8, 8), 2)):
        F = field(*dims, seed)
        assert gpu.LBM_InitializeFlowFromField(a, F) == 1
        assert gpu.LBM_InitializeFlowFromFieldNoAlloc(b, F) == 1, gpu.last_error()
        assert b.stepCount == 0
        for x, y in zip(state(gpu, a), state(gpu, b)):
            assert np.array_equal(x, y), dims
        gpu.LBM_StepN(a, 20, 0.05, 0, 0)
        gpu.LBM_StepN(b, 20, 0.05, 0, 0)
        for x, y in zip(state(gpu, a), state(gpu, b)):
            assert np.array_equal(x, y), dims
    gpu.LBM_Free(a)
    gpu.LBM_Free(b)


def test_no_alloc_field_init_refuses_what_would_allocate(gpu):
    g = make(gpu, 0)
    big = field(65, 64, 64, 3)  # 266,240 cells: 4.06 MiB > 4 MiB
    assert gpu.LBM_InitializeFlowFromFieldNoAlloc(g, big) == 0
    assert "exceeds" in gpu.last_error()
    assert gpu.LBM_InitializeFlowFromFieldNoAlloc(g, field(64, 64, 64, 4)) == 1  # exactly 4 MiB
    g.populationStorage = gpu.LBM_STORAGE_FP16
    assert gpu.LBM_InitializeFlowFromFieldNoAlloc(g, field(8, 4, 4, 5)) == 0
    assert "storage format" in gpu.last_error()
    g.populationStorage = gpu.LBM_STORAGE_FP32
    gpu.LBM_SetSolidSphere(g, 1.0, 0.0, 0.0, 0.4)
    assert gpu.LBM_InitializeFlowFromFieldNoAlloc(g, field(8, 4, 4, 6)) == 0
    assert "geometry" in gpu.last_error()
    assert gpu.LBM_InitializeFlowFromField(g, field(8, 4, 4, 6)) == 1
    assert gpu.LBM_InitializeFlowFromFieldNoAlloc(g, field(8, 4, 4, 6)) == 1
    gpu.LBM_Free(g)
