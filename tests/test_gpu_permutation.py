"""The in-place two-phase streaming is a pure permutation of the population
slots: every slot is written by exactly one lane per step and read by the
lane that owns it (the invariant the reference keeps with two buffers and
barriers, lbm.c:723-821).  With BGK at tau = 1e30 the collision is the
identity to the last bit (omega (f - feq) is below half an ulp of f), so after
n steps from a field of distinct equilibria f0 the stored post-collision
populations must be exactly f_new(c, i) = f0(c - (n - 1) e_i, i) and the
streamed ones f(c, i) = f0(c - n e_i, i) for every cell n + 1 cells or more
from the x faces (periodic y/z).  A slot written twice, skipped, or read
after a neighbour overwrote it in the same phase breaks the equality.  Both
phase-B forms run: whole rows per workgroup (row-aligned stores, nx = 64) and
per-cell stores (nx = 48), at every workgroup size and cache policy, in FP32
and in FP64 storage (where tau = 1e30 is the identity as well)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

E = np.asarray(O.D3Q19_E, dtype=np.int64)


def shifted(f0, dims, shift):
    """f0[c - shift * e_i, i] on the (x, y, z) torus, f0 as (nz, ny, nx, 19)."""
    out = np.empty_like(f0)
    for i in range(19):
        ex, ey, ez = (int(v) * shift for v in E[i])
        out[..., i] = np.roll(f0[..., i], (ez, ey, ex), axis=(0, 1, 2))
    return out


@pytest.mark.parametrize("nx", [64, 48])
@pytest.mark.parametrize("bs", [128, 256, 512])
@pytest.mark.parametrize("mode", [0, 3])
@pytest.mark.parametrize("storage", [0, 3])
def test_streaming_is_a_permutation_of_the_slots(gpu, nx, bs, mode, storage):
    ny, nz = 16, 12
    g = gpu.LBM_Create(nx, ny, nz, 0.1)
    g.periodicYZ = 1
    g.populationStorage = storage
    f64 = storage == gpu.LBM_STORAGE_FP64
    bits = np.uint64 if f64 else np.uint32
    rng = np.random.default_rng(nx + bs + mode)
    field = np.empty((nz, ny, nx, 4), np.float32)
    field[..., 0] = rng.uniform(0.03, 0.07, (nz, ny, nx))
    field[..., 1] = rng.uniform(-0.02, 0.02, (nz, ny, nx))
    field[..., 2] = rng.uniform(-0.02, 0.02, (nz, ny, nx))
    field[..., 3] = rng.uniform(0.98, 1.02, (nz, ny, nx))
    assert gpu.LBM_InitializeFlowFromField(g, field) == 1
    gpu.LBM_SetBlockSize(g, bs)
    gpu.LBM_SetStreamingAccess(g, mode)
    g.tau = 1e30
    f0 = gpu.LBM_CopyPopulationsToHost(g, f64=f64)[1].reshape(nz, ny, nx, 19)
    assert len(np.unique(f0)) > 0.9 * f0.size  # distinct values: a misplaced slot shows
    for n in range(1, 5):
        gpu.LBM_Step(g, 0.05, 0.0, 0.0)
        f, fn = (a.reshape(nz, ny, nx, 19) for a in gpu.LBM_CopyPopulationsToHost(g, f64=f64))
        inner = slice(n + 1, nx - n - 1)
        want_new = shifted(f0, (nx, ny, nz), n - 1)
        want = shifted(f0, (nx, ny, nz), n)
        bad = np.argwhere(fn[:, :, inner].view(bits) != want_new[:, :, inner].view(bits))
        assert bad.size == 0, ("f_new", n, bad[:4])
        bad = np.argwhere(f[:, :, inner].view(bits) != want[:, :, inner].view(bits))
        assert bad.size == 0, ("f", n, bad[:4])
    assert gpu.LBM_GetGuardResetCount(g) == 0
