"""The batched sweep as configured: 64 tunnels of uneven size (nx uniform in
[64, 192], ny = nz = nx / 2, 46 M cells in all), each with a seeded random
box, ellipsoid or slanted-rear body meshed and voxelised with Bouzidi links,
inlet u in [0.02, 0.08], tau in [0.51, 0.8]; MRT collision, FP16-shifted
storage, periodic y/z; advanced together by the batch step kernel replayed
from HIP graphs.

After 300 steps every tunnel's populations (both of the reference's
buffers), velocity field and FP64 force sums equal, bit for bit, those of the
same tunnel run alone through LBM_StepN -- whose FP16 results the 128x64x64 storage-format
and Taylor-Green tests tie to the oracle -- and its Cd/Cl are finite."""
import numpy as np
import pytest

from lattice_fluid_dynamics_amd import experiments

pytestmark = pytest.mark.gpu

STEPS = 300


def test_batch_of_64_tunnels_equals_single_grids(gpu):
    tunnels = experiments.sweep_tunnels(64)
    cells = [t["dims"][0] * t["dims"][1] * t["dims"][2] for t in tunnels]
    assert 40e6 < sum(cells) < 55e6
    b = gpu.Batch(storage=gpu.LBM_STORAGE_FP16, useMRT=1, useSmagorinsky=0, periodicYZ=1)
    kinds = []
    for t in tunnels:
        g, kind = experiments.sweep_geometry(t)
        assert b.add(g, t["tau"], t["u"]) == len(kinds)
        kinds.append(kind)
        g.free()
    assert len(set(kinds)) == 3  # boxes, ellipsoids and slanted-rear bodies
    b.initialize()
    assert b.ntunnels == 64 and b.total_cells == sum(cells)
    b.step(STEPS)  # HIP-graph replay, two steps per graph
    F, fixed = b.forces()
    coef = b.coefficients()
    print("64 tunnels, %d cells: Cd range %.3f-%.3f, |Cl| max %.4f, guard resets %d" % (
        b.total_cells, coef[:, 0].min(), coef[:, 0].max(), np.abs(coef[:, 1]).max(), b.guard_resets))
    assert np.all(np.isfinite(coef)) and np.all(coef[:, 0] > 0)
    assert b.guard_resets == 0
    for k, t in enumerate(tunnels):
        s, _ = experiments.sweep_geometry(t)
        s.tau = t["tau"]
        s.useMRT = 1
        s.periodicYZ = 1
        s.populationStorage = gpu.LBM_STORAGE_FP16
        assert gpu.LBM_InitializeFlow(s, t["u"], 0, 0) == 1
        gpu.LBM_StepN(s, STEPS, t["u"], 0, 0)
        fb, fnb = b.populations(k, cells[k])
        fs, fns = gpu.LBM_CopyPopulationsToHost(s, f64=True)
        assert np.array_equal(fb, fs) and np.array_equal(fnb, fns), (k, t)
        assert np.array_equal(b.velocity(k, cells[k]), gpu.LBM_CopyVelocityToHost(s)), (k, t)
        assert np.array_equal(F[k], gpu.LBM_ComputeDragForceF64(s)), (k, t)
        del fb, fnb, fs, fns
        gpu.LBM_Free(s)
