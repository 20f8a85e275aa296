"""Batched tunnels: every tunnel of a batch evolves bit-for-bit as the same
tunnel run alone through LBM_Step (which the parity tests tie to the oracle),
with and without HIP-graph replay, for every workgroup size and streaming
mode the autotuner can pick."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (dims, sphere, tau, inlet) -- uneven sizes, both Bouzidi branches
TUNNELS = [
    ((40, 20, 20), (-1.0, 0.1, 0.0, 0.7), 0.56, 0.05),
    ((64, 32, 32), (-0.5, -0.2, 0.1, 0.5), 0.62, 0.03),
    ((33, 17, 15), (0.3, 0.0, 0.2, 0.9), 0.80, 0.07),
]


def make_geometry(lbm, dims, sphere):
    g = lbm.LBM_Create(*dims, 0.05)
    lbm.LBM_SetSolidSphere(g, *sphere)
    return g


@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("storage", [0, 1, 2, 3])
def test_batch_equals_single_grids(gpu, periodic, storage):
    b = gpu.Batch(storage=storage, useMRT=1, useSmagorinsky=1, smagorinskyCs=0.1, periodicYZ=periodic)
    singles = []
    for dims, sph, tau, u in TUNNELS:
        geo = make_geometry(gpu, dims, sph)
        b.add(geo, tau, u)
        s = make_geometry(gpu, dims, sph)
        s.tau = tau
        s.useMRT = 1
        s.useSmagorinsky = 1
        s.periodicYZ = periodic
        s.populationStorage = storage
        assert gpu.LBM_InitializeFlow(s, u, 0, 0) == 1
        singles.append((s, u))
        geo.free()
    b.initialize()
    for n in (1, 6, 5):
        b.step(n)
        for s, u in singles:
            gpu.LBM_StepN(s, n, u, 0, 0)
    F, _ = b.forces()
    for k, (s, u) in enumerate(singles):
        cells = s.totalCells
        fb, fnb = b.populations(k, cells)
        fs, fns = gpu.LBM_CopyPopulationsToHost(s, f64=True)
        assert np.array_equal(fb, fs) and np.array_equal(fnb, fns), k
        assert np.array_equal(b.velocity(k, cells), gpu.LBM_CopyVelocityToHost(s)), k
        assert np.array_equal(F[k], gpu.LBM_ComputeDragForceF64(s)), k


def test_graph_replay_autotune_and_block_sizes(gpu):
    def build():
        b = gpu.Batch(storage=0, useMRT=1, periodicYZ=1)
        for dims, sph, tau, u in TUNNELS:
            geo = make_geometry(gpu, dims, sph)
            b.add(geo, tau, u)
            geo.free()
        b.initialize()
        return b

    ref = build()
    ref.set_block_size(256)
    for _ in range(24):  # eager single steps, no graph
        gpu.load_library().LBM_BatchStep(ref.ptr, 1)
    results = []
    for bs, mode, xo in ((128, 0, True), (512, 0, False), (256, 1, True), (512, 3, False), (128, 2, False)):
        b = build()
        b.set_block_size(bs)
        b.set_streaming(mode)  # cache policy and workgroup order only:
        b.set_block_order(xo)  # the fields stay bit-identical
        assert b.kernel_config() == (bs, mode, xo)
        b.step(24)
        results.append(b)
    with pytest.raises(ValueError):
        results[0].set_streaming(4)
    tuned = build()
    best, ms = tuned.autotune(4)
    assert best in (128, 256, 512) and np.all(ms > 0)
    assert tuned.kernel_config()[0] == best and tuned.kernel_config()[1] in (0, 1, 3)
    tuned.step(24)  # tuning restored the state, so this is still steps 1..24
    results.append(tuned)
    for b in results:
        for k, (dims, _, _, _) in enumerate(TUNNELS):
            n = int(np.prod(dims))
            fa, _ = ref.populations(k, n)
            fb, _ = b.populations(k, n)
            assert np.array_equal(fa, fb)


def test_batch_coefficients(gpu):
    b = gpu.Batch(storage=0, useMRT=1, periodicYZ=1)
    areas = []
    for dims, sph, tau, u in TUNNELS:
        geo = make_geometry(gpu, dims, sph)
        areas.append(gpu.LBM_ComputeProjectedArea(geo, 0))
        b.add(geo, tau, u)
        geo.free()
    b.initialize()
    b.step(200)
    c = b.coefficients()
    F, _ = b.forces()
    for k, (dims, sph, tau, u) in enumerate(TUNNELS):
        den = 0.5 * u * u * areas[k]
        assert c[k, 0] == pytest.approx(abs(F[k, 0]) / den, rel=1e-6)
        assert c[k, 1] == pytest.approx(F[k, 1] / den, rel=1e-5, abs=1e-7)
        assert c[k, 0] > 0


def test_batch_row_aligned_equals_per_cell_stores(gpu):
    """Tunnels whose rows fill whole workgroups take the row-aligned phase-B
    stores at workgroup 256 and per-cell stores at 128 (128 does not divide
    the 256-cell rows); both give the same bits, FP32, FP16 and BF16
    storage."""
    tunnels = [((64, 32, 32), (-1.0, 0.1, 0.0, 0.7), 0.56, 0.05),
               ((128, 16, 16), (-0.5, -0.2, 0.1, 0.5), 0.62, 0.03),
               ((256, 16, 8), (0.3, 0.0, 0.2, 0.9), 0.80, 0.07)]
    for storage in (0, 1, 2):
        res = []
        for bs in (256, 128):
            b = gpu.Batch(storage=storage, useRegularized=1, useMRT=0, useSmagorinsky=1, smagorinskyCs=0.1,
                          periodicYZ=1)
            for dims, sph, tau, u in tunnels:
                geo = make_geometry(gpu, dims, sph)
                b.add(geo, tau, u)
                geo.free()
            b.initialize()
            b.set_block_size(bs)
            b.step(9)
            res.append([b.populations(k, int(np.prod(d))) for k, (d, _, _, _) in enumerate(tunnels)])
        for (fa, fna), (fb, fnb) in zip(res[0], res[1]):
            assert np.array_equal(fa, fb) and np.array_equal(fna, fnb), storage
