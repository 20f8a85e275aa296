"""The reference's own solver tests that need a GPU (simulation/test/test_lbm.c:
369-592), run through the C ABI with the reference's sizes, parameters and
bounds.  These are the only physics-level checks the reference has:

* MRT flow steps with a nonzero drag (test_lbm.c:369-386);
* MRT + Smagorinsky steps without blow-up (test_lbm.c:388-401);
* sphere Cd at Re = 100 within (0.5, 3.5) after 4000 steps (test_lbm.c:403-456);
* box Cd samples positive and within 10x of each other (test_lbm.c:458-541);
* ground-plane cell counts and nonzero drag with the ground (test_lbm.c:543-592).

Each result is also run with FP64 storage and arithmetic (the GPU FP64 path,
itself equal to the FP64 CPU oracle to rel L2 1e-12) and the difference is
printed and bounded where the flow is steady enough for FP32 rounding not to
grow."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = (-0.3, -0.15, -0.15, 0.3, 0.15, 0.15)


def grid(lbm, dims, visc, storage=0, **flags):
    g = lbm.LBM_Create(*dims, visc)
    for k, v in flags.items():
        setattr(g, k, v)
    g.populationStorage = storage
    return g


def test_mrt_flow_step(gpu):  # test_lbm.c:369-386
    out = {}
    for storage in (0, gpu.LBM_STORAGE_FP64):
        g = grid(gpu, (16, 8, 8), 0.1, storage, useMRT=1)
        gpu.LBM_SetSolidAABB(g, *BOX)
        assert gpu.LBM_InitializeFlow(g, 0.05, 0, 0) == 1
        for _ in range(10):
            gpu.LBM_Step(g, 0.05, 0, 0)
        out[storage] = (gpu.LBM_ComputeDragForce(g), gpu.LBM_ComputeDragForceF64(g))
    (fx, fy, fz), F = out[0]
    _, F64 = out[gpu.LBM_STORAGE_FP64]
    print("MRT 16x8x8, 10 steps: fx %.6f (fixed point), FP64-sum Fx %.8f, FP64 storage %.8f" % (fx, F[0], F64[0]))
    assert abs(fx) > 1e-10
    assert abs(F[0] - F64[0]) <= 1e-5 * abs(F64[0])


def test_mrt_smagorinsky_step(gpu):  # test_lbm.c:388-401
    g = grid(gpu, (16, 8, 8), 0.01, useMRT=1, useSmagorinsky=1, smagorinskyCs=0.1)
    assert g.tau == pytest.approx(0.53)  # 0.5 + 3 * 0.01
    gpu.LBM_SetSolidAABB(g, *BOX)
    assert gpu.LBM_InitializeFlow(g, 0.05, 0, 0) == 1
    for _ in range(20):
        gpu.LBM_Step(g, 0.05, 0, 0)
    vel = gpu.LBM_CopyVelocityToHost(g)
    assert np.all(np.isfinite(vel))
    assert gpu.LBM_GetGuardResetCount(g) == 0


def test_sphere_cd_re100(gpu):  # test_lbm.c:403-456
    """128x64x64, sphere of diameter 1 world unit = 16 cells at the origin,
    U 0.05, nu = U D / 100 (tau 0.524), BGK, clamped y/z (LBM_Create's
    defaults), projected area taken before stepping.  The reference asserts
    0.5 < Cd < 3.5 after 4000 steps, as a guard against gross regressions
    (Cd blowing up or collapsing to 0), and observes Cd ~ 2.1.

    The reference scheme itself loses stability in this configuration: in
    the FP32 oracle (C restatement of the shaders) max |u| grows from 0.058
    at step 2000 to 0.10 at step 2500, starting at the outlet corner cell
    (127, 63, 63) where the Zou-He pressure outlet meets the clamped y and z
    faces, and from step ~3000 the stability guard resets cells and Cd
    wanders.  From then on Cd at a given step depends on rounding: at step
    4000 it is 0.73 in the FP32 oracle, 13.3 in FP64, and 0.4-2.6 in FP32
    storage on the GPU depending on the rounding of its reciprocal and
    square root.  So the test pins the flow while the scheme is stable: at
    step 1000, FP32 storage against FP64 storage within 0.1 % and no guard
    reset in either, and the reference's bound 0.5 < Cd < 3.5 at step 2000
    in both precisions (1.25 flow-throughs; FP32 and FP64 within 1 % there).
    Cd and the guard counts at step 4000 are printed."""
    U, D = 0.05, 1.0
    char_len = D * (128 / 8.0)
    visc = np.float32(U * char_len / 100.0)
    res = {}
    for storage in (0, gpu.LBM_STORAGE_FP64):
        g = grid(gpu, (128, 64, 64), float(visc), storage)
        assert g.tau == pytest.approx(0.524, abs=1e-6)
        assert (g.useMRT, g.useRegularized, g.useSmagorinsky, g.periodicYZ) == (0, 0, 0, 0)
        gpu.LBM_SetSolidSphere(g, 0.0, 0.0, 0.0, D / 2)
        A = gpu.LBM_ComputeProjectedArea(g, 0)
        assert gpu.LBM_InitializeFlow(g, U, 0, 0) == 1
        out = []
        for n in (1000, 1000, 2000):
            gpu.LBM_StepN(g, n, U, 0, 0)
            out.append((gpu.LBM_ComputeDragCoefficient(g, U, A), gpu.LBM_GetGuardResetCount(g)))
        res[storage] = (out, A)
    (c1, c2, c4), A = res[0]
    (d1, d2, d4), _ = res[gpu.LBM_STORAGE_FP64]
    print("sphere Re=100: area %.0f; step 1000: Cd %.5f (FP32), %.5f (FP64), rel diff %.2e, guard resets %d / %d; "
          "step 2000: Cd %.4f (FP32), %.4f (FP64), guard resets %d / %d; "
          "step 4000: Cd %.4f (FP32), %.4f (FP64), guard resets %d / %d"
          % (A, c1[0], d1[0], abs(c1[0] - d1[0]) / d1[0], c1[1], d1[1], c2[0], d2[0], c2[1], d2[1],
             c4[0], d4[0], c4[1], d4[1]))
    assert A > 150  # a 16-cell disc: ~ pi 8^2
    assert c1[1] == 0 and d1[1] == 0
    assert abs(c1[0] - d1[0]) <= 1e-3 * d1[0]
    assert 0.5 < c2[0] < 3.5 and 0.5 < d2[0] < 3.5
    assert abs(c2[0] - d2[0]) <= 1e-2 * d2[0]


def test_box_cd_convergence(gpu):  # test_lbm.c:458-541
    """128x64x64, MRT + Smagorinsky (Cs 0.1), nu = U 16 / 50 (tau 0.548), the
    small box, 2500 steps, then five Cd samples 100 steps apart: all
    positive, max/min < 10."""
    U = 0.05
    visc = float(np.float32(U * 16.0 / 50.0))
    res = {}
    for storage in (0, gpu.LBM_STORAGE_FP64):
        g = grid(gpu, (128, 64, 64), visc, storage, useMRT=1, useSmagorinsky=1, smagorinskyCs=0.1)
        gpu.LBM_SetSolidAABB(g, *BOX)
        A = gpu.LBM_ComputeProjectedArea(g, 0)
        assert gpu.LBM_InitializeFlow(g, U, 0, 0) == 1
        gpu.LBM_StepN(g, 2500, U, 0, 0)
        vel = gpu.LBM_CopyVelocityToHost(g).reshape(64, 64, 128, 4)
        samples = []
        for _ in range(5):
            gpu.LBM_StepN(g, 100, U, 0, 0)
            fx, fy, fz = gpu.LBM_ComputeDragForce(g)
            samples.append(abs(fx) / (0.5 * U * U * A) if A > 0 else 0.0)
        res[storage] = (np.array(samples), A, vel[32, 32, 0], vel[32, 32, 64])
    s, A, inlet, mid = res[0]
    s64 = res[gpu.LBM_STORAGE_FP64][0]
    print("box: area %.0f, inlet vel %s, mid vel %s" % (A, inlet, mid))
    print("Cd samples FP32 %s, FP64 %s, max rel diff %.2e" % (np.round(s, 4), np.round(s64, 4),
                                                              np.max(np.abs(s - s64) / s64)))
    assert A > 0
    assert np.all(s > 0)
    assert s.max() / max(s.min(), 0.001) < 10
    assert inlet[0] == pytest.approx(U, rel=1e-3)  # the Zou-He inlet holds u


def test_ground_plane(gpu):  # test_lbm.c:543-592
    """32x16x16, the small box, ground at world z = -0.15: lattice layer
    (int)((-0.15 + 2) 16 / 4) = 7, so layers 0-7 are ground except the body
    cells the box puts in layer 7; then 20 steps with a nonzero drag."""
    out = {}
    for storage in (0, gpu.LBM_STORAGE_FP64):
        g = grid(gpu, (32, 16, 16), 0.1, storage)
        gpu.LBM_SetSolidAABB(g, *BOX)
        gpu.LBM_AddGroundPlane(g, -0.15)
        s = gpu.LBM_CopySolidToHost(g).reshape(16, 16, 32)
        assert gpu.LBM_InitializeFlow(g, 0.05, 0, 0) == 1
        for _ in range(20):
            gpu.LBM_Step(g, 0.05, 0, 0)
        out[storage] = (s, gpu.LBM_ComputeDragForce(g), gpu.LBM_ComputeDragForceF64(g))
    s, (fx, fy, fz), F = out[0]
    F64 = out[gpu.LBM_STORAGE_FP64][2]
    body, ground, fluid = int(np.sum(s == 1)), int(np.sum(s == 2)), int(np.sum(s == 0))
    print("ground plane: body %d ground %d fluid %d; fx %.6f, FP64-sum Fx %.8f, FP64 storage %.8f"
          % (body, ground, fluid, fx, F[0], F64[0]))
    assert body > 0 and ground > 0 and fluid > 0
    assert ground == 8 * 32 * 16 - int(np.sum(s[:8] == 1))
    assert np.all(s[:8] != 0) and not np.any(s[8:] == 2)
    assert abs(fx) > 1e-10
    assert abs(F[0] - F64[0]) <= 1e-5 * abs(F64[0])
