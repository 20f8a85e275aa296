"""Population storage formats on the configuration they are meant for: MRT +
Smagorinsky (Cs 0.1), tau 0.52, inlet u 0.06, periodic y/z, the Ahmed-like
body (rounded front, 25-degree rear slant, ground clearance) with Bouzidi
links from the mesh.

* At 128x64x64 (the body at half its 64x24x18-cell size, so the CPU oracle
  finishes in seconds), after 300 steps: FP32 storage against the FP32
  oracle to the parity bar, FP64 storage against the FP64 oracle, and every
  format's error against the FP64 oracle stated and bounded.
* At the full 256x128x128 for the configuration's 20,000 steps, every
  format against the GPU's FP64 storage (which the test above and
  test_fp64_algorithm_exact_all_boundaries tie to the FP64 oracle): the Cd
  averaged over the last 4000 steps, finite fields, the total mass and the
  guard-reset count.

Bars for the 16-bit formats.  FP16S keeps 11 significant bits of 2^15 (f -
w_i), BF16S 8 bits of f - w_i; with |f - w_i| up to ~0.02 in these flows one
store rounds by up to ~5e-6 (FP16S) and ~4e-5 (BF16S) per population.  The
rounding enters every step and is damped by the collision, so the field
error settles near the per-store rounding over the flow's population
deviations: about 1e-3 of u (FP16S) and 1e-2 (BF16S) relative L2.  The
bars below are about twice the values measured on MI355X (printed by the
tests); BF16S's Cd is 9 % off after 20,000 steps, so it is a storage format
for throughput runs, not for drag figures.

The velocity error of the 16-bit formats is bounded away from the outlet
(x < nx - 8).  The Zou-He pressure outlet amplifies the storage rounding in
its last few planes by an amount that varies from run to run: with inlet
speeds 0.0599, 0.06, 0.0601 and 0.0602 the FP16S error over the whole
lattice after 300 steps scattered over 3.0e-4 - 1.4e-3, while the error
over x < nx - 8 stayed at 2.5e-4 - 3.3e-4 (BF16S 2.9e-3 - 4.0e-3).  The
whole-lattice error is printed and held to a looser bar."""
import numpy as np
import pytest

from lattice_fluid_dynamics_amd import geometry
from tests.helpers import compare, fields, make_oracle, make_pair, rel_l2

pytestmark = pytest.mark.gpu

TAU, U, CS = 0.52, 0.06, 0.1
FLAGS = dict(useMRT=1, useSmagorinsky=1, smagorinskyCs=CS, periodicYZ=1)
NAMES = {0: "fp32", 1: "fp16s", 2: "bf16s", 3: "fp64"}


def half_ahmed(nx, ny, nz):
    return geometry.ahmed_body(nx, ny, nz, cells=(32, 12, 9), front_radius_cells=3.0, clearance_cells=1.0)


def setup_mesh(tris, bbox):
    def f(g, og):
        if g is not None:
            from lattice_fluid_dynamics_amd import lbm
            lbm.LBM_SetSolidMesh(g, tris, bbox)
        og.set_solid_mesh(tris, bbox)
    return f


# stated per-format bars against the FP64 oracle after 300 steps at 128x64x64
# (measured on MI355X: FP32 u 1.8e-6 (x < nx - 8: 9.1e-7), rho 6.7e-8, Cd 8.1e-6;
# FP16S u 9.8e-4 (2.8e-4), rho 1.4e-5, Cd 3.2e-3; BF16S u 3.1e-3 (2.9e-3),
# rho 1.9e-4, Cd 1.2e-2)
BARS_300 = {0: {"u": 1e-5, "u_inner": 1e-5, "rho": 1e-6, "cd": 1e-4},
            1: {"u": 3e-3, "u_inner": 6e-4, "rho": 3e-5, "cd": 6e-3},
            2: {"u": 1e-2, "u_inner": 8e-3, "rho": 4e-4, "cd": 3.5e-2}}


def test_storage_formats_128x64x64_against_the_oracles(gpu):
    nx, ny, nz = 128, 64, 64
    tris, bbox = half_ahmed(nx, ny, nz)
    setup = setup_mesh(tris, bbox)
    steps = 300
    g32, o32 = make_pair(nx, ny, nz, (TAU - 0.5) / 3, setup, tau=TAU, u=U, **FLAGS)
    g64, o64 = make_pair(nx, ny, nz, (TAU - 0.5) / 3, setup, tau=TAU, u=U, precision=64,
                         storage=gpu.LBM_STORAGE_FP64, **FLAGS)
    A = gpu.LBM_ComputeProjectedArea(g32, 0)
    assert A == o32.projected_area(0) and A > 50
    grids = {0: g32, 3: g64}
    for st in (1, 2):
        g = gpu.LBM_Create(nx, ny, nz, (TAU - 0.5) / 3)
        g.tau = TAU
        for k, v in FLAGS.items():
            setattr(g, k, v)
        g.populationStorage = st
        gpu.LBM_SetSolidMesh(g, tris, bbox)
        assert gpu.LBM_InitializeFlow(g, U, 0, 0) == 1
        grids[st] = g
    for g in grids.values():
        gpu.LBM_StepN(g, steps, U, 0, 0)
    o32.step(U, 0, 0, n=steps)
    o64.step(U, 0, 0, n=steps)
    dyn = 0.5 * U * U * A

    # FP32 storage against the FP32 oracle: the parity bar holds for the
    # populations and density.  The velocity u = j / rho, with j a sum of
    # +-f over ten directions, keeps fewer significant bits: here the FP32
    # oracle's own u is 1.6-2e-5 (rel L2) from the FP64 oracle, and two FP32
    # evaluations in different operation order (the kernel's MRT applies
    # M^-1 S (m - meq) as a correction, and contracts multiply-adds) differ by
    # as much.  For u the bar is therefore that the GPU is no further from the
    # FP64 oracle than the FP32 oracle is, and within 3e-5 of it.
    d = compare(g32, o32)
    print("FP32 gpu vs FP32 oracle:", d)
    assert d["rho"] < 1e-5 and d["f"] < 1e-5 and d["f_new"] < 1e-5
    u_gpu64 = rel_l2(gpu.LBM_CopyVelocityToHost(g32)[:, :3], o64.vel[:, :3])
    u_or64 = rel_l2(o32.vel[:, :3], o64.vel[:, :3])
    print("u rel L2 vs FP64 oracle: gpu %.2e, FP32 oracle %.2e" % (u_gpu64, u_or64))
    assert d["u"] < 3e-5 and u_gpu64 <= u_or64
    ex32, fix32 = o32.force()
    fx, fy, fz, px, py, pz = gpu.LBM_ComputeDragForceDecomposed(g32)
    assert abs(fx - fix32[0] / 1e4) <= 1e-3 * abs(fix32[0] / 1e4)
    assert abs(px - fix32[4] / 1e4) <= 1e-3 * abs(fix32[4] / 1e4)
    # FP64 storage against the FP64 oracle: the algorithm, exact
    f64, fn64 = gpu.LBM_CopyPopulationsToHost(g64, f64=True)
    e64 = (rel_l2(f64, o64.f), rel_l2(fn64, o64.f_new))
    print("FP64 gpu vs FP64 oracle: f %.2e f_new %.2e" % e64)
    assert e64[0] < 1e-10 and e64[1] < 1e-10
    ex64, _ = o64.force()
    cd_ref = abs(ex64[0]) / dyn
    cl_ref = ex64[1] / dyn
    cdp_ref = abs(ex64[3]) / dyn
    # every format against the FP64 oracle
    ref_u, ref_rho = o64.vel[:, :3], o64.vel[:, 3]
    print("oracle FP32 vs FP64: u %.2e rho %.2e Cd %.2e" % (
        rel_l2(o32.vel[:, :3], ref_u), rel_l2(o32.vel[:, 3], ref_rho), abs(abs(ex32[0]) / dyn - cd_ref) / cd_ref))
    for st in (0, 1, 2):
        vel = gpu.LBM_CopyVelocityToHost(grids[st])
        F = gpu.LBM_ComputeDragForceF64(grids[st])
        inner = (np.arange(vel.shape[0]) % nx) < nx - 8
        err = {"u": rel_l2(vel[:, :3], ref_u), "u_inner": rel_l2(vel[inner, :3], ref_u[inner]),
               "rho": rel_l2(vel[:, 3], ref_rho),
               "cd": abs(abs(F[0]) / dyn - cd_ref) / cd_ref, "cl_abs": abs(F[1] / dyn - cl_ref),
               "cd_pressure": abs(abs(F[3]) / dyn - cdp_ref) / cdp_ref}
        print("%s vs FP64 oracle after %d steps: %s (Cd %.5f, FP64 %.5f; Cl FP64 %.5f)" % (
            NAMES[st], steps, {k: "%.2e" % v for k, v in err.items()}, abs(F[0]) / dyn, cd_ref, cl_ref))
        for k, bar in BARS_300[st].items():
            assert err[k] <= bar, (NAMES[st], k, err[k], bar)


# stated bars on the time-averaged Cd against FP64 storage over the full run
# (measured on MI355X: FP32 2.5e-6, FP16S 7.7e-3, BF16S 8.9e-2); FP32 is held
# to the 0.1 % parity bar
BARS_20K = {0: 1e-3, 1: 1.5e-2, 2: 1.8e-1}


def test_storage_fp16_256x128x128_20k_steps(gpu):
    """256x128x128 with the 64x24x18-cell Ahmed-like body (~20k triangles),
    20,000 steps in every storage format.  The flow sheds vortices (Re ~ 580
    on the body length), so instantaneous fields of different formats part
    ways; the comparison is on the Cd averaged over the last 4000 steps
    (samples every 200 steps), with Cl and the pressure/friction split
    printed.  Fields stay finite, the total mass stays within 1 % of its
    start (open inlet and outlet: it is not conserved exactly) and the
    stability guard never fires."""
    nx, ny, nz = 256, 128, 128
    tris, bbox = geometry.ahmed_body(nx, ny, nz)
    assert 15000 <= len(tris) <= 25000
    steps, interval, tail = 20000, 200, 20
    res = {}
    for st in (3, 0, 1, 2):
        g = gpu.LBM_Create(nx, ny, nz, (TAU - 0.5) / 3)
        g.tau = TAU
        for k, v in FLAGS.items():
            setattr(g, k, v)
        g.populationStorage = st
        gpu.LBM_SetSolidMesh(g, tris, bbox)
        A = gpu.LBM_ComputeProjectedArea(g, 0)
        assert gpu.LBM_InitializeFlow(g, U, 0, 0) == 1
        m0 = gpu.LBM_ComputeTotalMass(g)
        dyn = 0.5 * U * U * A
        cd, cl, cdp = [], [], []
        for _ in range(steps // interval):
            gpu.LBM_StepN(g, interval, U, 0, 0)
            F = gpu.LBM_ComputeDragForceF64(g)
            cd.append(abs(F[0]) / dyn)
            cl.append(F[1] / dyn)
            cdp.append(abs(F[3]) / dyn)
        vel = gpu.LBM_CopyVelocityToHost(g)
        res[st] = dict(cd=float(np.mean(cd[-tail:])), cl=float(np.mean(cl[-tail:])),
                       cdp=float(np.mean(cdp[-tail:])), finite=bool(np.all(np.isfinite(vel))),
                       mass=gpu.LBM_ComputeTotalMass(g) / m0, resets=gpu.LBM_GetGuardResetCount(g), area=A)
        gpu.LBM_Free(g)
    ref = res[3]["cd"]
    for st in (3, 0, 1, 2):
        r = res[st]
        r["cd_err"] = abs(r["cd"] - ref) / ref
        print("%s: Cd %.5f (rel err vs FP64 %.2e) Cl %.5f Cd_pressure %.5f Cd_friction %.5f mass ratio %.6f "
              "guard resets %d" % (NAMES[st], r["cd"], r["cd_err"], r["cl"], r["cdp"], max(r["cd"] - r["cdp"], 0),
                                    r["mass"], r["resets"]))
        assert r["finite"] and abs(r["mass"] - 1) < 1e-2 and r["resets"] == 0
    for st, bar in BARS_20K.items():
        assert res[st]["cd_err"] <= bar, (NAMES[st], res[st]["cd_err"], bar)
