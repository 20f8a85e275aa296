"""Measured workloads beyond the roofline bench, each printing JSON lines.

  python -m lattice_fluid_dynamics_amd.experiments storage [--steps 20000]
      Storage formats: 256x128x128 MRT + Smagorinsky (Cs 0.1), tau 0.52, inlet 0.06,
      Ahmed-like body (64x24x18 cells, front radius 6, 25 deg slant, 2-cell
      clearance, ~20k triangles, Bouzidi).  Cold starts with FP64, FP32, FP16
      and BF16 population storage; Cd/Cl with the pressure/friction split, and
      steps/seconds until the windowed Cd settles within 1 % of the FP64 value.

  python -m lattice_fluid_dynamics_amd.experiments warmstart [--train 384 --eval 32 --iters 3000]
      The learned warm start (see warmstart_experiment).

  python -m lattice_fluid_dynamics_amd.experiments sweep [--tunnels 64 --steps 5000]
      Batched sweep: 64 tunnels with nx ~ U[64,192], ny = nz = nx/2, a seeded random box,
      ellipsoid or slanted-rear body 10-20 % of ny, u ~ U[0.02,0.08],
      tau ~ U[0.51,0.8]; MRT + Bouzidi + FP16 storage in one batch (one launch per
      step, graph replay, autotuned workgroup); per-tunnel Cd/Cl.

  python -m lattice_fluid_dynamics_amd.experiments particles
      The particle and streamline kernels on the 512x256x256 car-proxy flow
      (MRT + Smagorinsky, FP32, 200 steps): 50,000 particles through the RK2
      trilinear update and the live update in each collision mode, and the
      256 x 64 RK4 streamlines; host-timed microseconds per call.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import convergence, geometry, lbm


def _emit(obj):
    print(json.dumps(obj), flush=True)


def storage_grid(storage: int):
    nx, ny, nz = 256, 128, 128
    g = lbm.LBM_Create(nx, ny, nz, (0.52 - 0.5) / 3)
    g.tau = 0.52
    g.useMRT = 1
    g.useSmagorinsky = 1
    g.smagorinskyCs = 0.1
    g.periodicYZ = 1
    g.populationStorage = storage
    tris, bbox = geometry.ahmed_body(nx, ny, nz)
    lbm.LBM_SetSolidMesh(g, tris, bbox)
    area = lbm.LBM_ComputeProjectedArea(g, 0)
    assert lbm.LBM_InitializeFlow(g, 0.06, 0, 0) == 1
    lbm.LBM_Autotune(g, 20, 0.06)  # workgroup and streaming mode; the state is restored
    return g, area, len(tris)


def storage_experiment(steps: int, interval: int = 200):
    names = {3: "fp64", 0: "fp32", 1: "fp16", 2: "bf16"}
    hist = {}
    for st in (3, 0, 1, 2):
        g, area, ntri = storage_grid(st)
        h = convergence.run_history(g, 0.06, area, steps, interval)
        hist[st] = h
        g.free()
    ref = hist[3].converged_value()
    for st, h in hist.items():
        t = convergence.time_to_within(h, ref, 0.01, window=10)
        tail = slice(-20, None)
        cd = float(np.mean(h.cd[tail]))
        cp = float(np.mean(h.cd_pressure[tail]))
        _emit({"experiment": "storage-256x128x128", "storage": names[st], "steps": steps, "triangles": ntri,
               "area": area, "cd": cd, "cl": float(np.mean(h.cl[tail])), "cd_pressure": cp,
               "cd_friction": max(cd - cp, 0.0), "cd_rel_err_vs_fp64": abs(cd - ref) / ref,
               "steps_to_1pct_of_fp64": t, "seconds_to_1pct": None if t is None else t * h.ms_per_step / 1e3,
               "ms_per_step": h.ms_per_step, "mlups": 256 * 128 * 128 / h.ms_per_step / 1e3,
               "cd_history_every_1000": [round(v, 5) for v in h.cd[4::5]]})


def sweep_bodies(seed: int, nx: int, ny: int, nz: int):
    """Seeded random box, ellipsoid or slanted-rear body, 10-20 % of ny tall
    (geometry.random_body; the body is given in world units)."""
    return geometry.random_body(seed)


def sweep_tunnels(ntun: int = 64, rng_seed: int = 12345):
    """The tunnels of the batched sweep: nx uniform in [64, 192], ny = nz =
    nx / 2, inlet u uniform in [0.02, 0.08], tau uniform in [0.51, 0.8], body
    seed = tunnel index.  A list of dicts (seed, dims, u, tau)."""
    rng = np.random.default_rng(rng_seed)
    out = []
    for s in range(ntun):
        nx = int(rng.integers(64, 193))
        u = float(rng.uniform(0.02, 0.08))
        tau = float(rng.uniform(0.51, 0.8))
        out.append({"seed": s, "dims": [nx, nx // 2, nx // 2], "u": u, "tau": tau})
    return out


def sweep_geometry(t):
    """A grid holding tunnel t's body (solid mask and Bouzidi links), for
    LBM_BatchAddTunnel or a single run; returns (grid, body kind)."""
    nx, ny, nz = t["dims"]
    g = lbm.LBM_Create(nx, ny, nz, (t["tau"] - 0.5) / 3)
    tris, bbox, kind = sweep_bodies(t["seed"], nx, ny, nz)
    lbm.LBM_SetSolidMesh(g, tris, bbox)
    return g, kind


def sweep_experiment(ntun: int, steps: int, storage: int = 1, tune_steps: int = 20):
    b = lbm.Batch(storage=storage, useMRT=1, useSmagorinsky=0, periodicYZ=1)
    meta = []
    t0 = time.time()
    for t in sweep_tunnels(ntun):
        g, kind = sweep_geometry(t)
        b.add(g, t["tau"], t["u"])
        meta.append({**t, "body": kind})
        g.free()
    b.initialize()
    setup_s = time.time() - t0
    cells = b.total_cells
    best, ms = b.autotune(tune_steps)
    ms_nograph = b.profile(20, use_graph=False) / 20
    ms_graph = b.profile(20, use_graph=True) / 20
    t = b.profile(steps, use_graph=True)
    coef = b.coefficients()
    _emit({"experiment": "sweep-summary", "tunnels": ntun, "total_cells": cells, "steps": steps,
           "storage": {0: "fp32", 1: "fp16", 2: "bf16"}[storage], "setup_s": round(setup_s, 1),
           "autotune_block": best, "autotune_streaming_mode": b.kernel_config()[1],
           "autotune_xcd_contiguous": b.kernel_config()[2],
           "autotune_ms_per_step": [round(float(v), 4) for v in ms],
           "ms_per_step_eager": round(ms_nograph, 4), "ms_per_step_graph": round(ms_graph, 4),
           "total_s": round(t / 1e3, 3), "aggregate_mlups": round(cells * steps / (t / 1e3) / 1e6, 1)})
    for k, m in enumerate(meta):
        m.update({"cd": float(coef[k, 0]), "cl": float(coef[k, 1]), "cd_pressure": float(coef[k, 2]),
                  "cd_friction": float(coef[k, 3])})
        _emit({"experiment": "sweep-tunnel", **m})


def _steps_to_1pct(h, ref, window):
    t = convergence.time_to_within(h, ref, 0.01, window=window)
    return None if t is None else int(t)


def learned_gain(dims=(128, 64, 64), factor: int = 4, n_train: int = 64, data_steps: int = 4000,
                 avg_steps: int = 1000, iters: int = 600, batch: int = 16, lr: float = 1e-3, n_eval: int = 8,
                 ref_steps: int = 8000, cold_steps: int = 8000, warm_steps: int = 4000, interval: int = 100,
                 window: int = 5, tail: int = 20, chunk: int = 256, log=None):
    """What the trained warm start gains over the cold start, end to end on
    the GPU: the FP32 solver generates the training fields (cold starts of
    the training tunnels, block-averaged to the coarse lattice and averaged
    over `avg_steps`), the network is trained on them with AdamW by the
    project's kernels, and on `n_eval` held-out tunnels (seeds 100000..) the
    time to a converged Cd is measured for the cold FP32 start and for the
    warm start with FP16 storage.

    Converged Cd: the tunnel's FP64-storage cold start (the GPU path that
    equals the FP64 oracle to 1e-12) averaged over its last `tail` samples of
    `ref_steps` steps.  Time to it: the first sample step after which the
    mean of the last `window` Cd samples (one every `interval` steps) stays
    within 1 % of it; seconds = steps times the batch's measured time per
    step divided by the tunnel count, plus the warm start's inference and
    initialisation time per tunnel.  Returns a dict with per-tunnel rows and
    the median speed-up (cold FP32 seconds over warm FP16 seconds)."""
    from . import warmstart as WS

    nx, ny, nz = dims
    t0 = time.time()
    specs = WS.random_specs(range(n_train))
    masks, fields = WS.solve_fields(specs, dims, factor, data_steps, avg_steps, interval, storage=0,
                                    chunk=min(n_train, chunk), log=log)
    gen_s = time.time() - t0
    ws = WS.WarmStart(nx // factor, ny // factor, nz // factor, max_batch=batch, seed=0)
    data = WS.DeviceDataset(masks, fields, specs, augment=True)
    t0 = time.time()
    losses = WS.train(ws, data, iters, batch=batch, lr=lr, wd=1e-4)
    train_s = time.time() - t0
    del data
    if log:
        log(f"learned gain: {n_train} training tunnels solved in {gen_s:.1f}s, {iters} AdamW steps in "
            f"{train_s:.1f}s (loss {np.mean(losses[:20]):.4f} -> {np.mean(losses[-20:]):.5f})")
    ev = WS.random_specs(range(100000, 100000 + n_eval))
    ref = WS.cd_histories(ev, dims, ref_steps, 3, None, factor, interval, log=log)
    cold = WS.cd_histories(ev, dims, cold_steps, 0, None, factor, interval, log=log)
    warm = WS.cd_histories(ev, dims, warm_steps, 1, ws, factor, interval, log=log)
    ws.free()
    rows = []
    for t, sp in enumerate(ev):
        cd_ref = ref["histories"][t].converged_value(tail=tail)
        row = {"seed": sp.seed, "u": round(sp.u, 5), "tau": round(sp.tau, 4), "cd_fp64": round(cd_ref, 5)}
        for name, r in (("cold_fp32", cold), ("warm_fp16", warm)):
            st = _steps_to_1pct(r["histories"][t], cd_ref, window)
            row[name + "_steps"] = st
            row[name + "_s"] = None if st is None else st * r["ms_per_step"] / n_eval / 1e3 + r["init_s"]
        row["speedup"] = (row["cold_fp32_s"] / row["warm_fp16_s"]
                          if row["cold_fp32_s"] is not None and row["warm_fp16_s"] is not None else None)
        rows.append(row)
    sp = [r["speedup"] for r in rows if r["speedup"] is not None]
    return {"rows": rows, "median_speedup": float(np.median(sp)) if sp else None,
            "min_speedup": float(np.min(sp)) if sp else None,
            "geomean_speedup": float(np.exp(np.mean(np.log(sp)))) if sp else None,
            "converged": {k: sum(r[k + "_steps"] is not None for r in rows) for k in ("cold_fp32", "warm_fp16")},
            "median_steps": {k: float(np.median([r[k + "_steps"] for r in rows if r[k + "_steps"] is not None]))
                             if any(r[k + "_steps"] is not None for r in rows) else None
                             for k in ("cold_fp32", "warm_fp16")},
            "ms_per_step_per_tunnel": {"cold_fp32": cold["ms_per_step"] / n_eval,
                                       "warm_fp16": warm["ms_per_step"] / n_eval,
                                       "ref_fp64": ref["ms_per_step"] / n_eval},
            "warm_init_s_per_tunnel": warm["init_s"], "data_gen_s": gen_s, "train_s": train_s,
            "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-20:])),
            "params": {"dims": list(dims), "factor": factor, "n_train": n_train, "data_steps": data_steps,
                       "avg_steps": avg_steps, "iters": iters, "batch": batch, "n_eval": n_eval,
                       "ref_steps": ref_steps, "cold_steps": cold_steps, "warm_steps": warm_steps,
                       "interval": interval, "window": window, "tail": tail}}


def gain_experiment(n_train: int, n_eval: int, iters: int, data_steps: int = 4000, steps: int = 8000,
                    warm_steps: int = 4000, dims=(128, 64, 64)):
    """learned_gain at the learned-start configuration's counts (e.g. 2000
    training and 200 held-out tunnels) at half resolution: one JSON line per
    held-out tunnel and a summary line."""
    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.time()
    r = learned_gain(dims=dims, n_train=n_train, n_eval=n_eval, iters=iters, data_steps=data_steps,
                     ref_steps=steps, cold_steps=steps, warm_steps=warm_steps, log=log)
    for row in r["rows"]:
        _emit({"experiment": "gain-tunnel", **row})
    sp = [row["speedup"] for row in r["rows"] if row["speedup"] is not None]
    _emit({"experiment": "gain-summary", "train_tunnels": n_train, "eval_tunnels": n_eval,
           "lattice": "%dx%dx%d" % tuple(dims), "converged": r["converged"], "median_steps": r["median_steps"],
           "speedup_median": r["median_speedup"], "speedup_geomean": r["geomean_speedup"],
           "speedup_min": r["min_speedup"],
           "speedup_p10_p90": [float(np.percentile(sp, 10)), float(np.percentile(sp, 90))] if sp else None,
           "share_at_least_3x": float(np.mean(np.array(sp) >= 3.0)) if sp else None,
           "ms_per_step_per_tunnel": r["ms_per_step_per_tunnel"], "warm_init_s_per_tunnel": r["warm_init_s_per_tunnel"],
           "data_gen_s": round(r["data_gen_s"], 1), "train_s": round(r["train_s"], 1),
           "loss_first": r["loss_first"], "loss_last": r["loss_last"], "params": r["params"],
           "wall_s": round(time.time() - t0, 1)})


def warmstart_experiment(n_train: int, n_eval: int, iters: int, data_steps: int = 6000, avg_steps: int = 2000,
                         cold_steps: int = 20000, warm_steps: int = 12000, batch: int = 16, lr: float = 1e-3,
                         weights_in=None, weights_out=None, chunk: int = 96, window: int = 20,
                         dims=(256, 128, 128), factor: int = 4, phase: str = "all", warm_fp32: bool = True,
                         reference: str = "fp32", eval_offset: int = 0, bootstrap: int = 0, boot_steps: int = 1200,
                         boot_avg: int = 1200, boot_iters: int = 1500):
    """The learned warm start at 256x128x128 (coarse lattice 64x32x32).

    Training tunnels use body seeds 0..n_train-1, held-out tunnels seeds
    100000..; inlet speed U[0.03, 0.07] and tau U[0.6, 0.9] per seed (MRT,
    periodic y/z, steady laminar wakes, Re ~ 3-50).  Targets: the FP32
    cold-start solution after `data_steps`, block-averaged to the coarse
    lattice and averaged over the next `avg_steps`.  Evaluation: Cd sampled
    every 100 steps; time to converged Cd = first step after which the mean
    of the last `window` samples stays within 1 % of the held-out tunnel's
    converged Cd: the mean of the last 6000 steps of its FP32 cold start, or
    with reference="fp64" of its cold start in FP64 storage (the GPU path
    that equals the FP64 oracle).  eval_offset shifts the held-out seeds, so
    that one evaluation can be split over several runs.

    bootstrap = N > 0 generates the training fields in two stages, so that
    many more tunnels fit one run: the first N tunnels from cold starts as
    above, a network trained on them (boot_iters AdamW steps), then every
    other training tunnel started from that network's prediction and solved
    in FP32 for boot_steps steps, its field averaged over the next boot_avg.
    The final network is trained from scratch on all the fields."""
    from . import warmstart as WS

    def log(m):
        print(m, file=sys.stderr, flush=True)

    nx, ny, nz = dims
    ws = WS.WarmStart(nx // factor, ny // factor, nz // factor, max_batch=batch, seed=0)
    if weights_in:
        ws.load(weights_in)
    if phase in ("all", "train") and not weights_in:
        specs = WS.random_specs(range(n_train))
        t0 = time.time()
        n_cold = bootstrap if 0 < bootstrap < n_train else n_train
        masks, fields = WS.solve_fields(specs[:n_cold], dims, factor, data_steps, avg_steps, 100, storage=0,
                                        chunk=chunk, log=log)
        if n_cold < n_train:
            boot = WS.WarmStart(nx // factor, ny // factor, nz // factor, max_batch=batch, seed=1)
            bdata = WS.DeviceDataset(masks, fields, specs[:n_cold], augment=True)
            WS.train(boot, bdata, boot_iters, batch=batch, lr=lr, wd=1e-4, log=log)
            del bdata
            m2, f2 = WS.solve_fields(specs[n_cold:], dims, factor, boot_steps, boot_avg, 100, storage=0,
                                     chunk=chunk, log=log, ws=boot)
            boot.free()
            masks, fields = np.concatenate([masks, m2]), np.concatenate([fields, f2])
        gen_s = time.time() - t0
        data = WS.DeviceDataset(masks, fields, specs, augment=True)
        t0 = time.time()
        losses = WS.train(ws, data, iters, batch=batch, lr=lr, wd=1e-4, log=log)
        train_s = time.time() - t0
        if weights_out:
            ws.save(weights_out)
        _emit({"experiment": "warmstart-train", "train_tunnels": n_train, "samples_with_mirrors": data.n,
               "cold_data_tunnels": n_cold, "bootstrap": {"steps": boot_steps, "avg_steps": boot_avg,
                                                          "iters": boot_iters} if n_cold < n_train else None,
               "lattice": f"{nx}x{ny}x{nz}", "coarse": f"{nx // factor}x{ny // factor}x{nz // factor}",
               "data_steps": data_steps, "avg_steps": avg_steps, "data_gen_s": round(gen_s, 1),
               "iters": iters, "batch": batch, "lr": lr, "train_s": round(train_s, 1),
               "ms_per_iter": round(train_s / max(iters, 1) * 1e3, 2), "params": ws.param_count,
               "loss_first": round(float(losses[0]), 5),
               "loss_last100": round(float(np.mean(losses[-100:])), 5),
               "loss_every_100": [round(float(np.mean(losses[k:k + 100])), 5) for k in range(0, len(losses), 100)]})
        del data
    if phase == "train":
        return
    specs = WS.random_specs(range(100000 + eval_offset, 100000 + eval_offset + n_eval))
    runs = {}
    ref_run = None
    if reference == "fp64":
        ref_run = WS.cd_histories(specs, dims, cold_steps, 3, None, factor, log=log)
    runs["cold_fp32"] = WS.cd_histories(specs, dims, cold_steps, 0, None, factor, avg_from=cold_steps - 6000, log=log)
    runs["cold_fp16"] = WS.cd_histories(specs, dims, warm_steps, 1, None, factor, log=log)
    runs["warm_fp16"] = WS.cd_histories(specs, dims, warm_steps, 1, ws, factor, log=log)
    if warm_fp32:
        runs["warm_fp32"] = WS.cd_histories(specs, dims, warm_steps, 0, ws, factor, log=log)
    conv = runs["cold_fp32"]["mean_field"]
    rows = []
    for t, sp in enumerate(specs):
        ref = (ref_run or runs["cold_fp32"])["histories"][t].converged_value(tail=60)
        row = {"seed": sp.seed, "u": round(sp.u, 5), "tau": round(sp.tau, 4), "cd_ref": round(ref, 5),
               "reference": reference}
        for name, r in runs.items():
            st = _steps_to_1pct(r["histories"][t], ref, window)
            per_tunnel_ms = r["ms_per_step"] / len(specs)
            row[name + "_steps"] = st
            row[name + "_s"] = None if st is None else round(st * per_tunnel_ms / 1e3 + r["init_s"], 4)
        fluid = runs["cold_fp32"]["masks"][t] < 0.5
        pred = runs["warm_fp16"]["predicted"][t]
        du = np.linalg.norm((pred[..., :3] - conv[t, ..., :3])[fluid]) / np.linalg.norm(conv[t, ..., :3][fluid])
        uni = np.zeros_like(conv[t, ..., :3])
        uni[..., 0] = sp.u
        du0 = np.linalg.norm((uni - conv[t, ..., :3])[fluid]) / np.linalg.norm(conv[t, ..., :3][fluid])
        for name in ("cold_fp32", "warm_fp16"):
            row[name + "_cd_first_30"] = [round(v, 4) for v in runs[name]["histories"][t].cd[:30]]
        row["pred_u_rel_l2"] = round(float(du), 4)
        row["uniform_u_rel_l2"] = round(float(du0), 4)
        rows.append(row)
        _emit({"experiment": "warmstart-tunnel", **row})

    def med(key):
        v = [r[key] for r in rows if r[key] is not None]
        return None if not v else float(np.median(v))

    ratios = [r["cold_fp32_s"] / r["warm_fp16_s"] for r in rows
              if r["cold_fp32_s"] is not None and r["warm_fp16_s"] is not None]
    _emit({"experiment": "warmstart-summary", "eval_tunnels": n_eval, "eval_seeds_from": 100000 + eval_offset,
           "reference": reference, "window_samples": window,
           "sample_interval": 100,
           **{k + "_ms_per_step_per_tunnel": round(r["ms_per_step"] / n_eval, 5) for k, r in runs.items()},
           **{k + "_converged": sum(r2[k + "_steps"] is not None for r2 in rows) for k in runs},
           **{k + "_median_steps": med(k + "_steps") for k in runs},
           **{k + "_median_s": med(k + "_s") for k in runs},
           "warm_init_s_per_tunnel": round(runs["warm_fp16"]["init_s"], 4),
           "speedup_cold_fp32_over_warm_fp16_median": None if not ratios else round(float(np.median(ratios)), 3),
           "speedup_geomean": None if not ratios else round(float(np.exp(np.mean(np.log(ratios)))), 3),
           "pred_u_rel_l2_median": med("pred_u_rel_l2"), "uniform_u_rel_l2_median": med("uniform_u_rel_l2")})


def particles_experiment(n: int = 50000, updates: int = 40):
    """Host-timed calls (after a synchronisation, up to the next one) of the
    particle and streamline kernels on a developed 512x256x256 flow."""
    from . import model as M

    g = lbm.LBM_Create(512, 256, 256, 0.005 / 3)
    g.tau, g.useMRT, g.useSmagorinsky, g.smagorinskyCs, g.periodicYZ = 0.505, 1, 1, 0.1, 1
    tris, bbox = geometry.car_proxy(seed=0)
    lbm.LBM_SetSolidMesh(g, tris, bbox)
    assert lbm.LBM_InitializeFlow(g, 0.05, 0.0, 0.0) == 1
    lbm.LBM_StepN(g, 200, 0.05, 0.0, 0.0)
    lbm.LBM_Synchronize(g)
    lo, hi = tuple(bbox[:3]), tuple(bbox[3:])
    center = tuple((a + b) / 2 for a, b in zip(lo, hi))
    grid = M.collision_grid(tris, bbox)
    keep = [lbm.DeviceBuffer(tris), lbm.DeviceBuffer(grid["cellStart"].astype(np.int32)),
            lbm.DeviceBuffer(grid["cellCount"].astype(np.int32)), lbm.DeviceBuffer(grid["triIndices"].astype(np.int32))]
    dgrid = {"cellStart_dev": keep[1].ptr, "cellCount_dev": keep[2].ptr,
             "triIndices_dev": (keep[3].ptr, grid["triIndices"].size), "origin": grid["origin"], "cell": grid["cell"]}
    rng = np.random.default_rng(1)
    p = np.zeros((n, 8), np.float32)
    p[:, 0], p[:, 1], p[:, 2] = rng.uniform(-4, 4, n), rng.uniform(-1.3, 1.3, n), rng.uniform(-1.3, 1.3, n)
    p[:, 7] = 1.0
    out = {"experiment": "particles", "lattice": "512x256x256", "particles": n, "updates": updates,
           "triangles": int(tris.shape[0])}

    def timed(fn, calls):
        fn()
        lbm.LBM_Synchronize(g)
        t0 = time.perf_counter()
        for k in range(calls):
            fn(k)
        lbm.LBM_Synchronize(g)
        return round((time.perf_counter() - t0) / calls * 1e6, 1)

    parts = lbm.DeviceParticles(n)
    parts.upload(p)
    out["rk2_trilinear_us"] = timed(lambda k=0: lbm.LBM_AdvectParticles(g, parts, 10.0, 8.0 / 512, 0, k), updates)
    for mode in range(4):
        flow = lbm.particle_flow((0.4, 0.0, 0.0), lo, hi, center, collision_mode=mode, use_lbm=1,
                                 triangles=(keep[0].ptr, tris.shape[0]), grid=dgrid)
        parts.upload(p)
        out[f"live_mode{mode}_us"] = timed(lambda k=0: lbm.LBM_AdvectParticlesLive(g, parts, 0.016, flow, 3, k),
                                           updates)
    parts.free()
    L = lbm.load_library()
    buf = L.LBM_DeviceAlloc(16 * 256 * 64)
    out["streamlines_256x64_us"] = timed(lambda k=0: L.LBM_TraceStreamlines(g.ptr, buf, 256, 64), 20)
    lbm.LBM_Synchronize(g)
    L.LBM_DeviceFree(buf)
    _emit(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["storage", "sweep", "warmstart", "gain", "particles"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--tunnels", type=int, default=64)
    ap.add_argument("--storage", type=int, default=1)
    ap.add_argument("--tune-steps", type=int, default=20, help="sweep: steps per autotuner candidate")
    ap.add_argument("--train", type=int, default=384, help="warmstart: training tunnels")
    ap.add_argument("--eval", type=int, default=32, help="warmstart: held-out tunnels")
    ap.add_argument("--iters", type=int, default=3000, help="warmstart: training iterations")
    ap.add_argument("--data-steps", type=int, default=None, help="warmstart: 6000; gain: 4000")
    ap.add_argument("--cold-steps", type=int, default=None, help="warmstart: 20000; gain: 8000 (also the FP64 run)")
    ap.add_argument("--warm-steps", type=int, default=None, help="warmstart: 12000; gain: 4000")
    ap.add_argument("--weights-in", default=None)
    ap.add_argument("--weights-out", default=None)
    ap.add_argument("--phase", choices=["all", "train", "eval"], default="all")
    ap.add_argument("--no-warm-fp32", action="store_true", help="warmstart: skip the warm-start FP32 runs")
    ap.add_argument("--reference", choices=["fp32", "fp64"], default="fp32",
                    help="warmstart: the cold start whose converged Cd is the target")
    ap.add_argument("--eval-offset", type=int, default=0, help="warmstart: first held-out seed - 100000")
    ap.add_argument("--bootstrap", type=int, default=0,
                    help="warmstart: cold-start only this many training tunnels, warm-start the rest")
    a = ap.parse_args(argv)
    if a.which == "storage":
        storage_experiment(a.steps or 20000)
    elif a.which == "sweep":
        sweep_experiment(a.tunnels, a.steps or 5000, a.storage, tune_steps=a.tune_steps)
    elif a.which == "particles":
        particles_experiment()
    elif a.which == "gain":
        gain_experiment(a.train, a.eval, a.iters, data_steps=a.data_steps or 4000, steps=a.cold_steps or 8000,
                        warm_steps=a.warm_steps or 4000)
    else:
        warmstart_experiment(a.train, a.eval, a.iters, data_steps=a.data_steps or 6000,
                             cold_steps=a.cold_steps or 20000, warm_steps=a.warm_steps or 12000,
                             weights_in=a.weights_in, weights_out=a.weights_out, phase=a.phase,
                             warm_fp32=not a.no_warm_fp32, reference=a.reference, eval_offset=a.eval_offset,
                             bootstrap=a.bootstrap)


if __name__ == "__main__":
    main(sys.argv[1:])
