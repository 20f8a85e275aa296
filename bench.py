#!/usr/bin/env python3
"""Benchmark of the D3Q19 step on one MI355X.

Workload (default): 512x256x256, MRT collision + Smagorinsky (Cs 0.1) -- the
operator the reference's application runs (simulation/src/main.c:503-515) --
FP32 populations, tau 0.505, inlet u 0.05, periodic y/z, a seeded (seed 0)
random ellipsoid-plus-box car proxy meshed with ~20k triangles and Bouzidi
bounce-back, 50,000 particles seeded on the inlet plane from seed 0 and
advected with RK2 trilinear sampling every 10 LBM steps (the substep count of
the reference's main loop).  The same lattice with regularized collision +
Smagorinsky (the 512x256x256 roofline configuration of the workload list) is
measured after it and reported under "secondary".

`time_to_converged_cd` is the metric's second half, measured end to end
after the throughput runs (bounded to ~20 s): the FP32 solver generates
training fields for 128 seeded bodies, the warm-start network is trained on
them by the project's kernels, and on 8 held-out bodies the seconds until Cd
settles within 1 % of the FP64 cold start are compared for the learned warm
start with FP16 storage and the plain FP32 cold start.  By default the
tunnels are 128x64x64 (the learned-start configuration at half resolution);
--ttc-dims 256x128x128 runs it at full resolution (several minutes).

A "step" is one LBM update of the whole lattice.  Before the timed region the
workload is set up, autotuned, run for --settle untimed steps (200: the
first steps after the autotuner's trials run measurably slower than the
steady state a long run settles into) and warmed up for --warmup steps.
`value` is lattice updates per second of the whole timed region (inputs resident in HBM; includes the
boundary prepass, the particle kernel and the host synchronisations).
`roofline` prices the fused step kernel alone at 152 B/cell/step (19 loads +
19 stores of FP32) against its average launch time from HIP events recorded on
the grid's stream around every launch of every 10th frame of the timed region,
counted back from its last frame, and read after the region (no host wait
inside it; the other frames run without events).  `cpu_baseline` times the oracle (the
C/OpenMP restatement of the reference shaders) on a bounded sample of the same
workload on the host cores.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {0: 152.0, 1: 76.0, 2: 76.0, 3: 304.0}

WORKLOADS = {
    "roofline-512x256x256-regularized-smagorinsky": dict(
        dims=(512, 256, 256), op="reg", smag=1, tau=0.505, u=0.05, particles=50000, storage=0),
    "512x256x256-mrt-smagorinsky": dict(
        dims=(512, 256, 256), op="mrt", smag=1, tau=0.505, u=0.05, particles=50000, storage=0),
    "256x128x128-mrt-smagorinsky-fp16": dict(
        dims=(256, 128, 128), op="mrt", smag=1, tau=0.52, u=0.06, particles=0, storage=1),
    "512x256x256-regularized-smagorinsky-fp16": dict(
        dims=(512, 256, 256), op="reg", smag=1, tau=0.505, u=0.05, particles=50000, storage=1),
    "512x256x256-mrt-smagorinsky-fp16": dict(
        dims=(512, 256, 256), op="mrt", smag=1, tau=0.505, u=0.05, particles=50000, storage=1),
}
DEFAULT = "512x256x256-mrt-smagorinsky"
SECONDARY = "roofline-512x256x256-regularized-smagorinsky"


def bytes_per_cell_storage(w):
    """Bytes of population storage per cell (the size of the probe copy)."""
    return {0: 76, 1: 38, 2: 38, 3: 152}[w["storage"]]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup_grid(lbm, geometry, w):
    nx, ny, nz = w["dims"]
    g = lbm.LBM_Create(nx, ny, nz, (w["tau"] - 0.5) / 3)
    g.tau = w["tau"]
    g.useRegularized = int(w["op"] == "reg")
    g.useMRT = int(w["op"] == "mrt")
    g.useSmagorinsky = w["smag"]
    g.smagorinskyCs = 0.1
    g.periodicYZ = 1
    g.populationStorage = w["storage"]
    t0 = time.time()
    tris, bbox = geometry.car_proxy(seed=0)
    lbm.LBM_SetSolidMesh(g, tris, bbox)
    area = lbm.LBM_ComputeProjectedArea(g, 0)
    if lbm.LBM_InitializeFlow(g, w["u"], 0.0, 0.0) != 1:
        raise RuntimeError("LBM_InitializeFlow failed: " + lbm.last_error())
    log(f"setup {time.time() - t0:.1f}s: {len(tris)} triangles, area {area}, "
        f"boundary cells {lbm.LBM_GetBoundaryCellCount(g)}")
    return g, tris, bbox, area


def run_frames(lbm, g, parts, w, nsteps, step0, event_every=0):
    """nsteps LBM steps in frames of 10 (particles advected after each frame).
    Frames are queued without host synchronisation (LBM_StepN), as the
    reference's main loop steps between samples.  With event_every = k > 0,
    every k-th frame counted back from the last runs through LBM_StepNTimed
    instead, whose launches are bracketed by HIP events on the grid's stream;
    nothing waits for them inside the region: LBM_CollectStepTimings reads
    them after its closing synchronisation.
    Returns (wall_s, main_ms_sum, boundary_ms_sum, timed_steps)."""
    nx = w["dims"][0]
    vel_scale = 8.0 / nx
    main = bnd = 0.0
    timed = 0
    lbm.LBM_Synchronize(g)
    t0 = time.perf_counter()
    done = 0
    frame = 0
    frames = (nsteps + 9) // 10
    while done < nsteps:
        n = min(10, nsteps - done)
        sample = event_every > 0 and (frames - 1 - frame) % event_every == 0
        if sample:
            lbm.LBM_StepNTimed(g, n, w["u"], 0.0, 0.0)
        else:
            lbm.LBM_StepN(g, n, w["u"], 0.0, 0.0)
        done += n
        frame += 1
        if parts is not None:
            lbm.LBM_AdvectParticles(g, parts, float(n), vel_scale, 0, step0 + frame)
    lbm.LBM_Synchronize(g)
    wall = time.perf_counter() - t0
    if event_every > 0:
        main, bnd, timed = lbm.LBM_CollectStepTimings(g)
    return wall, main, bnd, timed


def cpu_baseline(lbm, g, w, steps):
    """The FP32 oracle on the same lattice, geometry and operator (solid and q
    copied from the product), timed over `steps` steps on the host cores."""
    from oracle import oracle as O

    nx, ny, nz = w["dims"]
    og = O.OracleGrid(nx, ny, nz, (w["tau"] - 0.5) / 3, precision=32)
    og.tau = float(np.float32(w["tau"]))
    og.useRegularized = int(w["op"] == "reg")
    og.useMRT = int(w["op"] == "mrt")
    og.useSmagorinsky = w["smag"]
    og.smagorinskyCs = 0.1
    og.periodicYZ = 1
    og.solid[:] = lbm.LBM_CopySolidToHost(g)
    og.q[:] = lbm.LBM_CopyQToHost(g)
    og.initialize_flow(w["u"], 0.0, 0.0)
    og.step(w["u"], 0, 0, n=1)  # first touch of the buffers, untimed
    t0 = time.perf_counter()
    og.step(w["u"], 0, 0, n=steps)
    dt = time.perf_counter() - t0
    cells = nx * ny * nz
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(cells * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
            "sample": f"{steps} steps of the {nx}x{ny}x{nz} {w['op']}+Smagorinsky workload "
                      f"({dt:.1f} s), oracle C/OpenMP restatement of the reference shaders, FP32, AoS two-buffer"}


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (None if unreadable)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def load_traffic(workload):
    """HBM bytes per launch of the step kernel from the committed rocprofv3 PMC
    summary for this workload, if present: (bytes, source file) or (None, None).
    The counters are not collected by this run (a --pmc pass is a separate
    profiler run); the file holds the command that produced them."""
    rel = os.path.join("profiles", f"pmc_step_{workload}.json")
    p = os.path.join(ROOT, rel)
    if os.path.exists(p):
        try:
            return json.load(open(p)).get("hbm_bytes_per_launch"), rel
        except Exception:
            return None, None
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--secondary", default=SECONDARY, choices=sorted(WORKLOADS) + ["none"],
                    help="a second workload measured after the first and reported under 'secondary'")
    ap.add_argument("--ttc-dims", default="128x64x64",
                    help="tunnel size of the time-to-converged-Cd measurement ('none' = skip)")
    ap.add_argument("--settle", type=int, default=200,
                    help="untimed steps of the tuned configuration between the autotuner and the warm-up")
    ap.add_argument("--event-every", type=int, default=10,
                    help="time the kernel launches of every k-th 10-step frame with HIP events (0 = none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU slab "
              "decomposition is a design note, not part of this build", file=sys.stderr)
        sys.exit(2)
    # The library prints the reference's setup lines ("LBM Grid: ...") on the
    # process's stdout; keep stdout for the one JSON line by pointing file
    # descriptor 1 at stderr until the result is written.
    sys.stdout.flush()
    json_out = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)
    out = run(args)
    sys.stdout.flush()
    json_out.write(json.dumps(out) + "\n")
    json_out.flush()


def measure(lbm, geometry, name, args, cpu_steps):
    """Set-up, autotuning, warm-up and the timed region of one workload; the
    grid is freed on return.  Returns the fields of the JSON line."""
    w = WORKLOADS[name]
    g, tris, bbox, area = setup_grid(lbm, geometry, w)
    nx, ny, nz = w["dims"]
    cells = nx * ny * nz
    parts = None
    if w["particles"]:
        parts = lbm.DeviceParticles(w["particles"])
        lbm.LBM_SeedParticles(g, parts, 0)

    # each of the 18 candidates is timed over about 2^31 cell updates (at least
    # 20 steps), so a small lattice's choice does not rest on a millisecond
    tune_steps = max(20, min(1000, (1 << 31) // cells))
    tuned_block, tune_ms = lbm.LBM_Autotune(g, tune_steps, w["u"], 0.0, 0.0)
    _, streaming, xcd_order = lbm.LBM_GetKernelConfig(g)
    log(f"{name}: autotune: workgroup {tuned_block}, streaming mode {streaming}, XCD-contiguous order "
        f"{xcd_order} (best ms/step 128/256/512: {[round(float(v), 4) for v in tune_ms]})")
    # The autotuner's trials (each followed by a host wait and a restore of the
    # fields) leave the step kernel about 1.3 % slower for the next ~200 steps
    # of continuous load than in the steady state a long run settles into;
    # these untimed steps, like the tuner, are set-up before the warm-up.
    if args.settle:
        run_frames(lbm, g, parts, w, args.settle, 0)
    if args.warmup:
        run_frames(lbm, g, parts, w, args.warmup, 0)
    wall, main_ms, bnd_ms, timed_steps = run_frames(lbm, g, parts, w, args.steps, 10 ** 6,
                                                    event_every=args.event_every)
    mlups = cells * args.steps / wall / 1e6
    fx, fy, fz, px, py, pz = lbm.LBM_ComputeDragForceDecomposed(g)
    dyn = 0.5 * w["u"] ** 2 * area
    sane = np.isfinite(fx) and lbm.LBM_ComputeTotalMass(g) > 0.5 * cells
    if not sane:
        raise RuntimeError("simulation state is not finite after the timed region")

    copy_gbs = lbm.LBM_MeasureCopyBandwidth(int(bytes_per_cell_storage(w) * cells), 10)
    bpc = BYTES_PER_CELL[w["storage"]]
    roof = None
    if main_ms > 0:
        avg_s = main_ms / timed_steps / 1e3
        achieved = bpc * cells / avg_s / 1e9
        traffic, traffic_src = load_traffic(name)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": (f"{traffic_src}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of a separate "
                                   "profiler run, scaled by a copy kernel; not measured in this run")
                if traffic_src else None,
                "kernel": "k_step (fused stream+collide)", "bytes_per_cell": bpc,
                "avg_launch_ms": round(main_ms / timed_steps, 5), "timed_launches": timed_steps,
                "copy_kernel_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
                "workgroup": tuned_block, "streaming_mode": streaming, "xcd_contiguous": xcd_order,
                "boundary_prepass_avg_ms": round(bnd_ms / max(timed_steps, 1), 5)}
    cpu = None
    if cpu_steps > 0:
        log("timing the CPU oracle ...")
        cpu = cpu_baseline(lbm, g, w, cpu_steps)
    if parts is not None:
        parts.free()
    lbm.LBM_Free(g)
    storage = {0: "f32", 1: "f16 (shifted storage, f32 math)", 2: "bf16 (shifted storage, f32 math)", 3: "f64"}
    return {
        "value": round(mlups, 1), "ms_per_step": round(wall / args.steps * 1e3, 4),
        "dtype": storage[w["storage"]],
        "config": {"workload": name, "lattice": f"{nx}x{ny}x{nz}",
                   "collision": {"reg": "regularized", "mrt": "MRT"}[w["op"]] + "+Smagorinsky(Cs=0.1)",
                   "tau": w["tau"], "inlet_u": w["u"], "periodicYZ": 1,
                   "body": f"car proxy seed 0, {len(tris)} triangles, Bouzidi",
                   "particles": w["particles"], "particle_interval_steps": 10, "parallelism": "single GPU",
                   "settle_steps": args.settle},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cd": round(abs(fx) / dyn, 4) if dyn > 0 else None,
        "hbm_roofline_mlups": round(HBM_PEAK_GBS * 1e9 / bpc / 1e6, 1),
    }


def run(args):
    from lattice_fluid_dynamics_amd import geometry, lbm

    m = measure(lbm, geometry, args.workload, args, args.cpu_steps)
    out = {
        "metric": "MLUPS (D3Q19 lattice updates per second; % of HBM roofline at 152 B/cell/step FP32)",
        "value": m["value"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": m["dtype"], "data": "synthetic",
        "config": m["config"], "roofline": m["roofline"], "cpu_baseline": m["cpu_baseline"],
        "cd": m["cd"], "hbm_roofline_mlups": m["hbm_roofline_mlups"],
    }
    if args.secondary != "none" and args.secondary != args.workload:
        s = measure(lbm, geometry, args.secondary, args, 0)
        r = s["roofline"] or {}
        out["secondary"] = {
            "workload": args.secondary, "value": s["value"], "unit": "MLUPS", "ms_per_step": s["ms_per_step"],
            "collision": s["config"]["collision"], "roofline_frac": r.get("frac"),
            "avg_launch_ms": r.get("avg_launch_ms"), "achieved_gbs": r.get("achieved"),
            "traffic": r.get("traffic"), "traffic_source": r.get("traffic_source"),
            "workgroup": r.get("workgroup"), "streaming_mode": r.get("streaming_mode"), "cd": s["cd"]}
    if args.ttc_dims != "none":
        out["time_to_converged_cd"] = time_to_converged_cd(tuple(int(v) for v in args.ttc_dims.split("x")))
    return out


def time_to_converged_cd(dims):
    """Learned warm start + FP16 storage against the plain FP32 cold start:
    seconds to Cd within 1 % of the FP64 cold start (experiments.learned_gain)."""
    from lattice_fluid_dynamics_amd import experiments

    full = dims[0] >= 256
    # half resolution: 128 training tunnels and 1500 AdamW steps (about 16 s in
    # all); 64 tunnels and 600 steps gave a geometric-mean speed-up near 15x
    # against 20-22x with these
    kw = dict(data_steps=6000, avg_steps=2000, ref_steps=16000, cold_steps=16000, warm_steps=8000, iters=2000,
              window=10) if full else dict(n_train=128, iters=1500)
    t0 = time.time()
    r = experiments.learned_gain(dims=dims, log=log, **kw)
    return {"metric": "seconds to Cd within 1 % of the FP64 cold start: learned warm start + FP16 storage vs "
                      "plain FP32 cold start (median over held-out bodies)",
            "speedup_median": round(r["median_speedup"], 2) if r["median_speedup"] else None,
            "speedup_min": round(r["min_speedup"], 2) if r["min_speedup"] else None,
            "speedup_geomean": round(r["geomean_speedup"], 2) if r["geomean_speedup"] else None,
            "median_steps": r["median_steps"],
            "median_seconds_per_tunnel": {k: round(float(np.median([row[k + "_s"] for row in r["rows"]
                                                                    if row[k + "_s"] is not None])), 5)
                                          for k in ("cold_fp32", "warm_fp16")},
            "converged": r["converged"], "held_out_tunnels": len(r["rows"]),
            "ms_per_step_per_tunnel": {k: round(v, 5) for k, v in r["ms_per_step_per_tunnel"].items()},
            "warm_init_s_per_tunnel": round(r["warm_init_s_per_tunnel"], 5),
            "training": {"tunnels": r["params"]["n_train"], "data_gen_s": round(r["data_gen_s"], 1),
                         "adamw_steps": r["params"]["iters"], "train_s": round(r["train_s"], 1),
                         "loss_first": round(r["loss_first"], 5), "loss_last": round(r["loss_last"], 6)},
            "workload": f"{dims[0]}x{dims[1]}x{dims[2]} tunnels, coarse lattice 1/4, MRT, periodic y/z; "
                        "seeded random bodies, u U[0.03,0.07], tau U[0.6,0.9]; Cd every "
                        f"{r['params']['interval']} steps, window {r['params']['window']}",
            "wall_s": round(time.time() - t0, 1)}


if __name__ == "__main__":
    main()
