"""Time-to-converged-Cd measurement.

Runs a tunnel from a cold start (or from a warm-start field), samples the drag
coefficient every `interval` steps through the C drag monitor (the bookkeeping
of the reference main loop, simulation/src/main.c:1250-1391), and reports how
many steps and seconds it takes until the windowed mean Cd stays within a
tolerance of a reference value -- the converged Cd of the FP64 cold start.

Definitions:
  Cd(t)      |Fx| / (0.5 U^2 A), A = projected area, FP64 force sums.
  Cd_w(t)    mean of the last `window` samples (smooths vortex shedding).
  t_conv     first sample time after which |Cd_w - Cd_ref| <= tol * Cd_ref
             holds for every later sample of the run.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import lbm


@dataclass
class CdHistory:
    steps: List[int] = field(default_factory=list)
    cd: List[float] = field(default_factory=list)
    cl: List[float] = field(default_factory=list)
    cd_pressure: List[float] = field(default_factory=list)
    ms_per_step: float = 0.0

    def windowed(self, window: int) -> np.ndarray:
        c = np.asarray(self.cd, np.float64)
        out = np.empty_like(c)
        for k in range(len(c)):
            out[k] = c[max(0, k - window + 1):k + 1].mean()
        return out

    def converged_value(self, tail: int = 20) -> float:
        return float(np.mean(self.cd[-tail:]))


def run_history(g: lbm.Grid, u: float, area: float, steps: int, interval: int = 200,
                fixed_point: bool = False) -> CdHistory:
    """Advances `g` by `steps` steps, sampling Cd every `interval` steps."""
    g.forceFixedPoint = int(fixed_point)
    h = CdHistory()
    lbm.LBM_Synchronize(g)
    t_step = 0.0
    done = 0
    dyn = 0.5 * u * u * area
    while done < steps:
        n = min(interval, steps - done)
        t0 = time.perf_counter()
        lbm.LBM_StepN(g, n, u, 0.0, 0.0)
        lbm.LBM_Synchronize(g)
        t_step += time.perf_counter() - t0
        done += n
        F = lbm.LBM_ComputeDragForceF64(g)
        h.steps.append(done)
        h.cd.append(abs(F[0]) / dyn)
        h.cl.append(F[1] / dyn)
        h.cd_pressure.append(abs(F[3]) / dyn)
    h.ms_per_step = t_step / steps * 1e3
    return h


def time_to_within(h: CdHistory, ref: float, tol: float = 0.01, window: int = 10) -> Optional[int]:
    """Steps until the windowed Cd enters and stays in [ref(1-tol), ref(1+tol)];
    None if it never settles within the run."""
    w = h.windowed(window)
    ok = np.abs(w - ref) <= tol * abs(ref)
    if not ok[-1]:
        return None
    bad = np.nonzero(~ok)[0]
    k = 0 if len(bad) == 0 else bad[-1] + 1
    return int(h.steps[k])
