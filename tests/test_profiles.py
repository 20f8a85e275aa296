"""The committed rocprofv3 PMC summaries that bench.py reports as `traffic`:
each is internally consistent (per-phase counters scaled by the copy-kernel
calibration give the stated bytes per launch and ratio) and bench.py reads it
for its workload."""
import json
import os

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pmc_files():
    d = os.path.join(ROOT, "profiles")
    return sorted(f for f in os.listdir(d) if f.startswith("pmc_step_") and f.endswith(".json"))


@pytest.mark.parametrize("name", pmc_files())
def test_pmc_summary_is_consistent(name):
    p = json.load(open(os.path.join(ROOT, "profiles", name)))
    workload = name[len("pmc_step_"):-len(".json")]
    assert workload in bench.WORKLOADS
    w = bench.WORKLOADS[workload]
    nx, ny, nz = w["dims"]
    alg_kib = nx * ny * nz * bench.BYTES_PER_CELL[w["storage"]] / 1024
    assert p["algorithmic_KiB_per_launch"] == pytest.approx(alg_kib)
    # the calibration: FETCH_SIZE reads half of a known copy, WRITE_SIZE all of it
    assert p["calibration"]["fetch_scale"] == pytest.approx(2.0, rel=1e-3)
    assert p["calibration"]["write_scale"] == pytest.approx(1.0, rel=1e-3)
    a, b = p["phase_A_KiB"], p["phase_B_KiB"]
    mean_kib = (a["fetch"] + a["write"] + b["fetch"] + b["write"]) / 2
    assert p["hbm_bytes_per_launch"] == pytest.approx(mean_kib * 1024, rel=1e-6)
    assert p["traffic_over_algorithmic"] == pytest.approx(mean_kib / alg_kib, abs=1e-4)
    # HBM traffic can exceed the algorithmic bytes (mask, list, velocity
    # output, partial lines) but not fall below what the populations need
    assert 1.0 <= p["traffic_over_algorithmic"] < 1.15
    traffic, source = bench.load_traffic(workload)
    assert traffic == p["hbm_bytes_per_launch"]
    assert source == os.path.join("profiles", name)
