"""bench.py's fixed inputs, on the CPU: the algorithmic byte count the
roofline divides by, the update's byte model, the traffic it reads back from
the committed rocprofv3 summary, the timing segments inside the finite
window, and the refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

import bench
from cumc3d_amd.config import turbulence_3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_step_bytes_per_zone():
    # FP64 state of 8 words per zone: stage 1 reads U and writes U^1, stages 2
    # and 3 read U^(s-1) and U^0 and write U^s
    words = (2, 3, 3)
    assert bench.STAGE_BYTES_PER_ZONE == tuple(8.0 * 8 * w for w in words)
    assert bench.STEP_BYTES_PER_ZONE == sum(bench.STAGE_BYTES_PER_ZONE) == 512.0


def test_update_byte_model():
    cfg = turbulence_3d(256)
    n = cfg.nx * cfg.ny * cfg.nz
    # FP32 split: 24 doubles (U, u_old, face B and its lower faces, the new U)
    # and 31 floats (fluxes, edge EMFs, primitives, eps, cs, cell field, cell EMFs)
    assert bench.kernel_model_bytes(cfg, "update", 4) == 316 * n
    assert bench.kernel_model_bytes(cfg, "update", 8) == 432 * n


def test_traffic_read_from_the_committed_summary():
    # the mean of the three k_update<S, float, 0> rows of the summary's table
    path = bench.traffic_summary("turbulence", 256)
    rows = {}
    for line in open(path):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cols) > 5 and cols[0].startswith("k_update<") and "float" in cols[0]:
            rows[cols[0]] = float(cols[5])
    assert sorted(rows) == ["k_update<1, float, 0>", "k_update<2, float, 0>", "k_update<3, float, 0>"]
    got = bench.measured_traffic("update", 4, "turbulence", 256)
    assert got == pytest.approx(1e6 * sum(rows.values()) / 3, rel=1e-12)
    assert bench.measured_traffic("update", 4, "turbulence", 128) is None


@pytest.mark.parametrize("steps,seg,want", [(20, 21, [20]), (20, 20, [20]), (20, 8, [8, 8, 4]), (5, 1, [1] * 5)])
def test_segments(steps, seg, want):
    assert list(bench.segments(steps, seg)) == want


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one MI355X" in r.stdout
