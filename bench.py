#!/usr/bin/env python3
"""Throughput of the MHD update step on one MI355X.

A step is one FINDDT + one SSP-RK3 RUNGEKUTTA over the whole grid (the
reference's time-loop body, solver/mhd/main.f90:384-392), with the state
resident in HBM.  Workload: 3D magnetised blast at 256^3, periodic, HLLD +
WENO-Z, gamma=5/3, seeded IC (cumc3d_amd.ic.blast).

value = zone-updates/s = 256^3 * steps / wall time of the timed steps.

Prints one JSON line with, in addition, the roofline of the dominant kernel
(HIP-event durations on the library's stream, algorithmic bytes per launch)
and the CPU baseline (the reference Fortran, built from its own sources by
oracle/build_ref.py, on a bounded sample of the same physics).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s @256³ (FP32+learned vs FP64 plain), % HBM peak, L1 err"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(cfg, name: str) -> float:
    """Algorithmic HBM bytes of one launch (FP64 fields read once, written once)."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    w = 8.0
    if name == "flux_x":
        return w * (8 * (nx + 2) * ny * nz + 7 * (nx + 1) * ny * nz)
    if name == "flux_y":
        return w * (8 * nx * (ny + 2) * nz + 7 * nx * (ny + 1) * nz)
    if name == "flux_z":
        return w * (8 * nx * ny * (nz + 2) + 7 * nx * ny * (nz + 1))
    if name == "edge":
        return w * 15 * (nx + 1) * (ny + 1) * (nz + 1)
    if name == "update":
        return w * 50 * nx * ny * nz
    if name == "ghost":
        return w * 12 * ((nx + 6) * (ny + 6) * (nz + 6) - nx * ny * nz)
    raise KeyError(name)


def cpu_baseline(max_threads: int = 16):
    """The reference solver (oracle/_ref/bench_blast128) on 3 steps of the blast
    at 128^3, all host threads."""
    from cumc3d_amd import ic
    from oracle.cases import BENCH_CASES
    from oracle.refio import have_ref, run_ref
    name = "bench_blast128"
    case = BENCH_CASES[name]
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    sample = f"reference Fortran (flang -O2 -fopenmp), blast 128^3 HLLD+WENO-Z, {case.steps} steps"
    if not have_ref(name):
        return {"value": None, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
                "sample": sample + " (binary not built)"}
    st = ic.GENERATORS[case.ic](case.cfg)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ic.bin")
        ic.write_ic(p, case.cfg, st)
        ref = run_ref(name, p, case.steps, threads=threads, timeout=900)
    rate = case.cfg.zones * ref["nstep"] / ref["seconds"]
    return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
            "sample": sample + f", {ref['seconds']:.1f} s loop time"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(the halo-exchange design is described in DESIGN.md); run with --gpus 1.")
        return 2

    from cumc3d_amd import ic
    from cumc3d_amd.config import blast_3d
    from cumc3d_amd.solver import make_solver

    cfg = blast_3d(args.n)
    print(f"[bench] building IC {args.n}^3 ...", file=sys.stderr, flush=True)
    st = ic.blast(cfg)
    s = make_solver(cfg, st)
    del st
    print("[bench] warmup", file=sys.stderr, flush=True)
    s.run(args.warmup, total_time=1e300, sync_each_step=False)

    # kernel timing pass (HIP events between launches on the library stream)
    s.set_timing(True)
    s.run(2, total_time=1e300, sync_each_step=False)
    kt = s.timing()
    s.set_timing(False)

    print("[bench] timed steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    s.run(args.steps, total_time=1e300, sync_each_step=False)   # returns after a stream sync
    t1 = time.perf_counter()
    wall = t1 - t0
    zones = cfg.zones
    value = zones * args.steps / wall
    ms = 1e3 * wall / args.steps

    kernels = ["flux_x", "flux_y", "flux_z", "edge", "update", "ghost"]
    dom = max(kernels, key=lambda k: kt[k])
    dur = kt[dom] * 1e-3
    achieved = kernel_bytes(cfg, dom) / dur / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "kernel_ms": kt[dom], "kernel_ms_all": {k: kt[k] for k in kernels},
                "step_bytes_frac": 512.0 * value / (HBM_PEAK_GBS * 1e9)}

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (reference Fortran)", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline()
        except Exception as e:  # report, do not hide, a failed baseline
            cpu = {"value": None, "error": str(e)[:200]}

    line = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded blast IC, cumc3d_amd.ic.blast)",
        "config": {"workload": f"3D magnetised blast {args.n}^3 periodic, HLLD + WENO-Z, FP64 plain",
                   "zones": zones, "precision": "fp64", "riemann": "hlld", "recon": "weno"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
