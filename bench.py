#!/usr/bin/env python3
"""Throughput of the MHD update step on one MI355X.

A step is one FINDDT + one SSP-RK3 RUNGEKUTTA over the whole grid (the
reference's time-loop body, solver/mhd/main.f90:384-392), with the state
resident in HBM.  Workload (default): the 3D periodic Orszag-Tang vortex of
the 128^3 configuration run at 256^3 with HLLD + WENO-Z, gamma=5/3, seeded
+-1% density noise (cumc3d_amd.ic.orszag_tang).  The magnetised blast at 256^3
is available with --workload blast, but the reference scheme produces NaN in
it after 7 steps (the reference Fortran does the same at 128^3 after 8), so a
20-step timing of it would time NaN arithmetic.

The same steps are timed twice from the same initial state: the FP64 mode
(bit-exact with the reference) and the FP32 split (single-precision
reconstruction, Riemann solver and EMFs; double-precision RK accumulators and
energy-to-pressure recovery).  value = zone-updates/s of the FP32 split;
fp64 and the ratio are reported beside it, with the FP32 split's density L1
error against FP64 after the same steps with the same dt sequence.

The roofline object is for the dominant kernel of the FP32 run: HIP-event
durations on the library's stream, algorithmic bytes per launch.  The CPU
baseline is the reference Fortran (oracle/_ref, built from its own sources)
on a bounded sample of the same physics.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
                       [--workload ot|blast|turbulence] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s @256³ (FP32+learned vs FP64 plain), % HBM peak, L1 err"
HBM_PEAK_GBS = 8000.0
KERNELS = ["flux_x", "flux_y", "flux_z", "edge", "update", "ghost"]


def kernel_bytes(cfg, name: str, w: int) -> float:
    """Algorithmic HBM bytes of one launch: every field the kernel needs read
    once and every field it produces written once.  w = bytes of the working
    type (8 in the FP64 mode, 4 in the FP32 split); the face field, u_old and
    the FP32 split's conserved state are always 8-byte."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    n = nx * ny * nz
    split = w == 4
    if name in ("flux_x", "flux_y", "flux_z"):
        # 7 stencil fields per cell of the row (+2 halo cells), normal face
        # field and 7 fluxes per interface
        cells = {"flux_x": (nx + 2) * ny * nz, "flux_y": nx * (ny + 2) * nz, "flux_z": nx * ny * (nz + 2)}[name]
        faces = {"flux_x": (nx + 1) * ny * nz, "flux_y": nx * (ny + 1) * nz, "flux_z": nx * ny * (nz + 1)}[name]
        return 7 * w * cells + (8 + 7 * w) * faces
    if name == "edge":
        # EMF prep: v, bcell -> 3 cell EMFs on (n+2)^3; edge: 6 face-EMF
        # arrays + 3 cell-EMF arrays -> 3 edge EMFs on (n+1)^3
        return 9 * w * (nx + 2) * (ny + 2) * (nz + 2) + 12 * w * (nx + 1) * (ny + 1) * (nz + 1)
    if name == "update":
        # stage 2/3 of the fused update
        if split:
            return n * (8 * (5 + 8 + 3 + 3 + 5) + w * (8 + 15 + 3 + 5 + 1 + 1 + 3))
        return n * 8 * (5 + 3 + 1 + 8 + 3 + 15 + 3 + 5 + 1 + 1 + 3 + 3)
    if name == "ghost":
        g = (nx + 6) * (ny + 6) * (nz + 6) - n
        return g * (2 * w * 10 + 2 * 8 * 2)
    raise KeyError(name)


def make_workload(name: str, n: int):
    from cumc3d_amd import ic
    from cumc3d_amd.config import blast_3d, orszag_tang_3d, turbulence_3d
    if name == "ot":
        cfg = orszag_tang_3d(n, recon="weno", riemann="hlld")
        return cfg, ic.orszag_tang(cfg), f"3D periodic Orszag-Tang {n}^3, HLLD + WENO-Z"
    if name == "blast":
        cfg = blast_3d(n)
        return cfg, ic.blast(cfg), f"3D magnetised blast {n}^3 periodic, HLLD + WENO-Z"
    if name == "turbulence":
        cfg = turbulence_3d(n)
        return cfg, ic.turbulence(cfg), f"3D periodic decaying turbulence {n}^3, HLLD + WENO-Z"
    raise SystemExit(f"unknown workload {name}")


def time_precision(cfg, st, precision: str, warmup: int, steps: int):
    from cumc3d_amd.solver import make_solver
    s = make_solver(cfg.with_(precision=precision), st)
    s.run(warmup, total_time=1e300, sync_each_step=False)
    s.set_timing(True)
    s.run(2, total_time=1e300, sync_each_step=False)
    kt = s.timing()
    s.set_timing(False)
    t0 = time.perf_counter()
    s.run(steps, total_time=1e300, sync_each_step=False)   # returns after a stream sync
    wall = time.perf_counter() - t0
    del s
    return wall, kt


def l1_error(cfg, st, steps: int):
    """Density L1 (absolute, relative) and relative Linf of the FP32 split
    against FP64 after `steps` steps driven with FP64's dt sequence."""
    import numpy as np
    from cumc3d_amd.solver import make_solver
    a = make_solver(cfg, st)
    b = make_solver(cfg.with_(precision="fp32"), st)
    for _ in range(steps):
        dt = a.finddt()
        a.rungekutta(dt)
        b.finddt()
        b.rungekutta(dt)
    w = (slice(3, -3),) * 3
    x = a.state()["rho"][w]
    y = b.state()["rho"][w]
    d = np.abs(x - y)
    return float(d.mean()), float(d.mean() / np.abs(x).mean()), float(d.max() / np.abs(x).max()), a.t


def cpu_baseline(max_threads: int = 16):
    """The reference solver (oracle/_ref/bench_blast128) on 3 steps of the blast
    at 128^3 (steps the reference completes before its NaN), all host threads."""
    from cumc3d_amd import ic
    from oracle.cases import BENCH_CASES
    from oracle.refio import have_ref, run_ref
    import tempfile
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
    ap.add_argument("--workload", default="ot", choices=["ot", "blast", "turbulence"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-l1", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(the halo-exchange design is described in DESIGN.md); run with --gpus 1.")
        return 2

    cfg, st, label = make_workload(args.workload, args.n)
    zones = cfg.zones
    res = {}
    for prec in ("fp64", "fp32"):
        print(f"[bench] {label}: {prec} warmup {args.warmup}, timed {args.steps}", file=sys.stderr, flush=True)
        wall, kt = time_precision(cfg, st, prec, args.warmup, args.steps)
        res[prec] = (wall, kt)
        print(f"[bench] {prec}: {1e3 * wall / args.steps:.2f} ms/step", file=sys.stderr, flush=True)

    wall, kt = res["fp32"]
    value = zones * args.steps / wall
    fp64_value = zones * args.steps / res["fp64"][0]
    dom = max(KERNELS, key=lambda k: kt[k])
    achieved = kernel_bytes(cfg, dom, 4) / (kt[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "kernel_ms": kt[dom], "kernel_ms_all": {k: kt[k] for k in KERNELS},
                "kernel_ms_all_fp64": {k: res["fp64"][1][k] for k in KERNELS},
                "step_bytes_frac": 512.0 * value / (HBM_PEAK_GBS * 1e9)}

    err = None
    if not args.no_l1:
        nacc = args.warmup + 2 + args.steps
        print(f"[bench] L1 of fp32 vs fp64 after {nacc} steps", file=sys.stderr, flush=True)
        l1, l1rel, linf, t = l1_error(cfg, st, nacc)
        err = {"rho_l1": l1, "rho_l1_rel": l1rel, "rho_linf_rel": linf, "steps": nacc, "t": t}

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (reference Fortran)", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline()
        except Exception as e:  # report, do not hide, a failed baseline
            cpu = {"value": None, "error": str(e)[:200]}

    line = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": f"synthetic (seeded IC, cumc3d_amd.ic.{ {'ot': 'orszag_tang'}.get(args.workload, args.workload) })",
        "config": {"workload": label + ", FP32 split (FP64 accumulators)", "zones": zones,
                   "riemann": cfg.riemann, "recon": cfg.recon, "selector": None},
        "fp64": {"value": fp64_value, "ms_per_step": 1e3 * res["fp64"][0] / args.steps, "dtype": "f64"},
        "speedup_vs_fp64": value / fp64_value,
        "l1_vs_fp64": err,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
