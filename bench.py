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

The same steps are timed three times from the same initial state:
  fp64      the FP64 mode, bit-exact with the reference (WENO-Z everywhere);
  fp32      the FP32 split: single-precision reconstruction, Riemann solver and
            EMFs; double-precision RK accumulators and energy-to-pressure
            recovery (WENO-Z everywhere);
  fp32_sel  the FP32 split with the learned selector choosing TVD-MC, PPM or
            WENO-Z per cell and direction, retrained here on samples the solver
            generates from a 64^3 run of the same physics.
Each variant is timed --repeats times (default 5), every run from a fresh
solver on the same initial state with its own warmup; ms_per_step is the
median run's (all runs are listed).  value = zone-updates/s of fp32_sel; the
others and the ratio are reported beside it, with the density L1 error of
both FP32 variants against fp64 after the same steps with the same dt
sequence.

The roofline object is for the dominant kernel of the fp32_sel run:
HIP-event durations on the library's stream, algorithmic bytes per launch.
The CPU baseline is the reference Fortran (oracle/_ref, built from its own
sources) on a bounded sample of the same physics.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--repeats R] [--n 256]
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
        # edge: 6 face-EMF arrays + 3 cell-EMF arrays -> 3 edge EMFs on
        # (n+1)^3, plus the face-EMF fill of the (n+2)^3 box's shell (the
        # mapped interior face EMF read, the ghost one written: 4 per cell)
        shell = (nx + 2) * (ny + 2) * (nz + 2) - n
        return 12 * w * (nx + 1) * (ny + 1) * (nz + 1) + 4 * w * shell
    if name == "update":
        # stage 2/3 of the fused update
        if split:
            # U, u_old, face B (+ lower faces) in double; fluxes, edge EMFs
            # and the primitives written back in float (the stage-start
            # primitives are read only for the curvilinear/torus/RT sources)
            prim_in = 8 if (cfg.coordinate_flag != 0 or cfg.model != "riemann" or cfg.rt_gravity) else 0
            return n * (8 * (5 + 8 + 3 + 3 + 5) + w * (prim_in + 15 + 3 + 5 + 1 + 1 + 3 + 3))
        # reads: prim 5, bcell 3, eps, u_old 8, face B 3, fluxes 15, edge EMFs 3;
        # writes: prim 5, eps, cs, bcell 3, face B 3, cell EMFs 3
        return n * 8 * (5 + 3 + 1 + 8 + 3 + 15 + 3 + 5 + 1 + 1 + 3 + 3 + 3)
    if name == "ghost":
        g = (nx + 6) * (ny + 6) * (nz + 6) - n
        return g * (2 * w * 10 + 2 * 8 * 2)
    raise KeyError(name)


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_ot256_hlld.md")


def measured_traffic(kernel: str, w: int, n: int):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy of known size),
    averaged over the class's launches (the three RK stages of the update);
    None when the summary does not cover this workload."""
    if n != 256 or not os.path.isfile(TRAFFIC_SUMMARY):
        return None
    prefix = {"flux_x": "k_flux_x<", "flux_y": "k_flux_yz<", "flux_z": "k_flux_yz<", "edge": "k_edge<",
              "update": "k_update<", "ghost": "k_ghost<"}[kernel]
    typ = "float" if w == 4 else "double"
    vals = []
    for line in open(TRAFFIC_SUMMARY):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cols) < 6 or not cols[0].startswith(prefix) or typ not in cols[0]:
            continue
        if kernel in ("flux_y", "flux_z") and not cols[0].rstrip(">").endswith("1" if kernel == "flux_y" else "2"):
            continue
        vals.append(float(cols[5]) * 1e6)
    return sum(vals) / len(vals) if vals else None


def make_workload(name: str, n: int):
    from cumc3d_amd import ic
    from cumc3d_amd.config import blast_3d, orszag_tang_3d, torus_cyl, torus_sph, turbulence_3d
    if name == "ot":
        cfg = orszag_tang_3d(n, recon="weno", riemann="hlld")
        return cfg, ic.orszag_tang(cfg), f"3D periodic Orszag-Tang {n}^3, HLLD + WENO-Z"
    if name == "blast":
        cfg = blast_3d(n)
        return cfg, ic.blast(cfg), f"3D magnetised blast {n}^3 periodic, HLLD + WENO-Z"
    if name == "turbulence":
        cfg = turbulence_3d(n)
        return cfg, ic.turbulence(cfg), f"3D periodic decaying turbulence {n}^3, HLLD + WENO-Z"
    if name == "torus_sph":
        cfg = torus_sph(n, 3 * n // 4, n // 2, recon="weno", riemann="hll")
        return (cfg, ic.torus_sph(cfg),
                f"spherical torus (r, theta, phi) {cfg.nx}x{cfg.ny}x{cfg.nz} with FIXPOLE, HLL + WENO-Z")
    if name == "torus_cyl":
        cfg = torus_cyl(n, n // 2, n, recon="ppm")
        return cfg, ic.torus(cfg), f"cylindrical torus (r, phi, z) {cfg.nx}x{cfg.ny}x{cfg.nz}, HLLD + PPM"
    if name == "collapsar":
        from cumc3d_amd.config import collapsar
        # n = radial cells; n = 156 is the reference's shipped progenitor grid
        # (model/collapsar/profile: 156 x 44 x 128)
        cfg = collapsar(n, max(8, (44 * n) // 156), max(96, (128 * n) // 156), recon="ppm", riemann="hlld")
        return (cfg, ic.collapsar(cfg),
                f"collapsar (r, theta, phi) {cfg.nx}x{cfg.ny}x{cfg.nz} with self-gravity (SOR every "
                f"{cfg.n_pot} steps) and black-hole accretion, HLLD + PPM, FP64")
    raise SystemExit(f"unknown workload {name}")


def sor_kernel_bytes(cfg, colour: str) -> float:
    """Algorithmic bytes of one SOR kernel launch.  Red: rho and the three 3D
    coefficients of the red cells (16 B per cell), phi of every cell read
    once (8), the new red values written (4).  Black: the same for the black
    cells (16), the new red and old black values read (8), the black values
    written (4) and the old red values for the error terms (4)."""
    return (28.0 if colour == "red" else 32.0) * cfg.zones


def sor_bytes(cfg) -> float:
    """Algorithmic bytes of one SOR iteration (red and black sweeps,
    custom.f90:537-584): rho, the three 3D coefficients clp1/clm1/epsc and
    phi read once per cell, phi written once per cell; the 1D/2D
    coefficients and the 7-point neighbour re-reads are cache traffic."""
    return 48.0 * cfg.zones


def bench_collapsar(args):
    """The collapsar workload (FP64; self-gravity has no FP32 split): wall
    time of `steps` steps after `warmup` (the warmup holds step 0, whose
    potential solve starts from the initial update's single sweep), the
    per-kernel times and the Poisson solves of the same number of steps with
    device timing on."""
    from cumc3d_amd.solver import make_solver
    cfg, st, label = make_workload("collapsar", args.n)
    print(f"[bench] {label}: initial update + warmup {args.warmup}", file=sys.stderr, flush=True)
    s = make_solver(cfg, st)
    s.run(args.warmup, total_time=1e300, sync_each_step=False)
    s.set_timing(True)
    s.run(args.steps, total_time=1e300, sync_each_step=False)
    kt = s.timing()
    s.set_timing(False)
    print(f"[bench] timed {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    s.run(args.steps, total_time=1e300, sync_each_step=False)
    wall = time.perf_counter() - t0
    g = s.gravity()
    del s
    zones = cfg.zones
    value = zones * args.steps / wall
    solves = args.steps // cfg.n_pot
    sweep_ms = kt["poisson"] / kt["sor_sweeps"] if kt["sor_sweeps"] else None
    achieved = sor_bytes(cfg) / (sweep_ms * 1e-3) / 1e9 if sweep_ms else None
    dom = max(KERNELS, key=lambda k: kt[k])
    step_kernels = sum(kt[k] for k in KERNELS)
    line = {
        "metric": "zone-updates/s (collapsar with self-gravity, FP64)", "value": value, "unit": "zone-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (cumc3d_amd.ic.collapsar: power-law envelope around a cavity, rotating, uniform B)",
        "config": {"workload": label, "zones": zones, "n_pot": cfg.n_pot, "solves_per_window": solves},
        "gravity": {"poisson_ms_per_solve": kt["poisson"], "sor_sweeps_per_solve": kt["sor_sweeps"],
                    "sor_ms_per_sweep": sweep_ms, "m_bh": g["m_bh"], "last_solve_sweeps": g["iterations"],
                    "poisson_ms_per_step": kt["poisson"] * solves / args.steps if solves else 0.0},
        "roofline": {"bound": "hbm", "kernel": "sor sweep pair (k_sor_red + k_sor_black)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
                     "algorithmic_bytes": sor_bytes(cfg),
                     "step_kernels_ms": {k: kt[k] for k in KERNELS}, "step_kernel_sum_ms": 3 * step_kernels,
                     "dominant_step_kernel": dom},
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def profile_io(cfg, st, variant: str, sel_w=None, steps: int = 3):
    """Time print_hydroprofile's path at this size: cumc3d_profile (device
    pack + copy to the host) and the HDF5 write through libcumc3d_io."""
    import tempfile
    from cumc3d_amd import hydroprofile
    s = make_variant(cfg, st, variant, sel_w)
    s.run(steps, total_time=1e300, sync_each_step=False)
    t0 = time.perf_counter()
    p = s.profile()
    t1 = time.perf_counter()
    nbytes = sum(a.nbytes for a in p.values())
    with tempfile.TemporaryDirectory() as d:
        faces = hydroprofile.written_faces(s.grid)
        path = hydroprofile.profile_name(d, 1)
        t2 = time.perf_counter()
        hydroprofile.write_arrays(path, cfg, faces, s.t, p["primitive"], p["epsilon"], p["bfield"])
        t3 = time.perf_counter()
        size = os.path.getsize(path)
    del s
    return {"bytes": nbytes, "file_bytes": size, "pack_and_copy_ms": 1e3 * (t1 - t0),
            "pack_and_copy_GBs": nbytes / (t1 - t0) / 1e9, "hdf5_write_ms": 1e3 * (t3 - t2),
            "hdf5_write_GBs": size / (t3 - t2) / 1e9}


def train_selector(n: int, workload: str, thr_ppm: float, thr_weno: float, seed: int = 1, iters: int = 600):
    """Train the selector on samples the FP32 solver generates from a small run
    of the same physics (n^3, snapshots every 3 steps); returns its parameters
    and the training summary."""
    from cumc3d_amd.solver import make_solver
    cfg, st, _ = make_workload(workload, n)
    s = make_solver(cfg.with_(precision="fp32"), st)
    s.selector_init(capacity=1 << 20, seed=seed)
    counts = None
    for k in range(6):
        s.run(3, total_time=1e300)
        counts = s.selector_collect(stride=2, thr_ppm=thr_ppm, thr_weno=thr_weno, seed=k + 1)
    loss = s.selector_train(iters, blocks=128, lr=3e-3, class_weight=(1.0, 1.0, 2.0), seed=seed)
    w = s.selector_params()
    del s
    info = {"train_grid": f"{n}^3", "samples": int(counts[0]), "label_counts": counts[1:].tolist(),
            "thresholds": [thr_ppm, thr_weno], "loss_first": float(loss[:10].mean()),
            "loss_last": float(loss[-10:].mean()), "iters": iters}
    return w, info


SEL_EVERY = 1   # selector refresh interval in steps (set from --sel-every)


def calibration_copy(nbytes: int = 1 << 30):
    """One device-to-device copy of nbytes through the HIP runtime: rocprofv3
    counters of its copy kernel give the factor between FETCH_SIZE /
    WRITE_SIZE and bytes moved."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def make_variant(cfg, st, variant: str, sel_w=None):
    from cumc3d_amd.solver import make_solver
    prec = "fp64" if variant == "fp64" else "fp32"
    s = make_solver(cfg.with_(precision=prec), st)
    if variant == "fp32_sel":
        s.selector_init(capacity=1024)
        s.selector_params(sel_w)
        s.selector_enable(True, every=SEL_EVERY)
    return s


def time_variant(cfg, st, variant: str, warmup: int, steps: int, sel_w=None):
    s = make_variant(cfg, st, variant, sel_w)
    s.run(warmup, total_time=1e300, sync_each_step=False)
    s.set_timing(True)
    s.run(2, total_time=1e300, sync_each_step=False)
    kt = s.timing()
    s.set_timing(False)
    t0 = time.perf_counter()
    s.run(steps, total_time=1e300, sync_each_step=False)   # returns after a stream sync
    wall = time.perf_counter() - t0
    frac = None
    if variant == "fp32_sel":
        import numpy as np
        cnt = sum(np.bincount(s.selector_mask(d, refresh=False)[3:-3, 3:-3, 3:-3].ravel(), minlength=3)
                  for d in range(cfg.n_dim))
        frac = (cnt / cnt.sum()).tolist()
    del s
    return wall, kt, frac


def l1_errors(cfg, st, steps: int, sel_w=None):
    """Density L1 (absolute, relative) and relative Linf of the FP32 split and
    of FP32 + selector against FP64 after `steps` steps driven with FP64's dt
    sequence."""
    import numpy as np
    a = make_variant(cfg, st, "fp64")
    others = {"fp32": make_variant(cfg, st, "fp32")}
    if sel_w is not None:
        others["fp32_sel"] = make_variant(cfg, st, "fp32_sel", sel_w)
    for _ in range(steps):
        dt = a.finddt()
        a.rungekutta(dt)
        for b in others.values():
            b.finddt()
            b.rungekutta(dt)
    w = (slice(3, -3),) * 3
    sa = a.state()
    x = sa["rho"][w]
    out = {"fp64": {"divb_norm": divb_norm(cfg, sa)}}
    for k, b in others.items():
        sb = b.state()
        y = sb["rho"][w]
        d = np.abs(x - y)
        out[k] = {"rho_l1": float(d.mean()), "rho_l1_rel": float(d.mean() / np.abs(x).mean()),
                  "rho_linf_rel": float(d.max() / np.abs(x).max()), "divb_norm": divb_norm(cfg, sb)}
    out["steps"] = steps
    out["t"] = a.t
    return out


def divb_norm(cfg, st):
    """max |div B| * h / <|B|> over cells 2..n of a uniform Cartesian grid (the
    lowest face of each axis stays frozen in the reference, so cells 1 are left
    out)."""
    import numpy as np
    if cfg.coordinate_flag != 0:
        return None     # the curvilinear divergences are checked by the tests
    bx, by, bz = st["bx"], st["by"], st["bz"]
    h = (cfg.x_end - cfg.x_start) / cfg.nx
    i = (slice(4, -3),) * 3
    d = ((bx[i] - bx[4:-3, 4:-3, 3:-4]) + (by[i] - by[4:-3, 3:-4, 4:-3]) + (bz[i] - bz[3:-4, 4:-3, 4:-3])) / h
    bm = np.mean(np.sqrt(st["bcx"][3:-3, 3:-3, 3:-3] ** 2 + st["bcy"][3:-3, 3:-3, 3:-3] ** 2
                         + st["bcz"][3:-3, 3:-3, 3:-3] ** 2))
    return float(np.max(np.abs(d)) * h / bm)


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


def trace_summary(path: str, n: int, workload: str = "ot") -> str:
    """Markdown table of a rocprofv3 --kernel-trace CSV: per kernel, the
    dispatches at its largest grid (the n^3 workload; the selector's training
    run is smaller), with calls, mean/min/max duration and VGPRs, plus the
    algorithmic bytes and GB/s of the stage kernels (and, for the collapsar,
    of the SOR kernels: red 28 B and black 32 B per cell, see sor_kernel_bytes)."""
    import csv
    from collections import defaultdict
    from cumc3d_amd.config import orszag_tang_3d
    cfg = make_workload(workload, n)[0] if workload != "ot" else orszag_tang_3d(n)
    groups = defaultdict(list)
    for r in csv.DictReader(open(path)):
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        groups[r["Kernel_Name"]].append((grid, dur, int(r["VGPR_Count"]), int(r["Accum_VGPR_Count"])))
    rows = []
    for name, ds in groups.items():
        gmax = max(d[0] for d in ds)
        sel = [d for d in ds if d[0] == gmax]
        durs = [d[1] for d in sel]
        rows.append((sum(durs), name, len(sel), gmax, sum(durs) / len(durs), min(durs), max(durs), sel[0][2], sel[0][3]))
    rows.sort(reverse=True)
    cls = [("k_flux_x<", "flux_x"), ("k_flux_yz<", None), ("k_update<", "update")]
    out = ["| kernel | calls | grid (threads) | mean us | min us | max us | VGPR | AGPR | algorithmic MB | GB/s |",
           "|---|---|---|---|---|---|---|---|---|---|"]
    for tot, name, calls, grid, mean, mn, mx, vg, ag in rows:
        if calls < 2 and tot < 100:
            continue
        w = 4 if "<float" in name else 8
        kb = None
        if "k_flux_x<" in name:
            kb = kernel_bytes(cfg, "flux_x", w)
        elif "k_flux_yz<" in name:
            kb = kernel_bytes(cfg, "flux_y" if name.rstrip(")").split(",")[3].strip().startswith("1") else "flux_z", w)
        elif "k_update<" in name:
            kb = kernel_bytes(cfg, "update", w)
        elif "k_sor_red" in name or "k_sor_black" in name:
            kb = sor_kernel_bytes(cfg, "red" if "k_sor_red" in name else "black")
        mb = f"{kb / 1e6:.0f}" if kb else ""
        gbs = f"{kb / (mean * 1e-6) / 1e9:.0f}" if kb else ""
        short = name.replace("cumc3d::", "").replace("void ", "")
        short = short.split("(")[0]
        out.append(f"| `{short}` | {calls} | {grid} | {mean:.1f} | {mn:.1f} | {mx:.1f} | {vg} | {ag} | {mb} | {gbs} |")
    return "\n".join(out)


def main(argv=None):
    if argv is None and len(sys.argv) > 2 and sys.argv[1] == "--trace-summary":
        n = int(sys.argv[3]) if len(sys.argv) > 3 else 256
        print(trace_summary(sys.argv[2], n, sys.argv[4] if len(sys.argv) > 4 else "ot"))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed runs of each variant, each from a fresh solver on the same initial state; "
                         "value and ms_per_step are their median")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", default="ot",
                    choices=["ot", "blast", "turbulence", "torus_sph", "torus_cyl", "collapsar"])
    ap.add_argument("--profile-io", action="store_true",
                    help="also time the profile output path (device pack, copy, HDF5 write)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-l1", action="store_true")
    ap.add_argument("--train-n", type=int, default=64)
    ap.add_argument("--train-iters", type=int, default=600)
    ap.add_argument("--variants", default="fp64,fp32,fp32_sel")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="copy 1 GiB device to device first (the known-size kernel that scales the HBM counters)")
    ap.add_argument("--sel-every", type=int, default=8)
    ap.add_argument("--thr-ppm", type=float, default=1e-3)
    ap.add_argument("--thr-weno", type=float, default=1e-2)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one MI355X: the step has no multi-GPU path "
              "(no halo exchange between GPUs); run with --gpus 1.")
        return 2

    if args.workload == "collapsar":
        return bench_collapsar(args)
    cfg, st, label = make_workload(args.workload, args.n)
    zones = cfg.zones
    print(f"[bench] training the selector on a {args.train_n}^3 run", file=sys.stderr, flush=True)
    global SEL_EVERY
    SEL_EVERY = args.sel_every
    if args.pmc_calibrate:
        calibration_copy()
    sel_w, sel_info = train_selector(args.train_n, args.workload, args.thr_ppm, args.thr_weno,
                                     iters=args.train_iters)
    sel_info["refresh_every"] = SEL_EVERY
    print(f"[bench] selector: {sel_info}", file=sys.stderr, flush=True)
    res, walls = {}, {}
    for var in args.variants.split(","):
        runs = []
        for r in range(max(1, args.repeats)):
            print(f"[bench] {label}: {var} run {r + 1}: warmup {args.warmup}, timed {args.steps}",
                  file=sys.stderr, flush=True)
            runs.append(time_variant(cfg, st, var, args.warmup, args.steps, sel_w))
            print(f"[bench] {var}: {1e3 * runs[-1][0] / args.steps:.2f} ms/step", file=sys.stderr, flush=True)
        walls[var] = sorted(w for w, _, _ in runs)
        # the median run's wall time; its kernel timing and selector fractions
        # are the first run's
        res[var] = (walls[var][len(runs) // 2], runs[0][1], runs[0][2])

    def rate(v):
        return zones * args.steps / res[v][0]

    main_var = "fp32_sel" if "fp32_sel" in res else list(res)[-1]
    wall, kt, frac = res[main_var]
    value = rate(main_var)
    dom = max(KERNELS, key=lambda k: kt[k])
    achieved = kernel_bytes(cfg, dom, 8 if main_var == "fp64" else 4) / (kt[dom] * 1e-3) / 1e9
    traffic = measured_traffic(dom, 8 if main_var == "fp64" else 4, args.n) if args.workload == "ot" else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "algorithmic_bytes": kernel_bytes(cfg, dom, 8 if main_var == "fp64" else 4),
                "traffic_source": "profiles/traffic_ot256_hlld.md (rocprofv3 FETCH_SIZE + WRITE_SIZE)"
                if traffic else None,
                "kernel_ms": kt[dom],
                "kernel_ms_all": {v: {k: res[v][1][k] for k in KERNELS + ["selector"]} for v in res},
                "step_bytes_frac": 512.0 * value / (HBM_PEAK_GBS * 1e9)}

    err = None
    if not args.no_l1:
        nacc = args.warmup + 2 + args.steps
        print(f"[bench] L1 against FP64 after {nacc} steps", file=sys.stderr, flush=True)
        err = l1_errors(cfg, st, nacc, sel_w)

    io = None
    if args.profile_io:
        print("[bench] profile output path", file=sys.stderr, flush=True)
        io = profile_io(cfg, st, main_var, sel_w)

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
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64 (bf16 selector)",
        "data": f"synthetic (seeded IC, cumc3d_amd.ic.{ {'ot': 'orszag_tang', 'torus_cyl': 'torus'}.get(args.workload, args.workload) })",
        "config": {"workload": label + ", FP32 split + learned selector (FP64 accumulators)", "zones": zones,
                   "riemann": cfg.riemann, "recon": "selector: TVD-MC / PPM / WENO-Z per cell"},
        "variants": {v: {"value": rate(v), "ms_per_step": 1e3 * res[v][0] / args.steps,
                         "runs_ms_per_step": [round(1e3 * w / args.steps, 4) for w in walls[v]]} for v in res},
        "repeats": max(1, args.repeats),
        "speedup_vs_fp64": value / rate("fp64") if "fp64" in res else None,
        "selector": dict(sel_info, class_fractions=frac),
        "l1_vs_fp64": err,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if io is not None:
        line["output"] = io
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
