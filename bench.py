#!/usr/bin/env python3
"""Throughput of the MHD update step on one MI355X.

A step is one FINDDT + one SSP-RK3 RUNGEKUTTA over the whole grid (the
reference's time-loop body, solver/mhd/main.f90:98-113), with the state
resident in HBM.

Workload (default): 3D periodic decaying MHD turbulence at 256^3
(cumc3d_amd.ic.turbulence: rho=1, beta=1, rms Mach 1 solenoidal velocity with
|k|<=8 and E(k)~k^-2, mean B along z plus a 50% solenoidal dB), HLLD.  The
same steps are timed for:
  fp64        the FP64 mode, bit-exact with the reference (WENO-Z everywhere):
              the plain FP64 high-order path;
  fp32        the FP32 split (single-precision reconstruction, Riemann solver
              and EMFs; double-precision RK accumulators and energy-to-pressure
              recovery), WENO-Z everywhere;
  fp32_sel    the FP32 split with the learned selector choosing TVD-MC, PPM or
              WENO-Z per cell and direction, trained here on samples the
              solver generates from a 128^3 run of the same physics;
  fp32_tvdmc, fp32_ppm
              controls: the FP32 split with TVD-MC or PPM everywhere.
value = zone-updates/s of fp32_sel.

The reference scheme stays finite on this workload for a limited number of
steps only (26 at 256^3: the recovered pressure collapses in field-dominated
cells and NaN enters through c_s = sqrt(<0); tests/test_configs.py pins the
same on the CPU).  bench.py therefore measures that window first (the FP64
run from the initial state until FINDDT reports a non-finite cell) and times
every variant inside it: each timed run starts from the initial state with
its own `warmup` steps, and `steps` are timed in segments of at most
window - warmup steps (one segment for the defaults, 5 + 20 <= 26).  A timed
segment that ends non-finite is an error, not a number.  The density L1
errors are those of the FP32 variants against fp64 at the last finite FP64
step, every variant driven with FP64's dt sequence.

The roofline object is for the dominant kernel (the fused update) of the
fp32_sel run, on the algorithmic byte count of one RK3 step: 512 B per
zone-update (stage 1 reads U and writes U^1, 16 doubles; stages 2 and 3 read
U^(s-1) and U^0 and write U^s, 24 doubles each), i.e. 512/3 B per zone per
update launch on average, over the launch's mean duration measured with HIP
events on the library's stream during a timed run.  The same launch's
per-kernel byte model (every field it touches read or written once) is
reported beside it as kernel_model_bytes.
The CPU baseline is the reference Fortran (oracle/_ref, built from its own
sources with -O3 -ffp-contract=off -fopenmp) on a bounded sample of the same
physics and scheme, on the host share of cores.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--repeats R] [--n 256]
                       [--workload turbulence|ot|blast|torus_sph|torus_cyl|collapsar]
                       [--no-cpu] [--no-secondary]
       python bench.py --orbits 5 --n 192     (configuration 5's accuracy over 5 orbits)
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

METRIC = "zone-updates/s @256³ (FP32+learned vs FP64 plain), % HBM peak, L1 err"
HBM_PEAK_GBS = 8000.0
KERNELS = ["flux_x", "flux_y", "flux_z", "edge", "update", "ghost"]


def kernel_model_bytes(cfg, name: str, w: int) -> float:
    """The per-kernel byte model of one launch: every field the kernel needs
    read once and every field it produces written once.  w = bytes of the
    working type (8 in the FP64 mode, 4 in the FP32 split); the face field,
    u_old and the FP32 split's conserved state are always 8-byte.  This is a
    model of the kernel's own data movement, larger than the step's
    algorithmic bytes (step_bytes), because the decomposition into sweeps,
    edge EMFs and the update materialises fluxes, EMFs and primitives."""
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


# Algorithmic bytes of one RK3 step per zone-update: FP64 state of 8 words;
# stage 1 reads U and writes U^1 (16 words), stages 2 and 3 read U^(s-1) and
# U^0 and write U^s (24 words each), dt folded into stage 3.  The same for
# both precisions: the RK accumulators stay FP64.
STEP_BYTES_PER_ZONE = 512.0
STAGE_BYTES_PER_ZONE = (128.0, 192.0, 192.0)


def traffic_summary(workload: str, n: int) -> str:
    return os.path.join(ROOT, "profiles", f"traffic_{workload}{n}_hlld.md")


def measured_traffic(kernel: str, w: int, workload: str, n: int):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (calibrated against a copy
    of known size), averaged over the class's launches (the three RK stages
    of the update); None when no summary covers the workload."""
    path = traffic_summary(workload, n)
    if not os.path.isfile(path):
        return None
    prefix = {"flux_x": "k_flux_x<", "flux_y": "k_flux_yz<", "flux_z": "k_flux_yz<", "edge": "k_edge<",
              "update": "k_update<", "ghost": "k_ghost<"}[kernel]
    typ = "float" if w == 4 else "double"
    vals = []
    for line in open(path):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cols) < 6 or not cols[0].startswith(prefix) or typ not in cols[0]:
            continue
        if kernel in ("flux_y", "flux_z") and not cols[0].rstrip(">").endswith("1" if kernel == "flux_y" else "2"):
            continue
        try:
            vals.append(float(cols[5]) * 1e6)
        except ValueError:
            continue
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
        log = (lambda m: print(f"[bench] initial state: {m}", file=sys.stderr, flush=True)) if n > 256 else None
        return cfg, ic.turbulence(cfg, log=log), f"3D periodic decaying turbulence {n}^3, HLLD + WENO-Z"
    if name == "torus_sph":
        cfg = torus_sph(n, 3 * n // 4, n // 2, recon="weno", riemann="hll")
        return (cfg, ic.torus_sph(cfg),
                f"spherical torus (r, theta, phi) {cfg.nx}x{cfg.ny}x{cfg.nz} with FIXPOLE, HLL + WENO-Z")
    if name == "torus_cyl":
        cfg = torus_cyl(n, 2 * n // 3, n, recon="ppm")
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


def train_selector(n: int, workload: str, thr_ppm: float, thr_weno: float, seed: int = 1, iters: int = 600,
                   snapshots: int = 6, every: int = 3):
    """Train the selector on samples the FP32 solver generates from a smaller
    run of the same physics (n^3, a snapshot every `every` steps, sampled into
    a reservoir of 2^20); returns its parameters and the training summary.
    Snapshots stop at the run's first non-finite state (the reference scheme
    fails on the turbulence after a few tens of steps)."""
    from cumc3d_amd.solver import ENONFINITE, CumcError, make_solver
    cfg, st, _ = make_workload(workload, n)
    s = make_solver(cfg.with_(precision="fp32"), st)
    s.selector_init(capacity=1 << 20, seed=seed)
    counts, taken, offered = None, 0, 0
    stride = max(1, (3 * cfg.zones) // (1 << 19))     # about 2^19 candidates per snapshot
    for k in range(snapshots):
        try:
            s.run(every, total_time=1e300)
            s.finddt()
        except CumcError as e:
            if e.code != ENONFINITE:
                raise
            break
        counts = s.selector_collect(stride=stride, thr_ppm=thr_ppm, thr_weno=thr_weno, seed=k + 1)
        taken += 1
        offered += 3 * ((cfg.zones + stride - 1) // stride)
    if counts is None:
        raise RuntimeError("selector training: the training run went non-finite before its first snapshot")
    if np.min(counts[1:]) == 0:
        print(f"[bench] WARNING: selector training labels {counts[1:].tolist()} leave a class empty",
              file=sys.stderr, flush=True)
    loss = s.selector_train(iters, blocks=128, lr=3e-3, class_weight=(1.0, 1.0, 2.0), seed=seed)
    w = s.selector_params()
    del s
    info = {"train_grid": f"{cfg.nx}x{cfg.ny}x{cfg.nz}", "snapshots": taken, "snapshot_every": every, "stride": stride,
            "candidates_offered": offered, "samples_held": int(counts[0]), "label_counts": counts[1:].tolist(),
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


VARIANT_RECON = {"fp32_tvdmc": "tvdmc", "fp32_ppm": "ppm"}


def make_variant(cfg, st, variant: str, sel_w=None):
    from cumc3d_amd.solver import make_solver
    prec = "fp64" if variant == "fp64" else "fp32"
    c = cfg.with_(precision=prec)
    if variant in VARIANT_RECON:
        c = c.with_(recon=VARIANT_RECON[variant])
    s = make_solver(c, st)
    if variant == "fp32_sel":
        s.selector_init(capacity=1024)
        s.selector_params(sel_w)
        s.selector_enable(True, every=SEL_EVERY)
    return s


def restart(s, st, variant: str):
    """Put a variant's context back to the initial state: upload, INITIAL_UPDATE,
    and for the selector a mask refresh at the next step (as a fresh context)."""
    s.initial_model(st)
    s.initial_update()
    if variant == "fp32_sel":
        s.selector_enable(True, every=SEL_EVERY)


def finite_window(cfg, st, cap: int) -> int:
    """Steps after which the FP64 state is still finite, from the initial
    state (FINDDT raising CUMC3D_ENONFINITE marks the first non-finite state);
    `cap` if it stays finite that long."""
    from cumc3d_amd.solver import ENONFINITE, CumcError
    s = make_variant(cfg, st, "fp64")
    k = 0
    try:
        while k < cap:
            s.rungekutta(s.finddt())
            k += 1
        s.finddt()
    except CumcError as e:
        if e.code != ENONFINITE:
            raise
        k -= 1
    del s
    return k


def segments(steps: int, seg: int):
    out = []
    while steps > 0:
        out.append(min(seg, steps))
        steps -= out[-1]
    return out


def time_variant(s, st, variant: str, warmup: int, steps: int, seg: int, kernel_timing: bool = False):
    """Wall time of `steps` steps, in segments of at most `seg` steps, each
    from the initial state after `warmup` untimed steps; with kernel_timing
    the per-kernel HIP-event durations of the timed steps too."""
    wall, kt = 0.0, None
    for L in segments(steps, seg):
        restart(s, st, variant)
        s.run(warmup, total_time=1e300, sync_each_step=False)
        if kernel_timing:
            s.set_timing(True)
        t0 = time.perf_counter()
        s.run(L, total_time=1e300, sync_each_step=False)   # returns after a stream sync; raises if non-finite
        wall += time.perf_counter() - t0
        if kernel_timing:
            kt = kt or s.timing()
            s.set_timing(False)
    return wall, kt


def selector_fractions(s, cfg):
    cnt = sum(np.bincount(s.selector_mask(d, refresh=False)[3:-3, 3:-3, 3:-3].ravel(), minlength=3)
              for d in range(cfg.n_dim))
    return (cnt / cnt.sum()).tolist()


def l1_errors(cfg, st, steps: int, variants, sel_w=None):
    """Density L1 (absolute, relative) and relative Linf of each FP32 variant
    against FP64 after `steps` steps, every variant driven with FP64's dt
    sequence; max normalised div B of every variant."""
    from cumc3d_amd.solver import CumcError
    a = make_variant(cfg, st, "fp64")
    others = {v: make_variant(cfg, st, v, sel_w) for v in variants if v != "fp64"}
    failed = {}
    for k in range(steps):
        dt = a.finddt()
        a.rungekutta(dt)
        for v, b in others.items():
            if v in failed:
                continue
            try:
                b.finddt()
                b.rungekutta(dt)
            except CumcError as e:
                failed[v] = f"non-finite at step {k + 1}: {e}"
    w = (slice(3, -3),) * 3
    sa = a.state()
    x = sa["rho"][w]
    out = {"steps": steps, "t": a.t, "fp64": {"divb_norm": divb_norm(cfg, sa)}}
    for v, b in others.items():
        if v in failed:
            out[v] = {"error": failed[v]}
            continue
        sb = b.state()
        y = sb["rho"][w]
        d = np.abs(x - y)
        out[v] = {"rho_l1": float(d.mean()), "rho_l1_rel": float(d.mean() / np.abs(x).mean()),
                  "rho_linf_rel": float(d.max() / np.abs(x).max()), "divb_norm": divb_norm(cfg, sb),
                  "finite": bool(np.all(np.isfinite(y)))}
    return out


def divb_norm(cfg, st):
    """max |div B| * h / <|B|> over cells 2..n of a uniform Cartesian grid (the
    lowest face of each axis stays frozen in the reference, so cells 1 are left
    out)."""
    if cfg.coordinate_flag != 0:
        return None     # the curvilinear divergences are checked by the tests
    bx, by, bz = st["bx"], st["by"], st["bz"]
    h = (cfg.x_end - cfg.x_start) / cfg.nx
    i = (slice(4, -3),) * 3
    d = ((bx[i] - bx[4:-3, 4:-3, 3:-4]) + (by[i] - by[4:-3, 3:-4, 4:-3]) + (bz[i] - bz[3:-4, 4:-3, 4:-3])) / h
    bm = np.mean(np.sqrt(st["bcx"][3:-3, 3:-3, 3:-3] ** 2 + st["bcy"][3:-3, 3:-3, 3:-3] ** 2
                         + st["bcz"][3:-3, 3:-3, 3:-3] ** 2))
    return float(np.max(np.abs(d)) * h / bm)


def host_share() -> dict:
    """CPUs this process may use: the affinity mask, capped by a cgroup CPU
    quota and by OMP_NUM_THREADS when set; plus the CPU model."""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    threads = n if quota is None else min(n, quota)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        threads = min(threads, int(omp))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": threads, "affinity_cpus": n, "cgroup_quota_cpus": quota, "cpu_model": model}


BENCH_CPU_CASE = {"turbulence": "bench_turb128", "blast": "bench_blast128"}


def cpu_baseline(workload: str):
    """The reference solver (oracle/_ref/<bench case>, built from its own
    sources with -O3 -ffp-contract=off -fopenmp) on a bounded sample of the
    workload's physics and scheme, on the host share of cores, timed by its
    own loop clock (FINDDT + RUNGEKUTTA of every step)."""
    from cumc3d_amd import ic
    from oracle.cases import BENCH_CASES
    from oracle.refio import have_ref, run_ref
    import tempfile
    name = BENCH_CPU_CASE.get(workload)
    hs = host_share()
    if name is None:
        return None
    case = BENCH_CASES[name]
    c = case.cfg
    sample = (f"reference Fortran (flang -O3 -ffp-contract=off -fopenmp, {hs['threads']} threads), "
              f"the workload's physics and scheme at {c.nx}x{c.ny}x{c.nz} ({c.riemann.upper()} + "
              f"{ {'weno': 'WENO-Z', 'ppm': 'PPM', 'tvdmc': 'TVD-MC'}.get(c.recon, c.recon)}), "
              f"{case.steps} steps")
    base = {"unit": "zone-updates/s", "cores": hs["threads"], "kind": "reference", "cpu_model": hs["cpu_model"],
            "host_affinity_cpus": hs["affinity_cpus"], "host_cgroup_quota_cpus": hs["cgroup_quota_cpus"]}
    if not have_ref(name):
        return dict(base, value=None, sample=sample + " (binary not built)")
    st = ic.GENERATORS[case.ic](c)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ic.bin")
        ic.write_ic(p, c, st)
        ref = run_ref(name, p, case.steps, threads=hs["threads"], timeout=900)
    rate = c.zones * ref["nstep"] / ref["seconds"]
    return dict(base, value=rate, sample=sample + f", {ref['seconds']:.1f} s loop time")


def trace_summary(path: str, n: int, workload: str = "turbulence") -> str:
    """Markdown table of a rocprofv3 --kernel-trace CSV: per kernel, the
    dispatches at its largest grid (the n^3 workload; the selector's training
    run is smaller), with calls, mean/min/max duration and VGPRs, the kernel
    byte model (kernel_model_bytes) and its GB/s for the stage kernels, and
    for the update the step's algorithmic bytes per launch (128 / 192 / 192 B
    per zone in stages 1 / 2 / 3) with their fraction of the 8 TB/s peak (and,
    for the collapsar, the SOR kernels: red 28 B and black 32 B per cell, see
    sor_kernel_bytes)."""
    import csv
    from collections import defaultdict
    cfg = make_workload(workload, n)[0]
    groups = defaultdict(list)
    rows_in = list(csv.DictReader(open(path)))

    def grid_of(r):
        return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])

    # the selector's training run (a smaller grid) launches the y/z march
    # kernels with the same thread count as the n^3 run: leave out every
    # dispatch inside the span of the smaller-grid update launches
    upd = [r for r in rows_in if "k_update<" in r["Kernel_Name"]]
    gmax_upd = max((grid_of(r) for r in upd), default=0)
    small = [r for r in upd if grid_of(r) < gmax_upd]
    span = None
    if small:
        s0 = min(int(r["Start_Timestamp"]) for r in small)
        # from the end of the last full-size update before it (the training
        # run's first sweeps precede its first update)
        before = [int(r["End_Timestamp"]) for r in upd if grid_of(r) == gmax_upd and int(r["End_Timestamp"]) < s0]
        span = (max(before) + 1 if before else 0, max(int(r["End_Timestamp"]) for r in small))
    for r in rows_in:
        if span and span[0] <= int(r["Start_Timestamp"]) <= span[1]:
            continue
        grid = grid_of(r)
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        groups[r["Kernel_Name"]].append((grid, dur, int(r["VGPR_Count"]), int(r["Accum_VGPR_Count"])))
    rows = []
    for name, ds in groups.items():
        gmax = max(d[0] for d in ds)
        sel = [d for d in ds if d[0] == gmax]
        durs = [d[1] for d in sel]
        rows.append((sum(durs), name, len(sel), gmax, sum(durs) / len(durs), min(durs), max(durs), sel[0][2], sel[0][3]))
    rows.sort(reverse=True)
    out = ["| kernel | calls | grid (threads) | mean us | min us | max us | VGPR | AGPR | model MB | model GB/s "
           "| algorithmic MB | frac of 8 TB/s |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for tot, name, calls, grid, mean, mn, mx, vg, ag in rows:
        if calls < 2 and tot < 100:
            continue
        w = 4 if "<float" in name else 8
        kb, ab = None, None
        if "k_flux_x<" in name:
            kb = kernel_model_bytes(cfg, "flux_x", w)
        elif "k_flux_yz<" in name:
            kb = kernel_model_bytes(cfg, "flux_y" if name.rstrip(")").split(",")[3].strip().startswith("1") else "flux_z", w)
        elif "k_update<" in name:
            kb = kernel_model_bytes(cfg, "update", w)
            stage = int(name.split("k_update<")[1].split(",")[0])
            if 1 <= stage <= 3:
                ab = STAGE_BYTES_PER_ZONE[stage - 1] * cfg.zones
        elif "k_sor_red" in name or "k_sor_black" in name:
            kb = sor_kernel_bytes(cfg, "red" if "k_sor_red" in name else "black")
        mb = f"{kb / 1e6:.0f}" if kb else ""
        gbs = f"{kb / (mean * 1e-6) / 1e9:.0f}" if kb else ""
        amb = f"{ab / 1e6:.0f}" if ab else ""
        afr = f"{ab / (mean * 1e-6) / (HBM_PEAK_GBS * 1e9):.3f}" if ab else ""
        short = name.replace("cumc3d::", "").replace("void ", "")
        short = short.split("(")[0]
        out.append(f"| `{short}` | {calls} | {grid} | {mean:.1f} | {mn:.1f} | {mx:.1f} | {vg} | {ag} | {mb} | {gbs} "
                   f"| {amb} | {afr} |")
    return "\n".join(out)


def pmc_summary(paths, n: int, workload: str = "turbulence") -> str:
    """Markdown table of rocprofv3 --pmc counter collections (one CSV per
    counter group, all of the same program): per kernel at its largest grid
    (the selector's smaller training run left out, as in trace_summary), the
    mean per dispatch of each counter.  FETCH_SIZE / WRITE_SIZE (KiB) become
    MB through the counters of the run's 1 GiB device copy (bench.py
    --pmc-calibrate); FP64/FP32 flops become GFLOP/s over the dispatch's own
    duration; VALUBusy, VALUUtilization and OccupancyPercent are rocprofv3's
    derived metrics (%)."""
    import csv
    from collections import defaultdict
    cfg = make_workload(workload, n)[0]
    vals = defaultdict(lambda: defaultdict(list))    # kernel -> counter -> values
    durs = defaultdict(list)
    calib = {}

    def grid_of(r):
        return int(r["Grid_Size"])

    for path in paths:
        rows = list(csv.DictReader(open(path)))
        for r in rows:
            if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer") and grid_of(r) >= 131072:
                calib[r["Counter_Name"]] = float(r["Counter_Value"]) / (1 << 20)   # KiB per GiB moved
        upd = [r for r in rows if "k_update<" in r["Kernel_Name"]]
        gmax = max((grid_of(r) for r in upd), default=0)
        small = [r for r in upd if grid_of(r) < gmax]
        span = None
        if small:
            s0 = min(int(r["Start_Timestamp"]) for r in small)
            before = [int(r["End_Timestamp"]) for r in upd if grid_of(r) == gmax and int(r["End_Timestamp"]) < s0]
            span = (max(before) + 1 if before else 0, max(int(r["End_Timestamp"]) for r in small))
        gk = defaultdict(int)
        for r in rows:
            gk[r["Kernel_Name"]] = max(gk[r["Kernel_Name"]], grid_of(r))
        seen = set()
        for r in rows:
            if span and span[0] <= int(r["Start_Timestamp"]) <= span[1]:
                continue
            name = r["Kernel_Name"]
            if grid_of(r) != gk[name] or not name.startswith("void cumc3d::k_"):
                continue
            vals[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
            key = (path, r["Dispatch_Id"])
            if key not in seen:
                seen.add(key)
                durs[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    fs, ws = calib.get("FETCH_SIZE", 1.0), calib.get("WRITE_SIZE", 1.0)

    def mean(v):
        return sum(v) / len(v) if v else None

    out = [f"Calibration (1 GiB device copy): FETCH_SIZE factor {fs:.4f}, WRITE_SIZE factor {ws:.4f} "
           "(KiB counted per KiB moved).", "",
           "| kernel | dispatches per run | mean us (PMC run) | fetch MB | write MB | traffic MB | model MB | algorithmic MB "
           "| VALUBusy % | VALUUtil % | occupancy % | FP64 GFLOP/s | FP32 GFLOP/s | VALU insts (M) | VMEM rd/wr (M) |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    rows_out = []
    for name, cv in vals.items():
        short = name.replace("void cumc3d::", "").split("(")[0]
        d = mean(durs[name]) or 0.0
        fmb = mean(cv.get("FETCH_SIZE", [])) and mean(cv["FETCH_SIZE"]) * 1024 / fs / 1e6
        wmb = mean(cv.get("WRITE_SIZE", [])) and mean(cv["WRITE_SIZE"]) * 1024 / ws / 1e6
        w = 4 if "<float" in name else 8
        kb, ab = None, None
        if "k_flux_x<" in name:
            kb = kernel_model_bytes(cfg, "flux_x", w)
        elif "k_flux_yz<" in name:
            kb = kernel_model_bytes(cfg, "flux_y" if name.split("(")[0].rstrip(">").endswith("1") else "flux_z", w)
        elif "k_update<" in name:
            kb = kernel_model_bytes(cfg, "update", w)
            stg = int(name.split("k_update<")[1].split(",")[0])
            ab = STAGE_BYTES_PER_ZONE[stg - 1] * cfg.zones if 1 <= stg <= 3 else None
        elif "k_edge<" in name:
            kb = kernel_model_bytes(cfg, "edge", w)
        f64 = mean(cv.get("SQ_INSTS_VALU_FLOPS_FP64", []))
        f32 = mean(cv.get("SQ_INSTS_VALU_FLOPS_FP32", []))
        vi = mean(cv.get("SQ_INSTS_VALU", []))
        vr, vw = mean(cv.get("SQ_INSTS_VMEM_RD", [])), mean(cv.get("SQ_INSTS_VMEM_WR", []))

        def f(x, fmt):
            return format(x, fmt) if x is not None else ""
        rows_out.append(((fmb or 0) + (wmb or 0), f"| `{short}` | {len(durs[name]) // len(paths)} | {d * 1e6:.1f} | {f(fmb, '.0f')} "
                         f"| {f(wmb, '.0f')} | {f((fmb or 0) + (wmb or 0) if fmb is not None else None, '.0f')} "
                         f"| {f(kb and kb / 1e6, '.0f')} | {f(ab and ab / 1e6, '.0f')} "
                         f"| {f(mean(cv.get('VALUBusy', [])), '.1f')} | {f(mean(cv.get('VALUUtilization', [])), '.1f')} "
                         f"| {f(mean(cv.get('OccupancyPercent', [])), '.1f')} "
                         f"| {f(f64 / d / 1e9 if f64 is not None and d else None, '.0f')} "
                         f"| {f(f32 / d / 1e9 if f32 is not None and d else None, '.0f')} "
                         f"| {f(vi and vi / 1e6, '.1f')} | {f(vr and vr / 1e6, '.2f')} / {f(vw and vw / 1e6, '.2f')} |"))
    rows_out.sort(key=lambda t: -t[0])
    out += [r for _, r in rows_out]
    return "\n".join(out)


VARIANTS = "fp64,fp32,fp32_sel,fp32_tvdmc,fp32_ppm"


def measure(cfg, st, label, args, variants, sel_w, window, kernel_timing=True, repeats=None):
    """Time each variant (median of `repeats` runs) inside the FP64 window;
    per-kernel times from one more run with HIP events on."""
    seg = window - args.warmup
    if seg < 1:
        raise SystemExit(f"warmup {args.warmup} leaves no step inside the {window}-step finite window")
    repeats = max(1, args.repeats if repeats is None else repeats)
    from cumc3d_amd.solver import ENONFINITE, CumcError
    res, failed = {}, {}
    for var in variants:
        s = make_variant(cfg, st, var, sel_w)
        walls = []
        try:
            for r in range(repeats):
                print(f"[bench] {label}: {var} run {r + 1}: warmup {args.warmup}, timed {args.steps} "
                      f"in segments {segments(args.steps, seg)}", file=sys.stderr, flush=True)
                walls.append(time_variant(s, st, var, args.warmup, args.steps, seg)[0])
                print(f"[bench] {var}: {1e3 * walls[-1] / args.steps:.3f} ms/step", file=sys.stderr, flush=True)
            kt = None
            if kernel_timing:
                _, kt = time_variant(s, st, var, args.warmup, min(args.steps, seg), seg, kernel_timing=True)
        except CumcError as e:
            # a variant whose state goes non-finite inside the window is reported, not timed
            if e.code != ENONFINITE:
                raise
            print(f"[bench] {var}: non-finite inside the window: {e}", file=sys.stderr, flush=True)
            failed[var] = str(e)
            del s
            continue
        frac = selector_fractions(s, cfg) if var == "fp32_sel" else None
        del s
        walls.sort()
        res[var] = {"wall": walls[len(walls) // 2], "walls": walls, "kt": kt, "frac": frac}
    return res, failed


def main(argv=None):
    if argv is None and len(sys.argv) > 2 and sys.argv[1] == "--pmc-summary":
        # bench.py --pmc-summary N WORKLOAD csv...
        print(pmc_summary(sys.argv[4:], int(sys.argv[2]), sys.argv[3]))
        return 0
    if argv is None and len(sys.argv) > 2 and sys.argv[1] == "--trace-summary":
        n = int(sys.argv[3]) if len(sys.argv) > 3 else 256
        print(trace_summary(sys.argv[2], n, sys.argv[4] if len(sys.argv) > 4 else "turbulence"))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed runs of each variant, each from the initial state; "
                         "value and ms_per_step are their median")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", default="turbulence",
                    choices=["turbulence", "ot", "blast", "torus_sph", "torus_cyl", "collapsar"])
    ap.add_argument("--profile-io", action="store_true",
                    help="also time the profile output path (device pack, copy, HDF5 write)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-l1", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the secondary line (3D Orszag-Tang at the same size)")
    ap.add_argument("--train-n", type=int, default=0, help="selector training grid (default n/2)")
    ap.add_argument("--train-iters", type=int, default=600)
    ap.add_argument("--variants", default=VARIANTS)
    ap.add_argument("--window-cap", type=int, default=200,
                    help="longest FP64 run probed for the finite window")
    ap.add_argument("--window", type=int, default=0,
                    help="take the FP64 finite window as given instead of probing it (profiling runs)")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="copy 1 GiB device to device first (the known-size kernel that scales the HBM counters)")
    ap.add_argument("--sel-every", type=int, default=8)
    ap.add_argument("--selector-sweep", action="store_true",
                    help="instead of the bench line: the selector's time and L1 for several label thresholds "
                         "beside the fixed-scheme FP32 controls")
    ap.add_argument("--orbits", type=float, default=0.0,
                    help="instead of the bench line: the cylindrical torus (configuration 5) advanced this many "
                         "orbital periods at r = 10 in FP64, the FP32 split and FP32 + selector, with density L1 "
                         "against FP64 after each orbit")
    ap.add_argument("--orbit-wall", type=float, default=900.0,
                    help="with --orbits: stop (and report the orbits reached) after this many seconds")
    ap.add_argument("--thr-ppm", type=float, default=1e-3)
    ap.add_argument("--thr-weno", type=float, default=1e-2)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one MI355X: the step has no multi-GPU path "
              "(no halo exchange between GPUs); run with --gpus 1.")
        return 2

    if args.workload == "collapsar":
        return bench_collapsar(args)
    if args.selector_sweep:
        return selector_sweep(args)
    if args.orbits > 0:
        return orbit_accuracy(args)
    global SEL_EVERY
    SEL_EVERY = args.sel_every
    if args.pmc_calibrate:
        calibration_copy()
    cfg, st, label = make_workload(args.workload, args.n)
    zones = cfg.zones
    variants = args.variants.split(",")
    train_n = args.train_n or max(16, args.n // 2)

    print(f"[bench] {label}: FP64 finite window", file=sys.stderr, flush=True)
    window = args.window or finite_window(cfg, st, max(args.window_cap, args.warmup + 1))
    print(f"[bench] FP64 state finite for {window} steps", file=sys.stderr, flush=True)

    sel_w, sel_info = None, None
    if "fp32_sel" in variants:
        print(f"[bench] training the selector on a {train_n}^3 run", file=sys.stderr, flush=True)
        sel_w, sel_info = train_selector(train_n, args.workload, args.thr_ppm, args.thr_weno, iters=args.train_iters)
        sel_info["refresh_every"] = SEL_EVERY
        print(f"[bench] selector: {sel_info}", file=sys.stderr, flush=True)

    res, failed = measure(cfg, st, label, args, variants, sel_w, window)
    if "fp32_sel" in variants and "fp32_sel" not in res:
        raise SystemExit(f"fp32_sel went non-finite inside the FP64 window: {failed['fp32_sel']}")

    def rate(v):
        return zones * args.steps / res[v]["wall"]

    main_var = "fp32_sel" if "fp32_sel" in res else variants[-1]
    wall, kt = res[main_var]["wall"], res[main_var]["kt"]
    value = rate(main_var)
    dom = max(KERNELS, key=lambda k: kt[k])
    w = 8 if main_var == "fp64" else 4
    # the update runs once per stage: the step's algorithmic bytes over its three launches
    alg = STEP_BYTES_PER_ZONE * zones / 3.0 if dom == "update" else None
    achieved = alg / (kt[dom] * 1e-3) / 1e9 if alg else None
    traffic = measured_traffic(dom, w, args.workload, args.n)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "kernel": f"k_update (fused update, {'FP32 split' if w == 4 else 'FP64'}), mean of its 3 stage launches"
                if dom == "update" else dom,
                "algorithmic_bytes": alg,
                "algorithmic_basis": "512 B per zone-update per RK3 step (128/192/192 B per zone in stages 1/2/3), "
                                     "averaged over the update's 3 launches",
                "kernel_ms": kt[dom],
                "kernel_model_bytes": kernel_model_bytes(cfg, dom, w),
                "kernel_model_frac": kernel_model_bytes(cfg, dom, w) / (kt[dom] * 1e-3) / (HBM_PEAK_GBS * 1e9),
                "traffic_source": os.path.relpath(traffic_summary(args.workload, args.n), ROOT)
                + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, per launch)" if traffic else None,
                "kernel_ms_all": {v: {k: res[v]["kt"][k] for k in KERNELS + ["selector"]} for v in res if res[v]["kt"]},
                "step_bytes_frac": STEP_BYTES_PER_ZONE * value / (HBM_PEAK_GBS * 1e9)}

    err = None
    if not args.no_l1:
        print(f"[bench] L1 against FP64 after {window} steps (the last finite FP64 state)", file=sys.stderr,
              flush=True)
        err = l1_errors(cfg, st, window, variants, sel_w)

    secondary = None
    if not args.no_secondary and args.workload == "turbulence":
        print("[bench] secondary: Orszag-Tang", file=sys.stderr, flush=True)
        secondary = secondary_line(args)

    io = None
    if args.profile_io:
        print("[bench] profile output path", file=sys.stderr, flush=True)
        io = profile_io(cfg, st, main_var, sel_w)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (reference Fortran)", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(args.workload)
        except Exception as e:  # report, do not hide, a failed baseline
            cpu = {"value": None, "error": str(e)[:200]}

    seg = window - args.warmup
    line = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64 (bf16 selector MLP)",
        "data": f"synthetic (seeded IC, cumc3d_amd.ic.{ {'ot': 'orszag_tang', 'torus_cyl': 'torus'}.get(args.workload, args.workload) })",
        "config": {"workload": label + ", FP32 split + learned selector (FP64 accumulators)", "zones": zones,
                   "riemann": cfg.riemann, "recon": "selector: TVD-MC / PPM / WENO-Z per cell and direction",
                   "fp64_finite_steps": window,
                   "timing": f"each timed run starts from the initial state: {args.warmup} warmup steps, then "
                             f"{args.steps} timed steps in segments {segments(args.steps, seg)} (the FP64 state "
                             f"is finite for {window} steps)"},
        "variants": {v: {"value": rate(v), "ms_per_step": 1e3 * res[v]["wall"] / args.steps,
                         "runs_ms_per_step": [round(1e3 * x / args.steps, 4) for x in res[v]["walls"]]} for v in res},
        "variants_failed": failed or None,
        "repeats": max(1, args.repeats),
        "speedup_vs_fp64": value / rate("fp64") if "fp64" in res else None,
        "selector": dict(sel_info, class_fractions=res["fp32_sel"]["frac"]) if sel_info else None,
        "l1_vs_fp64": err,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if secondary is not None:
        line["secondary"] = secondary
    if io is not None:
        line["output"] = io
    print(json.dumps(line))
    return 0


def selector_sweep(args):
    """The selector's speed/accuracy trade-off on the workload: for each pair
    of label thresholds (thr_ppm, thr_weno) the selector is retrained on the
    half-size run, timed (median of 3 runs of warmup + steps), and its
    density L1 against FP64 measured after 8, 13, 20 and `window` steps with
    FP64's dt sequence, beside the fixed-scheme FP32 controls (TVD-MC, PPM,
    WENO-Z).  Prints one JSON object."""
    global SEL_EVERY
    SEL_EVERY = args.sel_every
    cfg, st, label = make_workload(args.workload, args.n)
    window = args.window or finite_window(cfg, st, max(args.window_cap, args.warmup + 1))
    seg = window - args.warmup
    pairs = [(1e-3, 1e-2), (1e-3, 3e-3), (3e-4, 1e-3), (1e-4, 3e-4), (1.0, 1e-3), (1.0, 3e-4)]
    runs = {v: (v, None) for v in ("fp32_tvdmc", "fp32_ppm", "fp32")}
    info = {}
    for tp, tw in pairs:
        w, inf = train_selector(args.train_n or max(16, args.n // 2), args.workload, tp, tw, iters=args.train_iters)
        key = f"sel_{tp:g}_{tw:g}"
        runs[key] = ("fp32_sel", w)
        info[key] = {"thresholds": [tp, tw], "label_counts": inf["label_counts"]}
    out = {"workload": label, "window": window, "steps": args.steps, "warmup": args.warmup, "variants": {}}
    for key, (var, w) in runs.items():
        s = make_variant(cfg, st, var, w)
        walls = sorted(time_variant(s, st, var, args.warmup, args.steps, seg)[0] for _ in range(3))
        frac = selector_fractions(s, cfg) if var == "fp32_sel" else None
        del s
        out["variants"][key] = dict(info.get(key, {}), ms_per_step=1e3 * walls[1] / args.steps, class_fractions=frac)
        print(f"[bench] {key}: {out['variants'][key]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    a = make_variant(cfg, st, "fp64")
    ss = {k: make_variant(cfg, st, v, w) for k, (v, w) in runs.items()}
    marks = sorted({8, 13, 20, window} & set(range(1, window + 1)))
    for step in range(1, window + 1):
        dt = a.finddt()
        a.rungekutta(dt)
        for b in ss.values():
            b.finddt()
            b.rungekutta(dt)
        if step in marks:
            x = a.state()["rho"][3:-3, 3:-3, 3:-3]
            for k, b in ss.items():
                y = b.state()["rho"][3:-3, 3:-3, 3:-3]
                out["variants"][k].setdefault("rho_l1", {})[step] = float(np.mean(np.abs(x - y)))
    print(json.dumps(out))
    return 0


def orbit_accuracy(args):
    """Configuration 5's accuracy check: the cylindrical torus (n x 2n/3 x n,
    HLLD + PPM, the pressure maximum at r = 10 with GM = 1, so one orbit there
    is T = 2 pi 10^1.5) advanced `args.orbits` orbital periods in FP64 (the
    reference scheme), the FP32 split and FP32 + selector (trained on a
    half-size run of the torus), every variant driven with FP64's dt
    sequence, for at most `args.orbit_wall` seconds.  It prints progress
    every 20 s; after each orbit (and where it stops) it records density L1
    against FP64: absolute over
    the whole grid, relative to the grid's mean density, and relative inside
    the torus (cells where FP64 has rho > 1e-3).  Prints one JSON object."""
    import math
    from cumc3d_amd.solver import CumcError
    global SEL_EVERY
    SEL_EVERY = args.sel_every
    cfg, st, label = make_workload("torus_cyl", args.n)
    t_orbit = 2.0 * math.pi * 10.0 ** 1.5
    t_end = args.orbits * t_orbit
    train_n = args.train_n or max(24, args.n // 2)
    print(f"[bench] {label}: training the selector on the {train_n}-cell run", file=sys.stderr, flush=True)
    sel_w, info = train_selector(train_n, "torus_cyl", args.thr_ppm, args.thr_weno, iters=args.train_iters)
    a = make_variant(cfg, st, "fp64")
    others = {v: make_variant(cfg, st, v, sel_w) for v in ("fp32", "fp32_sel")}
    marks = [min(t_end, (k + 1) * t_orbit) for k in range(int(math.ceil(args.orbits)))]
    out = {"workload": label, "orbits": args.orbits, "t_orbit": t_orbit, "t_end": t_end,
           "selector": {"label_counts": info["label_counts"], "refresh_every": SEL_EVERY}, "checks": []}
    failed = {}
    steps, t0, last_report = 0, time.time(), time.time()
    w = (slice(3, -3),) * 3

    def check():
        x = a.state()["rho"][w]
        torus = x > 1e-3
        row = {"t": a.t, "orbits": a.t / t_orbit, "steps": steps}
        for v, b in others.items():
            if v in failed:
                row[v] = {"error": failed[v]}
                continue
            d = np.abs(b.state()["rho"][w] - x)
            row[v] = {"rho_l1": float(d.mean()), "rho_l1_rel": float(d.mean() / x.mean()),
                      "rho_l1_rel_torus": float(d[torus].mean() / x[torus].mean()),
                      "rho_linf_rel": float(d.max() / x.max())}
            if v == "fp32_sel":
                row[v]["class_fractions"] = selector_fractions(b, cfg)
        out["checks"].append(row)
        print(f"[bench] t = {a.t:.2f} ({a.t / t_orbit:.3f} orbits, {steps} steps): "
              + ", ".join(f"{v} L1 {row[v].get('rho_l1', float('nan')):.3e}" for v in others),
              file=sys.stderr, flush=True)

    try:
        for t_mark in marks:
            while a.t < t_mark and time.time() - t0 < args.orbit_wall:
                dt = min(a.finddt(), t_mark - a.t)
                a.rungekutta(dt)
                steps += 1
                for v, b in others.items():
                    if v in failed:
                        continue
                    try:
                        b.finddt()
                        b.rungekutta(dt)
                    except CumcError as e:
                        failed[v] = f"non-finite at step {steps}: {e}"
                if time.time() - last_report > 20:
                    print(f"[bench] t = {a.t:.2f} ({a.t / t_orbit:.3f} orbits): {steps} steps, dt {dt:.3e}, "
                          f"{time.time() - t0:.0f} s", file=sys.stderr, flush=True)
                    last_report = time.time()
            check()
            if a.t < t_mark:
                out["stopped"] = f"wall-clock limit of {args.orbit_wall:.0f} s reached at t = {a.t:.3f}"
                break
    except CumcError as e:
        out["fp64_error"] = f"FP64 non-finite at step {steps} (t = {a.t:.3f}): {e}"
    out["steps"] = steps
    out["seconds"] = time.time() - t0
    print(json.dumps(out))
    return 0


def secondary_line(args):
    """3D Orszag-Tang at the same size (HLLD + WENO-Z): FP64 and FP32 +
    selector (trained on a half-size run of the vortex), 2 runs each."""
    cfg, st, label = make_workload("ot", args.n)
    window = finite_window(cfg, st, args.warmup + args.steps)
    sel_w, info = train_selector(max(16, args.n // 4), "ot", args.thr_ppm, args.thr_weno, iters=args.train_iters)
    res, _ = measure(cfg, st, label, args, ["fp64", "fp32_sel"], sel_w, window, kernel_timing=False, repeats=2)
    out = {"workload": label, "fp64_finite_steps_probed": window}
    for v in res:
        out[v] = {"value": cfg.zones * args.steps / res[v]["wall"], "ms_per_step": 1e3 * res[v]["wall"] / args.steps}
    out["speedup_vs_fp64"] = out["fp32_sel"]["value"] / out["fp64"]["value"]
    out["selector_label_counts"] = info["label_counts"]
    return out


if __name__ == "__main__":
    sys.exit(main())
