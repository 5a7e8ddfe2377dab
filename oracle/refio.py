"""Readers and runners for the reference oracle binaries (test infrastructure).

``run_ref`` executes oracle/_ref/<case>/cumc3d_ref (built by oracle/build_ref.py
from the reference's Fortran sources) on an IC file and returns its final state.
The dump layout is written by oracle/ref_harness/driver.f90.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def ref_binary(case: str) -> str:
    return os.path.join(HERE, "_ref", case, "cumc3d_ref")


def have_ref(case: str) -> bool:
    return os.path.isfile(ref_binary(case))


def read_dump(path: str) -> dict:
    with open(path, "rb") as f:
        nx, ny, nz, nstep, nthreads = (int(v) for v in np.fromfile(f, dtype=np.int32, count=5))
        t, maxdivb, seconds = np.fromfile(f, dtype=np.float64, count=3)
        shape = (nz + 6, ny + 6, nx + 6)
        g = int(np.prod(shape))
        prim = np.fromfile(f, dtype=np.float64, count=8 * g).reshape(shape + (8,))
        eps = np.fromfile(f, dtype=np.float64, count=g).reshape(shape)
        bcell = np.fromfile(f, dtype=np.float64, count=3 * g).reshape(shape + (3,))
        cs = np.fromfile(f, dtype=np.float64, count=g).reshape(shape)
        cons = np.fromfile(f, dtype=np.float64, count=8 * g).reshape(shape + (8,))
        dts = np.fromfile(f, dtype=np.float64, count=nstep)
    names = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")
    out = {n: np.ascontiguousarray(prim[..., i]) for i, n in enumerate(names)}
    out.update(eps=eps, cs=cs, bcx=np.ascontiguousarray(bcell[..., 0]),
               bcy=np.ascontiguousarray(bcell[..., 1]), bcz=np.ascontiguousarray(bcell[..., 2]),
               cons=cons, dts=dts, nstep=nstep, t=float(t), maxdivb=float(maxdivb),
               seconds=float(seconds), threads=nthreads)
    return out


GRID_ARRAYS = (("xF", -3, 0), ("yF", -3, 1), ("zF", -3, 2),
               ("dx", -2, 0), ("dy", -2, 1), ("dz", -2, 2), ("x", -2, 0), ("y", -2, 1), ("z", -2, 2),
               ("dx_cb", -1, 0), ("xbar", -1, 0), ("dcose", -1, 1), ("dsine", -1, 1), ("sine", -2, 1),
               ("sinf", -3, 1))


def read_grid(path: str) -> dict:
    """The GETGRID arrays written by the harness (OUT_FILE.grid), each
    returned over -3..n+3 at index j+3 with NaN where the reference's range
    does not reach."""
    with open(path, "rb") as f:
        n3 = [int(v) for v in np.fromfile(f, dtype=np.int32, count=3)]
        out = {}
        for name, lo, ax in GRID_ARRAYS:
            n = n3[ax]
            vals = np.fromfile(f, dtype=np.float64, count=n + 3 - lo + 1)
            a = np.full(n + 7, np.nan)
            a[lo + 3:] = vals
            out[name] = a
    return out


def dropin_binary(case: str, variant: str = "") -> str:
    """The reference program stepping through the Fortran drop-in on the GPU
    (variant "fp32" / "fp32sel": built with the drop-in's -DCUMC3D_USE_FP32 /
    -DCUMC3D_USE_FP32 -DCUMC3D_USE_SELECTOR)."""
    return os.path.join(HERE, "_ref", "dropin_" + variant if variant else "dropin", case, "cumc3d_dropin")


def write_grid_files(d: str, cfg) -> None:
    """The collapsar's CUSTOM_GRID input (model/collapsar/src/custom.f90:68-155):
    ./profile/r_grid.dat, th_grid.dat, phi_grid.dat with the faces -3..n+3 one
    per line, written as shortest round-trip decimals so the list-directed READ
    gets the same doubles."""
    from cumc3d_amd.grid import Grid
    g = Grid(cfg)
    os.makedirs(os.path.join(d, "profile"), exist_ok=True)
    for name, ax in (("r_grid.dat", g.x), ("th_grid.dat", g.y), ("phi_grid.dat", g.z)):
        with open(os.path.join(d, "profile", name), "w") as f:
            f.write("".join(repr(float(v)) + "\n" for v in ax.faces))


def read_grav(path: str, shape) -> dict:
    with open(path, "rb") as f:
        m_bh, r_sh = np.fromfile(f, dtype=np.float64, count=2)
        phi = np.fromfile(f, dtype=np.float64, count=int(np.prod(shape))).reshape(shape)
    return {"m_bh": float(m_bh), "r_sh": float(r_sh), "phi": phi}


def run_ref(case: str, ic_path: str, steps: int, threads: int = 1, timeout: float = 600,
            binary: str | None = None, extra_env: dict | None = None) -> dict:
    ic_path = os.path.abspath(ic_path)
    try:
        from oracle.cases import ALL_CASES
    except ImportError:     # run from oracle/ itself (build_ref.py)
        from cases import ALL_CASES
    cfg = ALL_CASES[case].cfg
    with tempfile.TemporaryDirectory() as d:
        # the program runs in a scratch directory; profiles go to ./outfile
        os.makedirs(os.path.join(d, "outfile"))
        if cfg.model == "collapsar":
            write_grid_files(d, cfg)
        out = os.path.join(d, "dump.bin")
        env = dict(os.environ)
        # Fresh heap memory filled with a fixed byte pattern, so that nothing
        # the program reads before writing can differ between runs.
        env["MALLOC_PERTURB_"] = "255"
        env.update(extra_env or {})
        r = subprocess.run([os.path.abspath(binary or ref_binary(case)), ic_path, out, str(steps), str(threads)],
                           capture_output=True, text=True, timeout=timeout, env=env, cwd=d)
        if r.returncode != 0:
            raise RuntimeError(f"reference oracle failed: {r.stderr[-2000:]}")
        res = read_dump(out)
        if os.path.exists(out + ".grid"):
            res["grid"] = read_grid(out + ".grid")
        if os.path.exists(out + ".grav"):
            res["grav"] = read_grav(out + ".grav", cfg.shape)
        if os.path.exists(out + ".lrk"):
            # l_rk(1:8, -2:n+3, ...) of one more SPATIAL on the final state
            lrk = np.fromfile(out + ".lrk", dtype=np.float64).reshape(tuple(cfg.shape) + (8,))
            res["lrk"] = {f"lrk{i + 1}": np.ascontiguousarray(lrk[..., i]) for i in range(8)}
        if os.path.exists(out + ".divb"):
            # FIND_DIVB's log record, WRITE(999,*) global_time, maxDivB
            with open(out + ".divb") as f:
                res["divb_log"] = f.read()
        profiles = sorted(os.listdir(os.path.join(d, "outfile")))
        if profiles:
            from cumc3d_amd.hydroprofile import read_profile
            res["profiles"] = {p: read_profile(os.path.join(d, "outfile", p)) for p in profiles}
        return res
