#!/usr/bin/env python3
"""Build the reference oracle binaries from the reference's own Fortran sources.

Recipe (test infrastructure only):
  * the reference sources are compiled where they lie under /root/reference;
    nothing is copied into this repository.  Each case gets a scratch
    directory outside the repository holding symbolic links to the sources and
    the case's generated ``parameter.h`` (the reference's compile-time
    configuration file, solver/mhd/parameter.h), because a Fortran INCLUDE
    searches the including file's directory first;
  * getgrid.f90:133,142,151 use the PGI extension NOT(logical); the flag
    -DNOT=.NOT. maps it to the standard operator without editing the file;
  * main.f90 (HDF5 output) and dataio.f90 are replaced by the harness driver
    oracle/ref_harness/driver.f90, and the model's GET_MODEL by a file-fed
    initial model (oracle/ref_harness/ic_plain.f90 / ic_torus.f90 /
    ic_collapsar.f90); the collapsar's CUSTOM_GRID reads the case's faces
    from ./profile/*_grid.dat, which oracle/refio.py writes next to the run;
  * flang -O2 -ffp-contract=off -fopenmp for the parity cases; the CPU-baseline
    samples (BENCH_CASES) with -O3 -ffp-contract=off -fopenmp.
Outputs: oracle/_ref/<case>/cumc3d_ref.

For the cases in DROPIN_CASES it also links oracle/_ref/dropin/<case>/cumc3d_dropin:
the same reference program with solver/mhd/rungekutta.f90 replaced by the
Fortran drop-in of this repository (cumc3d_amd/fortran/*.f90) over
libcumc3d_hip.so, i.e. the reference's own driver stepping on the GPU, and
writing its profiles through the drop-in print_hydroprofile (libcumc3d_io.so).

usage: python oracle/build_ref.py [case ...]     (default: every case)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from cases import ALL_CASES, BENCH_CASES  # noqa: E402

REF = os.environ.get("CUMC3D_REFERENCE", "/root/reference")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
SCRATCH = os.environ.get("CUMC3D_REF_SCRATCH", "/tmp/cumc3d_ref_build")
OUT = os.path.join(HERE, "_ref")
REPO = os.path.dirname(HERE)
HIPLIB_DIR = os.path.join(REPO, "cumc3d_amd")
DROPIN_SRC = [os.path.join(REPO, "cumc3d_amd", "fortran", f)
              for f in ("cumc3d_hip.f90", "rungekutta_hip.f90", "cumc3d_io.f90", "dataio_hip.f90")]
DROPIN_CASES = ("c1_briowu", "ot_hlld_weno", "ot_bc_mixed", "torus_hlld_ppm", "ot2d_hlld_ppm",
                "sph_torus_ppm", "sph_blast_weno", "collapsar2d_ppm", "collapsar_ppm")
# drop-in variants with the arithmetic switches of rungekutta_hip.f90:
# oracle/_ref/dropin_<variant>/<case>/cumc3d_dropin
DROPIN_VARIANTS = {"fp32": ["-DCUMC3D_USE_FP32"], "fp32sel": ["-DCUMC3D_USE_FP32", "-DCUMC3D_USE_SELECTOR"]}
DROPIN_VARIANT_CASES = ("ot_hlld_weno", "c4_turb24_weno")
FFLAGS = ["-cpp", "-O2", "-ffp-contract=off", "-fopenmp", "-DDIVB", "-DFIXPOLE"]
FFLAGS_BENCH = ["-cpp", "-O3", "-ffp-contract=off", "-fopenmp", "-DDIVB", "-DFIXPOLE"]

# solver/mhd/source.list minus main.f90, dataio.f90 (HDF5) and openacc.f90
# (device data regions only), in module dependency order.
SOLVER = ["definition.f90", "rsolver_module.f90", "mhd_module.f90", "weno_module.f90",
          "ppmc_module.f90", "tvd_module.f90", "boundary.f90", "fromrvetou.f90",
          "getgrid.f90", "initial.f90", "rungekutta.f90", "setup.f90", "spatial.f90",
          "update.f90"]
MODEL_SRC = {
    "riemann": ("model/riemann/src", ["custom.f90", "findpressure.f90", "getepsilon.f90",
                                      "soundspeed.f90"]),
    "torus": ("model/torus-cyn/src", ["def.f90", "custom.f90", "findpressure.f90",
                                      "getepsilon.f90", "soundspeed.f90"]),
    "torus_sph": ("model/torus-sph/src", ["def.f90", "custom.f90", "findpressure.f90",
                                          "getepsilon.f90", "soundspeed.f90"]),
    "collapsar": ("model/collapsar/src", ["def.f90", "custom.f90", "poisson.f90", "findpressure.f90",
                                          "getepsilon.f90", "soundspeed.f90"]),
}
TORUS_MODELS = ("torus", "torus_sph")


def fort_real(v: float) -> str:
    s = repr(float(v))
    if "e" in s:
        m, e = s.split("e")
        return f"{m}D{int(e)}"
    return s + "D0"


def fort_bool(b: bool) -> str:
    return ".true." if b else ".false."


def parameter_h(cfg) -> str:
    with open(os.path.join(HERE, "ref_harness", "parameter.h.in")) as f:
        t = f.read()
    rep = {
        "@COORD@": str(cfg.coordinate_flag), "@NDIM@": str(cfg.n_dim),
        "@BC@": ",".join(str(b) for b in cfg.boundary_flag),
        "@X0@": fort_real(cfg.x_start), "@X1@": fort_real(cfg.x_end),
        "@Y0@": fort_real(cfg.y_start), "@Y1@": fort_real(cfg.y_end),
        "@Z0@": fort_real(cfg.z_start), "@Z1@": fort_real(cfg.z_end),
        "@NX@": str(cfg.nx), "@NY@": str(cfg.ny), "@NZ@": str(cfg.nz),
        "@CFL@": fort_real(cfg.cfl),
        "@LF@": fort_bool(cfg.riemann == "lf"), "@HLL@": fort_bool(cfg.riemann == "hll"),
        "@HLLC@": fort_bool(cfg.riemann == "hllc"), "@HLLD@": fort_bool(cfg.riemann == "hlld"),
        "@TVDMM@": fort_bool(cfg.recon == "tvdmm"), "@TVDMC@": fort_bool(cfg.recon == "tvdmc"),
        "@TVDVL@": fort_bool(cfg.recon == "tvdvl"), "@PPM@": fort_bool(cfg.recon == "ppm"),
        "@WENO@": fort_bool(cfg.recon == "weno"),
    }
    for k, v in rep.items():
        t = t.replace(k, v)
    return t


def torus_param_h(cfg) -> str:
    # Names are those of model/torus-cyn/src/param.h:9-61 (model/torus-sph's
    # param.h declares the same names); custom.f90 uses r_sh (CUSTOM_SOURCE)
    # and refine_grid (CUSTOM_GRID); the rest only feed GET_MODEL, which the
    # file-fed initial model replaces.
    return "\n".join([
        "! torus-cyn model parameters for the oracle build",
        "REAL*8, PARAMETER :: p_beta = 100.0D0",
        "LOGICAL, PARAMETER :: normalize_by_vol = .false.",
        "LOGICAL, PARAMETER :: normalize_by_minbeta = .true.",
        f"REAL*8, PARAMETER :: r_sh = {fort_real(cfg.r_sh)}",
        "REAL*8, PARAMETER :: gamma = (5.0d0/3.0d0)",
        "REAL*8, PARAMETER :: q_grad = 2.0d0",
        "REAL*8, PARAMETER :: s_in = 6.0d0",
        "REAL*8, PARAMETER :: s_max = 10.0d0",
        "LOGICAL, PARAMETER :: refine_grid = .false.",
        "LOGICAL, PARAMETER :: corona = .false.",
        "REAL*8, PARAMETER :: a_corn = 1.0d0",
        "REAL*8, PARAMETER :: a_eta = 1.0d-4",
        "REAL*8, PARAMETER :: rho_cut = 2.0d-1",
        "REAL*8, PARAMETER :: rho_fac = 1.0d-6",
        "REAL*8, PARAMETER :: rho_max = 1.0d0",
        ""])


def collapsar_param_h(cfg) -> str:
    # Names and values of model/collapsar/src/param.h; the Poisson settings
    # (n_pot, relax_max, tolerance) come from the configuration.  The unit
    # and rotation settings only feed GET_MODEL, which the file-fed initial
    # model replaces; the SOR weight omega_weight is a PARAMETER of
    # model/collapsar/src/def.f90:31 (1.90), which the configuration must match.
    if cfg.sor_omega != 1.90:
        raise ValueError("the reference's SOR weight is fixed at 1.90 (def.f90:31)")
    return "\n".join([
        "! collapsar model parameters for the oracle build",
        "REAL*8, PARAMETER :: clight = 2.99792458D10",
        "REAL*8 :: max_alv = 5.0d0*clight",
        "INTEGER, PARAMETER :: rotation_rule = 2",
        "REAL*8 :: s_0 = (5.0D3)*(1.0D5)",
        "REAL*8 :: w_0 = 10.0D0",
        "REAL*8 :: b_0 = 1.0D10",
        "REAL*8 :: v_crit = 0.01D0*clight",
        f"INTEGER, PARAMETER :: n_pot = {cfg.n_pot}",
        f"INTEGER, PARAMETER :: relax_max = {cfg.relax_max}",
        f"REAL*8, PARAMETER :: tolerance = {fort_real(cfg.sor_tolerance)}",
        ""])


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_case(name: str) -> str:
    case = ALL_CASES[name]
    cfg = case.cfg
    work = os.path.join(SCRATCH, name)
    if os.path.isdir(work):
        shutil.rmtree(work)
    os.makedirs(work)
    with open(os.path.join(work, "parameter.h"), "w") as f:
        f.write(parameter_h(cfg))
    mdir, msrc = MODEL_SRC[cfg.model]
    if cfg.model in TORUS_MODELS:
        with open(os.path.join(work, "param.h"), "w") as f:
            f.write(torus_param_h(cfg))
    elif cfg.model == "collapsar":
        with open(os.path.join(work, "param.h"), "w") as f:
            f.write(collapsar_param_h(cfg))
    else:
        # model/riemann/src/custom.f90 includes param.h for test_model; 0 selects
        # no extra gravity source (custom.f90:143).
        with open(os.path.join(work, "param.h"), "w") as f:
            f.write("INTEGER, PARAMETER :: test_model = 0\n")
    fflags = FFLAGS_BENCH if name in BENCH_CASES else FFLAGS
    srcs = [(os.path.join(REF, "solver/mhd", s), s) for s in SOLVER]
    srcs += [(os.path.join(REF, mdir, s), "model_" + s) for s in msrc]
    ic_src = {"torus": "ic_torus.f90", "torus_sph": "ic_torus.f90", "collapsar": "ic_collapsar.f90"}
    harness = ["harness_io.f90", ic_src.get(cfg.model, "ic_plain.f90"), "driver.f90"]
    dflags = ["-DCUMC3D_COLLAPSAR"] if cfg.model == "collapsar" else []
    objs = []
    for path, link in srcs:
        if not os.path.isfile(path):
            raise FileNotFoundError(path)
        os.symlink(path, os.path.join(work, link))
        extra = ["-DNOT=.NOT."] if link == "getgrid.f90" else []
        obj = link.replace(".f90", ".o")
        run([FLANG, *fflags, *extra, "-c", link, "-o", obj], work)
        objs.append(obj)
    for h in harness:
        obj = h.replace(".f90", ".o")
        run([FLANG, *fflags, *(dflags if h == "driver.f90" else []), "-c", os.path.join(HERE, "ref_harness", h),
             "-o", obj], work)
        objs.append(obj)
    outdir = os.path.join(OUT, name)
    os.makedirs(outdir, exist_ok=True)
    libdir = os.path.join(os.path.dirname(os.path.dirname(FLANG)), "lib")
    run([FLANG, "-fopenmp", "-o", os.path.join(outdir, "cumc3d_ref"), *objs,
         f"-Wl,-rpath,{libdir}"], work)
    if all(os.path.exists(os.path.join(HIPLIB_DIR, f)) for f in ("libcumc3d_hip.so", "libcumc3d_io.so")):
        if name in DROPIN_CASES:
            build_dropin(name, cfg, work, objs, libdir, dflags)
        if name in DROPIN_VARIANT_CASES:
            for v, vflags in DROPIN_VARIANTS.items():
                build_dropin(name, cfg, work, objs, libdir, dflags, vflags, "dropin_" + v)
    return outdir


def build_dropin(name, cfg, work, objs, libdir, dflags=(), vflags=(), subdir="dropin"):
    """Link the reference program with the Fortran drop-in instead of rungekutta.f90
    (vflags: the drop-in's arithmetic switches, into oracle/_ref/<subdir>/<case>)."""
    extra = {"torus": ["-DCUMC3D_TORUS"], "torus_sph": ["-DCUMC3D_TORUS", "-DCUMC3D_TORUS_SPH"],
             "collapsar": ["-DCUMC3D_COLLAPSAR"]}.get(cfg.model, []) + list(vflags)
    tag = subdir + "_"
    dobjs = [o for o in objs if o not in ("rungekutta.o", "driver.o")]
    for src in DROPIN_SRC:
        obj = tag + os.path.basename(src).replace(".f90", ".o")
        run([FLANG, *FFLAGS, *extra, "-c", src, "-o", obj], work)
        dobjs.append(obj)
    run([FLANG, *FFLAGS, "-DHIPDROP", *dflags, "-c", os.path.join(HERE, "ref_harness", "driver.f90"),
         "-o", "driver_hipdrop.o"], work)
    dobjs.append("driver_hipdrop.o")
    outdir = os.path.join(OUT, subdir, name)
    os.makedirs(outdir, exist_ok=True)
    # the library is found relative to the binary, so the tree can move
    run([FLANG, "-fopenmp", "-o", os.path.join(outdir, "cumc3d_dropin"), *dobjs,
         f"-L{HIPLIB_DIR}", "-lcumc3d_hip", "-lcumc3d_io", f"-Wl,-rpath,{libdir}",
         "-Wl,-rpath,$ORIGIN/../../../../cumc3d_amd"], work)


def available() -> bool:
    return os.path.isdir(os.path.join(REF, "solver", "mhd")) and os.path.exists(FLANG)


def main(argv):
    names = argv or list(ALL_CASES)
    if not available():
        print("reference sources or flang not found; oracle/_ref not built")
        return 0
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = {ex.submit(build_case, n): n for n in names}
        for fu in cf.as_completed(futs):
            print("built", futs[fu], "->", fu.result())
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
