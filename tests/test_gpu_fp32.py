"""The FP32 precision split on the GPU: single-precision reconstruction,
Riemann solver and EMFs with double-precision RK accumulators and
total-energy-to-pressure recovery, judged against the FP64 mode (which is
itself bit-exact with the reference).  Needs an MI355X.

Tolerances (stated here, checked below): after the listed steps with the same
dt sequence, the density L1 error against FP64 is at most 1e-4 (the bound the
FP32 split is specified to), and as a tighter regression guard the relative
L1 at most 2e-5 and the relative Linf at most 2e-3 (limiter schemes such as
TVD-MC amplify rounding differences at their branch points, so they sit an
order of magnitude above WENO-Z and PPM); div B of the face field on cells
2..n at FP64 round-off (normalised by <|B|>/h, <= 1e-12).  Cells 1 are excluded from div B because
the reference keeps the lowest face of every axis frozen (fromrvetou.f90:201-207),
which both precisions reproduce."""
import numpy as np
import pytest

from cumc3d_amd import ic
from cumc3d_amd.config import blast_3d, orszag_tang_3d, torus_cyl, torus_sph
from cumc3d_amd.solver import make_solver

pytestmark = pytest.mark.gpu

L1_ABS_TOL = 1e-4
L1_TOL = 2e-5
LINF_TOL = 2e-3


def run_pair(cfg, ic_name, steps):
    st = ic.GENERATORS[ic_name](cfg)
    a = make_solver(cfg, st)
    b = make_solver(cfg.with_(precision="fp32"), st)
    dts = []
    for _ in range(steps):
        dt = a.finddt()
        dts.append(dt)
        a.rungekutta(dt)
        b.finddt()
        b.rungekutta(dt)      # the FP64 dt sequence, so both states sit at the same time
    return a, b, dts


def interior(x):
    return x[3:-3, 3:-3, 3:-3]


def rel_errors(a, b, f):
    x, y = interior(a.state()[f]), interior(b.state()[f])
    scale = np.mean(np.abs(x))
    return np.mean(np.abs(x - y)) / scale, np.max(np.abs(x - y)) / np.max(np.abs(x))


def abs_l1(a, b, f):
    x, y = interior(a.state()[f]), interior(b.state()[f])
    return np.mean(np.abs(x - y))


def divb_inner(s, cfg):
    """max |div B| over cells 2..n (Cartesian, uniform grid), normalised by <|B|>/h."""
    st = s.state()
    bx, by, bz = st["bx"], st["by"], st["bz"]
    h = 1.0 / cfg.nx
    sl = (slice(4, -3),) * 3
    d = (bx[sl] - bx[4:-3, 4:-3, 3:-4]) + (by[sl] - by[4:-3, 3:-4, 4:-3]) + (bz[sl] - bz[3:-4, 4:-3, 4:-3])
    bmag = np.mean(np.sqrt(interior(st["bcx"]) ** 2 + interior(st["bcy"]) ** 2 + interior(st["bcz"]) ** 2))
    return np.max(np.abs(d)) / h / (bmag / h)


@pytest.mark.parametrize("cfg,ic_name,steps", [
    (orszag_tang_3d(32, recon="weno"), "orszag_tang", 10),
    (orszag_tang_3d(24, recon="ppm", riemann="hllc"), "orszag_tang", 10),
    (orszag_tang_3d(24, recon="tvdmc"), "orszag_tang", 10),
    (orszag_tang_3d(24, recon="tvdmm"), "orszag_tang", 10),
    (orszag_tang_3d(24, recon="tvdvl", riemann="hll"), "orszag_tang", 10),
    (blast_3d(32), "blast", 4),
], ids=["ot32_weno_hlld", "ot24_ppm_hllc", "ot24_tvdmc_hlld", "ot24_tvdmm_hlld", "ot24_tvdvl_hll", "blast32_weno_hlld"])
def test_fp32_close_to_fp64(cfg, ic_name, steps):
    print(f"case {cfg.nx}^3 {cfg.recon}/{cfg.riemann} fp32 vs fp64", flush=True)
    a, b, _ = run_pair(cfg, ic_name, steps)
    for f in ("rho", "p", "vx", "by"):
        l1, linf = rel_errors(a, b, f)
        print(f"  {f}: rel L1 {l1:.2e}  rel Linf {linf:.2e}", flush=True)
        if f == "rho":
            assert abs_l1(a, b, f) <= L1_ABS_TOL
            assert l1 <= L1_TOL and linf <= LINF_TOL, (l1, linf)
    assert divb_inner(b, cfg) <= 1e-12


def test_fp32_torus_runs_close():
    cfg = torus_cyl(32, 16, 32, recon="ppm")
    print("case torus 32x16x32 fp32 vs fp64", flush=True)
    a, b, _ = run_pair(cfg, "torus", 4)
    l1, linf = rel_errors(a, b, "rho")
    print(f"  rho: rel L1 {l1:.2e}  rel Linf {linf:.2e}", flush=True)
    assert l1 <= L1_TOL


def test_fp32_configuration5_grid():
    """Configuration 5's own grid, the cylindrical torus at 192 x 128 x 192
    (HLLD + PPM): 200 steps of the FP32 split against FP64, with FP64's dt
    sequence.  The density L1 bound is 1e-4; measured 6.7e-8, so 1e-6 guards
    against a regression."""
    cfg = torus_cyl(192, 128, 192, recon="ppm")
    print("case torus 192x128x192 fp32 vs fp64, 200 steps", flush=True)
    a, b, _ = run_pair(cfg, "torus", 200)
    l1a = abs_l1(a, b, "rho")
    l1, linf = rel_errors(a, b, "rho")
    print(f"  rho: L1 {l1a:.2e}  rel L1 {l1:.2e}  rel Linf {linf:.2e}", flush=True)
    assert l1a <= L1_ABS_TOL and l1a <= 1e-6


def test_fp32_uniform_state_is_stationary():
    """A uniform moving state has zero flux divergence in any precision: every
    cell does the same arithmetic, so the state stays exactly uniform, and it
    changes by no more than the rounding of the RK combine.  The field is zero:
    with a uniform field the reference's frozen lowest face (fromrvetou.f90:201-207)
    keeps its initial value while the RK combine (1/3, 2/3 weights) moves the
    other faces by an ulp, which makes cells 1 differ at round-off in FP64
    too."""
    cfg = orszag_tang_3d(16, recon="weno").with_(precision="fp32")
    st = ic.orszag_tang(cfg)
    for f in ("rho", "p", "eps"):
        st[f][...] = {"rho": 1.0, "p": 0.6, "eps": 0.6 / (2.0 / 3.0)}[f]
    st["vx"][...], st["vy"][...], st["vz"][...] = 0.3, -0.2, 0.1
    st["bx"][...], st["by"][...], st["bz"][...] = 0.0, 0.0, 0.0
    s = make_solver(cfg, st)
    before = s.state()
    s.run(3)
    after = s.state()
    for f in ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz"):
        x, y = interior(before[f]), interior(after[f])
        assert np.ptp(y) == 0.0, f
        assert np.max(np.abs(y - x)) <= 2e-7 * np.max(np.abs(x)), f


def test_fp32_deterministic():
    cfg = orszag_tang_3d(24, recon="weno").with_(precision="fp32")
    st = ic.orszag_tang(cfg)
    a = make_solver(cfg, st)
    b = make_solver(cfg, st)
    da, db = a.run(5), b.run(5)
    assert np.array_equal(da, db)
    for f in ("rho", "p", "bx"):
        assert np.array_equal(a.state()[f], b.state()[f])


def test_fp32_spherical_torus_runs_close():
    """Spherical geometry with FIXPOLE in the FP32 split (the polar EMF
    average accumulates in single precision in both modes)."""
    cfg = torus_sph(32, 24, 16, recon="ppm")
    print("case spherical torus 32x24x16 fp32 vs fp64", flush=True)
    a, b, _ = run_pair(cfg, "torus_sph", 4)
    l1, linf = rel_errors(a, b, "rho")
    print(f"  rho: rel L1 {l1:.2e}  rel Linf {linf:.2e}", flush=True)
    assert abs_l1(a, b, "rho") <= L1_ABS_TOL
    assert l1 <= L1_TOL


def test_fp32_stages_and_spatial():
    """cumc3d_stage and cumc3d_spatial in the FP32 split: three stages give
    exactly cumc3d_rk3_step (same kernels), and SPATIAL's l_rk of the initial
    state is within single-precision rounding of the FP64 mode's: 1e-4 of the
    largest value of the variable (of the vector, for momentum and field),
    since the FP32 fluxes are rounded to float before they are differenced."""
    cfg = orszag_tang_3d(24, recon="weno")
    st = ic.orszag_tang(cfg)
    s32 = make_solver(cfg.with_(precision="fp32"), st)
    whole = make_solver(cfg.with_(precision="fp32"), st)
    s64 = make_solver(cfg, st)
    l32, l64 = s32.spatial(), s64.spatial()
    whole.spatial()   # SPATIAL advances the HLLD carry; keep both FP32 runs in step
    w = (slice(3, -3),) * 3
    # vector components against the largest component of their vector
    # (B_z of this vortex, and so its l_rk, is zero to round-off)
    groups = {1: (1,), 2: (2, 3, 4), 3: (2, 3, 4), 4: (2, 3, 4), 5: (5,), 6: (6, 7, 8), 7: (6, 7, 8), 8: (6, 7, 8)}
    for i in range(1, 9):
        a, b = l32[f"lrk{i}"][w], l64[f"lrk{i}"][w]
        scale = max(np.max(np.abs(l64[f"lrk{g}"][w])) for g in groups[i])
        assert scale > 0 and np.max(np.abs(a - b)) <= 1e-4 * scale, (i, np.max(np.abs(a - b)) / scale)
    for _ in range(2):
        dt = s32.finddt()
        assert dt == whole.finddt()
        for k in (1, 2, 3):
            s32.stage(k, dt)
        whole.rungekutta(dt)
    x, y = s32.state(), whole.state()
    for f in ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps"):
        assert np.array_equal(x[f], y[f]), f
