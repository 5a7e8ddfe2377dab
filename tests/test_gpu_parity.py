"""FP64 parity of the HIP library with the reference: the reference's own
outputs (golden fixtures written by the reference Fortran) and the C oracle,
bit for bit.  Needs an MI355X."""
import numpy as np
import pytest

from conftest import check_lrk, check_profile, interior_window, load_golden, lrk_mask, zero_signs_equal
from cumc3d_amd import hydroprofile
from cumc3d_amd import ic
from cumc3d_amd.config import Config, collapsar, orszag_tang_3d, torus_cyl, torus_sph
from cumc3d_amd.solver import make_solver
from oracle import oracle as O
from oracle.cases import CASES

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "bcx", "bcy", "bcz")


def assert_same(got, want, w=None, fields=FIELDS, label=""):
    for f in fields:
        a = got[f] if w is None else got[f][w]
        b = want[f] if w is None else want[f][w]
        if not np.array_equal(a, b, equal_nan=True):
            d = np.abs(a - b)
            raise AssertionError(f"{label}{f}: {np.count_nonzero(d)} cells differ, max {np.nanmax(d):.3e}")
        if not zero_signs_equal(a, b):
            raise AssertionError(f"{label}{f}: the sign of a zero differs")


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_matches_reference_fixture(name):
    """Every parity case bit for bit against the reference Fortran's output."""
    case = CASES[name]
    gold = load_golden(name)
    print(f"case {name}", flush=True)
    s = make_solver(case.cfg, ic.GENERATORS[case.ic](case.cfg))
    dts = s.run(case.steps)
    assert np.array_equal(dts, gold["dts"])
    w = interior_window(case.cfg)
    st = s.state()
    assert_same({f: st[f][w] for f in FIELDS}, gold, label=name + ": ")
    assert s.find_divb() == float(gold["maxdivb"])
    if "phi" in gold:   # collapsar: black hole and potential after the last UPDATE
        g = s.gravity()
        assert g["m_bh"] == float(gold["m_bh"]) and g["r_sh"] == float(gold["r_sh"])
        cfg = case.cfg
        box = (slice(2, cfg.nz + 4), slice(2, cfg.ny + 4), slice(2, cfg.nx + 4))
        assert np.array_equal(g["phi"][box], gold["phi"]), "phi differs"
    # SPATIAL once more on the final state (cumc3d_spatial)
    check_lrk(s.spatial(), gold, case.cfg, label=name + ": ")


@pytest.mark.parametrize("name", ["ot_hlld_ppm", "ot2d_hlld_ppm", "torus_hlld_ppm", "sph_torus_ppm",
                                  "collapsar2d_npot2"])
def test_gpu_stages_match_oracle(name):
    """cumc3d_stage: the state after every stage bit for bit against the
    oracle's stages, the three stages in a row against cumc3d_rk3_step, and
    cumc3d_spatial between two stages against the oracle's SPATIAL there."""
    case = CASES[name]
    cfg = case.cfg
    st0 = ic.GENERATORS[case.ic](cfg)
    print(f"case {name} stage by stage", flush=True)
    s, whole, o = make_solver(cfg, st0), make_solver(cfg, st0), O.Oracle(cfg, st0)
    w = interior_window(cfg)
    for step in range(3):
        dt = s.finddt()
        assert dt == o.finddt() and dt == whole.finddt()
        for k in (1, 2, 3):
            s.stage(k, dt)
            o.stage(k, dt)
            assert_same(s.state(), o.state(), w, label=f"step {step} stage {k}: ")
            if step == 2 and k == 1:
                # the carry SPATIAL advances is the same on both sides
                lr, lo = s.spatial(), o.spatial()
                check_lrk(lr, {f: v[w] for f, v in lo.items()}, cfg, label="mid-step SPATIAL: ")
        if step < 2:
            whole.rungekutta(dt)
            assert_same(s.state(), whole.state(), label=f"step {step} vs cumc3d_rk3_step: ")


def test_gpu_stage_order_and_spatial_keeps_the_rest():
    """Stages run 1, 2, 3; a whole step or run cannot start in the middle of
    one.  cumc3d_spatial leaves the elements it does not form as given and
    does not change the state."""
    from cumc3d_amd.solver import CumcError
    case = CASES["ot_hlld_ppm"]
    cfg = case.cfg
    s = make_solver(cfg, ic.GENERATORS[case.ic](cfg))
    dt = s.finddt()
    with pytest.raises(CumcError):
        s.stage(2, dt)
    with pytest.raises(CumcError):
        s.stage(4, dt)
    s.stage(1, dt)
    with pytest.raises(CumcError):
        s.rungekutta(dt)
    with pytest.raises(CumcError):
        s.run(1)
    with pytest.raises(CumcError):
        s.stage(3, dt)
    before = s.state()
    fill = np.full(cfg.shape + (8,), 7.25)
    lr = s.spatial(fill)
    assert_same(s.state(), before, label="state after SPATIAL: ")
    w = interior_window(cfg)
    for i in range(1, 9):
        outside = lr[f"lrk{i}"].copy()
        inside = np.zeros(cfg.shape, dtype=bool)
        inside[w] = lrk_mask(cfg, i)
        assert np.all(outside[~inside] == 7.25), i
        assert np.all(np.isfinite(outside[inside])) and np.any(outside[inside] != 7.25), i
    s.stage(2, dt)
    s.stage(3, dt)
    s.rungekutta(s.finddt())


@pytest.mark.parametrize("threads", [1, 37])
@pytest.mark.parametrize("name", ["ot_hlld_weno", "ot_hlld_tvdmc", "ot_hllc_ppm", "ot_bc_mixed",
                                  "blast_weno", "ot2d_hlld_ppm", "torus_hlld_ppm", "torus_sym_tvdmc",
                                  "sph_torus_ppm", "sph_blast_weno", "collapsar_ppm"])
def test_gpu_long_march_segments(name, threads, monkeypatch):
    """The y/z sweeps with long segments: CUMC3D_MARCH_THREADS=1 gives one
    segment per column (every thread marches the whole column), 37 an uneven
    split; the result must not depend on the segmentation."""
    monkeypatch.setenv("CUMC3D_MARCH_THREADS", str(threads))
    case = CASES[name]
    gold = load_golden(name)
    print(f"case {name} march_threads={threads}", flush=True)
    s = make_solver(case.cfg, ic.GENERATORS[case.ic](case.cfg))
    dts = s.run(case.steps)
    assert np.array_equal(dts, gold["dts"])
    w = interior_window(case.cfg)
    st = s.state()
    assert_same({f: st[f][w] for f in FIELDS}, gold, label=f"{name}/{threads}: ")


@pytest.mark.parametrize("cfg,ic_name,steps", [
    (orszag_tang_3d(32, recon="weno"), "orszag_tang", 4),
    (orszag_tang_3d(24, recon="ppm").with_(nx=40, boundary_flag=(1, 2, 0, 0, 2, 1)), "orszag_tang", 4),
    (torus_cyl(32, 16, 32, recon="ppm"), "torus", 4),
    (torus_sph(40, 32, 16, recon="weno", riemann="hll"), "torus_sph", 4),
    (CASES["sph_blast_weno"].cfg.with_(nx=40, ny=24, nz=20), "sph_blast", 6),
], ids=["ot32_weno", "ot_aniso_ppm", "torus_ppm", "sph_torus_weno_hll", "sph_blast_hll"])
def test_gpu_matches_oracle_larger(cfg, ic_name, steps):
    """Bigger and anisotropic grids (several waves per row, uneven x tiling) against the C oracle."""
    print(f"case {cfg.nx}x{cfg.ny}x{cfg.nz} {cfg.recon}/{cfg.riemann}", flush=True)
    st = ic.GENERATORS[ic_name](cfg)
    s = make_solver(cfg, st)
    o = O.Oracle(cfg, st)
    dg = s.run(steps)
    do = o.run(steps)
    assert np.array_equal(dg, do)
    assert_same(s.state(), o.state(), w=(slice(2, -2),) * 3)


@pytest.mark.parametrize("mode,omode", [("zero", 0), ("field", 1)])
def test_gpu_hlld_bn_modes(mode, omode):
    case = CASES["ot_hlld_ppm"]
    st = ic.GENERATORS[case.ic](case.cfg)
    s = make_solver(case.cfg, st, hlld_bn_mode=mode)
    o = O.Oracle(case.cfg, st, hlld_bn_mode=omode)
    assert np.array_equal(s.run(6), o.run(6))
    assert_same(s.state(), o.state(), w=(slice(2, -2),) * 3)


def test_gpu_host_dt_path_equals_device_loop():
    """finddt + rungekutta from the host equals the device-side time loop."""
    case = CASES["ot_hlld_weno"]
    st = ic.GENERATORS[case.ic](case.cfg)
    a = make_solver(case.cfg, st)
    b = make_solver(case.cfg, st)
    for _ in range(3):
        a.rungekutta(a.finddt())
    b.run(3)
    assert_same(a.state(), b.state())


def test_gpu_run_stops_at_total_time():
    """Steps queued past total_time are no-ops (main.f90:98 loop condition) and the
    last dt is clipped (main.f90:104-106)."""
    cfg = CASES["ot_hlld_ppm"].cfg
    st = ic.orszag_tang(cfg)
    a = make_solver(cfg, st)
    T = 0.011
    dts = a.run(100, total_time=T, sync_each_step=True)
    assert a.t == pytest.approx(T, abs=0)
    b = make_solver(cfg, st)
    b.run(len(dts) + 7, total_time=T, sync_each_step=False)
    assert b.t == a.t
    assert_same(a.state(), b.state())
    o = O.Oracle(cfg, st)
    od = o.run(100, total_time=T)
    assert np.array_equal(od, dts)


def test_gpu_nan_state_is_reported():
    """A NaN cell gives a NaN CFL step: FINDDT and the device loop return
    CUMC3D_ENONFINITE instead of stepping on with the other cells' minimum."""
    from cumc3d_amd.solver import CumcError
    case = CASES["ot_hlld_ppm"]
    st = ic.GENERATORS[case.ic](case.cfg)
    st["rho"][6, 7, 8] = np.nan
    with pytest.raises(CumcError, match="NaN"):
        make_solver(case.cfg, st).finddt()
    with pytest.raises(CumcError, match="NaN"):
        make_solver(case.cfg, st).run(2)
    good = make_solver(case.cfg, ic.GENERATORS[case.ic](case.cfg))
    assert good.finddt() > 0.0


def test_gpu_deterministic():
    case = CASES["ot_hlld_weno"]
    st = ic.GENERATORS[case.ic](case.cfg)
    a = make_solver(case.cfg, st)
    b = make_solver(case.cfg, st)
    a.run(5)
    b.run(5)
    assert_same(a.state(), b.state())


@pytest.mark.parametrize("name,precision", [("ot_hlld_ppm", "fp64"), ("sph_torus_ppm", "fp64"),
                                            ("ot2d_hlld_ppm", "fp64"), ("ot_hlld_ppm", "fp32")])
def test_gpu_profile_file(name, precision, tmp_path):
    """print_hydroprofile through the device pack (cumc3d_profile) and
    libcumc3d_io: the file written for a run holds exactly the state's
    sections, in every precision (the FP32 split widens its single-precision
    primitives exactly)."""
    case = CASES[name]
    print(f"case {name} {precision} profile", flush=True)
    s = make_solver(case.cfg.with_(precision=precision), ic.GENERATORS[case.ic](case.cfg))
    s.run(3)
    path = hydroprofile.write_profile(s, str(tmp_path / "outfile"), 4)
    assert path.endswith("outfile/rkiter-4-nm.hdf5")
    check_profile(hydroprofile.read_profile(path), case.cfg, s.state(), s.t)


# ---------------------------------------------------------------------------
# collapsar: self-gravity and black hole
# ---------------------------------------------------------------------------
def assert_gravity(s, o, cfg):
    g, h = s.gravity(), o.gravity()
    assert g["m_bh"] == h["m_bh"] and g["r_sh"] == h["r_sh"]
    assert g["iterations"] == h["iterations"]
    box = (slice(2, cfg.nz + 4), slice(2, cfg.ny + 4), slice(2, cfg.nx + 4))
    assert np.array_equal(g["phi"][box], h["phi"][box]), "phi differs"


@pytest.mark.parametrize("cfg,steps", [
    (collapsar(48, 32, n_dim=2, recon="weno", riemann="hll"), 21),
    (collapsar(20, 12, 96, recon="ppm", riemann="hlld"), 11),
    (collapsar(24, 16, n_dim=2, recon="tvdmc", riemann="hlld").with_(n_pot=3), 10),
    # odd nx: padded colour slots, and a black SOR block boundary inside a row
    (collapsar(41, 30, n_dim=2, recon="ppm", riemann="hllc"), 8),
], ids=["2d_48x32_weno_hll", "3d_20x12x96_ppm", "2d_npot3_tvdmc", "2d_41x30_odd_ppm_hllc"])
def test_gpu_collapsar_matches_oracle(cfg, steps):
    """Several Poisson solves (the initial update, steps 0, n_pot, ...) and the
    accretion of every stage against the C oracle: state, dt, m_bh, r_sh,
    potential and the sweep count of the last solve."""
    print(f"case collapsar {cfg.nx}x{cfg.ny}x{cfg.nz} {cfg.recon}/{cfg.riemann} n_pot={cfg.n_pot}", flush=True)
    st = ic.collapsar(cfg)
    s = make_solver(cfg, st)
    o = O.Oracle(cfg, st)
    assert_gravity(s, o, cfg)
    assert np.array_equal(s.run(steps), o.run(steps))
    assert_same(s.state(), o.state(), w=(slice(2, -2),) * 3 if cfg.n_dim == 3 else (slice(3, 4), slice(2, -2), slice(2, -2)))
    assert_gravity(s, o, cfg)


@pytest.mark.parametrize("name", ["ot2d_hlld_ppm", "collapsar2d_ppm", "sph2d_torus_tvdmc"])
def test_gpu_ghost_fill_deterministic_on_inactive_axes(name):
    """On the single-cell z axis of a 2D run the periodic ghost chain of the
    reference resolves to cell 1; the one-pass ghost fill reads that cell
    directly, so the whole downloaded box, outer ghost layers included, is the
    same on every run (reading ghosts being written made them vary)."""
    case = CASES[name]
    st = ic.GENERATORS[case.ic](case.cfg)
    a = make_solver(case.cfg, st)
    b = make_solver(case.cfg, st)
    a.run(case.steps)
    b.run(case.steps)
    assert_same(a.state(), b.state())
    za = a.state()
    for f in ("rho", "vy", "by"):     # copied fields (p is re-derived against each cell's own floor)
        for l in (0, 1, 2, 4, 5, 6):             # l = -2..0 and 2..4 copy cell 1 (l index 3)
            assert np.array_equal(za[f][l, 3:-3, 3:-3], za[f][3, 3:-3, 3:-3]), (f, l)


def test_gpu_collapsar_host_step_path():
    """finddt + rungekutta from the host (the drop-in's path) count steps for
    the n_pot schedule as the device loop does."""
    case = CASES["collapsar2d_npot2"]
    st = ic.GENERATORS[case.ic](case.cfg)
    a = make_solver(case.cfg, st)
    b = make_solver(case.cfg, st)
    for _ in range(5):
        a.rungekutta(a.finddt())
    b.run(2)
    b.run(3)
    assert_same(a.state(), b.state())
    ga, gb = a.gravity(), b.gravity()
    assert ga["m_bh"] == gb["m_bh"] and np.array_equal(ga["phi"], gb["phi"])


def test_gpu_collapsar_stops_where_the_reference_stops():
    """The two STOPs of CUSTOM_UPDATE come back as CUMC3D_ESTOP: the horizon
    at the inner face (custom.f90:482-484) and a solve converging on its
    last allowed sweep (:644-647); a solve that runs out goes on."""
    from cumc3d_amd.solver import CumcError, Solver
    cfg = collapsar(16, 12, n_dim=2)
    st = ic.collapsar(cfg)
    big = dict(st, gravity=st["gravity"].copy())
    big["gravity"][9] = 0.5 * cfg.faces[0][3] * (ic.CLIGHT * st["gravity"][10]) ** 2 * 1.0001
    s = Solver(cfg)
    s.initial_model(big)
    with pytest.raises(CumcError, match="event horizon"):
        s.initial_update()
    with pytest.raises(CumcError):
        s.rungekutta(1.0)        # no step after the stop
    s = Solver(cfg.with_(relax_max=1))
    s.initial_model(st)
    with pytest.raises(CumcError, match="poisson"):
        s.initial_update()
    s = make_solver(cfg.with_(relax_max=3), st)
    s.run(1)
    assert s.gravity()["iterations"] == 4
