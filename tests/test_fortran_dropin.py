"""The Fortran drop-in: the reference program itself (solver modules, model
hooks, its time loop) with RUNGEKUTTA/FINDDT replaced by
cumc3d_amd/fortran/rungekutta_hip.f90 over the C-ABI, against the reference
fixtures bit for bit.  Needs an MI355X and the prebuilt drop-in binaries
(oracle/_ref/dropin, linked by oracle/build_ref.py).

The same binaries write two profiles through the drop-in print_hydroprofile
(cumc3d_amd/fortran/dataio_hip.f90): rkiter-0 from the host arrays after
INITIAL_UPDATE, rkiter-1 packed on the device after the last step.  Both
must hold exactly the sections dataio.f90:178-233 writes of that state."""
import os
import tempfile

import numpy as np
import pytest

from conftest import check_lrk, check_profile, interior_window, load_golden, zero_signs_equal
from cumc3d_amd import ic
from oracle import oracle as O
from oracle.build_ref import DROPIN_CASES
from oracle.cases import CASES
from oracle.refio import dropin_binary, run_ref

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "bcx", "bcy", "bcz", "cs")


@pytest.mark.parametrize("name", DROPIN_CASES)
def test_dropin_matches_reference(name):
    binary = dropin_binary(name)
    if not os.path.isfile(binary):
        pytest.skip(f"{binary} not built (oracle/build_ref.py needs the reference sources)")
    case = CASES[name]
    gold = load_golden(name)
    print(f"case {name} (Fortran drop-in)", flush=True)
    st = ic.GENERATORS[case.ic](case.cfg)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ic.bin")
        ic.write_ic(p, case.cfg, st)
        got = run_ref(name, p, case.steps, threads=1, timeout=300, binary=binary)
    assert np.array_equal(got["dts"], gold["dts"])
    w = interior_window(case.cfg)
    for f in FIELDS:
        if f not in gold:
            continue
        a, b = got[f][w], gold[f]
        assert np.array_equal(a, b, equal_nan=True), f"{name}: {f} differs, max {np.nanmax(np.abs(a - b)):.3e}"
        assert zero_signs_equal(a, b), f"{name}: {f}: the sign of a zero differs"
    assert got["maxdivb"] == float(gold["maxdivb"])
    if "phi" in gold:   # collapsar: CUSTOM_DEF's phi, m_bh, r_sh synced back from the device
        cfg = case.cfg
        g = got["grav"]
        assert g["m_bh"] == float(gold["m_bh"]) and g["r_sh"] == float(gold["r_sh"])
        assert np.array_equal(g["phi"][2:cfg.nz + 4, 2:cfg.ny + 4, 2:cfg.nx + 4], gold["phi"])
    # CUMC3D_HIP_SPATIAL on the final state into the module's l_rk
    check_lrk(got["lrk"], gold, case.cfg, label=f"{name} (CUMC3D_HIP_SPATIAL): ")
    check_profile(got["profiles"]["rkiter-1-nm.hdf5"], case.cfg, got, got["t"])
    o = O.Oracle(case.cfg, st)         # the state after INITIAL_UPDATE
    check_profile(got["profiles"]["rkiter-0-nm.hdf5"], case.cfg, {f: o.get(f) for f in FIELDS[:9]}, 0.0)



# ---------------------------------------------------------------------------
# the drop-in's arithmetic switches: -DCUMC3D_USE_FP32 (the FP32 split) and
# -DCUMC3D_USE_SELECTOR (the learned selector, parameters from the file named
# by CUMC3D_SELECTOR_PARAMS, refreshed every CUMC3D_SELECTOR_EVERY steps)
# ---------------------------------------------------------------------------
from oracle.build_ref import DROPIN_VARIANT_CASES  # noqa: E402


def python_run(case, st, fp32=True, sel_w=None, every=8):
    from cumc3d_amd.solver import make_solver
    s = make_solver(case.cfg.with_(precision="fp32" if fp32 else "fp64"), st)
    if sel_w is not None:
        s.selector_init(capacity=128, seed=1)
        s.selector_params(sel_w)
        s.selector_enable(True, every=every)
    dts = s.run(case.steps)
    return dts, s.state()


@pytest.mark.parametrize("name", DROPIN_VARIANT_CASES)
def test_dropin_fp32_switch(name):
    """-DCUMC3D_USE_FP32: the reference's own driver steps the FP32 split.  The
    result is bit for bit the Python mirror's FP32 run of the same case (the
    same library and step sequence).  Against the reference's FP64 output
    (the fixture), the Orszag-Tang case keeps the FP32 split's density bound,
    absolute L1 <= 1e-4 on cells 1..n; the turbulence case, whose 11 steps
    end one step before the reference scheme's first non-finite state, only
    has to stay finite."""
    binary = dropin_binary(name, "fp32")
    if not os.path.isfile(binary):
        pytest.skip(f"{binary} not built (oracle/build_ref.py needs the reference sources)")
    case = CASES[name]
    gold = load_golden(name)
    print(f"case {name} (Fortran drop-in, -DCUMC3D_USE_FP32)", flush=True)
    st = ic.GENERATORS[case.ic](case.cfg)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ic.bin")
        ic.write_ic(p, case.cfg, st)
        got = run_ref(name, p, case.steps, threads=1, timeout=300, binary=binary)
    dts, want = python_run(case, st)
    assert np.array_equal(got["dts"], dts)
    inner = (slice(3, -3),) * 3
    for f in FIELDS:
        assert np.array_equal(got[f][inner], want[f][inner]), f
    w1 = tuple(slice(1, -1) for _ in range(3))       # cells 1..n of the fixture window 0..n+1
    rho, ref = got["rho"][inner], gold["rho"][w1]
    assert np.all(np.isfinite(rho))
    l1 = float(np.mean(np.abs(rho - ref)))
    print(f"  density L1 against the reference's FP64: {l1:.2e}", flush=True)
    if name == "ot_hlld_weno":
        assert l1 <= 1e-4


@pytest.mark.parametrize("name", DROPIN_VARIANT_CASES)
def test_dropin_selector_switch(name):
    """-DCUMC3D_USE_FP32 -DCUMC3D_USE_SELECTOR: the drop-in reads the selector's
    parameters from CUMC3D_SELECTOR_PARAMS and refreshes its masks every
    CUMC3D_SELECTOR_EVERY steps.  With a fixed random parameter vector (every
    class occurs) the run is bit for bit the Python mirror's selector run."""
    from cumc3d_amd.solver import SEL_NPARAM
    binary = dropin_binary(name, "fp32sel")
    if not os.path.isfile(binary):
        pytest.skip(f"{binary} not built (oracle/build_ref.py needs the reference sources)")
    case = CASES[name]
    print(f"case {name} (Fortran drop-in, -DCUMC3D_USE_FP32 -DCUMC3D_USE_SELECTOR)", flush=True)
    st = ic.GENERATORS[case.ic](case.cfg)
    w = (np.random.default_rng(11).standard_normal(SEL_NPARAM) * 0.4).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ic.bin")
        ic.write_ic(p, case.cfg, st)
        wpath = os.path.join(d, "sel.bin")
        w.tofile(wpath)
        got = run_ref(name, p, case.steps, threads=1, timeout=300, binary=binary,
                      extra_env={"CUMC3D_SELECTOR_PARAMS": wpath, "CUMC3D_SELECTOR_EVERY": "2"})
    dts, want = python_run(case, st, sel_w=w, every=2)
    assert np.array_equal(got["dts"], dts)
    inner = (slice(3, -3),) * 3
    for f in FIELDS:
        assert np.array_equal(got[f][inner], want[f][inner]), f
