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

from conftest import check_lrk, check_profile, interior_window, load_golden
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

