#!/usr/bin/env python3
"""Write the golden fixtures: outputs of the reference solver on the parity cases.

Runs oracle/_ref/<case>/cumc3d_ref (the reference Fortran compiled from its own
sources by oracle/build_ref.py) single-threaded with zero-filled allocations
(oracle/refio.run_ref) on each case of oracle/cases.py and stores, per case,
the dt sequence and the final state on cells 0..n+1 of every axis:
rho, vx, vy, vz, p, bx, by, bz (face), eps, bcx, bcy, bcz; lrk1..lrk8, the
l_rk of one more SPATIAL on the final state (the harness calls it after its
dump), on the same cells; for spherical
cases also the GETGRID metric arrays the step reads (grid_*); for the
collapsar also m_bh, r_sh and the potential phi on cells 0..n+1.

usage: python tests/golden/make_golden.py [case ...]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cumc3d_amd import ic  # noqa: E402
from oracle.cases import CASES  # noqa: E402
from oracle.refio import have_ref, run_ref  # noqa: E402

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "bcx", "bcy", "bcz")


def window(cfg):
    """cells 0..n+1 on the active axes, cell 1 on inactive ones"""
    sl = []
    for n, active in ((cfg.nz, cfg.n_dim > 2), (cfg.ny, cfg.n_dim > 1), (cfg.nx, True)):
        sl.append(slice(2, n + 4) if active else slice(3, 4))
    return tuple(sl)


def main(names):
    for name in names or list(CASES):
        case = CASES[name]
        if not have_ref(name):
            print("skip", name, "(reference binary not built)")
            continue
        st = ic.GENERATORS[case.ic](case.cfg)
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "ic.bin")
            ic.write_ic(p, case.cfg, st)
            ref = run_ref(name, p, case.steps, threads=1)
        w = window(case.cfg)
        out = {f: ref[f][w] for f in FIELDS}
        out["dts"] = ref["dts"]
        out["t"] = np.array(ref["t"])
        out["maxdivb"] = np.array(ref["maxdivb"])
        if case.cfg.coordinate_flag == 2:
            # the reference's GETGRID arrays (after FIXPOLE), -3..n+3 at j+3
            for k, v in ref["grid"].items():
                out["grid_" + k] = v
        if "lrk" in ref:
            # l_rk of one more SPATIAL on the final state, same cells
            for k, v in ref["lrk"].items():
                out[k] = v[w]
        if "grav" in ref:
            cfg = case.cfg
            out["m_bh"] = np.array(ref["grav"]["m_bh"])
            out["r_sh"] = np.array(ref["grav"]["r_sh"])
            # phi on cells 0..n+1 of every axis (the solved box, custom.f90:497-631)
            out["phi"] = ref["grav"]["phi"][2:cfg.nz + 4, 2:cfg.ny + 4, 2:cfg.nx + 4]
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {ref['nstep']} steps -> {path} ({os.path.getsize(path) // 1024} KB)")


if __name__ == "__main__":
    main(sys.argv[1:])
