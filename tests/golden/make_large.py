#!/usr/bin/env python3
"""Write the production-size parity fixtures: the reference solver's state on
the LARGE_CASES of oracle/cases.py (128^3 grids), stored as digests.

Runs oracle/_ref/<case>/cumc3d_ref (the reference Fortran compiled from its own
sources by oracle/build_ref.py) single-threaded (oracle/refio.run_ref) and
stores in tests/golden/large_<case>.json:
  * the dt sequence (hex floats, exact);
  * t and max div B after the last step (hex floats);
  * per field (rho, vx, vy, vz, p, bx, by, bz, eps, bcx, bcy, bcz), on cells
    0..n+1 of every axis as C-order float64: the SHA-256 of the bytes, and the
    sum over each z plane (hex floats, to locate a difference);
  * for the collapsar, the potential phi on cells 0..n+1 (digests) and the
    black hole's m_bh and r_sh (hex floats);
  * the same digests of l_rk from one more SPATIAL on the final state, over
    the elements the library forms (conftest.lrk_mask: l_rk(1:5) on cells
    1..n, l_rk(6:8) on the faces the update advances), flattened.
The arrays themselves (hundreds of MB at 128^3) are not stored.

usage: python tests/golden/make_large.py [case ...]
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from conftest import interior_window, lrk_mask  # noqa: E402
from cumc3d_amd import ic  # noqa: E402
from oracle.cases import LARGE_CASES  # noqa: E402
from oracle.refio import have_ref, run_ref  # noqa: E402

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "bcx", "bcy", "bcz")


def window(cfg):
    """cells 0..n+1 of every active axis, cell 1 of an inactive one (1D/2D)"""
    return interior_window(cfg)


def digest(a: np.ndarray) -> dict:
    """SHA-256 of the C-order float64 bytes and the sum over each leading-axis
    plane (of blocks of 4096 elements for a flattened array)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    parts = a if a.ndim > 1 else [a[i:i + 4096] for i in range(0, a.size, 4096)]
    planes = [float(p.sum()).hex() for p in parts]
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "plane_sums": planes}


def main(names):
    for name in names or list(LARGE_CASES):
        case = LARGE_CASES[name]
        if not have_ref(name):
            print("skip", name, "(reference binary not built)")
            continue
        st = ic.GENERATORS[case.ic](case.cfg)
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "ic.bin")
            ic.write_ic(p, case.cfg, st)
            ref = run_ref(name, p, case.steps, threads=1, timeout=4 * 3600)
        w = window(case.cfg)
        out = {"case": name, "steps": int(ref["nstep"]), "dts": [float(v).hex() for v in ref["dts"]],
               "t": float(ref["t"]).hex(), "maxdivb": float(ref["maxdivb"]).hex(),
               "fields": {f: digest(ref[f][w]) for f in FIELDS},
               "lrk": {k: digest(v[w][lrk_mask(case.cfg, int(k[3:]))]) for k, v in ref["lrk"].items()},
               "finite": bool(all(np.all(np.isfinite(ref[f][w])) for f in FIELDS)),
               **({"phi": digest(ref["grav"]["phi"][w]), "m_bh": float(ref["grav"]["m_bh"]).hex(),
                   "r_sh": float(ref["grav"]["r_sh"]).hex()} if "grav" in ref else {}),
               "reference_seconds": float(ref["seconds"])}
        path = os.path.join(HERE, f"large_{name}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=0)
        print(f"{name}: {ref['nstep']} steps, finite={out['finite']} -> {path} ({os.path.getsize(path) // 1024} KB)")


if __name__ == "__main__":
    main(sys.argv[1:])
