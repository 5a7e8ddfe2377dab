"""FP64 parity at production sizes.  Needs an MI355X.

* Production sizes against the reference: the reference Fortran's state
  after the last step of each LARGE_CASES case of oracle/cases.py
  (configuration 2 exactly: 128^3 Orszag-Tang, HLLD + PPM; the vortex on a
  150 x 130 x 110 box and in 2D at 512^2; Brio-Wu on 8192 cells; the vortex with
  HLLC + WENO-Z and LF + van Leer and the magnetised blast with HLL + MC at
  128^3; configuration 4's
  turbulence at 128^3 and at its own 256^3, HLLD + WENO-Z; configuration 5's
  cylindrical torus at
  96 x 64 x 96, HLLD + PPM, with its floors, point-mass gravity and
  no-inflow clamps; the spherical torus with FIXPOLE at 128 x 96 x 64; the
  collapsar with self-gravity at 156 x 44 x 128, its potential and black hole
  included), stored as SHA-256 digests of every field on cells 0..n+1 plus
  per-plane sums (tests/golden/large_*.json, written by
  tests/golden/make_large.py).  The dt sequence, every field, max div B and
  the l_rk of one more SPATIAL must match bit for bit.  At this size the
  sweeps run their full march segment counts, the loads their 32-bit offsets
  across ~150 MB fields and the blocks spread over every XCD.
* 256^3 (configuration 4, HLLD + WENO-Z, FP64) by size-independent
  properties: 20 steps (inside the 26 the reference scheme stays finite)
  leave a finite state, the normalised div B on cells 2..n stays at
  round-off, and a second run from the same initial state is bit-identical.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, interior_window, lrk_mask
from cumc3d_amd import ic
from cumc3d_amd.config import turbulence_3d
from cumc3d_amd.solver import make_solver
from oracle.cases import LARGE_CASES

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "bcx", "bcy", "bcz")


def window(cfg):
    """cells 0..n+1 of every active axis, cell 1 of an inactive one (1D/2D)"""
    return interior_window(cfg)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def first_plane_off(a, planes):
    a = np.ascontiguousarray(a, dtype=np.float64)
    for i, p in enumerate(planes):
        if float(a[i].sum()).hex() != p:
            return i
    return None


@pytest.mark.parametrize("name", list(LARGE_CASES))
def test_gpu_matches_reference_at_production_size(name):
    path = os.path.join(GOLDEN, f"large_{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"fixture large_{name}.json missing")
    gold = json.load(open(path))
    case = LARGE_CASES[name]
    print(f"case {name}: {case.cfg.nx}x{case.cfg.ny}x{case.cfg.nz}, {gold['steps']} steps", flush=True)
    s = make_solver(case.cfg, ic.GENERATORS[case.ic](case.cfg))
    dts = s.run(gold["steps"])
    assert [float(v).hex() for v in dts] == gold["dts"]
    st = s.state()
    w = window(case.cfg)
    bad = []
    for f in FIELDS:
        if sha(st[f][w]) != gold["fields"][f]["sha256"]:
            bad.append((f, first_plane_off(st[f][w], gold["fields"][f]["plane_sums"])))
    assert not bad, f"fields differing (field, first differing z plane of cells 0..n+1): {bad}"
    assert float(s.find_divb()).hex() == gold["maxdivb"]
    if "phi" in gold:   # collapsar: the potential and the black hole after the last step
        g = s.gravity()
        assert float(g["m_bh"]).hex() == gold["m_bh"] and float(g["r_sh"]).hex() == gold["r_sh"]
        assert sha(g["phi"][w]) == gold["phi"]["sha256"], "phi differs"
    lrk = s.spatial()
    bad = [k for k, v in gold["lrk"].items() if sha(lrk[k][w][lrk_mask(case.cfg, int(k[3:]))]) != v["sha256"]]
    assert not bad, f"l_rk differs: {bad}"


def divb_norm(cfg, st):
    bx, by, bz = st["bx"], st["by"], st["bz"]
    h = 1.0 / cfg.nx
    i = (slice(4, -3),) * 3
    d = ((bx[i] - bx[4:-3, 4:-3, 3:-4]) + (by[i] - by[4:-3, 3:-4, 4:-3]) + (bz[i] - bz[3:-4, 4:-3, 4:-3])) / h
    bm = np.mean(np.sqrt(st["bcx"][3:-3, 3:-3, 3:-3] ** 2 + st["bcy"][3:-3, 3:-3, 3:-3] ** 2
                         + st["bcz"][3:-3, 3:-3, 3:-3] ** 2))
    return float(np.max(np.abs(d)) * h / bm)


def test_gpu_turbulence_256_properties():
    cfg = turbulence_3d(256)
    st0 = ic.turbulence(cfg)
    digests = []
    for run in range(2):
        print(f"turbulence 256^3 FP64 run {run + 1}: 20 steps", flush=True)
        s = make_solver(cfg, st0)
        dts = s.run(20)
        assert len(dts) == 20 and np.all(np.isfinite(dts)) and np.all(dts > 0)
        st = s.state()
        del s
        inner = (slice(3, -3),) * 3
        for f in ("rho", "p", "vx", "vy", "vz", "bx", "by", "bz", "eps"):
            assert np.all(np.isfinite(st[f][inner])), f
        assert np.min(st["rho"][inner]) > 0.0 and np.min(st["p"][inner]) > 0.0
        dn = divb_norm(cfg, st)
        print(f"  normalised div B on cells 2..n: {dn:.3e}", flush=True)
        assert dn < 1e-12
        digests.append([dts.tobytes()] + [sha(st[f][inner]) for f in FIELDS])
        del st
    assert digests[0] == digests[1], "two runs from the same state differ"
