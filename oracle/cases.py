"""Parity cases shared by the oracle recipes and the tests (test infrastructure).

Each case names a configuration, an initial-condition generator and a step
count.  The reference oracle (oracle/_ref/<case>/cumc3d_ref, built by
oracle/build_ref.py from the reference sources) and the C restatement
(oracle/mhd_oracle.c) both run these cases; golden fixtures in tests/golden/
hold the reference's outputs for them.
"""
from __future__ import annotations

import dataclasses
import math
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from cumc3d_amd.config import (Config, blast_3d, briowu_1d, collapsar, orszag_tang_3d,  # noqa: E402
                               torus_cyl, torus_sph, turbulence_3d)


@dataclasses.dataclass
class Case:
    name: str
    cfg: Config
    ic: str
    steps: int


def _ot(n, recon, riemann, bc=(0, 0, 0, 0, 0, 0)):
    return Config(nx=n, ny=n, nz=n, n_dim=3, cfl=0.4, ggas=5.0 / 3.0,
                  riemann=riemann, recon=recon, boundary_flag=bc)


CASES = {c.name: c for c in [
    # config 1 exactly: Brio-Wu, 800 cells, HLLD + TVD-MC, gamma=2, CFL 0.4, t=0.1
    Case("c1_briowu", briowu_1d(800), "briowu", 100000),
    # 3D periodic Orszag-Tang variant (config 2 physics at 12^3), every reconstruction
    Case("ot_hlld_tvdmm", _ot(12, "tvdmm", "hlld"), "orszag_tang", 12),
    Case("ot_hlld_tvdvl", _ot(12, "tvdvl", "hlld"), "orszag_tang", 12),
    Case("ot_hlld_tvdmc", _ot(12, "tvdmc", "hlld"), "orszag_tang", 12),
    Case("ot_hlld_ppm", _ot(12, "ppm", "hlld"), "orszag_tang", 12),
    Case("ot_hlld_weno", _ot(12, "weno", "hlld"), "orszag_tang", 12),
    # every Riemann solver
    Case("ot_lf_ppm", _ot(12, "ppm", "lf"), "orszag_tang", 12),
    Case("ot_hll_ppm", _ot(12, "ppm", "hll"), "orszag_tang", 12),
    Case("ot_hllc_ppm", _ot(12, "ppm", "hllc"), "orszag_tang", 12),
    # outflow and reflecting boundaries in 3D
    Case("ot_bc_mixed", _ot(12, "weno", "hlld", bc=(2, 2, 1, 1, 2, 1)), "orszag_tang", 10),
    # decaying turbulence (config 4 physics at 24^3, HLLD + WENO-Z): the last
    # step before the reference scheme's first non-finite state (step 12)
    Case("c4_turb24_weno", turbulence_3d(24, recon="weno", riemann="hlld"), "turbulence", 11),
    # magnetised blast (config 3 physics at 16^3)
    Case("blast_weno", Config(nx=16, ny=16, nz=16, n_dim=3, cfl=0.4, ggas=5.0 / 3.0,
                              riemann="hlld", recon="weno", boundary_flag=(0,) * 6),
         "blast", 10),
    # 2D (n_dim=2, nz=1) exercises the 2D face-EMF fill (mhd_module.f90:343-356)
    Case("ot2d_hlld_ppm", Config(nx=24, ny=24, nz=1, n_dim=2, cfl=0.4, ggas=5.0 / 3.0,
                                 riemann="hlld", recon="ppm", boundary_flag=(0, 0, 0, 0, 0, 0)),
         "orszag_tang", 20),
    # cylindrical torus with floors, gravity, no-inflow clamps (config 5 physics, small)
    Case("torus_hlld_ppm", Config(nx=24, ny=8, nz=24, n_dim=3, coordinate_flag=1,
                                  x_start=2.0, x_end=26.0, y_start=0.0, y_end=2.0 * math.pi,
                                  z_start=-12.0, z_end=12.0, cfl=0.4, ggas=5.0 / 3.0,
                                  riemann="hlld", recon="ppm", model="torus",
                                  boundary_flag=(1, 1, 0, 0, 1, 1)),
         "torus", 10),
    # 100 steps (the length the FP64 parity criterion is stated for): the
    # 128^3 Orszag-Tang configuration's physics (HLLD + PPM) at 24^3, and the
    # cylindrical torus
    Case("ot24_ppm_100", _ot(24, "ppm", "hlld"), "orszag_tang", 100),
    Case("torus_ppm_100", Config(nx=24, ny=8, nz=24, n_dim=3, coordinate_flag=1,
                                 x_start=2.0, x_end=26.0, y_start=0.0, y_end=2.0 * math.pi,
                                 z_start=-12.0, z_end=12.0, cfl=0.4, ggas=5.0 / 3.0,
                                 riemann="hlld", recon="ppm", model="torus",
                                 boundary_flag=(1, 1, 0, 0, 1, 1)),
         "torus", 100),
    # cylindrical axis (3) and equatorial (4) symmetry flags
    Case("torus_sym_tvdmc", Config(nx=16, ny=8, nz=12, n_dim=3, coordinate_flag=1,
                                   x_start=2.0, x_end=26.0, y_start=0.0, y_end=2.0 * math.pi,
                                   z_start=0.0, z_end=12.0, cfl=0.4, ggas=5.0 / 3.0,
                                   riemann="hllc", recon="tvdmc", model="torus",
                                   boundary_flag=(3, 1, 0, 0, 4, 1)),
         "torus", 8),
    # spherical (r, theta, phi) with FIXPOLE: the torus-sph hooks (gravity,
    # floors, no-inflow clamp) on the full polar range with axis boundaries,
    # in 3D and axisymmetric 2D (one phi cell: the polar EMF average then
    # rounds it to single precision, mhd_module.f90:293,443-449).  WENO-Z
    # with HLLC or HLLD turns the polar cells to NaN within two steps in the
    # reference itself on these grids, so WENO-Z is paired with HLL here.
    Case("sph_torus_ppm", torus_sph(20, 16, 8, recon="ppm"), "torus_sph", 10),
    Case("sph2d_torus_tvdmc", torus_sph(24, 24, recon="tvdmc", n_dim=2), "torus_sph", 20),
    # magnetised blast next to the polar axis (plain hooks), field threading the pole
    Case("sph_blast_weno", Config(nx=16, ny=16, nz=12, n_dim=3, coordinate_flag=2,
                                  x_start=0.4, x_end=2.0, y_start=0.0, y_end=math.pi,
                                  z_start=0.0, z_end=2.0 * math.pi, cfl=0.4, ggas=5.0 / 3.0,
                                  riemann="hll", recon="weno", boundary_flag=(1, 1, 3, 3, 0, 0)),
         "sph_blast", 10),
    Case("sph_blast_100", Config(nx=16, ny=16, nz=12, n_dim=3, coordinate_flag=2,
                                 x_start=0.4, x_end=2.0, y_start=0.0, y_end=math.pi,
                                 z_start=0.0, z_end=2.0 * math.pi, cfl=0.4, ggas=5.0 / 3.0,
                                 riemann="hllc", recon="ppm", boundary_flag=(1, 1, 3, 3, 0, 0)),
         "sph_blast", 100),
    # collapsar: self-gravity (the red-black SOR potential at the initial update
    # and every n_pot steps), black-hole accretion through the inner face and
    # the atmosphere floor, on a logarithmic radial grid the reference reads
    # from ./profile/*_grid.dat.  The reference's SOR, whose periodic phi ghost
    # values lag one sweep behind, diverges with fewer than about 64 phi
    # cells, so the 3D case has 96; axisymmetric 2D converges quickly.
    Case("collapsar2d_ppm", collapsar(24, 16, n_dim=2, recon="ppm", riemann="hlld"), "collapsar", 20),
    Case("collapsar2d_weno_hll", collapsar(24, 16, n_dim=2, recon="weno", riemann="hll"), "collapsar", 20),
    Case("collapsar2d_npot2", collapsar(16, 12, n_dim=2, recon="tvdmc", riemann="hllc").with_(n_pot=2),
         "collapsar", 9),
    Case("collapsar_ppm", collapsar(12, 8, 96, recon="ppm", riemann="hlld"), "collapsar", 12),
]}


# Production-size parity cases: the reference run single-threaded (its HLLD
# carry is per-thread under OpenMP), its state after the last step stored as
# per-field SHA-256 digests plus per-plane sums (tests/golden/large_*.json,
# written by tests/golden/make_large.py), not as arrays.
LARGE_CASES = {c.name: c for c in [
    # configuration 2 exactly: 128^3 Orszag-Tang, HLLD + PPM, FP64
    Case("c2_ot128_ppm", orszag_tang_3d(128, recon="ppm", riemann="hlld"), "orszag_tang", 20),
    # configuration 4's physics (256^3 decaying turbulence) at 128^3, HLLD + WENO-Z
    Case("c4_turb128_weno", turbulence_3d(128, recon="weno", riemann="hlld"), "turbulence", 12),
    # configuration 4's benchmark grid itself (256^3 turbulence, HLLD + WENO-Z)
    Case("c4_turb256_weno", turbulence_3d(256, recon="weno", riemann="hlld"), "turbulence", 5),
    # the other Riemann solvers and limiters at 128^3: HLLC + WENO-Z and LF +
    # van Leer on the vortex, HLL + MC on the magnetised blast (strong shocks)
    Case("ot128_hllc_weno", orszag_tang_3d(128, recon="weno", riemann="hllc"), "orszag_tang", 10),
    Case("ot128_lf_tvdvl", orszag_tang_3d(128, recon="tvdvl", riemann="lf"), "orszag_tang", 10),
    Case("blast128_hll_tvdmc", blast_3d(128, recon="tvdmc", riemann="hll"), "blast", 6),
    # a box of unequal sides that are not multiples of the 64-lane wave
    # (150 x 130 x 110), so rows, march segments and waves do not line up
    Case("ot150x130x110_ppm", orszag_tang_3d(16, recon="ppm").with_(nx=150, ny=130, nz=110), "orszag_tang", 10),
    # 2D at 512^2 (the y march over whole columns, no z sweep) and configuration
    # 1's Brio-Wu tube on 8192 cells (one row of 8192 interfaces, 128 waves)
    Case("ot2d_512_ppm", Config(nx=512, ny=512, nz=1, n_dim=2, cfl=0.4, ggas=5.0 / 3.0,
                                riemann="hlld", recon="ppm", boundary_flag=(0, 0, 0, 0, 0, 0)),
         "orszag_tang", 40),
    Case("briowu_8192", briowu_1d(8192), "briowu", 400),
    # configuration 5's cylindrical torus at half its resolution (96 x 64 x 96),
    # HLLD + PPM: geometric sources, floors, point-mass gravity, no-inflow clamps
    Case("c5_torus96_ppm", torus_cyl(96, 64, 96, recon="ppm", riemann="hlld"), "torus", 40),
    # the spherical torus with FIXPOLE at 128 x 96 x 64, HLLD + PPM
    Case("sph_torus128_ppm", torus_sph(128, 96, 64, recon="ppm"), "torus_sph", 30),
    # and with the scheme of bench.py's spherical-torus workload, HLL + WENO-Z
    Case("sph_torus128_hll_weno", torus_sph(128, 96, 64, recon="weno", riemann="hll"), "torus_sph", 30),
    # the collapsar at the reference's shipped progenitor grid size (156 x 44 x
    # 128): self-gravity solved at the initial update and at step 10
    Case("collapsar156_ppm", collapsar(156, 44, 128, recon="ppm", riemann="hlld"), "collapsar", 12),
]}

# CPU-baseline samples for bench.py: the reference solver on the benchmark
# physics at a grid that one host run covers in about 10-30 s.  They are
# built with -O3 (oracle/build_ref.py), the parity cases with -O2.
BENCH_CASES = {c.name: c for c in [
    # the headline workload's physics and scheme (256^3 decaying turbulence,
    # HLLD + WENO-Z) at 128^3
    Case("bench_turb128", turbulence_3d(128, recon="weno", riemann="hlld"), "turbulence", 2),
    Case("bench_blast128", Config(nx=128, ny=128, nz=128, n_dim=3, cfl=0.4, ggas=5.0 / 3.0,
                                  total_time=0.02, riemann="hlld", recon="weno",
                                  boundary_flag=(0,) * 6), "blast", 3),
    # the collapsar workload of bench.py --workload collapsar (156 x 44 x 128,
    # the reference's shipped progenitor grid size)
    Case("bench_collapsar156", collapsar(156, 44, 128, recon="ppm", riemann="hlld"), "collapsar", 22),
]}

ALL_CASES = {**CASES, **LARGE_CASES, **BENCH_CASES}
