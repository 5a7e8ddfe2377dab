#!/usr/bin/env python3
"""Write tests/golden/listdirected.json: how flang's runtime prints REAL*8
values with list-directed output (the format of FIND_DIVB's log record,
solver/mhd/mhd_module.f90:199).

1. A probe program of this repository's own (below, not reference code) is
   compiled with the image's flang; it reads bit patterns and writes each
   value with WRITE(*,*).  2000 seeded values: uniform, log-uniform over
   1e-30..1e30, random bit patterns, ratios of integers, and edge cases
   (zeros, subnormals, powers of ten, 0.1-boundary values).
2. Where oracle/_ref is built, the reference harness's own FIND_DIVB records
   (one per case) with the global_time and maxDivB of its dump.

usage: python tests/golden/make_listdirected.py
"""
import json
import os
import random
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

PROBE = """program listdirected_probe
  implicit none
  integer*8 :: bits
  real*8 :: v
  integer :: ios
  do
    read(*,*,iostat=ios) bits
    if (ios /= 0) exit
    v = transfer(bits, v)
    write(*,*) v
  end do
end program
"""


def bits_of(v):
    return struct.unpack("<q", struct.pack("<d", v))[0]


def values():
    rng = random.Random(20261202)
    edge = [0.0, -0.0, 1.0, -1.0, 0.1, 0.09, 0.095, 0.099, 0.0999999999999999, 0.05, 10.0, 1e14, 1e15, 1e16,
            999999999999999.9, 123456789012345.6, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
            0.1 * 3.0, 2.0 / 3.0 * 0.1, 2.0 ** 60, 6.02e23, -3.25e-7, 12345.678]
    out = list(edge)
    while len(out) < 2000:
        k = len(out) % 4
        if k == 0:
            v = rng.random()
        elif k == 1:
            v = rng.uniform(-1.0, 1.0) * 10.0 ** rng.randint(-30, 30)
        elif k == 2:
            v = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(63)))[0]
        else:
            v = rng.randint(1, 10 ** 6) / rng.randint(1, 10 ** 6)
        if v == v and abs(v) != float("inf"):
            out.append(v)
    return out


def probe(vals):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.f90")
        exe = os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write(PROBE)
        libdir = os.path.join(os.path.dirname(os.path.dirname(FLANG)), "lib")
        subprocess.run([FLANG, "-O0", src, "-o", exe, f"-Wl,-rpath,{libdir}"], check=True)
        inp = "\n".join(str(bits_of(v)) for v in vals) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()
    assert len(lines) == len(vals)
    return lines


def reference_divb_records():
    from cumc3d_amd import ic
    from oracle.cases import CASES
    from oracle.refio import have_ref, run_ref
    recs = []
    for name in ("c1_briowu", "ot_hlld_ppm", "torus_hlld_ppm", "sph_blast_100", "blast_weno"):
        if not have_ref(name):
            continue
        case = CASES[name]
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "ic.bin")
            ic.write_ic(p, case.cfg, ic.GENERATORS[case.ic](case.cfg))
            ref = run_ref(name, p, 5)
        recs.append({"case": name, "time_bits": bits_of(ref["t"]), "maxdivb_bits": bits_of(ref["maxdivb"]),
                     "record": ref["divb_log"]})
    return recs


def main():
    vals = values()
    lines = probe(vals)
    data = {"flang": FLANG, "values": [[bits_of(v), line] for v, line in zip(vals, lines)],
            "divb_records": reference_divb_records()}
    path = os.path.join(HERE, "listdirected.json")
    with open(path, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print(f"{len(vals)} values, {len(data['divb_records'])} reference records -> {path}")


if __name__ == "__main__":
    main()
