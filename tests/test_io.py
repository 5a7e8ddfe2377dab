"""The output files without a GPU: libcumc3d_io.so's HDF5 profile layout
(print_hydroprofile, solver/mhd/dataio.f90:7-249) and flang's list-directed
REAL*8 format of FIND_DIVB's log record (solver/mhd/mhd_module.f90:199).

Pins:
* the list-directed format against 2000 values printed by flang's own runtime
  and against the records the reference harness itself wrote
  (tests/golden/listdirected.json, tests/golden/make_listdirected.py);
* the profile layout against print_hydroprofile's dataset list and array
  sections, read back through the HDF5 library and cross-read by the HDF5
  tools (h5ls), and against the reading convention of the reference's own
  analysis script (analysis/plot/plotprofile-1d.py:39-63: primitive.T[v, j, k, l])."""
import json
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from cumc3d_amd import hydroprofile as H
from cumc3d_amd import ic
from cumc3d_amd.config import Config
from cumc3d_amd.grid import Grid
from oracle import oracle as O
from oracle.cases import CASES

HEADER = os.path.join(ROOT, "include", "cumc3d_io.h")


def d_of(bits):
    return struct.unpack("<d", struct.pack("<q", bits))[0]


def test_io_library_exports_every_declared_symbol():
    L = H.load_io()
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(cumc3d_io_[a-z_0-9]+)\s*\(", text)))
    assert set(names) == set(H.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def test_list_directed_real_matches_flang():
    data = json.load(open(os.path.join(GOLDEN, "listdirected.json")))
    bad = [(b, t) for b, t in data["values"] if " " + H.format_real(d_of(b)) != t]
    assert len(data["values"]) == 2000
    assert not bad, bad[:5]


def test_divb_record_matches_reference_harness():
    data = json.load(open(os.path.join(GOLDEN, "listdirected.json")))
    assert data["divb_records"], "no reference records in the fixture"
    for r in data["divb_records"]:
        assert H.divb_record(d_of(r["time_bits"]), d_of(r["maxdivb_bits"])) == r["record"], r["case"]


def test_divb_log_appends(tmp_path):
    p = str(tmp_path / "divb.dat")
    H.append_divb(p, 0.0, 0.0, truncate=True)
    H.append_divb(p, 0.25, 1.5e-14)
    assert open(p).read() == " 0. 0.\n .25 1.5E-14\n"


def test_profile_name():
    assert H.profile_name("./outfile", 7) == "./outfile/rkiter-7-nm.hdf5"


def sections(cfg, st):
    """print_hydroprofile's arrays of a ghosted-box state in its order:
    built element by element from the Fortran section ranges."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    names = ("rho", "vx", "vy", "vz", "p")
    prim = np.empty(5 * (nx + 2) * (ny + 2) * (nz + 2))
    eps = np.empty((nx + 2) * (ny + 2) * (nz + 2))
    bf = np.empty(3 * (nx + 3) * (ny + 3) * (nz + 3))
    i = 0
    for l in range(0, nz + 2):
        for k in range(0, ny + 2):
            for j in range(0, nx + 2):
                for v in range(5):
                    prim[i] = st[names[v]][l + 2, k + 2, j + 2]
                    i += 1
    i = 0
    for l in range(0, nz + 2):
        for k in range(0, ny + 2):
            for j in range(0, nx + 2):
                eps[i] = st["eps"][l + 2, k + 2, j + 2]
                i += 1
    i = 0
    for l in range(-1, nz + 2):
        for k in range(-1, ny + 2):
            for j in range(-1, nx + 2):
                for d, b in enumerate(("bx", "by", "bz")):
                    bf[i] = st[b][l + 2, k + 2, j + 2]
                    i += 1
    return (prim.reshape(nz + 2, ny + 2, nx + 2, 5), eps.reshape(nz + 2, ny + 2, nx + 2),
            bf.reshape(nz + 3, ny + 3, nx + 3, 3))


def test_profile_layout(tmp_path):
    case = CASES["torus_sym_tvdmc"]
    cfg = case.cfg
    o = O.Oracle(cfg, ic.GENERATORS[case.ic](cfg))
    o.run(2)
    st = o.state()
    prim, eps, bf = sections(cfg, st)
    g = Grid(cfg)
    path = H.write_arrays(str(tmp_path / "p.hdf5"), cfg, (g.x.faces, g.y.faces, g.z.faces), o.t, prim, eps, bf)
    assert H.datasets(path) == sorted(H.DATASETS)
    got = H.read_profile(path)
    assert got["time"].tolist() == [o.t]
    assert got["dimension"].tolist() == [3.0]
    assert got["coordinate"].tolist() == [2.0]               # coordinate_flag + 1
    assert np.array_equal(got["x-interface"], g.x.faces[2:cfg.nx + 5])    # xF(-1:nx+1)
    assert np.array_equal(got["y-interface"], g.y.faces[2:cfg.ny + 5])
    assert np.array_equal(got["z-interface"], g.z.faces[2:cfg.nz + 5])
    assert got["primitive"].shape == (cfg.nz + 2, cfg.ny + 2, cfg.nx + 2, 5)
    assert got["epsilon"].shape == (cfg.nz + 2, cfg.ny + 2, cfg.nx + 2)
    assert got["bfield"].shape == (cfg.nz + 3, cfg.ny + 3, cfg.nx + 3, 3)
    assert np.array_equal(got["primitive"], prim)
    assert np.array_equal(got["epsilon"], eps)
    assert np.array_equal(got["bfield"], bf)
    # the analysis script's convention: primitive.T[v, j, k, l] = prim(v+1, j, k, l)
    P = got["primitive"].T
    B = got["bfield"].T
    for (j, k, l) in ((0, 0, 0), (3, 5, 7), (cfg.nx + 1, cfg.ny + 1, cfg.nz + 1)):
        assert P[0, j, k, l] == st["rho"][l + 2, k + 2, j + 2]
        assert P[4, j, k, l] == st["p"][l + 2, k + 2, j + 2]
        assert got["epsilon"].T[j, k, l] == st["eps"][l + 2, k + 2, j + 2]
        assert B[1, j + 1, k + 1, l + 1] == st["by"][l + 2, k + 2, j + 2]
    assert B[0, 0, 0, 0] == st["bx"][1, 1, 1]                 # face -1 on every axis


def test_profile_read_by_hdf5_tools(tmp_path):
    """An independent reader (the HDF5 command-line tool) sees the nine
    datasets with the reversed Fortran dimensions."""
    h5ls = shutil.which("h5ls") or "/opt/conda/bin/h5ls"
    if not os.path.exists(h5ls):
        pytest.skip("h5ls not installed")
    cfg = Config(nx=6, ny=5, nz=4, n_dim=3, boundary_flag=(0,) * 6)
    st = ic.empty_state(cfg)
    st["rho"][:] = 1.0
    prim, eps, bf = sections(cfg, st)
    g = Grid(cfg)
    path = H.write_arrays(str(tmp_path / "q.hdf5"), cfg, (g.x.faces, g.y.faces, g.z.faces), 0.5, prim, eps, bf)
    out = subprocess.run([h5ls, path], capture_output=True, text=True, check=True).stdout
    dims = {ln.split()[0]: re.search(r"\{(.*)\}", ln).group(1) for ln in out.splitlines() if "Dataset" in ln}
    assert dims == {"bfield": "7, 8, 9, 3", "coordinate": "1", "dimension": "1", "epsilon": "6, 7, 8",
                    "primitive": "6, 7, 8, 5", "time": "1", "x-interface": "9", "y-interface": "8",
                    "z-interface": "7"}
