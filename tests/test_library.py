"""The C-ABI boundary without a GPU: the library loads, exports every entry
point declared in include/cumc3d_hip.h and validates descriptors on the host."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from cumc3d_amd import solver
from cumc3d_amd.config import Config, briowu_1d

HEADER = os.path.join(ROOT, "include", "cumc3d_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cumc3d_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = solver.load_library()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(solver.EXPORTS)


def make_desc(**kw):
    cfg = briowu_1d(16)
    faces = [np.arange(-3, n + 4, dtype=np.float64) / max(n, 1) for n in (16, 1, 1)]
    d = solver.Desc()
    d.nx, d.ny, d.nz, d.n_dim = 16, 1, 1, 1
    for i in range(6):
        d.boundary_flag[i] = 1
    d.riemann, d.recon, d.model, d.hlld_bn_mode, d.precision = 3, 2, 0, 2, 0
    d.cfl, d.ggas = cfg.cfl, cfg.ggas
    d.xF, d.yF, d.zF = (f.ctypes.data_as(solver.PD) for f in faces)
    for k, v in kw.items():
        if k == "boundary_flag":
            for i, b in enumerate(v):
                d.boundary_flag[i] = b
        else:
            setattr(d, k, v)
    return d, faces


@pytest.mark.parametrize("field,value", [
    ("nx", 0), ("n_dim", 0), ("n_dim", 4), ("ny", 2), ("coordinate_flag", 3), ("riemann", 4),
    ("recon", 5), ("model", 3), ("hlld_bn_mode", 3), ("precision", 7),
    # each torus model's hooks belong to its geometry (model/torus-cyn, model/torus-sph)
    ("model", 1), ("model", 2),
])
def test_init_rejects_invalid_descriptor(field, value):
    L = solver.load_library()
    d, keep = make_desc(**{field: value})
    h = C.c_void_p()
    assert L.cumc3d_init(C.byref(d), C.byref(h)) == 1
    assert not h.value
    assert L.cumc3d_last_error().decode()


@pytest.mark.parametrize("flags", [(4, 1, 1, 1, 1, 1), (1, 4, 1, 1, 1, 1), (1, 1, 1, 1, 3, 1), (1, 1, 1, 1, 1, 3),
                                   (2, 1, 1, 1, 1, 1)])
def test_init_rejects_boundary_combinations_like_setup_eqns(flags):
    """SETUP_EQNS STOPs on equatorial x-boundaries and axial z-boundaries
    (setup.f90:120-145), and mirror boundaries need 3 cells on the axis; the
    library returns CUMC3D_EINVAL instead."""
    L = solver.load_library()
    d, keep = make_desc(boundary_flag=flags)
    if flags[0] == 2:
        d.nx = 2
    h = C.c_void_p()
    assert L.cumc3d_init(C.byref(d), C.byref(h)) == 1


def test_error_strings():
    L = solver.load_library()
    for code in range(6):
        assert L.cumc3d_error_string(code)


def test_config_validation():
    with pytest.raises(ValueError):
        Config(nx=8, ny=2, n_dim=1)
    with pytest.raises(ValueError):
        Config(nx=8, riemann="roe")
    with pytest.raises(ValueError):
        Config(nx=8, coordinate_flag=3)
    with pytest.raises(ValueError):
        Config(nx=8, model="torus_sph")
    with pytest.raises(ValueError):
        Config(nx=8, coordinate_flag=2, model="torus")
    Config(nx=8, coordinate_flag=2, model="torus_sph")


FORTRAN = os.path.join(ROOT, "cumc3d_amd", "fortran")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def test_fortran_module_binds_every_declared_symbol():
    """The Fortran drop-in's module carries a bind(C) interface for every entry
    point of each header: cumc3d_hip.f90 for include/cumc3d_hip.h,
    cumc3d_io.f90 for the writer entries the drop-in print_hydroprofile uses."""
    bound = set(re.findall(r"NAME='(cumc3d_[a-z_0-9]+)'", open(os.path.join(FORTRAN, "cumc3d_hip.f90")).read()))
    assert set(declared_functions()) <= bound
    io = set(re.findall(r"NAME='(cumc3d_io_[a-z_0-9]+)'", open(os.path.join(FORTRAN, "cumc3d_io.f90")).read()))
    assert {"cumc3d_io_write_profile", "cumc3d_io_last_error"} <= io


def test_fortran_modules_compile(tmp_path):
    """The bindings and the stand-alone drop-in sources compile with the image's flang."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not installed")
    import subprocess
    for f in ("cumc3d_hip.f90", "cumc3d_io.f90"):
        r = subprocess.run([FLANG, "-c", os.path.join(FORTRAN, f), "-o", str(tmp_path / (f + ".o"))],
                           cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
