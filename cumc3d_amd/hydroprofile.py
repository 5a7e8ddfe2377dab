"""The reference's output files, written through libcumc3d_io.so
(include/cumc3d_io.h): the HDF5 profile of print_hydroprofile
(solver/mhd/dataio.f90:7-249) and FIND_DIVB's log record
(solver/mhd/mhd_module.f90:199), plus a reader for the profiles (h5py is
not needed).

``write_profile(solver, "outfile", n_iter)`` packs the arrays on the device
(``Solver.profile``) and writes ``outfile/rkiter-<n_iter>-nm.hdf5`` with the
reference's nine datasets.  ``read_profile(path)`` returns them as stored
(C order, the Fortran dimensions reversed) -- ``primitive.T[v, j, k, l]`` is
prim(v+1, j, k, l) as in analysis/plot/plotprofile-1d.py:39-41.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB = os.path.join(HERE, "libcumc3d_io.so")
PD = C.POINTER(C.c_double)

# print_hydroprofile's datasets, in the order it writes them
DATASETS = ("time", "dimension", "coordinate", "x-interface", "y-interface", "z-interface",
            "primitive", "epsilon", "bfield")
EXPORTS = ("cumc3d_io_write_profile", "cumc3d_io_profile_name", "cumc3d_io_dataset_shape",
           "cumc3d_io_read_dataset", "cumc3d_io_list", "cumc3d_io_format_real", "cumc3d_io_format_divb",
           "cumc3d_io_append_divb", "cumc3d_io_last_error")


class Meta(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_dim", C.c_int),
                ("coordinate_flag", C.c_int), ("time", C.c_double), ("xF", PD), ("yF", PD), ("zF", PD)]


_lib = None


def load_io():
    """Load libcumc3d_io.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(IO_LIB):
        raise RuntimeError(f"{IO_LIB} is missing: build it with `python -m cumc3d_amd.build`")
    L = C.CDLL(IO_LIB)
    L.cumc3d_io_write_profile.argtypes = [C.c_char_p, C.POINTER(Meta), PD, PD, PD]
    L.cumc3d_io_profile_name.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    L.cumc3d_io_dataset_shape.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.cumc3d_io_read_dataset.argtypes = [C.c_char_p, C.c_char_p, PD, C.c_longlong]
    L.cumc3d_io_list.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.cumc3d_io_format_real.argtypes = [C.c_double, C.c_char_p, C.c_int]
    L.cumc3d_io_format_divb.argtypes = [C.c_double, C.c_double, C.c_char_p, C.c_int]
    L.cumc3d_io_append_divb.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int]
    L.cumc3d_io_last_error.restype = C.c_char_p
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise RuntimeError(load_io().cumc3d_io_last_error().decode())


def _p(a):
    return a.ctypes.data_as(PD)


def profile_name(directory: str, n_iter: int) -> str:
    buf = C.create_string_buffer(4096)
    _check(load_io().cumc3d_io_profile_name(directory.encode(), n_iter, buf, len(buf)))
    return buf.value.decode()


def written_faces(grid):
    """The face coordinates print_hydroprofile writes: GETGRID's after the
    FIXPOLE override (getgrid.f90:91-98) moved the polar faces to 0 and pi."""
    yF = grid.y.faces
    if grid.cfg.coordinate_flag == 2 and grid.cfg.fixpole:
        yF = grid.spherical()["yF"]
    return grid.x.faces, yF, grid.z.faces


def write_arrays(path: str, cfg, faces, time: float, primitive, epsilon, bfield) -> str:
    """Write one profile file from arrays in print_hydroprofile's layout
    (as Solver.profile returns them); faces are xF, yF, zF over -3..n+3."""
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in faces]
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (primitive, epsilon, bfield)]
    shapes = [(cfg.nz + 2, cfg.ny + 2, cfg.nx + 2, 5), (cfg.nz + 2, cfg.ny + 2, cfg.nx + 2),
              (cfg.nz + 3, cfg.ny + 3, cfg.nx + 3, 3)]
    for a, s in zip(arrs, shapes):
        if a.shape != s:
            raise ValueError(f"array of shape {a.shape}, expected {s}")
    m = Meta(cfg.nx, cfg.ny, cfg.nz, cfg.n_dim, cfg.coordinate_flag, float(time), _p(f[0]), _p(f[1]), _p(f[2]))
    _check(load_io().cumc3d_io_write_profile(path.encode(), C.byref(m), *(_p(a) for a in arrs)))
    return path


def write_profile(solver, directory: str, n_iter: int) -> str:
    """print_hydroprofile for a Solver: <directory>/rkiter-<n_iter>-nm.hdf5."""
    os.makedirs(directory, exist_ok=True)
    path = profile_name(directory, n_iter)
    faces = written_faces(solver.grid)
    p = solver.profile()
    return write_arrays(path, solver.cfg, faces, solver.t, p["primitive"], p["epsilon"], p["bfield"])


def datasets(path: str) -> list:
    buf = C.create_string_buffer(1 << 16)
    _check(load_io().cumc3d_io_list(path.encode(), buf, len(buf)))
    return [n for n in buf.value.decode().split("\n") if n]


def read_dataset(path: str, name: str) -> np.ndarray:
    L = load_io()
    rank = C.c_int()
    dims = (C.c_longlong * 8)()
    _check(L.cumc3d_io_dataset_shape(path.encode(), name.encode(), C.byref(rank), dims))
    shape = tuple(dims[i] for i in range(rank.value))
    out = np.empty(shape)
    _check(L.cumc3d_io_read_dataset(path.encode(), name.encode(), _p(out), out.size))
    return out


def read_profile(path: str) -> dict:
    return {n: read_dataset(path, n) for n in datasets(path)}


def format_real(x: float) -> str:
    """REAL*8 as flang's list-directed output writes it."""
    buf = C.create_string_buffer(64)
    _check(load_io().cumc3d_io_format_real(float(x), buf, len(buf)))
    return buf.value.decode()


def divb_record(time: float, maxdivb: float) -> str:
    """FIND_DIVB's log record WRITE(999,*) global_time, maxDivB."""
    buf = C.create_string_buffer(160)
    _check(load_io().cumc3d_io_format_divb(float(time), float(maxdivb), buf, len(buf)))
    return buf.value.decode()


def append_divb(path: str, time: float, maxdivb: float, truncate: bool = False) -> None:
    _check(load_io().cumc3d_io_append_divb(path.encode(), float(time), float(maxdivb), int(truncate)))
