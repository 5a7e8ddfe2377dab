import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def interior_window(cfg):
    """cells 0..n+1 on active axes, cell 1 on inactive axes (matches the golden fixtures)"""
    sl = []
    for n, active in ((cfg.nz, cfg.n_dim > 2), (cfg.ny, cfg.n_dim > 1), (cfg.nx, True)):
        sl.append(slice(2, n + 4) if active else slice(3, 4))
    return tuple(sl)


def lrk_mask(cfg, i):
    """Where the library forms l_rk(i) (1..8), on the fixture window of
    interior_window: l_rk(1:5) on cells 1..n, l_rk(6:8) on the faces the update
    advances (0..n along the component's own axis, 1..n across).  The reference
    also forms the face values on the ghost rows across (spatial.f90:639-702),
    which no stage reads back."""
    own = {6: 2, 7: 1, 8: 0}.get(i)    # array axis (z, y, x) of the face normal
    shape, sl = [], []
    for a, (n, active) in enumerate(((cfg.nz, cfg.n_dim > 2), (cfg.ny, cfg.n_dim > 1), (cfg.nx, True))):
        if not active:
            shape.append(1)
            sl.append(slice(0, 1))
        else:
            shape.append(n + 2)
            sl.append(slice(0 if a == own else 1, n + 1))
    m = np.zeros(shape, dtype=bool)
    m[tuple(sl)] = True
    return m


def check_lrk(got, gold, cfg, label=""):
    """got: lrk1..lrk8 over the ghosted box; gold: over the fixture window."""
    w = interior_window(cfg)
    for i in range(1, 9):
        m = lrk_mask(cfg, i)
        a, b = got[f"lrk{i}"][w][m], gold[f"lrk{i}"][m]
        if not np.array_equal(a, b):
            d = np.abs(a - b)
            raise AssertionError(f"{label}l_rk({i}): {np.count_nonzero(d)} differ, max {np.nanmax(d):.3e}")


def load_golden(name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} missing")
    with np.load(path) as d:
        return {k: d[k] for k in d.files}


def check_profile(prof, cfg, state, t):
    """A profile file (read back as a dict) holds print_hydroprofile's
    datasets of `state` (ghosted-box arrays) at time t
    (solver/mhd/dataio.f90:36-233): primitive = prim(1:5, 0:n+1, ...),
    epsilon(0:n+1, ...), bfield = prim(ibx:ibz, -1:n+1, ...), the interfaces
    xF(-1:n+1) (yF after FIXPOLE), stored with the Fortran dimensions reversed."""
    from cumc3d_amd.grid import Grid
    from cumc3d_amd.hydroprofile import written_faces
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    c = (slice(2, nz + 4), slice(2, ny + 4), slice(2, nx + 4))      # cells 0..n+1
    f = (slice(1, nz + 4), slice(1, ny + 4), slice(1, nx + 4))      # faces -1..n+1
    faces = written_faces(Grid(cfg))
    assert prof["time"].tolist() == [t]
    assert prof["dimension"].tolist() == [float(cfg.n_dim)]
    assert prof["coordinate"].tolist() == [float(cfg.coordinate_flag + 1)]
    for name, fa, n in (("x-interface", faces[0], nx), ("y-interface", faces[1], ny), ("z-interface", faces[2], nz)):
        assert np.array_equal(prof[name], fa[2:n + 5]), name
    prim = prof["primitive"]
    assert prim.shape == (nz + 2, ny + 2, nx + 2, 5)
    for v, name in enumerate(("rho", "vx", "vy", "vz", "p")):
        assert np.array_equal(prim[..., v], state[name][c], equal_nan=True), name
    assert np.array_equal(prof["epsilon"], state["eps"][c], equal_nan=True)
    bf = prof["bfield"]
    assert bf.shape == (nz + 3, ny + 3, nx + 3, 3)
    for d, name in enumerate(("bx", "by", "bz")):
        assert np.array_equal(bf[..., d], state[name][f], equal_nan=True), name


def zero_signs_equal(a, b) -> bool:
    """The cells holding zero carry the same sign in a and b (np.array_equal
    counts -0.0 == +0.0; the reference's MIN/MAX and floors leave signed zeros
    that a bit-exact result must reproduce)."""
    a, b = np.asarray(a), np.asarray(b)
    return np.array_equal(np.signbit(a) & (a == 0), np.signbit(b) & (b == 0))
