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


def load_golden(name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} missing")
    with np.load(path) as d:
        return {k: d[k] for k in d.files}
