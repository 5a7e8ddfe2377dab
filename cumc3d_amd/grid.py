"""Grid metrics, computed the way the reference's GetGrid does
(solver/mhd/getgrid.f90:26-71): uniform face coordinates
``xF(j) = x_start + j*dx_ini`` for j=-3..n+3, then ``dx(j) = xF(j)-xF(j-1)``
and ``x(j) = (xF(j)+xF(j-1))/2`` for j=-2..n+3.

Arrays are returned with an offset: ``xF[j+3]`` and ``x[j+2]``.  The MI355X
library derives the same metrics from the face coordinates in
``cumc3d_init`` (cumc3d_amd/csrc/grid.cpp).
"""
from __future__ import annotations

import numpy as np

from .config import Config


class Axis:
    """Face and centre coordinates of one axis with 3 ghost cells per side."""

    def __init__(self, start: float, end: float, n: int, faces: np.ndarray | None = None):
        self.n = n
        if faces is None:
            d_ini = (end - start) / float(n)                  # getgrid dx_ini (parameter.h:75)
            j = np.arange(-3, n + 4, dtype=np.float64)
            faces = start + j * d_ini                          # getgrid.f90:26-28
        faces = np.ascontiguousarray(faces, dtype=np.float64)
        if faces.shape != (n + 7,):
            raise ValueError("face array needs n+7 entries (-3..n+3)")
        self.faces = faces                                     # xF(-3..n+3)
        self.d = faces[1:] - faces[:-1]                        # dx(-2..n+3), getgrid.f90:44-46
        self.c = (faces[1:] + faces[:-1]) / 2                  # x(-2..n+3),  getgrid.f90:55-57

    def centre(self, j):
        """Centre coordinate of cell j (j in -2..n+3)."""
        return self.c[np.asarray(j) + 2]

    def face(self, j):
        """Coordinate of face j+1/2, i.e. xF(j) (j in -3..n+3)."""
        return self.faces[np.asarray(j) + 3]

    def width(self, j):
        return self.d[np.asarray(j) + 2]


class Grid:
    def __init__(self, cfg: Config, xfaces=None, yfaces=None, zfaces=None):
        self.cfg = cfg
        self.x = Axis(cfg.x_start, cfg.x_end, cfg.nx, xfaces)
        self.y = Axis(cfg.y_start, cfg.y_end, cfg.ny, yfaces)
        self.z = Axis(cfg.z_start, cfg.z_end, cfg.nz, zfaces)

    def centres(self):
        """Cell-centre coordinates broadcast over the ghosted box, shape (nz+6, ny+6, nx+6)."""
        X = self.x.c[None, None, :]
        Y = self.y.c[None, :, None]
        Z = self.z.c[:, None, None]
        return X, Y, Z
