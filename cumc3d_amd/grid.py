"""Grid metrics, computed the way the reference's GetGrid does
(solver/mhd/getgrid.f90:26-71): uniform face coordinates
``xF(j) = x_start + j*dx_ini`` for j=-3..n+3, then ``dx(j) = xF(j)-xF(j-1)``
and ``x(j) = (xF(j)+xF(j-1))/2`` for j=-2..n+3.

Arrays are returned with an offset: ``xF[j+3]`` and ``x[j+2]``.  The MI355X
library derives the same metrics from the face coordinates in
``cumc3d_init`` (cumc3d_amd/csrc/capi.hip).

Spherical grids (coordinate_flag 2) also need the metrics of
getgrid.f90:68-98: ``dx_cb = xF(j)^3 - xF(j-1)^3``, ``xbar``, ``dcose``,
``dsine``, ``sine`` and ``sinf``, with the FIXPOLE override (polar faces at
theta = 0 and pi, sinf = 0 there) applied after the others are taken.
"""
from __future__ import annotations

import math

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
        if cfg.faces is not None:                              # CUSTOM_GRID faces
            xfaces = xfaces if xfaces is not None else cfg.faces[0]
            yfaces = yfaces if yfaces is not None else cfg.faces[1]
            zfaces = zfaces if zfaces is not None else cfg.faces[2]
        self.x = Axis(cfg.x_start, cfg.x_end, cfg.nx, xfaces)
        self.y = Axis(cfg.y_start, cfg.y_end, cfg.ny, yfaces)
        self.z = Axis(cfg.z_start, cfg.z_end, cfg.nz, zfaces)

    def spherical(self) -> dict:
        """getgrid.f90:68-98 metrics, each over -3..n+3 at index j+3 (entries
        outside the reference's loop ranges are 0), FIXPOLE applied when the
        configuration asks for it."""
        nx, ny = self.cfg.nx, self.cfg.ny
        xF, yF = self.x.faces, self.y.faces
        dx = np.zeros(nx + 7); dx[1:] = self.x.d
        xc = np.zeros(nx + 7); xc[1:] = self.x.c
        yc = np.zeros(ny + 7); yc[1:] = self.y.c
        dx_cb = np.zeros(nx + 7)
        xbar = np.zeros(nx + 7)
        dx_cb[2:] = xF[2:] * xF[2:] * xF[2:] - xF[1:-1] * xF[1:-1] * xF[1:-1]       # j = -1..nx+3
        xbar[2:] = (dx_cb[2:] / 3.0) / (xc[2:] * dx[2:])
        dcose = np.zeros(ny + 7)
        dsine = np.zeros(ny + 7)
        dcose[2:] = -(np.cos(yF[2:]) - np.cos(yF[1:-1]))
        dsine[2:] = np.sin(yF[2:]) - np.sin(yF[1:-1])
        sine = np.zeros(ny + 7); sine[1:] = np.sin(yc[1:])
        sinf = np.sin(yF)
        yFp = yF.copy()
        if self.cfg.coordinate_flag == 2 and self.cfg.fixpole:
            yFp[3] = 0.0
            yFp[ny + 3] = math.pi
            sinf[3] = 0.0
            sinf[ny + 3] = 0.0
        return dict(dx_cb=dx_cb, xbar=xbar, dcose=dcose, dsine=dsine, sine=sine, sinf=sinf, yF=yFp)

    def centres(self):
        """Cell-centre coordinates broadcast over the ghosted box, shape (nz+6, ny+6, nx+6)."""
        X = self.x.c[None, None, :]
        Y = self.y.c[None, :, None]
        Z = self.z.c[:, None, None]
        return X, Y, Z
