"""Run configuration for the MI355X MHD update step.

The reference fixes its configuration at compile time through Fortran
PARAMETERs in solver/mhd/parameter.h:34-115 (geometry, dimension, boundary
flags, box, grid size, CFL, Riemann-solver and reconstruction switches) plus a
model parameter file (model/riemann/src/param.h, model/torus-cyn/src/param.h).
Here the same settings are a runtime value, passed to the C-ABI through
``cumc3d_desc`` (include/cumc3d_hip.h).  Field names follow the reference's
PARAMETER names where one exists.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple

# Riemann solvers (solver/mhd/parameter.h:91-101 LF_flag/HLL_flag/HLLC_flag/HLLD_flag)
RIEMANN = {"lf": 0, "hll": 1, "hllc": 2, "hlld": 3}
# Reconstructions (parameter.h:103-115 tvdmm/tvdmc/tvdvl/ppmc/weno flags)
RECON = {"tvdmm": 0, "tvdvl": 1, "tvdmc": 2, "ppm": 3, "weno": 4}
# In-step model hooks (model/riemann, model/torus-cyn, model/torus-sph, model/collapsar)
MODELS = {"riemann": 0, "torus": 1, "torus_sph": 2, "collapsar": 3}
# Arithmetic of the spatial operator
PRECISION = {"fp64": 0, "fp32": 1}


@dataclasses.dataclass
class Config:
    nx: int
    ny: int = 1
    nz: int = 1
    n_dim: int = 1
    coordinate_flag: int = 0              # 0 Cartesian, 1 cylindrical (r, phi, z), 2 spherical (r, theta, phi)
    boundary_flag: Tuple[int, ...] = (1, 1, 1, 1, 1, 1)
    x_start: float = 0.0
    x_end: float = 1.0
    y_start: float = 0.0
    y_end: float = 1.0
    z_start: float = 0.0
    z_end: float = 1.0
    cfl: float = 0.4
    ggas: float = 5.0 / 3.0
    total_time: float = 1.0e30
    riemann: str = "hlld"
    recon: str = "tvdmc"
    model: str = "riemann"
    # model/riemann/src/custom.f90:143-155: constant 0.1 downward gravity for
    # the Rayleigh-Taylor tests (test_model 6 and 18).
    rt_gravity: bool = False
    # model/torus-cyn/src/param.h:19 Paczynski-Wiita radius (0 = Newtonian, GM=1)
    r_sh: float = 0.0
    precision: str = "fp64"
    # -DFIXPOLE (make.inc:8-21): on spherical grids the polar faces sit at
    # theta = 0 and pi with sin = 0 (getgrid.f90:91-98), and the radial edge
    # EMFs on them are averaged over phi (mhd_module.f90:437-461)
    fixpole: bool = True
    # Face coordinates xF, yF, zF (-3..n+3, n+7 values each) of a custom grid
    # (CUSTOM_GRID, e.g. model/collapsar/src/custom.f90:68-155 reading
    # ./profile/*_grid.dat); None on an axis keeps the uniform faces of
    # getgrid.f90:26-34.
    faces: Optional[Tuple[Optional[Tuple[float, ...]], ...]] = None
    # Self-gravity of the collapsar model (model/collapsar/src/param.h:37-44,
    # def.f90:31): potential solved every n_pot steps by red-black SOR with
    # weight sor_omega, to relative tolerance sor_tolerance, at most relax_max sweeps
    n_pot: int = 10
    relax_max: int = 100000
    sor_tolerance: float = 1.0e-10
    sor_omega: float = 1.90

    def __post_init__(self):
        self.boundary_flag = tuple(int(b) for b in self.boundary_flag)
        if len(self.boundary_flag) != 6:
            raise ValueError("boundary_flag needs 6 entries")
        if self.riemann not in RIEMANN:
            raise ValueError(f"unknown Riemann solver {self.riemann!r}")
        if self.recon not in RECON:
            raise ValueError(f"unknown reconstruction {self.recon!r}")
        if self.model not in MODELS:
            raise ValueError(f"unknown model {self.model!r}")
        if self.precision not in PRECISION:
            raise ValueError(f"unknown precision {self.precision!r}")
        if self.coordinate_flag not in (0, 1, 2):
            raise ValueError("coordinate_flag must be 0 (Cartesian), 1 (cylindrical) or 2 (spherical)")
        if self.model == "torus" and self.coordinate_flag != 1:
            raise ValueError("the torus-cyn model needs cylindrical coordinates")
        if self.model == "torus_sph" and self.coordinate_flag != 2:
            raise ValueError("the torus-sph model needs spherical coordinates")
        if self.model == "collapsar" and self.coordinate_flag != 2:
            raise ValueError("the collapsar model needs spherical coordinates")
        if self.n_pot < 1 or self.relax_max < 1:
            raise ValueError("n_pot and relax_max must be positive")
        if self.faces is not None:
            if len(self.faces) != 3:
                raise ValueError("faces needs (xF, yF, zF)")
            faces = []
            for f, n in zip(self.faces, (self.nx, self.ny, self.nz)):
                if f is not None:
                    f = tuple(float(v) for v in f)
                    if len(f) != n + 7:
                        raise ValueError("a face array needs n+7 entries (-3..n+3)")
                faces.append(f)
            self.faces = tuple(faces)
        if self.n_dim == 1 and (self.ny != 1 or self.nz != 1):
            raise ValueError("n_dim=1 needs ny=nz=1")
        if self.n_dim == 2 and self.nz != 1:
            raise ValueError("n_dim=2 needs nz=1")

    # ghosted extents: -2..n+3 on each axis (solver/mhd/setup.f90:348-351)
    @property
    def shape(self):
        return (self.nz + 6, self.ny + 6, self.nx + 6)

    @property
    def zones(self) -> int:
        return self.nx * self.ny * self.nz

    def with_(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


# ---------------------------------------------------------------------------
# Named workloads.  Sizes and physics follow the five measurement configurations
# of this project; small variants of them are used as parity cases.
# ---------------------------------------------------------------------------

def briowu_1d(nx: int = 800) -> Config:
    """Brio-Wu shock tube: 800 cells, gamma=2, HLLD + TVD-MC, CFL 0.4, t=0.1
    (model/riemann/src/riemann1d.f90:229-250, test_model=10)."""
    return Config(nx=nx, n_dim=1, cfl=0.4, ggas=2.0, total_time=0.1,
                  riemann="hlld", recon="tvdmc", boundary_flag=(1, 1, 1, 1, 1, 1))


def orszag_tang_3d(n: int = 128, recon: str = "ppm", riemann: str = "hlld") -> Config:
    """3D periodic Orszag-Tang variant, gamma=5/3, HLLD + PPM."""
    return Config(nx=n, ny=n, nz=n, n_dim=3, cfl=0.4, ggas=5.0 / 3.0,
                  riemann=riemann, recon=recon, boundary_flag=(0, 0, 0, 0, 0, 0))


def blast_3d(n: int = 256, recon: str = "weno", riemann: str = "hlld") -> Config:
    """3D magnetised blast in a periodic unit box, HLLD + WENO-Z, t=0.02."""
    return Config(nx=n, ny=n, nz=n, n_dim=3, cfl=0.4, ggas=5.0 / 3.0, total_time=0.02,
                  riemann=riemann, recon=recon, boundary_flag=(0, 0, 0, 0, 0, 0))


def turbulence_3d(n: int = 256, recon: str = "weno", riemann: str = "hlld") -> Config:
    """3D periodic decaying MHD turbulence, t=0.5."""
    return Config(nx=n, ny=n, nz=n, n_dim=3, cfl=0.4, ggas=5.0 / 3.0, total_time=0.5,
                  riemann=riemann, recon=recon, boundary_flag=(0, 0, 0, 0, 0, 0))


def torus_cyl(nr: int = 192, nphi: int = 128, nz: int = 192, recon: str = "ppm",
              riemann: str = "hlld") -> Config:
    """Cylindrical (r, phi, z) accretion torus in a point-mass potential (GM=1)."""
    return Config(nx=nr, ny=nphi, nz=nz, n_dim=3, coordinate_flag=1,
                  x_start=2.0, x_end=26.0, y_start=0.0, y_end=2.0 * math.pi,
                  z_start=-12.0, z_end=12.0, cfl=0.4, ggas=5.0 / 3.0,
                  riemann=riemann, recon=recon, model="torus", r_sh=0.0,
                  boundary_flag=(1, 1, 0, 0, 1, 1))


def torus_sph(nr: int = 128, nth: int = 96, nphi: int = 64, recon: str = "ppm",
              riemann: str = "hlld", n_dim: int = 3) -> Config:
    """Spherical (r, theta, phi) accretion torus in the Paczynski-Wiita
    potential with r_sh = 1 (model/torus-sph/src/param.h:19): the full polar
    range with axis boundaries at both poles, periodic in phi.  With n_dim=2
    the run is axisymmetric (one phi cell)."""
    return Config(nx=nr, ny=nth, nz=nphi if n_dim == 3 else 1, n_dim=n_dim, coordinate_flag=2,
                  x_start=2.0, x_end=14.0, y_start=0.0, y_end=math.pi,
                  z_start=0.0, z_end=2.0 * math.pi, cfl=0.4, ggas=5.0 / 3.0,
                  riemann=riemann, recon=recon, model="torus_sph", r_sh=1.0,
                  boundary_flag=(1, 1, 3, 3, 0, 0))


def log_faces(n: int, r_in: float, r_out: float) -> Tuple[float, ...]:
    """Radial faces xF(-3..n+3) growing by a constant ratio from xF(0) = r_in
    to xF(n) = r_out (a logarithmic grid like the collapsar's r_grid.dat)."""
    q = (r_out / r_in) ** (1.0 / n)
    return tuple(r_in * q ** j for j in range(-3, n + 4))


def collapsar(nr: int = 156, nth: int = 44, nphi: int = 128, recon: str = "ppm",
              riemann: str = "hlld", n_dim: int = 3, faces=None, ggas: float = 4.0 / 3.0,
              r_in: float = 32.0, r_out: float = 6.4e4) -> Config:
    """Collapsar: a rotating, magnetised progenitor around a black hole, with
    self-gravity (model/collapsar).  Spherical (r, theta, phi) with the axis
    at both poles, periodic in phi, outflow in r, on a custom grid (CUSTOM_GRID,
    model/collapsar/src/custom.f90:68-155).  The default size is the
    reference's shipped progenitor grid (model/collapsar/profile: 156 radial,
    44 polar, 128 azimuthal cells).  Without `faces` the radial faces are
    logarithmic from r_in to r_out and theta, phi are uniform over [0, pi] and
    [0, 2 pi]; cumc3d_amd.ic.read_collapsar_profile gives the faces of a
    profile directory."""
    nz = nphi if n_dim == 3 else 1
    if faces is None:
        faces = (log_faces(nr, r_in, r_out), None, None)
    return Config(nx=nr, ny=nth, nz=nz, n_dim=n_dim, coordinate_flag=2,
                  x_start=r_in, x_end=r_out, y_start=0.0, y_end=math.pi,
                  z_start=0.0, z_end=2.0 * math.pi, cfl=0.4, ggas=ggas,
                  riemann=riemann, recon=recon, model="collapsar",
                  boundary_flag=(1, 1, 3, 3, 0, 0), faces=faces)
