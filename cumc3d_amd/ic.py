"""Initial conditions and the binary state-file format.

A state is a dict of float64 arrays of shape (nz+6, ny+6, nx+6) over the
ghosted box (-2..n+3 per axis, solver/mhd/setup.f90:348-351):

    rho, vx, vy, vz, p      cell-centred primitives (prim(irho..itau))
    bx, by, bz              face-centred field: bx[l,k,j] is Bx on face j+1/2
                            (prim(ibx..ibz), solver/mhd/fromrvetou.f90:54-56)
    eps                     specific internal energy (epsilon)
    rho_floor, p_floor, eps_floor   (torus models only, model/torus-cyn/src/def.f90)
    gravity                 (collapsar only) prim_a(1:8), eps_a, m_bh, vel2code

The generators build the five measurement workloads; the Brio-Wu tube follows
model/riemann/src/riemann1d.f90:229-250.  Random perturbations use a counter
hash (splitmix64 of seed and interior cell index) so every consumer of an IC
sees the same numbers.

``write_ic`` emits the file read by the reference oracle's file-fed GET_MODEL
(oracle/ref_harness/ic_plain.f90) and by the C-ABI upload path: a stream of
int32[6] header, float64[2] (ggas, total_time), prim in Fortran order
(8, nx+6, ny+6, nz+6), epsilon, then the three floor arrays when bit 0 of
header[5] is set and the collapsar's gravity record (prim_a(1:8), eps_a,
m_bh, vel2code) when bit 1 is.
"""
from __future__ import annotations

import math

import numpy as np

from .config import Config
from .grid import Grid

PRIM_ORDER = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")
MAGIC = 0x4D433344  # 'MC3D'
# st["gravity"] of the collapsar: prim_a(1:8), eps_a, m_bh, vel2code
GRAVITY_LEN = 11


# ---------------------------------------------------------------------------
# counter-based random numbers
# ---------------------------------------------------------------------------

def splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, index: np.ndarray) -> np.ndarray:
    """U[0,1) from the hash of (seed, index)."""
    key = (np.uint64(seed & 0xFFFFFFFF) << np.uint64(32)) ^ np.asarray(index, dtype=np.uint64)
    return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def interior_uniform(cfg: Config, seed: int) -> np.ndarray:
    """One U[0,1) number per interior cell, broadcast on the ghosted box (ghosts get 0.5)."""
    out = np.full(cfg.shape, 0.5)
    l, k, j = np.meshgrid(np.arange(cfg.nz), np.arange(cfg.ny), np.arange(cfg.nx), indexing="ij")
    idx = j + cfg.nx * (k + cfg.ny * l)
    out[3:cfg.nz + 3, 3:cfg.ny + 3, 3:cfg.nx + 3] = uniform01(seed, idx.ravel()).reshape(idx.shape)
    return out


def empty_state(cfg: Config) -> dict:
    return {name: np.zeros(cfg.shape) for name in PRIM_ORDER + ("eps",)}


# ---------------------------------------------------------------------------
# file format
# ---------------------------------------------------------------------------

def write_ic(path: str, cfg: Config, st: dict) -> None:
    has_floor = int("rho_floor" in st) + 2 * int("gravity" in st)
    prim = np.stack([st[n] for n in PRIM_ORDER], axis=-1)      # (nz+6, ny+6, nx+6, 8)
    with open(path, "wb") as f:
        np.array([MAGIC, 1, cfg.nx, cfg.ny, cfg.nz, has_floor], dtype=np.int32).tofile(f)
        np.array([cfg.ggas, cfg.total_time], dtype=np.float64).tofile(f)
        np.ascontiguousarray(prim, dtype=np.float64).tofile(f)
        np.ascontiguousarray(st["eps"], dtype=np.float64).tofile(f)
        if has_floor & 1:
            for n in ("rho_floor", "p_floor", "eps_floor"):
                np.ascontiguousarray(st[n], dtype=np.float64).tofile(f)
        if has_floor & 2:
            np.ascontiguousarray(st["gravity"], dtype=np.float64).tofile(f)


def read_ic(path: str):
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype=np.int32, count=6)
        if hdr[0] != MAGIC:
            raise ValueError("not a cumc3d IC file")
        nx, ny, nz, has_floor = (int(v) for v in hdr[2:6])
        ggas, total_time = np.fromfile(f, dtype=np.float64, count=2)
        shape = (nz + 6, ny + 6, nx + 6)
        g = int(np.prod(shape))
        prim = np.fromfile(f, dtype=np.float64, count=8 * g).reshape(shape + (8,))
        st = {n: np.ascontiguousarray(prim[..., i]) for i, n in enumerate(PRIM_ORDER)}
        st["eps"] = np.fromfile(f, dtype=np.float64, count=g).reshape(shape)
        if has_floor & 1:
            for n in ("rho_floor", "p_floor", "eps_floor"):
                st[n] = np.fromfile(f, dtype=np.float64, count=g).reshape(shape)
        if has_floor & 2:
            st["gravity"] = np.fromfile(f, dtype=np.float64, count=GRAVITY_LEN)
    return (nx, ny, nz, float(ggas), float(total_time)), st


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------

def brio_wu(cfg: Config) -> dict:
    """Brio-Wu tube (riemann1d.f90:229-250): cells j=0..nx of every row get the
    left or right state by x(j) <= 0.5; Bx = 0.75 on those faces."""
    g = Grid(cfg)
    st = empty_state(cfg)
    xc = g.x.c                                     # x(-2..nx+3)
    sl = slice(2, cfg.nx + 3)                      # j = 0..nx
    left = xc[sl] <= 0.5
    st["rho"][..., sl] = np.where(left, 1.0, 0.125)
    st["p"][..., sl] = np.where(left, 1.0, 0.1)
    st["by"][..., sl] = np.where(left, 1.0, -1.0)
    st["bx"][..., sl] = 0.75
    st["eps"][..., sl] = st["p"][..., sl] / st["rho"][..., sl] / (cfg.ggas - 1.0)
    return st


def _face_field_from_az(cfg: Config, g: Grid, az_fn):
    """Bx, By on faces from a z-edge vector potential Az(xF, yF) (Cartesian):
    Bx(j,k) = (Az(j,k) - Az(j,k-1))/dy(k),  By(j,k) = -(Az(j,k) - Az(j-1,k))/dx(j),
    which has zero discrete divergence."""
    xf = g.x.faces  # xF(-3..nx+3)
    yf = g.y.faces
    AZ = az_fn(xf[None, :], yf[:, None])       # AZ[k+3, j+3] = Az(xF(j), yF(k))
    dy = g.y.d[:, None]                         # dy(-2..ny+3)
    dx = g.x.d[None, :]
    bx = (AZ[1:, 1:] - AZ[:-1, 1:]) / dy        # index [k+2, j+2] for k,j in -2..n+3
    by = -(AZ[1:, 1:] - AZ[1:, :-1]) / dx
    return bx, by


def orszag_tang(cfg: Config, seed: int = 20261202) -> dict:
    """3D Orszag-Tang variant: rho=25/(36 pi)(1 +- 1% noise), p=5/(12 pi),
    v=(-sin 2pi y, sin 2pi x, 0.1 sin 2pi z), face B from
    Az=(cos 4pi x/(4pi) + cos 2pi y/(2pi))/sqrt(4 pi)."""
    g = Grid(cfg)
    X, Y, Z = g.centres()
    st = empty_state(cfg)
    u = interior_uniform(cfg, seed)
    st["rho"][:] = 25.0 / (36.0 * math.pi) * (1.0 + 0.02 * (u - 0.5))
    st["p"][:] = 5.0 / (12.0 * math.pi)
    st["vx"][:] = -np.sin(2 * math.pi * Y) + 0 * X
    st["vy"][:] = np.sin(2 * math.pi * X) + 0 * Y
    st["vz"][:] = 0.1 * np.sin(2 * math.pi * Z) + 0 * X
    s4 = math.sqrt(4 * math.pi)
    bx, by = _face_field_from_az(cfg, g, lambda x, y: (np.cos(4 * math.pi * x) / (4 * math.pi)
                                                       + np.cos(2 * math.pi * y) / (2 * math.pi)) / s4)
    st["bx"][:] = bx[None, :, :]
    st["by"][:] = by[None, :, :]
    st["eps"][:] = st["p"] / st["rho"] / (cfg.ggas - 1.0)
    return st


def blast(cfg: Config, seed: int = 20261202, edge_cells: float = 1.0) -> dict:
    """Magnetised blast: rho=1 (+-1% noise), p=0.1 outside and 10 inside r<0.1
    around the box centre, v=0, uniform B=(10/sqrt2, 10/sqrt2, 0)/sqrt(4 pi).

    The pressure edge is a tanh of width ``edge_cells`` cells;
    ``edge_cells=0`` gives the sharp step the configuration specifies (p=10
    where r<0.1).  With the sharp step, WENO-Z reconstruction of the
    primitives (weno_module.f90:315-408) overshoots to negative face
    pressures at the 100x jump and the HLLD solver returns NaN within the
    first step, in the reference algorithm as in this one (pinned by
    tests/test_oracle.py); the default smooths the edge over one cell."""
    g = Grid(cfg)
    X, Y, Z = g.centres()
    st = empty_state(cfg)
    u = interior_uniform(cfg, seed)
    st["rho"][:] = 1.0 + 0.02 * (u - 0.5)
    cx = 0.5 * (cfg.x_start + cfg.x_end)
    cy = 0.5 * (cfg.y_start + cfg.y_end)
    cz = 0.5 * (cfg.z_start + cfg.z_end)
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2)
    if edge_cells > 0:
        w = edge_cells * (cfg.x_end - cfg.x_start) / cfg.nx
        st["p"][:] = 0.1 + 9.9 * 0.5 * (1.0 - np.tanh((r - 0.1) / w))
    else:
        st["p"][:] = np.where(r < 0.1, 10.0, 0.1)
    b = 10.0 / math.sqrt(2.0) / math.sqrt(4 * math.pi)
    st["bx"][:] = b
    st["by"][:] = b
    st["eps"][:] = st["p"] / st["rho"] / (cfg.ggas - 1.0)
    return st


def _solenoidal_modes(n: int, kmax: float, slope: float, seed: int, rng_stream: int, log=None):
    """Random-phase solenoidal vector field on an n^3 periodic grid with
    |k| <= kmax and amplitude |a_k| ~ k**slope; returns real arrays (3, n, n, n)
    indexed [comp, l, k, j].  log(msg), if given, hears after each component
    (a 512^3 field takes about a minute)."""
    kk = np.fft.fftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(kk, kk, kk, indexing="ij")
    K2 = KX * KX + KY * KY + KZ * KZ
    kmag = np.sqrt(K2)
    mask = (kmag > 0) & (kmag <= kmax)
    idx = np.arange(n ** 3, dtype=np.uint64).reshape(n, n, n)
    field = []
    for c in range(3):
        ph = uniform01(seed + 7919 * (3 * rng_stream + c), idx)
        am = uniform01(seed + 104729 * (3 * rng_stream + c) + 1, idx)
        g = np.sqrt(-2.0 * np.log(np.maximum(am, 1e-300)))     # Rayleigh amplitude
        f = np.zeros((n, n, n), dtype=np.complex128)
        f[mask] = g[mask] * np.exp(2j * math.pi * ph[mask]) * kmag[mask] ** slope
        field.append(f)
        if log:
            log(f"solenoidal field {rng_stream}: component {c + 1} of 3 drawn")
    fx, fy, fz = field
    # project out the compressive part: f - k (k.f)/k^2
    with np.errstate(invalid="ignore", divide="ignore"):
        kdotf = np.where(mask, (KX * fx + KY * fy + KZ * fz) / np.where(K2 > 0, K2, 1), 0)
    fx = fx - KX * kdotf
    fy = fy - KY * kdotf
    fz = fz - KZ * kdotf
    out = np.stack([np.fft.ifftn(f).real for f in (fx, fy, fz)])
    if log:
        log(f"solenoidal field {rng_stream}: transformed")
    return out


def _periodic_ghost(a: np.ndarray) -> np.ndarray:
    """Embed an (nz, ny, nx) periodic interior into the ghosted (nz+6, ny+6, nx+6) box."""
    return np.pad(a, 3, mode="wrap")


def turbulence(cfg: Config, seed: int = 20261202, log=None) -> dict:
    """Decaying turbulence: rho=1, beta=1, rms Mach 1 solenoidal velocity with
    |k|<=8 and E(k) ~ k^-2, mean field B0 along z plus a solenoidal dB at 50%
    of B0 built from an edge vector potential (discrete div B = 0).  log(msg),
    if given, hears the progress of the two random fields."""
    n = cfg.nx
    assert cfg.ny == n and cfg.nz == n, "turbulence needs a cubic grid"
    g = Grid(cfg)
    gam = cfg.ggas
    p0 = 1.0 / gam                          # c_s = 1
    b0 = math.sqrt(2.0 * p0)                # beta = 2 p / B0^2 = 1
    v = _solenoidal_modes(n, 8.0, -2.0, seed, 0, log)     # amplitude ~ k^-2 gives E(k) ~ k^-2
    vrms = math.sqrt(float(np.mean(np.sum(v * v, axis=0))))
    v *= 1.0 / vrms                          # rms Mach 1 with c_s = 1
    st = empty_state(cfg)
    st["rho"][:] = 1.0
    st["p"][:] = p0
    st["vx"][:] = _periodic_ghost(v[0])
    st["vy"][:] = _periodic_ghost(v[1])
    st["vz"][:] = _periodic_ghost(v[2])
    # vector potential on edges: A_x at (x_j, y_{k+1/2}, z_{l+1/2}), etc. (same
    # periodic array indices), then face B = curl A.
    a = _solenoidal_modes(n, 8.0, -3.0, seed, 1, log)
    ax, ay, az = (np.asarray(c) for c in a)
    h = 1.0 / n
    # Bx(face j, k, l) = (Az(j,k) - Az(j,k-1))/dy - (Ay(j,l) - Ay(j,l-1))/dz
    bx = (az - np.roll(az, 1, axis=1)) / h - (ay - np.roll(ay, 1, axis=0)) / h
    by = (ax - np.roll(ax, 1, axis=0)) / h - (az - np.roll(az, 1, axis=2)) / h
    bz = (ay - np.roll(ay, 1, axis=2)) / h - (ax - np.roll(ax, 1, axis=1)) / h
    brms = math.sqrt(float(np.mean(bx * bx + by * by + bz * bz)))
    s = 0.5 * b0 / brms
    st["bx"][:] = _periodic_ghost(bx * s)
    st["by"][:] = _periodic_ghost(by * s)
    st["bz"][:] = _periodic_ghost(bz * s + b0)
    del g
    st["eps"][:] = st["p"] / st["rho"] / (gam - 1.0)
    return st


def torus(cfg: Config, seed: int = 20261202, r_in: float = 6.0, r_max: float = 10.0,
          rho_floor: float = 1.0e-6, beta: float = 100.0, a_cut: float = 0.2,
          p_noise: float = 0.01) -> dict:
    """Constant-angular-momentum torus in the point-mass potential Phi=-1/R
    (GM=1, r_sh=0) with inner edge r_in and pressure maximum at r_max,
    polytropic p=K rho^gamma, rho_max=1.  Poloidal loops from
    A_phi ~ max(rho - a_cut rho_max, 0) normalised to minimum beta, 1% pressure
    noise.  Floors follow model/torus-cyn/src/model.f90:266-276 (interior
    cells; the ghost cells of the floor arrays are zero)."""
    assert cfg.coordinate_flag == 1
    g = Grid(cfg)
    gam = cfg.ggas
    R = g.x.c[None, None, :]
    Zc = g.z.c[:, None, None]
    phi = lambda r, z: -1.0 / np.sqrt(r * r + z * z)
    l2 = r_max                                # Keplerian specific angular momentum^2 at r_max
    c_const = phi(r_in, 0.0) + l2 / (2.0 * r_in * r_in)
    h_max = c_const - phi(r_max, 0.0) - l2 / (2.0 * r_max * r_max)
    k_poly = h_max * (gam - 1.0) / gam        # rho_max = 1
    # the enthalpy only where the radius is positive: ghost cells of a grid
    # starting near the axis can sit at r <= 0, where the potential and the
    # centrifugal term are infinite (their state is the floor either way)
    Rp = np.where(R > 0.0, R, 1.0)
    h = np.where(R > 0.0, c_const - phi(Rp, Zc) - l2 / (2.0 * Rp * Rp), 0.0)
    h = np.maximum(h, 0.0)
    rho_t = (h * (gam - 1.0) / (gam * k_poly)) ** (1.0 / (gam - 1.0))
    inside = (rho_t > rho_floor) & (R >= r_in)
    shape = cfg.shape
    st = empty_state(cfg)
    st["rho"][:] = np.broadcast_to(np.where(inside, rho_t, rho_floor), shape)
    st["vy"][:] = np.broadcast_to(np.where(inside, math.sqrt(l2) / R, 0.0), shape)
    st["p"][:] = k_poly * st["rho"] ** gam
    u = interior_uniform(cfg, seed)
    st["p"][:] = np.where(np.broadcast_to(inside, shape),
                          st["p"] * (1.0 + 2.0 * p_noise * (u - 0.5)), st["p"])
    st["eps"][:] = st["p"] / st["rho"] / (gam - 1.0)
    # A_phi at the (r, z) corners (xF(j), zF(l)); axisymmetric
    xf = g.x.faces[None, :]
    zf = g.z.faces[:, None]
    xfp = np.where(xf > 0.0, xf, 1.0)      # A_phi = 0 on corners at r <= 0 (ghost faces)
    hc = np.where(xf > 0.0, np.maximum(c_const - phi(xfp, zf) - l2 / (2.0 * xfp * xfp), 0.0), 0.0)
    rc = (hc * (gam - 1.0) / (gam * k_poly)) ** (1.0 / (gam - 1.0))
    A = np.maximum(rc - a_cut, 0.0)           # A[l+3, j+3]
    dz = g.z.d[:, None]
    xc = g.x.c[None, :]
    dx = g.x.d[None, :]
    # model/torus-cyn/src/model.f90:144-145 (curl of A_phi in cylindrical coordinates)
    br = -(A[1:, 1:] - A[:-1, 1:]) / dz                                   # [l+2, j+2]
    bz = (xf[:, 1:] * A[1:, 1:] - xf[:, :-1] * A[1:, :-1]) / (xc * dx)
    # normalise to minimum beta over cells with field
    brc = 0.5 * (br[:, 1:] + br[:, :-1])
    bzc = 0.5 * (bz[1:, :] + bz[:-1, :])
    b2 = brc[1:, :] ** 2 + bzc[:, 1:] ** 2
    pin = k_poly * np.maximum(st["rho"][1:, 0, 1:], 0) ** gam
    with np.errstate(divide="ignore"):
        betas = np.where(b2 > 0, 2.0 * pin / np.where(b2 > 0, b2, 1.0), np.inf)
    bscale = math.sqrt(float(np.min(betas)) / beta) if np.isfinite(np.min(betas)) else 0.0
    st["bx"][:] = (br * bscale)[:, None, :]
    st["bz"][:] = (bz * bscale)[:, None, :]
    fl = np.zeros(shape)
    fl[3:-3, 3:-3, 3:-3] = rho_floor
    st["rho_floor"] = fl
    st["p_floor"] = k_poly * fl ** gam
    with np.errstate(divide="ignore", invalid="ignore"):
        st["eps_floor"] = np.where(fl > 0, k_poly * fl ** (gam - 1.0) / (gam - 1.0), 0.0)
    return st


def _sph_face_field(cfg: Config, g: Grid, m: dict, ac: np.ndarray):
    """Face Bx (radial) and By (polar) on faces j,k = 0..n from a corner
    vector potential A_phi(xF(j), yF(k)) given as ac[k+1, j+1] for j,k = -1..n,
    by the spherical curl of model/torus-sph/src/model.f90:146-152:
    Bx = (sinf(k) A(j,k) - sinf(k-1) A(j,k-1)) / (xF(j) dcose(k) + tiny),
    By = -(xF(j) A(j,k) - xF(j-1) A(j-1,k)) / (x(j) dx(j))."""
    nx, ny = cfg.nx, cfg.ny
    tiny = np.finfo(np.float64).tiny                        # parameter.h:24 small_num
    xF = g.x.faces[2:nx + 4]                                # xF(-1..nx)
    xc = g.x.c[2:nx + 3][None, :]                           # x(0..nx)
    dx = g.x.d[2:nx + 3][None, :]
    sinf = m["sinf"][2:ny + 4][:, None]                     # sinf(-1..ny)
    dcose = m["dcose"][3:ny + 4][:, None]                   # dcose(0..ny)
    bx = (sinf[1:] * ac[1:, 1:] - sinf[:-1] * ac[:-1, 1:]) / (xF[None, 1:] * dcose + tiny)
    by = -(xF[None, 1:] * ac[1:, 1:] - xF[None, :-1] * ac[1:, :-1]) / (xc * dx)
    return bx, by                                           # [k, j] for k,j = 0..n


def torus_sph(cfg: Config, seed: int = 20261202, p_beta: float = 350.0, q_grad: float = 2.0,
              s_in: float = 3.0, s_max: float = 4.7, rho_cut: float = 0.5, rho_fac: float = 1.0e-4,
              rho_max: float = 1.0, noise: float = 1.0e-4) -> dict:
    """Spherical (r, theta, phi) torus of model/torus-sph/src/model.f90:57-301:
    a power-law angular-momentum torus (q = 2) in the Paczynski-Wiita
    potential -1/(r - r_sh), polytropic with rho_max = 1 at s_max, inner edge
    s_in, atmosphere rho_fac*rho_max; poloidal loops from
    A_phi = max(rho - rho_cut, 0)/rho_max at cell corners, normalised to
    minimum plasma beta p_beta; density perturbed by 1e-4 relative noise inside
    the torus.  Parameters default to model/torus-sph/src/param.h."""
    assert cfg.coordinate_flag == 2
    g = Grid(cfg)
    m = g.spherical()
    gam, rsh = cfg.ggas, cfg.r_sh
    nx, ny = cfg.nx, cfg.ny
    psi = lambda r: -1.0 / (r - rsh)
    e = 2.0 - 2.0 * q_grad
    lkep2 = s_max ** (2.0 * q_grad - 1.0) / (s_max - rsh) ** 2
    c_const = psi(s_in) - lkep2 * s_in ** e / e
    h_max = -psi(s_max) + lkep2 * s_max ** e / e + c_const
    k_poly = h_max * (gam - 1.0) / gam / rho_max ** (gam - 1.0)
    if k_poly < 0.0:
        raise ValueError("the polytropic constant is negative")
    R = g.x.c[3:nx + 3][None, :]                            # x(1..nx)
    s_eq = R * m["sine"][4:ny + 4][:, None]                 # x(j) sine(k), k = 1..ny
    h = np.maximum(-psi(R) + lkep2 * s_eq ** e / e + c_const, 0.0)
    rho_loc = (h * (gam - 1.0) / gam / k_poly) ** (1.0 / (gam - 1.0))
    rho_atm = rho_fac * rho_max
    inside = (rho_loc > rho_atm) & (s_eq >= s_in)
    rho = np.where(inside, rho_loc, rho_atm)
    vphi = np.where(inside, math.sqrt(lkep2) * s_eq ** (-q_grad) * s_eq, 0.0)
    A = np.zeros((ny + 2, nx + 2))                          # A_phi(j,k), j,k = 0..n+1 at [k, j]
    A[1:ny + 1, 1:nx + 1] = np.where(rho >= rho_cut, (rho - rho_cut) / rho_max, 0.0)

    def faces(scale):
        Av = A * scale
        ac = np.zeros((ny + 2, nx + 2))                     # A_corner(j,k), j,k = -1..n at [k+1, j+1]
        ac[2:, 2:] = 0.25 * (Av[1:ny + 1, 1:nx + 1] + Av[1:ny + 1, 2:] + Av[2:, 1:nx + 1] + Av[2:, 2:])
        return _sph_face_field(cfg, g, m, ac)

    bx, by = faces(1.0)
    # minimum-beta normalisation over cells with field (model.f90:160-180)
    xF = g.x.faces
    dcb, dxx = m["dx_cb"][4:nx + 4][None, :], g.x.d[3:nx + 3][None, :]
    bcx = 1.5 * dxx / dcb * (xF[4:nx + 4] ** 2 * bx[1:, 1:] + xF[3:nx + 3] ** 2 * bx[1:, :-1])
    dyy, dce = g.y.d[3:ny + 3][:, None], m["dcose"][4:ny + 4][:, None]
    sf = m["sinf"]
    bcy = 0.5 * dyy / dce * (sf[4:ny + 4][:, None] * by[1:, 1:] + sf[3:ny + 3][:, None] * by[:-1, 1:])
    b2 = bcx ** 2 + bcy ** 2
    p = k_poly * rho ** gam
    if np.any(b2 > 0):
        beta_min = float(np.min(2.0 * p[b2 > 0] / b2[b2 > 0]))
        bx, by = faces(math.sqrt(beta_min / p_beta))
    st = empty_state(cfg)
    I = (slice(3, cfg.nz + 3), slice(3, ny + 3), slice(3, nx + 3))
    F = (slice(2, cfg.nz + 3), slice(2, ny + 3), slice(2, nx + 3))   # faces 0..n
    st["rho"][I] = rho
    st["vz"][I] = vphi
    st["p"][I] = p
    st["eps"][I] = k_poly * rho ** (gam - 1.0) / (gam - 1.0)
    st["bx"][F] = bx
    st["by"][F] = by
    fl = np.zeros(cfg.shape)
    fl[I] = rho_atm
    st["rho_floor"] = fl
    st["p_floor"] = k_poly * fl ** gam
    st["eps_floor"] = np.where(fl > 0, k_poly * fl ** (gam - 1.0) / (gam - 1.0), 0.0)
    u = interior_uniform(cfg, seed)
    r = st["rho"]
    st["rho"] = np.where(r > fl, r + r * (u - 0.5) / 0.5 * noise, r)
    return st


def sph_blast(cfg: Config, seed: int = 20261202, centre=(1.0, 0.35, 1.0), radius: float = 0.3,
              p_in: float = 10.0, p_out: float = 0.1, b0: float = 1.0) -> dict:
    """Magnetised blast on a spherical grid near the polar axis: uniform
    density 1, an over-pressured ball (centre given as r, theta, phi) in a
    uniform field b0 along the polar axis, set from the corner vector
    potential A_phi = b0 r sin(theta)/2 by the same spherical curl as the
    torus, so the face field is divergence-free to round-off."""
    assert cfg.coordinate_flag == 2
    g = Grid(cfg)
    m = g.spherical()
    nx, ny = cfg.nx, cfg.ny
    X, Y, Z = g.centres()                                    # r, theta, phi of the cell centres
    r0, t0, f0 = centre
    cx = np.array([r0 * math.sin(t0) * math.cos(f0), r0 * math.sin(t0) * math.sin(f0), r0 * math.cos(t0)])
    px = X * np.sin(Y) * np.cos(Z) - cx[0]
    py = X * np.sin(Y) * np.sin(Z) - cx[1]
    pz = X * np.cos(Y) - cx[2]
    inside = np.sqrt(px * px + py * py + pz * pz) < radius
    st = empty_state(cfg)
    st["rho"][:] = 1.0
    st["p"][:] = np.where(inside, p_in, p_out)
    st["eps"][:] = st["p"] / st["rho"] / (cfg.ggas - 1.0)
    xF = g.x.faces[2:nx + 4]                                 # xF(-1..nx)
    sinf = m["sinf"][2:ny + 4]                               # sinf(-1..ny)
    ac = 0.5 * b0 * xF[None, :] * sinf[:, None]
    bx, by = _sph_face_field(cfg, g, m, ac)
    F = (slice(2, cfg.nz + 3), slice(2, ny + 3), slice(2, nx + 3))
    st["bx"][F] = bx
    st["by"][F] = by
    return st


# model/collapsar/profile/par.dat: ggas, black-hole mass (code units) and the
# cgs-to-code factors of mass, length and time (G = c = M_sun = 1)
COLLAPSAR_UNITS = {"m_bh": 1.9027781052250767, "mass": 5.028992139685286e-34,
                   "length": 6.771995159595613e-06, "time": 203019.30744592712}
CLIGHT = 2.99792458e10                                       # model/collapsar/src/param.h:10


def read_collapsar_profile(path: str) -> dict:
    """The collapsar's profile directory (model/collapsar/profile): face
    coordinates r_grid.dat, th_grid.dat, phi_grid.dat (one value per line,
    -3..n+3; model/collapsar/src/custom.f90:68-155), the radial rho, p of
    hydro.dat over -2..nx+3 and par.dat (ggas, m_bh, mass, length and time
    factors; model/collapsar/src/model.f90:44-99)."""
    import os

    def col(name, ncol=1):
        with open(os.path.join(path, name)) as f:
            rows = [ln.split() for ln in f if ln.strip()]
        return np.array([[float(v) for v in r[:ncol]] for r in rows])

    par = col("par.dat", 5)[0]
    hydro = col("hydro.dat", 2)
    out = {"xF": col("r_grid.dat")[:, 0], "yF": col("th_grid.dat")[:, 0], "rho": hydro[:, 0], "p": hydro[:, 1],
           "ggas": float(par[0]), "m_bh": float(par[1]), "mass": float(par[2]), "length": float(par[3]),
           "time": float(par[4])}
    if os.path.exists(os.path.join(path, "phi_grid.dat")):
        out["zF"] = col("phi_grid.dat")[:, 0]
    if len(out["rho"]) != len(out["xF"]) - 1:
        raise ValueError("hydro.dat needs nx+6 rows for nx+7 radial faces")
    return out


def collapsar(cfg: Config, seed: int = 20261202, profile: dict | None = None, rho_env: float = 2.0e-11,
              rho_atm: float = 2.6656378907450677e-15, k_env: float = 1.2, k_atm: float = 2.6,
              slope: float = 2.5, cavity: float = 3.0, b_gauss: float = 1.0e10, w_cgs: float = 10.0,
              noise: float = 1.0e-4) -> dict:
    """Initial model of model/collapsar/src/model.f90:6-252 on cfg's grid:
    the radial rho, p profile copied to every (theta, phi); epsilon =
    p/rho/(ggas-1) taken before the density noise; differential rotation
    v_phi = w_0 s_core^2/(s_core^2 + s^2) s (rotation_rule 2) where rho
    exceeds the atmosphere rho(nx), with s_core the first cell above it; the
    uniform field of A_phi = xF sinf/2 scaled to b_0 (1e10 G in code units);
    1e-4 relative density noise in the star.  The atmosphere prim_a = (rho,
    0, 0, 0, p) of cell nx, eps_a, m_bh and vel2code go into st["gravity"].

    `profile` is a read_collapsar_profile() dict on cfg's faces; without it the
    radial profile is synthetic: an envelope rho_env (r/r_c)^-slope with
    p = k_env rho^(4/3) from r_c = cavity * xF(0) out to where it meets the
    atmosphere rho_atm (p = k_atm rho_atm^(4/3)), atmosphere inside r_c."""
    assert cfg.model == "collapsar" and cfg.coordinate_flag == 2
    g = Grid(cfg)
    m = g.spherical()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    gam = cfg.ggas
    units = dict(COLLAPSAR_UNITS)
    if profile is not None:
        rho1, p1 = np.asarray(profile["rho"], float), np.asarray(profile["p"], float)
        units.update({k: profile[k] for k in ("m_bh", "mass", "length", "time")})
    else:
        r = g.x.c                                            # x(-2..nx+3)
        r_c = cavity * g.x.faces[3]
        env = rho_env * (r / r_c) ** (-slope)
        star = (r >= r_c) & (env > rho_atm)
        star[nx + 2:] = False                                # cell nx and beyond: atmosphere
        rho1 = np.where(star, env, rho_atm)
        p1 = np.where(star, k_env * rho1 ** (4.0 / 3.0), k_atm * rho_atm ** (4.0 / 3.0))
    if rho1.shape != (nx + 6,):
        raise ValueError("the radial profile needs nx+6 values (-2..nx+3)")
    vel2code = units["length"] / units["time"]
    gauss2code = (units["mass"] / units["length"]) ** 0.5 / units["time"]
    st = empty_state(cfg)
    st["rho"][:] = rho1[None, None, :]
    st["p"][:] = p1[None, None, :]
    st["eps"][:] = st["p"] / st["rho"] / (gam - 1.0)
    rho_nx = rho1[nx + 2]
    above = rho1 > rho_nx
    s_core = None
    for j in range(1, nx + 1):                               # model.f90:106-112
        if rho1[j + 1] <= rho_nx and rho1[j + 2] > rho_nx:
            s_core = g.x.c[j + 2]
            break
    I = (slice(3, nz + 3), slice(3, ny + 3), slice(3, nx + 3))
    if s_core is not None:
        w_0 = w_cgs / units["time"]
        s_eq = g.x.c[3:nx + 3][None, :] * m["sine"][4:ny + 4][:, None]
        omega = w_0 * s_core ** 2 / (s_core ** 2 + s_eq ** 2)
        st["vz"][I] = np.where(above[3:nx + 3][None, :], omega * s_eq, 0.0)[None]
    xF = g.x.faces[2:nx + 4]                                 # xF(-1..nx)
    sinf = m["sinf"][2:ny + 4]                               # sinf(-1..ny)
    ac = xF[None, :] * sinf[:, None] * 0.5                   # A_phi(j,k) = xF(j) sinf(k)/2
    bx, by = _sph_face_field(cfg, g, m, ac)
    b_0 = b_gauss * gauss2code
    F = (slice(2, nz + 3), slice(2, ny + 3), slice(2, nx + 3))
    st["bx"][F] = bx * b_0
    st["by"][F] = by * b_0
    u = interior_uniform(cfg, seed)
    r3 = st["rho"]
    inside = np.zeros(cfg.shape, bool)
    inside[I] = True
    st["rho"] = np.where(inside & (r3 > rho_nx), r3 * (1.0 + noise * (u - 0.5) / 0.5), r3)
    prim_a = np.zeros(8)
    prim_a[0], prim_a[4] = rho_nx, p1[nx + 2]
    eps_a = prim_a[4] / prim_a[0] / (gam - 1.0)
    st["gravity"] = np.concatenate([prim_a, [eps_a, units["m_bh"], vel2code]])
    return st


GENERATORS = {
    "briowu": brio_wu,
    "orszag_tang": orszag_tang,
    "blast": blast,
    "turbulence": turbulence,
    "torus": torus,
    "torus_sph": torus_sph,
    "sph_blast": sph_blast,
    "collapsar": collapsar,
}
