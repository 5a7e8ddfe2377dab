"""ctypes binding of the C restatement oracle/mhd_oracle.c (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; it is the checker, never the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

from cumc3d_amd.config import MODELS, RECON, RIEMANN, Config  # noqa: E402
from cumc3d_amd.grid import Grid  # noqa: E402
from cumc3d_amd.ic import CLIGHT, PRIM_ORDER  # noqa: E402

PD = C.POINTER(C.c_double)


class _Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_dim", C.c_int),
                ("coordinate_flag", C.c_int), ("boundary_flag", C.c_int * 6),
                ("riemann", C.c_int), ("recon", C.c_int), ("model", C.c_int),
                ("rt_gravity", C.c_int), ("hlld_bn_mode", C.c_int),
                ("cfl", C.c_double), ("ggas", C.c_double), ("r_sh", C.c_double),
                ("xF", PD), ("yF", PD), ("zF", PD), ("fixpole", C.c_int),
                ("prim_a", C.c_double * 4), ("eps_a", C.c_double), ("m_bh", C.c_double),
                ("c_code", C.c_double), ("omega", C.c_double), ("tolerance", C.c_double),
                ("relax_max", C.c_int), ("n_pot", C.c_int), ("sor_fold", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(_Desc)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_load.argtypes = [C.c_void_p, PD, PD, PD]
        L.oracle_initial_update.argtypes = [C.c_void_p]
        L.oracle_finddt.argtypes = [C.c_void_p]
        L.oracle_finddt.restype = C.c_double
        L.oracle_rungekutta.argtypes = [C.c_void_p, C.c_double]
        L.oracle_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.oracle_spatial.argtypes = [C.c_void_p]
        L.oracle_run.argtypes = [C.c_void_p, C.c_int, C.c_double, PD, PD]
        L.oracle_run.restype = C.c_int
        L.oracle_find_divb.argtypes = [C.c_void_p]
        L.oracle_find_divb.restype = C.c_double
        L.oracle_get.argtypes = [C.c_void_p, C.c_char_p, PD]
        L.oracle_grid.argtypes = [C.c_void_p, C.c_char_p, PD]
        L.oracle_gravity.argtypes = [C.c_void_p, PD, PD, PD, C.POINTER(C.c_int)]
        L.oracle_gravity.restype = C.c_int
        L.oracle_poisson.argtypes = [C.c_void_p, C.c_char_p, PD]
        L.oracle_poisson.restype = C.c_long
        L.oracle_cells.argtypes = [C.c_void_p]
        L.oracle_cells.restype = C.c_long
        L.oracle_tvd.argtypes = [C.c_int, PD, C.c_double, C.c_double, C.c_double, PD, PD]
        L.oracle_ppm.argtypes = [PD, PD, PD, PD]
        L.oracle_weno.argtypes = [PD, PD, PD, PD]
        L.oracle_ppm_weight.argtypes = [C.c_double] * 5 + [PD]
        L.oracle_weno_weight.argtypes = [C.c_double] * 5 + [PD]
        L.oracle_riemann_point.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, PD, PD, PD]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(PD)


class Oracle:
    """One CPU instance of the reference algorithm on a configuration."""

    def __init__(self, cfg: Config, state: dict, hlld_bn_mode: int = 2, grid: Grid | None = None,
                 fixpole: int | None = None, sor_fold: int = 0):
        self.cfg = cfg
        self.grid = grid or Grid(cfg)
        L = lib()
        self._faces = [np.ascontiguousarray(a.faces) for a in (self.grid.x, self.grid.y, self.grid.z)]
        d = _Desc()
        d.nx, d.ny, d.nz, d.n_dim = cfg.nx, cfg.ny, cfg.nz, cfg.n_dim
        d.coordinate_flag = cfg.coordinate_flag
        for i, b in enumerate(cfg.boundary_flag):
            d.boundary_flag[i] = b
        d.riemann = RIEMANN[cfg.riemann]
        d.recon = RECON[cfg.recon]
        d.model = MODELS[cfg.model]
        d.rt_gravity = int(cfg.rt_gravity)
        d.hlld_bn_mode = hlld_bn_mode
        d.cfl, d.ggas, d.r_sh = cfg.cfl, cfg.ggas, cfg.r_sh
        d.xF, d.yF, d.zF = (_p(f) for f in self._faces)
        d.fixpole = int(cfg.fixpole) if fixpole is None else fixpole
        if cfg.model == "collapsar":
            gv = state["gravity"]
            for i in range(4):
                d.prim_a[i] = gv[i]
            d.eps_a, d.m_bh = gv[8], gv[9]
            d.c_code = CLIGHT * gv[10]                       # clight*vel2code, custom.f90:481
            d.omega, d.tolerance = cfg.sor_omega, cfg.sor_tolerance
            d.relax_max, d.n_pot = cfg.relax_max, cfg.n_pot
            d.sor_fold = sor_fold
        self._h = L.oracle_create(C.byref(d))
        if not self._h:
            raise ValueError("invalid boundary flag combination")
        prim = np.ascontiguousarray(np.stack([state[n] for n in PRIM_ORDER], axis=-1))
        eps = np.ascontiguousarray(state["eps"])
        floors = None
        if "rho_floor" in state:
            floors = np.ascontiguousarray(np.stack([state["rho_floor"], state["p_floor"],
                                                    state["eps_floor"]]))
        L.oracle_load(self._h, _p(prim), _p(eps), _p(floors) if floors is not None else None)
        L.oracle_initial_update(self._h)
        self.t = 0.0
        self.nstep = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_destroy(h)
            self._h = None

    def finddt(self) -> float:
        return lib().oracle_finddt(self._h)

    def rk3(self, dt: float) -> None:
        lib().oracle_rungekutta(self._h, dt)

    def stage(self, s: int, dt: float) -> None:
        """One stage of RUNGEKUTTA; stages run 1, 2, 3 in order."""
        lib().oracle_stage(self._h, s, dt)

    def spatial(self) -> dict:
        """SPATIAL on the current state: l_rk as lrk1..lrk8 on the ghosted box."""
        lib().oracle_spatial(self._h)
        return {f"lrk{i}": self.get(f"lrk{i}") for i in range(1, 9)}

    def run(self, max_steps: int, total_time: float | None = None) -> np.ndarray:
        T = self.cfg.total_time if total_time is None else total_time
        dts = np.zeros(max(max_steps, 1))
        t = C.c_double(self.t)
        n = lib().oracle_run(self._h, max_steps, T, C.byref(t), _p(dts))
        self.t = t.value
        self.nstep += n
        return dts[:n]

    def gravity(self) -> dict:
        """Collapsar: potential phi (ghosted box), m_bh, r_sh, the SOR
        iterations of the last solve and the status (0 running, 1 the horizon
        reached xF(0), 2 the solve converged on its last allowed sweep)."""
        phi = np.empty(self.cfg.shape)
        m, r, it = C.c_double(), C.c_double(), C.c_int()
        status = lib().oracle_gravity(self._h, _p(phi), C.byref(m), C.byref(r), C.byref(it))
        if status < 0:
            raise ValueError("not a collapsar run")
        return {"phi": phi, "m_bh": m.value, "r_sh": r.value, "iterations": it.value, "status": status}

    def poisson(self, name: str) -> np.ndarray:
        """poisson_coef coefficient `name`: ajp1/ajm1 as [j], bkp1/bkm1 as
        [k, j], clp1/clm1/epsc on the ghosted box."""
        cfg = self.cfg
        size = {"ajp1": cfg.nx + 1, "ajm1": cfg.nx + 1, "bkp1": (cfg.nx + 1) * (cfg.ny + 1),
                "bkm1": (cfg.nx + 1) * (cfg.ny + 1)}.get(name, int(np.prod(cfg.shape)))
        out = np.zeros(size)
        if lib().oracle_poisson(self._h, name.encode(), _p(out)) < 0:
            raise KeyError(name)
        if name in ("bkp1", "bkm1"):
            return out.reshape(cfg.ny + 1, cfg.nx + 1)
        if name in ("clp1", "clm1", "epsc"):
            return out.reshape(cfg.shape)
        return out

    def divb(self) -> float:
        return lib().oracle_find_divb(self._h)

    def grid_metric(self, name: str) -> np.ndarray:
        """GETGRID array `name` over -3..n+3 (index j+3)."""
        out = np.zeros(max(self.cfg.nx, self.cfg.ny, self.cfg.nz) + 7)
        n = lib().oracle_grid(self._h, name.encode(), _p(out))
        if n < 0:
            raise KeyError(name)
        return out[:n]

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.cfg.shape)
        if lib().oracle_get(self._h, name.encode(), _p(out)) != 0:
            raise KeyError(name)
        return out

    def state(self) -> dict:
        names = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "cs", "bcx", "bcy", "bcz")
        return {n: self.get(n) for n in names}


# point kernels -------------------------------------------------------------

def tvd(limiter: int, w, vm, vc, vp):
    w = np.ascontiguousarray(w, dtype=np.float64)
    a, b = C.c_double(), C.c_double()
    lib().oracle_tvd(limiter, _p(w), vm, vc, vp, C.byref(a), C.byref(b))
    return a.value, b.value


def ppm(w, v):
    w = np.ascontiguousarray(w, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    a, b = C.c_double(), C.c_double()
    lib().oracle_ppm(_p(w), _p(v), C.byref(a), C.byref(b))
    return a.value, b.value


def weno(w, v):
    w = np.ascontiguousarray(w, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    a, b = C.c_double(), C.c_double()
    lib().oracle_weno(_p(w), _p(v), C.byref(a), C.byref(b))
    return a.value, b.value


def ppm_weight(d):
    w = np.zeros(14)
    lib().oracle_ppm_weight(*[float(x) for x in d], _p(w))
    return w


def weno_weight(d):
    w = np.zeros(10)
    lib().oracle_weno_weight(*[float(x) for x in d], _p(w))
    return w


def riemann_point(solver: str, direction: int, ggas: float, qL, qR, hlld_bn_mode: int = 0):
    """qL, qR: 8 primitives (rho, vx, vy, vz, p, bx, by, bz); returns 8 fluxes."""
    a = np.zeros(9)
    b = np.zeros(9)
    a[1:] = qL
    b[1:] = qR
    f = np.zeros(9)
    lib().oracle_riemann_point(RIEMANN[solver], direction, ggas, hlld_bn_mode, _p(a), _p(b), _p(f))
    return f[1:].copy()
