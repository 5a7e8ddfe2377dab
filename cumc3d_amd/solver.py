"""Host-side binding of libcumc3d_hip.so (include/cumc3d_hip.h).

``Solver`` mirrors the reference's driver sequence (solver/mhd/main.f90:50-139):
``initial_model`` (INITIAL_MODEL + POPULATE_DEVICE), ``initial_update``
(INITIAL_UPDATE), ``finddt`` (FINDDT), ``rungekutta`` (RUNGEKUTTA), ``run`` (the
time loop), ``find_divb`` (FIND_DIVB) and ``state`` (the host update done by
print_hydroprofile).  Everything runs in the HIP library; this module only
marshals arguments.  It raises when the library is not built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .config import MODELS, PRECISION, RECON, RIEMANN, Config
from .grid import Grid
from .ic import PRIM_ORDER

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcumc3d_hip.so")

PD = C.POINTER(C.c_double)

# symbols declared in include/cumc3d_hip.h
EXPORTS = ("cumc3d_init", "cumc3d_upload", "cumc3d_initial_update", "cumc3d_finddt",
           "cumc3d_rk3_step", "cumc3d_stage", "cumc3d_spatial", "cumc3d_run", "cumc3d_run_noalloc",
           "cumc3d_download", "cumc3d_profile", "cumc3d_gravity",
           "cumc3d_divb",
           "cumc3d_finalize", "cumc3d_error_string", "cumc3d_last_error",
           "cumc3d_set_timing", "cumc3d_get_timing",
           "cumc3d_selector_init", "cumc3d_selector_set_params", "cumc3d_selector_get_params",
           "cumc3d_selector_collect", "cumc3d_selector_train", "cumc3d_selector_enable",
           "cumc3d_selector_mlp", "cumc3d_selector_mask")

SEL_NPARAM = 1699

BN_MODES = {"zero": 0, "field": 1, "carry": 2}


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_dim", C.c_int),
                ("coordinate_flag", C.c_int), ("boundary_flag", C.c_int * 6),
                ("riemann", C.c_int), ("recon", C.c_int), ("model", C.c_int),
                ("rt_gravity", C.c_int), ("hlld_bn_mode", C.c_int), ("precision", C.c_int),
                ("cfl", C.c_double), ("ggas", C.c_double), ("r_sh", C.c_double),
                ("xF", PD), ("yF", PD), ("zF", PD), ("fixpole", C.c_int), ("metrics", C.c_void_p),
                ("gravity", C.c_void_p)]


class GravityParams(C.Structure):
    """cumc3d_gravity_params: the collapsar's CUSTOM_DEF values after GET_MODEL."""
    _fields_ = [("prim_a", C.c_double * 4), ("eps_a", C.c_double), ("m_bh", C.c_double),
                ("c_code", C.c_double), ("omega", C.c_double), ("tolerance", C.c_double),
                ("relax_max", C.c_int), ("n_pot", C.c_int)]


def gravity_params(cfg: Config, gravity) -> GravityParams:
    """From a state's "gravity" record (prim_a(1:8), eps_a, m_bh, vel2code)."""
    from .ic import CLIGHT
    g = GravityParams()
    for i in range(4):
        g.prim_a[i] = float(gravity[i])
    g.eps_a, g.m_bh = float(gravity[8]), float(gravity[9])
    g.c_code = CLIGHT * float(gravity[10])          # clight*vel2code (model/collapsar/src/custom.f90:481)
    g.omega, g.tolerance = cfg.sor_omega, cfg.sor_tolerance
    g.relax_max, g.n_pot = cfg.relax_max, cfg.n_pot
    return g


class CumcError(RuntimeError):
    """A non-zero return of the library; ``code`` is the CUMC3D_E* value."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


ENONFINITE = 5   # CUMC3D_ENONFINITE: a cell's CFL time step was NaN or not positive


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libcumc3d_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CumcError(f"{path} is missing: build it with `python -m cumc3d_amd.build`")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.cumc3d_init.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
    L.cumc3d_upload.argtypes = [vp, PD, PD, PD]
    L.cumc3d_initial_update.argtypes = [vp]
    L.cumc3d_finddt.argtypes = [vp, PD]
    L.cumc3d_rk3_step.argtypes = [vp, C.c_double]
    L.cumc3d_stage.argtypes = [vp, C.c_int, C.c_double]
    L.cumc3d_spatial.argtypes = [vp, PD]
    L.cumc3d_run.argtypes = [vp, C.c_int, C.c_double, C.c_int, PD, PD, C.POINTER(C.c_int)]
    L.cumc3d_run_noalloc.argtypes = [vp, C.c_int, C.c_double, C.c_int, PD, PD, C.POINTER(C.c_int)]
    L.cumc3d_download.argtypes = [vp, PD, PD, PD, PD]
    L.cumc3d_profile.argtypes = [vp, PD, PD, PD]
    L.cumc3d_gravity.argtypes = [vp, PD, PD, PD, C.POINTER(C.c_int)]
    L.cumc3d_divb.argtypes = [vp, PD]
    L.cumc3d_finalize.argtypes = [vp]
    L.cumc3d_error_string.argtypes = [C.c_int]
    L.cumc3d_error_string.restype = C.c_char_p
    L.cumc3d_last_error.restype = C.c_char_p
    L.cumc3d_set_timing.argtypes = [vp, C.c_int]
    L.cumc3d_get_timing.argtypes = [vp, PD, C.c_int]
    PF = C.POINTER(C.c_float)
    L.cumc3d_selector_init.argtypes = [vp, C.c_int, C.c_uint]
    L.cumc3d_selector_set_params.argtypes = [vp, PF, C.c_int]
    L.cumc3d_selector_get_params.argtypes = [vp, PF, C.c_int]
    L.cumc3d_selector_collect.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_uint, C.POINTER(C.c_int)]
    L.cumc3d_selector_train.argtypes = [vp, C.c_int, C.c_int, C.c_double, PD, C.c_uint, PF]
    L.cumc3d_selector_enable.argtypes = [vp, C.c_int, C.c_int]
    L.cumc3d_selector_mlp.argtypes = [vp, PF, C.c_int, PF]
    L.cumc3d_selector_mask.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_ubyte)]
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(PD)


def _pf(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _check(rc: int):
    if rc != 0:
        L = load_library()
        raise CumcError(f"{L.cumc3d_error_string(rc).decode()}: {L.cumc3d_last_error().decode()}", rc)


class Solver:
    """One device context for a configuration.  The collapsar model needs the
    gravity record of its initial model (``state["gravity"]``) at init: pass
    it as ``gravity`` or the context is created by ``initial_model``."""

    def __init__(self, cfg: Config, hlld_bn_mode: str = "carry", grid: Grid | None = None, gravity=None):
        self.cfg = cfg
        self.grid = grid or Grid(cfg)
        self._bn_mode = hlld_bn_mode
        self._h = None
        self.t = 0.0
        self.nstep = 0
        if cfg.model == "collapsar" and gravity is None:
            return
        self._create(gravity)

    def _create(self, gravity):
        cfg = self.cfg
        hlld_bn_mode = self._bn_mode
        L = load_library()
        self._faces = [np.ascontiguousarray(a.faces) for a in (self.grid.x, self.grid.y, self.grid.z)]
        d = Desc()
        d.nx, d.ny, d.nz, d.n_dim = cfg.nx, cfg.ny, cfg.nz, cfg.n_dim
        d.coordinate_flag = cfg.coordinate_flag
        for i, b in enumerate(cfg.boundary_flag):
            d.boundary_flag[i] = b
        d.riemann = RIEMANN[cfg.riemann]
        d.recon = RECON[cfg.recon]
        d.model = MODELS[cfg.model]
        d.rt_gravity = int(cfg.rt_gravity)
        d.hlld_bn_mode = BN_MODES[hlld_bn_mode]
        d.precision = PRECISION[cfg.precision]
        d.cfl, d.ggas, d.r_sh = cfg.cfl, cfg.ggas, cfg.r_sh
        d.xF, d.yF, d.zF = (_p(f) for f in self._faces)
        d.fixpole = int(cfg.fixpole)
        d.metrics = None
        self._gp = None
        if cfg.model == "collapsar":
            self._gp = gravity_params(cfg, gravity)
            d.gravity = C.cast(C.pointer(self._gp), C.c_void_p)
        h = C.c_void_p()
        _check(L.cumc3d_init(C.byref(d), C.byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            load_library().cumc3d_finalize(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference driver sequence ---------------------------------------
    def initial_model(self, state: dict):
        """Upload GET_MODEL output (prim, epsilon, floors) to the device."""
        cfg = self.cfg
        if self._h is None:
            self._create(state["gravity"])
        prim = np.ascontiguousarray(np.stack([state[n] for n in PRIM_ORDER], axis=-1), dtype=np.float64)
        if prim.shape != cfg.shape + (8,):
            raise ValueError(f"state arrays must have shape {cfg.shape}")
        eps = np.ascontiguousarray(state["eps"], dtype=np.float64)
        floors = None
        if "rho_floor" in state:
            floors = np.ascontiguousarray(np.stack([state["rho_floor"], state["p_floor"], state["eps_floor"]]),
                                          dtype=np.float64)
        _check(load_library().cumc3d_upload(self._h, _p(prim), _p(eps), _p(floors)))
        self.t = 0.0
        self.nstep = 0

    def initial_update(self):
        _check(load_library().cumc3d_initial_update(self._h))

    def finddt(self) -> float:
        dt = C.c_double()
        _check(load_library().cumc3d_finddt(self._h, C.byref(dt)))
        return dt.value

    def rungekutta(self, dt: float):
        _check(load_library().cumc3d_rk3_step(self._h, dt))
        self.t += dt
        self.nstep += 1

    def stage(self, stage: int, dt: float):
        """One stage of RUNGEKUTTA (1, 2, 3 in order, the step's dt each time);
        stage 3 ends the step."""
        _check(load_library().cumc3d_stage(self._h, stage, dt))
        if stage == 3:
            self.t += dt
            self.nstep += 1

    def spatial(self, l_rk: np.ndarray | None = None) -> dict:
        """SPATIAL on the current state: l_rk as lrk1..lrk8 (ghosted box
        arrays).  l_rk(1:5) is formed on cells 1..n and l_rk(6:8) on the faces
        the update advances; elsewhere the values of `l_rk` (shape + (8,),
        zeros if None) are kept."""
        cfg = self.cfg
        buf = np.zeros(cfg.shape + (8,)) if l_rk is None else np.array(l_rk, dtype=np.float64, order="C")
        if buf.shape != cfg.shape + (8,):
            raise ValueError(f"l_rk must have shape {cfg.shape + (8,)}")
        _check(load_library().cumc3d_spatial(self._h, _p(buf)))
        return {f"lrk{i + 1}": np.ascontiguousarray(buf[..., i]) for i in range(8)}

    def run(self, max_steps: int, total_time: float | None = None, sync_each_step: bool = True,
            noalloc: bool = False) -> np.ndarray:
        """main.f90's loop on the device; returns the dt of every step taken.
        noalloc: through cumc3d_run_noalloc (no device allocation in the call)."""
        T = self.cfg.total_time if total_time is None else total_time
        t = C.c_double(self.t)
        dts = np.zeros(max(max_steps, 1))
        n = C.c_int()
        L = load_library()
        fn = L.cumc3d_run_noalloc if noalloc else L.cumc3d_run
        _check(fn(self._h, max_steps, T, int(sync_each_step), C.byref(t), _p(dts), C.byref(n)))
        self.t = t.value
        self.nstep += n.value
        return dts[:n.value]

    def gravity(self) -> dict:
        """Collapsar: potential phi over the ghosted box, m_bh, r_sh and the
        iteration count of the last Poisson solve."""
        phi = np.empty(self.cfg.shape)
        m, r, it = C.c_double(), C.c_double(), C.c_int()
        _check(load_library().cumc3d_gravity(self._h, _p(phi), C.byref(m), C.byref(r), C.byref(it)))
        return {"phi": phi, "m_bh": m.value, "r_sh": r.value, "iterations": it.value}

    def find_divb(self) -> float:
        m = C.c_double()
        _check(load_library().cumc3d_divb(self._h, C.byref(m)))
        return m.value

    def state(self) -> dict:
        """Host copy of prim, epsilon, bcell and cs (ghosted box arrays)."""
        cfg = self.cfg
        prim = np.empty(cfg.shape + (8,))
        bcell = np.empty(cfg.shape + (3,))
        eps = np.empty(cfg.shape)
        cs = np.empty(cfg.shape)
        _check(load_library().cumc3d_download(self._h, _p(prim), _p(eps), _p(bcell), _p(cs)))
        out = {n: np.ascontiguousarray(prim[..., i]) for i, n in enumerate(PRIM_ORDER)}
        out.update(eps=eps, cs=cs, bcx=np.ascontiguousarray(bcell[..., 0]),
                   bcy=np.ascontiguousarray(bcell[..., 1]), bcz=np.ascontiguousarray(bcell[..., 2]))
        return out

    def profile(self) -> dict:
        """print_hydroprofile's arrays (dataio.f90:178-233) packed on the
        device: primitive (nz+2, ny+2, nx+2, 5), epsilon (nz+2, ny+2, nx+2)
        and bfield (nz+3, ny+3, nx+3, 3) -- C-order views of the reference's
        Fortran sections prim(1:5, 0:nx+1, ...), epsilon(0:nx+1, ...),
        prim(ibx:ibz, -1:nx+1, ...)."""
        cfg = self.cfg
        nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
        prim = np.empty((nz + 2, ny + 2, nx + 2, 5))
        eps = np.empty((nz + 2, ny + 2, nx + 2))
        bf = np.empty((nz + 3, ny + 3, nx + 3, 3))
        _check(load_library().cumc3d_profile(self._h, _p(prim), _p(eps), _p(bf)))
        return {"primitive": prim, "epsilon": eps, "bfield": bf}

    # -- measurement -------------------------------------------------------
    def set_timing(self, on: bool):
        _check(load_library().cumc3d_set_timing(self._h, int(on)))

    def timing(self) -> dict:
        ms = np.zeros(10)
        _check(load_library().cumc3d_get_timing(self._h, _p(ms), 10))
        keys = ("flux_x", "flux_y", "flux_z", "edge", "update", "ghost", "step", "selector",
                "poisson", "sor_sweeps")
        return dict(zip(keys, ms.tolist()))

    # -- learned reconstruction selector ---------------------------------------
    def selector_init(self, capacity: int = 1 << 20, seed: int = 1):
        _check(load_library().cumc3d_selector_init(self._h, capacity, seed))

    def selector_params(self, w: np.ndarray | None = None) -> np.ndarray:
        """Set (if w is given) and return the FP32 parameter vector."""
        L = load_library()
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float32)
            _check(L.cumc3d_selector_set_params(self._h, _pf(w), w.size))
        out = np.zeros(SEL_NPARAM, dtype=np.float32)
        _check(L.cumc3d_selector_get_params(self._h, _pf(out), SEL_NPARAM))
        return out

    def selector_collect(self, stride: int = 4, thr_ppm: float = 1e-3, thr_weno: float = 1e-2,
                         seed: int = 1) -> np.ndarray:
        """Append samples from the current state; returns (held, n0, n1, n2)."""
        c = (C.c_int * 4)()
        _check(load_library().cumc3d_selector_collect(self._h, stride, thr_ppm, thr_weno, seed, c))
        return np.array(list(c))

    def selector_train(self, iters: int, blocks: int = 64, lr: float = 3e-3, class_weight=(1.0, 1.0, 1.0),
                       seed: int = 1) -> np.ndarray:
        cw = np.ascontiguousarray(class_weight, dtype=np.float64)
        loss = np.zeros(max(iters, 1), dtype=np.float32)
        _check(load_library().cumc3d_selector_train(self._h, iters, blocks, lr, _p(cw), seed, _pf(loss)))
        return loss[:iters]

    def selector_enable(self, on: bool = True, every: int = 1):
        _check(load_library().cumc3d_selector_enable(self._h, int(on), every))

    def selector_mlp(self, feat: np.ndarray) -> np.ndarray:
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, 16)
        out = np.zeros((feat.shape[0], 3), dtype=np.float32)
        _check(load_library().cumc3d_selector_mlp(self._h, _pf(feat), feat.shape[0], _pf(out)))
        return out

    def selector_mask(self, direction: int, refresh: bool = True) -> np.ndarray:
        out = np.zeros(self.cfg.shape, dtype=np.uint8)
        _check(load_library().cumc3d_selector_mask(self._h, direction, int(refresh),
                                                   out.ctypes.data_as(C.POINTER(C.c_ubyte))))
        return out


def make_solver(cfg: Config, state: dict, **kw) -> Solver:
    s = Solver(cfg, **kw)
    s.initial_model(state)
    s.initial_update()
    return s
