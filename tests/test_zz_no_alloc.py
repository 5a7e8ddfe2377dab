"""cumc3d_run_noalloc, the time loop that allocates nothing, against
cumc3d_run through the C-ABI.  Needs an MI355X.

Two calls in a row on fresh contexts, the second on a smaller grid with
fewer steps: the dt log (which cumc3d_run_noalloc keeps in a static device
array shared by every call on the GPU) and the whole downloaded state must be
bit-identical to cumc3d_run's.  Also: max_steps beyond the array's size with
a dt log is rejected before anything runs, and kernel timing is refused."""
import ctypes as C

import numpy as np
import pytest

from cumc3d_amd import ic
from cumc3d_amd.config import orszag_tang_3d
from cumc3d_amd.solver import CumcError, load_library, make_solver, _p

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz", "eps", "bcx", "bcy", "bcz", "cs")


def run_pair(cfg, steps, sync):
    st = ic.orszag_tang(cfg)
    a = make_solver(cfg, st)
    b = make_solver(cfg, st)
    da = a.run(steps, total_time=1e300, sync_each_step=sync)
    db = b.run(steps, total_time=1e300, sync_each_step=sync, noalloc=True)
    assert len(da) == steps and np.array_equal(da, db), (da, db)
    assert a.t == b.t and a.nstep == b.nstep == steps
    sa, sb = a.state(), b.state()
    for f in FIELDS:
        assert np.array_equal(sa[f], sb[f]), f


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_run_noalloc_matches_run(precision):
    print(f"noalloc vs run: 24^3 WENO {precision}, 10 steps, then 12^3 PPM, 4 steps", flush=True)
    run_pair(orszag_tang_3d(24, recon="weno").with_(precision=precision), 10, False)
    run_pair(orszag_tang_3d(12, recon="ppm").with_(precision=precision), 4, True)


def test_run_noalloc_limits():
    cfg = orszag_tang_3d(12, recon="ppm")
    s = make_solver(cfg, ic.orszag_tang(cfg))
    L = load_library()
    t = C.c_double(0.0)
    n = C.c_int(-1)
    big = np.zeros(65537)
    rc = L.cumc3d_run_noalloc(s._h, 65537, 1e300, 0, C.byref(t), _p(big), C.byref(n))
    assert rc == 1 and n.value == -1 and t.value == 0.0        # CUMC3D_EINVAL, nothing ran
    s.set_timing(True)
    with pytest.raises(CumcError) as e:
        s.run(2, total_time=1e300, noalloc=True)
    assert e.value.code == 4                                    # CUMC3D_ESTATE
    s.set_timing(False)
    # without a dt log any step count runs (here 3)
    rc = L.cumc3d_run_noalloc(s._h, 3, 1e300, 0, C.byref(t), None, C.byref(n))
    assert rc == 0 and n.value == 3 and t.value > 0.0
