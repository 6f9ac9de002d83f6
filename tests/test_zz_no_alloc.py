"""fr2d_step_dev_noalloc (no reference counterpart: fr2d_step for callers that cannot allocate)
through the C ABI against the existing entry point fr2d_step (fr2d_set_state_dev / fr2d_step /
fr2d_get_state_dev) on the same seeded inputs: bitwise equal, over two calls in a row, the second
with a smaller state (the scratch keeps the first call's contents), for every supported scheme and
both boundary modes; oversize states and RK4 are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import frhip
from frhip import _lib, ics

pytestmark = pytest.mark.gpu

SCRATCH_MAX = 6291456  # doubles per stage buffer


def _vortex(nx, ny, bc, seed):
    ps = frhip.FRPSpace2D(0.0, 10.0, nx, 0.0, 10.0, ny, 3, 1, 1)
    u = ics.perturbed(ics.isentropic_vortex(ps, 1.4, strength=5.0), seed=seed, amplitude=0.01)
    return ps, u, frhip.FREuler2D(ps, 1.4, "rusanov", bc)


def _reference(s, u, scheme, dt, nsteps):
    """The existing entry point on a device copy of u (reference layout in and out)."""
    import torch

    ud = torch.from_numpy(u.ravel(order="F").copy()).cuda()
    L = _lib.lib()
    torch.cuda.synchronize()
    _lib.check(L.fr2d_set_state_dev(s.h, C.c_void_p(ud.data_ptr())))
    _lib.check(L.fr2d_step(s.h, _lib.RK[scheme], float(dt), int(nsteps)))
    out = torch.empty_like(ud)
    _lib.check(L.fr2d_get_state_dev(s.h, C.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _noalloc(s, u, scheme, dt, nsteps):
    import torch

    ud = torch.from_numpy(u.ravel(order="F").copy()).cuda()
    torch.cuda.synchronize()
    rc = _lib.lib().fr2d_step_dev_noalloc(s.h, C.c_void_p(ud.data_ptr()), _lib.RK[scheme], float(dt), int(nsteps))
    _lib.check(rc)
    torch.cuda.synchronize()
    return ud.cpu().numpy()


@pytest.mark.parametrize("scheme", ["ssprk3", "midpoint", "euler"])
@pytest.mark.parametrize("bc", ["periodic", "ghost"])
def test_noalloc_equals_fr2d_step_twice(scheme, bc):
    for k, (nx, ny, nsteps) in enumerate(((96, 80, 7), (40, 24, 4))):  # the second call is smaller
        print(f"noalloc {scheme} {bc} {nx}x{ny} p=3, {nsteps} steps", flush=True)
        ps, u, s = _vortex(nx, ny, bc, seed=10 + k)
        dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
        ref = _reference(s, u, scheme, dt, nsteps)
        got = _noalloc(s, u, scheme, dt, nsteps)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
        s.close()


def test_noalloc_refuses_oversize_and_rk4():
    import torch

    L = _lib.lib()
    ps = frhip.FRPSpace2D(0.0, 10.0, 400, 0.0, 10.0, 400, 3, 1, 1)
    s = frhip.FREuler2D(ps, 1.4, "rusanov", "periodic")
    assert s.len > SCRATCH_MAX
    d = torch.zeros(8, dtype=torch.float64, device="cuda")  # never touched: refused before a launch
    assert L.fr2d_step_dev_noalloc(s.h, C.c_void_p(d.data_ptr()), 0, 1e-3, 1) == 5  # FR_E_UNSUPPORTED
    assert b"scratch" in L.fr_last_error()
    s.close()
    ps2, u2, s2 = _vortex(16, 8, "periodic", seed=3)
    ud = torch.from_numpy(u2.ravel(order="F").copy()).cuda()
    assert L.fr2d_step_dev_noalloc(s2.h, C.c_void_p(ud.data_ptr()), _lib.RK["rk4"], 1e-3, 1) == 5
    assert L.fr2d_step_dev_noalloc(s2.h, C.c_void_p(ud.data_ptr()), _lib.RK["tsit5"], 1e-3, 1) == 5
    assert np.array_equal(ud.cpu().numpy(), u2.ravel(order="F"))
    s2.close()
