"""GPU parity of the point-metric kernels (fr2d_set_metrics; an FRPSpace2D over a curvilinear
mesh, src/struct.jl:130-142, with the residual of example/shock-vortex.jl:26-118 reading
iJ[i, j][k, l] per point) against the CPU oracle on the same seeded inputs.

Tolerances: residual ||du_gpu - du_cpu||_inf / ||du_cpu||_inf <= 1e-12; states after stepping
||u_gpu - u_cpu||_2 / ||u_cpu||_2 <= 1e-10 (FP64)."""
import numpy as np
import pytest

import frhip
from oracle import oracle as O
from test_oracle_metrics import G, oracle_for, polar_space, smooth_state

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
STATE_TOL = 1e-10


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("flux", ["hll", "rusanov", "roe"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_polar_residual_parity(deg, flux, bc):
    # 19 x 7 leaves partial tiles in both directions
    ps = polar_space(19, 7, deg)
    u = smooth_state(ps, seed=deg)
    s = frhip.FREuler2D(ps, G, flux, bc)
    du = s.rhs(u)
    ref = oracle_for(ps, flux, bc).rhs(u)
    assert rel_inf(du, ref) <= RES_TOL
    assert np.all(du[0] == 0) and np.all(du[-1] == 0) and np.all(du[:, 0] == 0) and np.all(du[:, -1] == 0)
    s.close()


@pytest.mark.parametrize("nr,nth", [(30, 40), (14, 33), (64, 5)])
def test_polar_residual_shapes_and_layouts(nr, nth):
    """nr + 2 = 32 is its own padded pitch (one metric copy serves both layouts); the others pad."""
    ps = polar_space(nr, nth, 3)
    u = smooth_state(ps, seed=nr)
    s = frhip.FREuler2D(ps, G, "hll", "ghost")
    orc = oracle_for(ps, "hll", "ghost")
    assert rel_inf(s.rhs(u), orc.rhs(u)) <= RES_TOL
    # the stage path (padded arrays, second metric copy): one explicit Euler stage of the state
    dt = 1e-4
    s.set_state(u)
    s.step(dt, 1, "euler")
    ref = orc.stage(u, u, 0.0, 1.0, dt)
    assert rel_l2(s.get_state(), ref) <= STATE_TOL
    s.close()


def test_uniform_point_metric_matches_the_rectangle_kernel():
    nx, ny, deg = 33, 9, 3
    ps = frhip.FRPSpace2D(0.0, 10.0, nx, 0.0, 10.0, ny, deg, 1, 1)
    u = smooth_state(ps, seed=7)
    rect = frhip.FREuler2D(ps, G, "rusanov", "periodic")
    pt = frhip.FREuler2D(ps, G, "rusanov", "periodic")
    iJp = np.zeros((nx + 2, ny + 2, deg + 1, deg + 1, 2, 2), order="F")
    iJp[...] = ps.iJ_elem
    pt.set_metrics(iJp)
    a, b = rect.rhs(u), pt.rhs(u)
    assert rel_inf(b, a) <= 1e-14
    rect.close()
    pt.close()


@pytest.mark.parametrize("scheme,nsteps", [("ssprk3", 6), ("rk4", 3), ("midpoint", 4), ("tsit5", 2)])
def test_polar_stepping_parity(scheme, nsteps):
    ps = polar_space(24, 20, 2)
    u = smooth_state(ps, seed=5)
    s = frhip.FREuler2D(ps, G, "hll", "ghost")
    orc = oracle_for(ps, "hll", "ghost")
    dt = 2e-4
    s.set_state(u)
    s.step(dt, nsteps, scheme)
    ref = u.copy(order="F")
    getattr(orc, scheme)(ref, dt, nsteps)
    got = s.get_state()
    assert rel_l2(got, ref) <= STATE_TOL
    # graph replay of the same steps from the same state is bitwise the eager result
    s.set_state(u)
    s.step_graph(dt, nsteps, 1, scheme)
    assert np.array_equal(s.get_state(), got)
    s.close()


def test_polar_dropin_residual_on_device_arrays():
    import torch

    ps = polar_space(40, 24, 3)
    u = smooth_state(ps, seed=8)
    s = frhip.FREuler2D(ps, G, "rusanov", "ghost")
    ud = torch.from_numpy(u.ravel(order="F").copy()).cuda()
    dud = torch.full_like(ud, float("nan"))
    s.dudt(dud, ud)
    torch.cuda.synchronize()
    got = dud.cpu().numpy().reshape(u.shape, order="F")
    assert rel_inf(got, oracle_for(ps, "rusanov", "ghost").rhs(u)) <= RES_TOL
    s.close()


def test_set_metrics_rejects_bad_input():
    ps = polar_space(8, 6, 2)
    s = frhip.FREuler2D(ps, G, "hll", "ghost")
    bad = np.array(ps.iJ, order="F")
    bad[3, 2, 1, 1, 0, 0] = np.nan
    with pytest.raises(frhip.FRError):
        s.set_metrics(bad)
    flip = np.array(ps.iJ, order="F")
    flip[4, 4, 0, 2] = flip[4, 4, 0, 2][::-1]  # rows swapped: negative determinant
    with pytest.raises(frhip.FRError):
        s.set_metrics(flip)
    with pytest.raises(ValueError):
        s.set_metrics(np.asarray(ps.iJ)[:, :, :2])
    # a ghost cell's entries are not read: NaN there is accepted and changes nothing
    ghost = np.array(ps.iJ, order="F")
    ghost[0, :] = np.nan
    u = smooth_state(ps, seed=1)
    ref = s.rhs(u)
    s.set_metrics(ghost)
    assert np.array_equal(s.rhs(u), ref)
    s.close()
