"""GPU parity of the 2D Euler residual, stages and dissipation kernels (libfrhip, gfx950)
against the CPU oracle on the same seeded inputs.

Tolerances: residual ||du_gpu - du_cpu||_inf / ||du_cpu||_inf <= 1e-12; solution after the
stepping runs ||u_gpu - u_cpu||_2 / ||u_cpu||_2 <= 1e-10 (FP64 throughout).
"""
import numpy as np
import pytest

import frhip
from frhip import ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
STATE_TOL = 1e-10


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def setup(nx, ny, deg, flux, bc, L=10.0, seed=11, amp=0.01):
    ps = frhip.FRPSpace2D(0.0, L, nx, 0.0, L, ny, deg, 1, 1)
    u = ics.isentropic_vortex(ps, 1.4, strength=5.0)
    if amp:
        u = ics.perturbed(u, seed=seed, amplitude=amp)
    solver = frhip.FREuler2D(ps, 1.4, flux, bc)
    orc = O.Fr2d(nx, ny, deg, L / nx, L / ny, 1.4, flux, bc)
    return ps, u, solver, orc


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("flux", ["hll", "rusanov", "roe"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_residual_parity(deg, flux, bc):
    # 19 x 7 leaves partial tiles in both directions
    ps, u, s, orc = setup(19, 7, deg, flux, bc)
    du = s.rhs(u)
    ref = orc.rhs(u)
    assert rel_inf(du, ref) <= RES_TOL
    assert np.all(du[0] == 0) and np.all(du[-1] == 0) and np.all(du[:, 0] == 0) and np.all(du[:, -1] == 0)


@pytest.mark.parametrize("nx,ny", [(1, 1), (16, 4), (17, 5), (32, 12), (48, 3)])
def test_residual_parity_shapes(nx, ny):
    ps, u, s, orc = setup(nx, ny, 3, "hll", "ghost", seed=nx * 100 + ny)
    assert rel_inf(s.rhs(u), orc.rhs(u)) <= RES_TOL


def test_residual_and_stage_parity_random_shapes():
    """Seeded sweep over grid shapes (1-53 x 1-21 cells: full, partial and single tiles in both
    directions), degrees 1-5, the three fluxes and both boundary modes: the residual (drop-in,
    reference layout) and one SSP-RK3 step (padded stage arrays) against the oracle, 24 draws."""
    rng = np.random.default_rng(2027)
    for draw in range(24):
        nx, ny = int(rng.integers(1, 54)), int(rng.integers(1, 22))
        deg = int(rng.integers(1, 6))
        flux = ["hll", "rusanov", "roe"][draw % 3]
        bc = ["ghost", "periodic"][(draw // 3) % 2]
        ps, u, s, orc = setup(nx, ny, deg, flux, bc, seed=draw)
        case = (nx, ny, deg, flux, bc)
        assert rel_inf(s.rhs(u), orc.rhs(u)) <= RES_TOL, case
        dt = 0.02 * min(ps.dx0, 10.0 / ny) / 4.0
        s.set_state(u)
        s.step(dt, 1, "ssprk3")
        uc = u.copy(order="F")
        orc.ssprk3(uc, dt, 1)
        assert rel_l2(s.get_state(), uc) <= STATE_TOL, case
        s.close()


def test_general_geometry_parity_random_shapes():
    """The same kind of seeded sweep for the two general-geometry kernels: sheared parallelogram
    cells (one 2x2 metric, face frames) and polar meshes (a metric per solution point), shapes
    1-45 x 1-17, degrees 1-5, fluxes and boundary modes; residual against the oracle."""
    from test_oracle_metrics import G, oracle_for, polar_space, smooth_state

    rng = np.random.default_rng(2028)
    for draw in range(16):
        nx, ny = int(rng.integers(1, 46)), int(rng.integers(1, 18))
        deg = int(rng.integers(1, 6))
        flux = ["hll", "rusanov", "roe"][draw % 3]
        bc = ["ghost", "periodic"][(draw // 3) % 2]
        case = (nx, ny, deg, flux, bc)
        ps, u, s, _ = setup(nx, ny, deg, flux, bc, seed=100 + draw)
        dx, dy, iJ, n1, n2 = _sheared_geometry(nx, ny, shear=0.2 + 0.05 * draw)
        s.set_geometry(iJ, n1, n2)
        orc = O.Fr2d(nx, ny, deg, dx, dy, 1.4, flux, bc, iJ=iJ, n1=n1, n2=n2)
        assert rel_inf(s.rhs(u), orc.rhs(u)) <= RES_TOL, ("parallelogram",) + case
        s.close()
        pp = polar_space(nx, ny, deg)
        up = smooth_state(pp, seed=draw)
        sp = frhip.FREuler2D(pp, G, flux, bc)
        assert rel_inf(sp.rhs(up), oracle_for(pp, flux, bc).rhs(up)) <= RES_TOL, ("polar",) + case
        sp.close()


@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_wide_grid_strip_order(bc):
    """1100 x 6: 69 tiles across, so the workgroups take column strips of 16 tiles with a last
    strip of 5, and the tile rows are partial; residual and SSP-RK3 steps against the oracle."""
    ps, u, s, orc = setup(1100, 6, 3, "rusanov", bc, L=10.0, seed=5)
    assert rel_inf(s.rhs(u), orc.rhs(u)) <= RES_TOL
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    s.set_state(u)
    s.step(dt, 3)
    uc = u.copy(order="F")
    orc.ssprk3(uc, dt, 3)
    assert rel_l2(s.get_state(), uc) <= STATE_TOL


@pytest.mark.parametrize("nx,ny", [(19, 7), (30, 5), (46, 3)])
def test_padded_device_state(nx, ny):
    """The device-resident state keeps the reference layout with rows padded to a multiple of 16
    elements: read through fr2d_state_dev with the pitch of fr2d_state_pitch it equals
    get_state, and set_state / fr2d_get_state_dev convert device arrays both ways."""
    import ctypes as C

    import torch

    from frhip import _lib

    ps, u, s, orc = setup(nx, ny, 2, "hll", "ghost", seed=nx)
    dev = torch.from_numpy(u.ravel(order="F").copy()).cuda()
    s.set_state(dev)
    back = torch.empty_like(dev)
    _lib.check(_lib.lib().fr2d_get_state_dev(s.h, C.c_void_p(back.data_ptr())))
    torch.cuda.synchronize()
    assert torch.equal(back, dev)
    pitch, plane = s.state_pitch()
    assert pitch % 16 == 0 and nx + 2 <= pitch < nx + 2 + 16 and plane == pitch * (ny + 2)
    # the padded array itself, read on the device through the library's row contraction with
    # the 1 x 1 identity (a copy: x * 1.0 is exact) from the raw device pointer of the state
    from frhip import ops

    n = plane * 9 * 4
    padded = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    one = np.ones((1, 1))
    torch.cuda.synchronize()
    _lib.check(ops._bind().fr_poly_derivative_dev(C.c_void_p(s.state_ptr()), n, 1, 1, _lib.dptr(one),
                                                  C.c_void_p(padded.data_ptr()), None))
    torch.cuda.synchronize()
    a = padded.cpu().numpy().reshape((pitch, ny + 2, 3, 3, 4), order="F")[: nx + 2]
    assert np.array_equal(a, u)


def test_residual_garbage_du_overwritten():
    ps, u, s, orc = setup(20, 9, 3, "rusanov", "periodic")
    du = np.full(s.shape, np.nan, order="F")
    s.dudt(du, u, None, 0.0)
    assert np.all(np.isfinite(du))
    assert rel_inf(du, orc.rhs(u)) <= RES_TOL


def test_residual_device_entry_point():
    import torch

    ps, u, s, orc = setup(40, 24, 3, "rusanov", "periodic")
    s.use_torch_stream()
    ud = torch.from_numpy(u.ravel(order="F").copy()).cuda()
    dud = torch.full_like(ud, 7.0)
    s.dudt(dud, ud, None, 0.0)
    torch.cuda.synchronize()
    du = dud.cpu().numpy().reshape(s.shape, order="F")
    assert rel_inf(du, orc.rhs(u)) <= RES_TOL


def test_torch_tensors_ordered_without_stream_selection():
    """dudt, set_state and primitives on torch tensors, with the context on its own stream (no
    use_torch_stream): each call waits for what torch's current stream queued before it (a NaN
    fill of du, a copy into u) and torch's stream waits for the call (a copy out of du queued just
    after it), on the default stream and on a side stream."""
    import torch

    ps, u, s, orc = setup(128, 96, 3, "rusanov", "periodic")
    ref = orc.rhs(u)
    host = torch.from_numpy(u.ravel(order="F").copy())
    side = torch.cuda.Stream()
    for st in (torch.cuda.current_stream(), side):
        with torch.cuda.stream(st):
            ud = torch.empty(host.numel(), dtype=torch.float64, device="cuda")
            ud.copy_(host, non_blocking=False)
            ud.mul_(1.0)  # queued work that writes u right before the call
            dud = torch.empty_like(ud)
            dud.fill_(float("nan"))
            s.dudt(dud, ud)
            out = dud.clone()  # queued right after the call on torch's stream
        torch.cuda.synchronize()
        du = out.cpu().numpy().reshape(s.shape, order="F")
        assert np.all(np.isfinite(du))
        assert rel_inf(du, ref) <= RES_TOL
        with torch.cuda.stream(st):
            ud2 = torch.zeros_like(ud)
            ud2.add_(ud)  # u written on torch's stream just before set_state
            s.set_state(ud2)
        assert np.array_equal(s.get_state(), u)
        n = ps.nx * ps.ny * 16 * 4
        with torch.cuda.stream(st):
            prim = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
            s.primitives(out=prim)
            pc = prim.clone()
        torch.cuda.synchronize()
        assert np.all(np.isfinite(pc.cpu().numpy()))


def test_c2_full_size_residual():
    """256 x 256, p = 3, Rusanov, periodic: the benchmark configuration."""
    ps, u, s, orc = setup(256, 256, 3, "rusanov", "periodic", amp=0.0)
    du = s.rhs(u)
    ref = orc.rhs(u)
    assert rel_inf(du, ref) <= RES_TOL
    # discrete conservation on the periodic grid: sum_k,l wp du = 0 for every variable
    w = ps.wp
    tot = np.einsum("ijklm,kl->m", du[1:-1, 1:-1], w)
    scale = np.einsum("ijklm,kl->m", np.abs(du[1:-1, 1:-1]), w)
    assert np.all(np.abs(tot) <= 1e-12 * scale)


def test_staggered_start_tuned_and_results_unchanged():
    """The first stepping call tunes the staggered start of single-round grids (256 x 256, p = 3:
    1024 tiles, four per CU) to one of its candidates and leaves it off elsewhere; the steps still
    match the oracle (the delay changes the timing only)."""
    ps, u, s, orc = setup(256, 256, 3, "rusanov", "periodic", amp=0.0)
    assert s.stagger() == -1
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    s.set_state(u)
    s.step(dt, 3)
    assert s.stagger() in (0, 3, 4, 5, 6, 7)
    uc = u.copy(order="F")
    orc.ssprk3(uc, dt, 3)
    assert rel_l2(s.get_state(), uc) <= STATE_TOL
    ps2, u2, s2, _ = setup(24, 20, 3, "rusanov", "periodic", amp=0.0)
    s2.set_state(u2)
    s2.step(dt, 1)
    assert s2.stagger() == 0
    # a fixed delay replaces the tuning and gives the same bits as no delay, eager and replayed
    got = {}
    for d in (0, 5):
        fx = frhip.FREuler2D(ps, 1.4, "rusanov", "periodic")
        fx.set_stagger(d)
        fx.set_state(u)
        fx.step(dt, 3)
        assert fx.stagger() == d
        got[d] = fx.get_state()
        fx.set_state(u)
        fx.step_graph(dt, 3, 1)
        assert np.array_equal(fx.get_state(), got[d])
        fx.close()
    assert np.array_equal(got[0], got[5])
    with pytest.raises(frhip.FRError):
        s.set_stagger(17)


@pytest.mark.parametrize("scheme", ["ssprk3", "rk4", "midpoint", "euler", "tsit5"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_stepping_parity(scheme, bc):
    ps, u, s, orc = setup(24, 20, 3, "rusanov", bc, amp=0.0)
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    s.set_state(u)
    s.step(dt, 10, scheme)
    ug = s.get_state()
    uc = u.copy(order="F")
    if scheme == "ssprk3":
        orc.ssprk3(uc, dt, 10)
    elif scheme == "rk4":
        orc.rk4(uc, dt, 10)
    elif scheme == "midpoint":
        orc.midpoint(uc, dt, 10)
    elif scheme == "tsit5":
        orc.tsit5(uc, dt, 10)
    else:
        for _ in range(10):
            uc = orc.stage(uc, uc, 0.0, 1.0, dt)
        if bc == "periodic":
            orc.fill_periodic(uc)
    assert rel_l2(ug, uc) <= STATE_TOL


@pytest.mark.parametrize("scheme", ["ssprk3", "rk4", "midpoint", "euler", "tsit5"])
def test_zero_dt_step_keeps_state(scheme):
    """dt = 0 takes the general stage form (b = 0): every scheme returns the state (SSP-RK3 up to
    the rounding of 1/3 + 2/3, the others bitwise)."""
    ps, u, s, orc = setup(20, 8, 3, "hll", "periodic")
    s.set_state(u)
    s.step(0.0, 3, scheme)
    ug = s.get_state()
    if scheme == "ssprk3":
        assert np.allclose(ug[1:-1, 1:-1], u[1:-1, 1:-1], rtol=1e-15, atol=0.0)
    else:
        assert np.array_equal(ug[1:-1, 1:-1], u[1:-1, 1:-1])


def test_graph_replay_matches_eager():
    """The context keeps its captured step graph: every replay (with the host reading the state,
    eager steps on another context and a second replay count in between) equals the eager
    launches bitwise; a change of dt recaptures."""
    ps, u, s, orc = setup(32, 16, 3, "rusanov", "periodic", amp=0.0)
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    s.set_state(u)
    s2 = frhip.FREuler2D(ps, 1.4, "rusanov", "periodic")
    s2.set_state(u)
    for rep in (1, 1, 2, 1):
        s.step(dt, 2 * rep)
        eager = s.get_state()
        s2.step_graph(dt, 2, rep)
        assert np.array_equal(eager, s2.get_state())
    s.step(0.5 * dt, 2)
    s2.step_graph(0.5 * dt, 2, 1)
    assert np.array_equal(s.get_state(), s2.get_state())


@pytest.mark.parametrize("scheme", ["ssprk3", "midpoint", "euler"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_state_alternation(scheme, bc):
    """SSP-RK3, midpoint and Euler steps of one call alternate the state between two device
    arrays and the call ends with the state in its own array: any split of 7 steps into calls
    (single steps, 3 + 4, one call of 7, a 2-step graph replayed with an odd eager step after it)
    gives the same state bitwise, and the ghost layer stays that of the initial state."""
    ps, u, s, orc = setup(20, 12, 3, "hll", bc, amp=0.01, seed=9)
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    states = []
    for calls in ((1,) * 7, (3, 4), (7,)):
        s.set_state(u)
        for k in calls:
            s.step(dt, k, scheme)
        states.append(s.get_state())
    s.set_state(u)
    s.step_graph(dt, 2, 3, scheme)
    s.step(dt, 1, scheme)
    states.append(s.get_state())
    for st in states[1:]:
        assert np.array_equal(st, states[0])
    uc = u.copy(order="F")
    if scheme == "euler":
        for _ in range(7):
            uc = orc.stage(uc, uc, 0.0, 1.0, dt)
        if bc == "periodic":
            orc.fill_periodic(uc)
    else:
        getattr(orc, scheme)(uc, dt, 7)
    assert rel_l2(states[0], uc) <= STATE_TOL
    if bc == "ghost":
        assert np.array_equal(states[0][0], u[0]) and np.array_equal(states[0][:, -1], u[:, -1])


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_sensor_filter_parity(deg):
    ps = frhip.FRPSpace2D(0.0, 2.0, 40, 0.0, 1.0, 20, deg, 1, 1)
    u = ics.shock_vortex(ps, 1.4)
    u = ics.perturbed(u, seed=5, amplitude=0.02)
    s = frhip.FREuler2D(ps, 1.4, "hll", "ghost")
    orc = O.Fr2d(40, 20, deg, 2.0 / 40, 1.0 / 20, 1.4, "hll", "ghost")
    s.set_state(u)
    n_gpu = s.sensor_filter("modal", 5e-4, kappa=9.0)
    uc = u.copy(order="F")
    n_cpu, _ = orc.sensor_filter(uc, 5e-4, kappa=9.0)
    assert n_gpu == n_cpu and n_cpu > 0
    assert rel_inf(s.get_state(), uc) <= 1e-13


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_positivity_parity(deg):
    ps = frhip.FRPSpace2D(0.0, 1.0, 12, 0.0, 1.0, 10, deg, 1, 1)
    u = ics.isentropic_vortex(ps, 1.4, strength=5.0, center=(0.5, 0.5))
    rng = np.random.default_rng(3)
    n = deg + 1
    # push a few solution points below zero density / pressure
    for _ in range(6):
        i, j, k, l = rng.integers(1, 11), rng.integers(1, 9), rng.integers(0, n), rng.integers(0, n)
        u[i, j, k, l, 0] *= -0.2
        u[i, j, k, l, 3] *= 0.05
    s = frhip.FREuler2D(ps, 1.4, "hll", "ghost")
    orc = O.Fr2d(12, 10, deg, 1 / 12, 1 / 10, 1.4, "hll", "ghost")
    s.set_state(u)
    s.positivity()
    uc = u.copy(order="F")
    bad = orc.positivity(uc)
    assert bad == 0
    assert not np.array_equal(uc, u)  # the limiter corrected cells
    assert rel_inf(s.get_state(), uc) <= 1e-12
    s.check_state()


def test_limiter_failure_is_reported():
    """An element whose mean density is not positive fails the limiter's assert
    (src/dissipation.jl:143-145): the plain and the fused limiter report FR_E_LIMITER and leave
    that element as it was, like the oracle's count of failed elements."""
    ps = frhip.FRPSpace2D(0.0, 1.0, 12, 0.0, 1.0, 10, 3, 1, 1)
    u = ics.isentropic_vortex(ps, 1.4, strength=5.0, center=(0.5, 0.5))
    u[4, 5, :, :, 0] = -0.5
    orc = O.Fr2d(12, 10, 3, 1 / 12, 1 / 10, 1.4, "hll", "ghost")
    assert orc.positivity(u.copy(order="F")) == 1
    for fused in (False, True):
        s = frhip.FREuler2D(ps, 1.4, "hll", "ghost")
        s.set_state(u)
        with pytest.raises(frhip.FRError) as e:
            if fused:
                s.filter_refresh_limit("off", rule="outflow")
            else:
                s.positivity()
        assert e.value.code == 7
        assert np.array_equal(s.get_state()[4, 5], u[4, 5])
        s.close()


def test_check_state_flags_negative_and_nan():
    ps, u, s, orc = setup(16, 8, 2, "hll", "ghost", amp=0.0)
    bad = u.copy(order="F")
    bad[3, 3, 0, 0, 0] = -1.0
    s.set_state(bad)
    with pytest.raises(frhip.FRError) as e:
        s.check_state()
    assert e.value.code == 4
    bad[3, 3, 0, 0, 0] = np.nan
    s.set_state(bad)
    with pytest.raises(frhip.FRError) as e:
        s.check_state()
    assert e.value.code == 3


def _ghost_rule_numpy(u, rule):
    """The ghost refresh rules written out with numpy slices (host loops of the reference scripts)."""
    v = u.copy(order="F")
    nx, ny = v.shape[0] - 2, v.shape[1] - 2
    if rule == "periodic":  # example/euler2d_wave.jl:125-135
        v[0, 1:ny + 1] = v[nx, 1:ny + 1]
        v[nx + 1, 1:ny + 1] = v[1, 1:ny + 1]
        v[:, 0] = v[:, ny]
        v[:, ny + 1] = v[:, 1]
        v[0, 0], v[0, ny + 1], v[nx + 1, 0], v[nx + 1, ny + 1] = v[nx, ny], v[nx, 1], v[1, ny], v[1, 1]
    elif rule == "outflow":
        v[0, 1:ny + 1] = v[1, 1:ny + 1]
        v[nx + 1, 1:ny + 1] = v[nx, 1:ny + 1]
        v[:, 0] = v[:, 1]
        v[:, ny + 1] = v[:, ny]
        v[0, 0], v[0, ny + 1], v[nx + 1, 0], v[nx + 1, ny + 1] = v[1, 1], v[1, ny], v[nx, 1], v[nx, ny]
    elif rule == "wave":  # example/euler2d_wave.jl:127-132, in its order (variable 3 is index 2)
        v[0] = v[nx]
        v[nx + 1] = v[1]
        v[:, 0] = v[:, ny]
        v[:, 0, :, :, 2] *= -1
        v[:, ny + 1] = v[:, 1]
        v[:, ny + 1, :, :, 2] *= -1
    elif rule == "wave_y":  # example/euler2d_wave.jl:153-158, in its order (variable 2 is index 1)
        v[:, 0] = v[:, ny]
        v[:, ny + 1] = v[:, 1]
        v[0] = v[nx]
        v[0, :, :, :, 1] *= -1
        v[nx + 1] = v[1]
        v[nx + 1, :, :, :, 1] *= -1
    else:  # example/shock-vortex.jl:322-325, in its order
        v[:, 0] = v[:, 1]
        v[:, ny + 1] = v[:, ny]
        v[nx + 1, :] = v[nx, :]
    return v


@pytest.mark.parametrize("rule", ["periodic", "outflow", "inflow_left", "wave", "wave_y"])
def test_fill_ghost_rules(rule):
    nx, ny, deg = 13, 9, 3
    ps = frhip.FRPSpace2D(0.0, 1.0, nx, 0.0, 1.0, ny, deg, 1, 1)
    rng = np.random.default_rng(3)
    u = np.asfortranarray(rng.random((nx + 2, ny + 2, deg + 1, deg + 1, 4)) + 1.0)
    s = frhip.FREuler2D(ps, 1.4, "hll", "ghost")
    s.set_state(u)
    s.fill_ghost(rule)
    assert np.array_equal(s.get_state(), _ghost_rule_numpy(u, rule))


@pytest.mark.parametrize("deg", [1, 3, 4, 5])
def test_primitives_parity(deg):
    """Post-processing of example/shock-vortex.jl:333-352: coordinates and (rho, U, V, 1/lambda)
    of every solution point against the oracle's loop restatement (host copy and device output)."""
    import torch

    ps, u, s, orc = setup(12, 7, deg, "hll", "ghost")
    s.set_state(u)
    x, y, sol = s.primitives()
    xo, yo, so = O.solution_field2d(u, ps.xpg, 1.4)
    assert np.array_equal(x, xo) and np.array_equal(y, yo)
    for m in range(4):
        assert rel_inf(sol[..., m], so[..., m]) <= 1e-14
    s.use_torch_stream()
    d = torch.full((sol.size,), np.nan, dtype=torch.float64, device="cuda")
    s.primitives(out=d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().reshape(sol.shape, order="F"), sol)


@pytest.mark.parametrize("scheme", ["ssprk3", "rk4", "midpoint", "tsit5"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_degree5_stepping(scheme, bc):
    """Degree 5 (nsp = 6, one workgroup of six wavefronts per CU): steps against the oracle, graph
    replay bitwise equal to eager; the space of test/runtests.jl:16-20 (FRPSpace2D(0, 1, 20, 0, 1,
    20, 5, 1, 1)) builds a solver."""
    ps, u, s, orc = setup(21, 9, 5, "hll", bc, seed=21)
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    s.set_state(u)
    s.step(dt, 3, scheme)
    got = s.get_state()
    uc = u.copy(order="F")
    getattr(orc, scheme)(uc, dt, 3)
    assert rel_l2(got, uc) <= STATE_TOL
    s.set_state(u)
    s.step_graph(dt, 3, 1, scheme)
    assert np.array_equal(s.get_state(), got)
    rt = frhip.FRPSpace2D(0.0, 1.0, 20, 0.0, 1.0, 20, 5, 1, 1)
    frhip.FREuler2D(rt, 1.4, "hll", "ghost").close()


# ----------------------------------------------------------------- parallelogram cells
def _sheared_geometry(nx, ny, L=10.0, shear=0.4):
    dx, dy = L / nx, L / ny
    J = np.array([[dx / 2, shear * dx / 2], [0.0, dy / 2]])  # a = (dx, 0), b = (shear dx, dy)
    iJ = np.linalg.inv(J)
    return dx, dy, iJ, iJ[0] / np.linalg.norm(iJ[0]), iJ[1] / np.linalg.norm(iJ[1])


@pytest.mark.parametrize("deg", [1, 3, 4])
@pytest.mark.parametrize("flux", ["hll", "rusanov", "roe"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_parallelogram_residual_parity(deg, flux, bc):
    """Sheared cells (dev/parallelogram.jl:78-162 interface terms) on 19 x 7 (partial tiles)."""
    nx, ny = 19, 7
    ps, u, s, _ = setup(nx, ny, deg, flux, bc)
    dx, dy, iJ, n1, n2 = _sheared_geometry(nx, ny)
    s.set_geometry(iJ)  # normals default to grad r, grad s
    orc = O.Fr2d(nx, ny, deg, dx, dy, 1.4, flux, bc, iJ=iJ, n1=n1, n2=n2)
    assert rel_inf(s.rhs(u), orc.rhs(u)) <= RES_TOL


def test_parallelogram_rotated_grid_on_device():
    """The rectangular kernel's residual, turned, equals the parallelogram kernel's residual on the
    turned grid and state (rotation invariance of the Euler equations)."""
    nx, ny, deg = 20, 9, 3
    ps, u, s, _ = setup(nx, ny, deg, "rusanov", "periodic")
    du = s.rhs(u)
    c, sn = np.cos(0.7), np.sin(0.7)
    dx, dy = 10.0 / nx, 10.0 / ny
    s2 = frhip.FREuler2D(ps, 1.4, "rusanov", "periodic")
    s2.set_geometry([[2 * c / dx, 2 * sn / dx], [-2 * sn / dy, 2 * c / dy]], (c, sn), (-sn, c))

    def rot(a):
        v = a.copy(order="F")
        v[..., 1] = c * a[..., 1] - sn * a[..., 2]
        v[..., 2] = sn * a[..., 1] + c * a[..., 2]
        return v

    ref = rot(du)
    assert rel_inf(s2.rhs(rot(u)), ref) <= RES_TOL


def test_parallelogram_stepping_parity():
    nx, ny, deg = 24, 12, 3
    ps, u, s, _ = setup(nx, ny, deg, "hll", "periodic", amp=0.0)
    dx, dy, iJ, n1, n2 = _sheared_geometry(nx, ny)
    s.set_geometry(iJ, n1, n2)
    orc = O.Fr2d(nx, ny, deg, dx, dy, 1.4, "hll", "periodic", iJ=iJ, n1=n1, n2=n2)
    dt = 0.02 * dx
    s.set_state(u)
    s.step(dt, 10)
    uc = u.copy(order="F")
    orc.ssprk3(uc, dt, 10)
    assert rel_l2(s.get_state(), uc) <= STATE_TOL
    # the captured graph follows a geometry change
    s.set_state(u)
    s.step_graph(dt, 5, 2)
    s.set_geometry([[2 / dx, 0.0], [0.0, 2 / dy]])
    s.set_state(u)
    s.step_graph(dt, 5, 2)
    ur = u.copy(order="F")
    O.Fr2d(nx, ny, deg, dx, dy, 1.4, "hll", "periodic").ssprk3(ur, dt, 10)
    assert rel_l2(s.get_state(), ur) <= STATE_TOL


@pytest.mark.parametrize("scheme", ["ssprk3", "rk4", "midpoint", "tsit5"])
def test_parallelogram_stepping_ghost_layer(scheme):
    """Sheared cells with the ghost layer as data: the first stage of every step copies the ghost
    layer into the stage buffers (parallelogram kernels, buffer addressing), so every scheme
    matches the oracle, which holds the ghost layer fixed through the step."""
    nx, ny, deg = 21, 10, 3
    ps, u, s, _ = setup(nx, ny, deg, "hll", "ghost", amp=0.01, seed=4)
    dx, dy, iJ, n1, n2 = _sheared_geometry(nx, ny)
    s.set_geometry(iJ, n1, n2)
    orc = O.Fr2d(nx, ny, deg, dx, dy, 1.4, "hll", "ghost", iJ=iJ, n1=n1, n2=n2)
    dt = 0.02 * dx
    s.set_state(u)
    s.step(dt, 6, scheme)
    uc = u.copy(order="F")
    getattr(orc, scheme)(uc, dt, 6)
    ug = s.get_state()
    assert rel_l2(ug, uc) <= STATE_TOL
    assert np.array_equal(ug[0], u[0]) and np.array_equal(ug[:, -1], u[:, -1])  # ghosts held


def test_set_geometry_rejects_bad_input():
    ps, u, s, _ = setup(8, 4, 2, "hll", "periodic")
    with pytest.raises(frhip.FRError):
        s.set_geometry([[1.0, 2.0], [2.0, 1.0]])  # negative determinant
    with pytest.raises(frhip.FRError):
        s.set_geometry([[1.0, 0.0], [0.0, 1.0]], n1=(0.0, 0.0))


def test_flat_addressing_large_grid():
    """State arrays of 4 GiB and more take the kernel's flat 64-bit addressing instead of buffer
    resources.  A 32 x 32 periodic vortex tiled 91 x 91 times (2912^2 cells, 4.35 GB per state)
    must give the tiled residual of the small problem bitwise: the arithmetic per element is the
    same, only the addressing differs."""
    import types

    import torch

    n0, K, deg = 32, 91, 3
    ps, u, s, _ = setup(n0, n0, deg, "rusanov", "periodic")
    du_small = s.rhs(u)
    N = n0 * K
    big_ps = types.SimpleNamespace(nx=N, ny=N, ngx=1, ngy=1,
                                   **{k: getattr(ps, k) for k in ("deg", "ll", "lr", "dl", "dhl", "dhr", "V", "iV",
                                                                  "iJ_elem")})
    sb = frhip.FREuler2D(big_ps, 1.4, "rusanov", "periodic")
    assert sb.len * 8 >= 2 ** 32
    # the NaN fill and the launch share torch's default stream (HIP's null stream), so the
    # launch is ordered after the fill
    n = deg + 1
    small = torch.from_numpy(u.ravel(order="F").copy()).cuda().view(4, n, n, n0 + 2, n0 + 2)
    big = torch.zeros((4, n, n, N + 2, N + 2), dtype=torch.float64, device="cuda")
    big[:, :, :, 1:-1, 1:-1] = small[:, :, :, 1:-1, 1:-1].repeat(1, 1, 1, K, K)
    dbig = torch.full_like(big, float("nan"))
    sb.use_torch_stream()
    sb.dudt(dbig, big)
    torch.cuda.synchronize()
    dsm = torch.from_numpy(du_small.ravel(order="F").copy()).cuda().view(4, n, n, n0 + 2, n0 + 2)
    for m in range(4):
        assert torch.equal(dbig[m, :, :, 1:-1, 1:-1], dsm[m, :, :, 1:-1, 1:-1].repeat(1, 1, K, K))
        ring = torch.cat([dbig[m, :, :, 0, :].flatten(), dbig[m, :, :, -1, :].flatten(),
                          dbig[m, :, :, :, 0].flatten(), dbig[m, :, :, :, -1].flatten()])
        assert torch.all(ring == 0)
    sb.close()


def test_stream_selection_orders_against_torch():
    """fr2d_set_stream(NULL) is HIP's null stream (torch's default stream), a torch side stream
    is honoured, and FR_OWN_STREAM returns to the context's stream: in every case a fill queued
    just before the launch on the same stream is overwritten by the residual."""
    import torch

    ps, u, s, orc = setup(64, 64, 3, "rusanov", "periodic")
    ref = orc.rhs(u)
    ud = torch.from_numpy(u.ravel(order="F").copy()).cuda()
    side = torch.cuda.Stream()
    for mode in ("default", "side", "own"):
        dud = torch.empty_like(ud)
        if mode == "own":
            s.use_own_stream()
            dud.fill_(float("nan"))
            torch.cuda.synchronize()
            s.dudt(dud, ud)
        else:
            st = torch.cuda.current_stream() if mode == "default" else side
            s.use_torch_stream(st)
            with torch.cuda.stream(st):
                dud.fill_(float("nan"))
                s.dudt(dud, ud)
        torch.cuda.synchronize()
        du = dud.cpu().numpy().reshape(s.shape, order="F")
        assert np.all(np.isfinite(du)), mode
        assert rel_inf(du, ref) <= RES_TOL


def test_checkpoint_restart_is_bitwise(tmp_path):
    """Checkpoint / resume: 7 steps, the state written to a file in the reference layout
    (frhip.io.save_state), a fresh context loaded from it and 5 more steps give bitwise the state
    of 12 uninterrupted steps."""
    from frhip import io

    ps, u, s, _ = setup(24, 17, 3, "hll", "ghost", seed=13)
    dt, _ = ics.cfl_steps(u, 1.4, ps.dx0, 0.05, 1.0)
    s.set_state(u)
    s.step(dt, 12)
    straight = s.get_state()
    s.set_state(u)
    s.step(dt, 7)
    path = str(tmp_path / "ckpt.npz")
    io.save_state(path, s.get_state(), nx=24, ny=17, deg=3, gamma=1.4, t=7 * dt)
    s.close()
    v, meta = io.load_state(path)
    r = frhip.FREuler2D(frhip.FRPSpace2D(0.0, 10.0, meta["nx"], 0.0, 10.0, meta["ny"], meta["deg"], 1, 1),
                        meta["gamma"], "hll", "ghost")
    r.set_state(v)
    r.step(dt, 5)
    assert np.array_equal(r.get_state(), straight)
    r.close()


@pytest.mark.parametrize("direction", ["x", "y"])
def test_euler2d_wave_host_loop(direction):
    """example/euler2d_wave.jl as written: PSpace2D(0, 1, 20, 0, 1, 30, 1, 1), p = 2, HLL, a density
    wave (prim (rho, 1, 0, rho) along x, (rho, 0, 1, rho) along y), and per iteration the script's
    ghost rule of that direction (:127-132, y-momentum negated on the y ghost rows, for the wave
    along x; :153-158, x-momentum negated on the x ghost columns, for the wave along y) then one
    Midpoint step of dt = 0.002, 250 iterations to t = 0.5; against the oracle's loop of the same
    calls.  gamma 5/3 (the script takes KitBase's Gas default, which is not part of the reference
    tree).  The ghost layer is refreshed between steps only, as the script does, so inside a step
    the boundary data lags; the wave still arrives half a period on: the mean density error
    against 1 + 0.1 sin(2 pi (c - 0.5)) is below 1.5e-3 (the oracle's: 6.6e-4 along x, 6.5e-4
    along y)."""
    g, dt = 5.0 / 3.0, 0.002
    rule = "wave" if direction == "x" else "wave_y"
    ps = frhip.FRPSpace2D(0.0, 1.0, 20, 0.0, 1.0, 30, 2, 1, 1)
    c = ps.xpg[..., 0] if direction == "x" else ps.xpg[..., 1]
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * c)
    u0 = np.zeros(ps.xpg.shape[:4] + (4,), order="F")
    for idx in np.ndindex(*rho.shape):
        prim = [rho[idx], 1.0, 0.0, rho[idx]] if direction == "x" else [rho[idx], 0.0, 1.0, rho[idx]]
        u0[idx] = O.prim_conserve2(prim, g)
    s = frhip.FREuler2D(ps, g, "hll", "ghost")
    orc = O.Fr2d(20, 30, 2, 1.0 / 20, 1.0 / 30, g, "hll", "ghost")
    s.set_state(u0)
    ref = u0.copy(order="F")
    for _ in range(250):
        s.fill_ghost(rule)
        s.step(dt, 1, "midpoint")
        ref = _ghost_rule_numpy(ref, rule)
        orc.midpoint(ref, dt, 1)
    got = s.get_state()
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= STATE_TOL
    exact = 1.0 + 0.1 * np.sin(2 * np.pi * (c - 0.5))
    assert np.mean(np.abs(got[1:-1, 1:-1, :, :, 0] - exact[1:-1, 1:-1])) < 1.5e-3
    s.close()


def test_parallelogram_script_host_loop():
    """dev/parallelogram.jl as written: its sheared vertices (30 x 15 cells, dx = 1/30, dy = 1/30,
    each row shifted by 0.5 / ny), FRPSpace2D over them (the inverse Jacobian [60 -60; 0 60] in
    every cell), p = 1, HLL, x faces in the frame of n1 = (cos(-pi/4), sin(-pi/4)) and y faces in
    that of n2 = (0, 1) (:175-186), a density wave per cell index rho = 1 + 0.1 sin(2 pi i / nx)
    with prim (rho, 1, 0, rho), and per iteration the script's periodic ghost copy then one
    forward-Euler step of dt = 0.001, nt / 2 = 500 iterations; against the oracle's loop of the same
    calls (rel L2 <= 1e-10).  gamma 5/3 (KitBase's Gas default, outside the reference tree)."""
    nx, ny, g, dt = 30, 15, 5.0 / 3.0, 0.001
    Dx, Dy = 1.0 / nx, 0.5 / ny
    V = np.zeros((nx + 2, ny + 2, 4, 2), order="F")
    for j in range(ny + 2):
        for i in range(nx + 2):  # the script's OffsetArray indices 0..nx+1
            V[i, j, 0, 0] = 0.5 * (j - 1) / ny + (i - 1) * Dx
            V[i, j, 1, 0] = V[i, j, 0, 0] + Dx
            V[i, j, 2, 0] = V[i, j, 1, 0] + Dy
            V[i, j, 3, 0] = V[i, j, 2, 0] - Dx
            V[i, j, 0, 1] = 0.5 * (j - 1) * Dy
            V[i, j, 1, 1] = V[i, j, 0, 1]
            V[i, j, 2, 1] = V[i, j, 0, 1] + Dy
            V[i, j, 3, 1] = V[i, j, 2, 1]
    _, iJp = frhip.rs_jacobi_mesh(V, frhip.legendre_point(1))
    iJ = iJp[1, 1, 0, 0]
    assert np.allclose(iJ, [[60.0, -60.0], [0.0, 60.0]], rtol=1e-12, atol=1e-10)
    u0 = np.zeros((nx + 2, ny + 2, 2, 2, 4), order="F")
    for i in range(nx + 2):
        rho = 1.0 + 0.1 * np.sin(2 * np.pi * i / nx)
        u0[i] = O.prim_conserve2([rho, 1.0, 0.0, rho], g)
    n1, n2 = (np.cos(-np.pi / 4), np.sin(-np.pi / 4)), (0.0, 1.0)
    ps = frhip.FRPSpace2D(0.0, 1.0, nx, 0.0, 0.5, ny, 1, 1, 1)
    s = frhip.FREuler2D(ps, g, "hll", "ghost")
    s.set_geometry(iJ, n1, n2)
    orc = O.Fr2d(nx, ny, 1, Dx, Dy, g, "hll", "ghost", iJ=iJ, n1=n1, n2=n2)
    s.set_state(u0)
    ref = u0.copy(order="F")
    for _ in range(500):
        s.fill_ghost("periodic")
        s.step(dt, 1, "euler")
        ref = orc.stage(_ghost_rule_numpy(ref, "periodic"), _ghost_rule_numpy(ref, "periodic"), 0.0, 1.0, dt)
    got = s.get_state()
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= STATE_TOL
    s.close()


def test_shock_vortex_script_host_loop():
    """example/shock-vortex.jl's run as written: FRPSpace2D(0, 2, 100, 0, 1, 50, 2, 1, 1), gamma 5/3,
    Ma 1.12, the script's initial condition (frhip.ics.shock_vortex_script), HLL, and 100 iterations
    of its loop (:296-326): positivity limiter with weights wp / 4, one Midpoint step of dt =
    0.001, the modal sensor (S0 = -3 log10(deg), kappa 9) + l2 filter of strength 5e-4 on every
    cell, the ghost copy of bottom / top / right.  The device loop call by call against the
    oracle's (rel L2 <= 1e-10, the same filtered-cell count, no limiter failure), and the native
    loop (fr2d_run_host_loop) bitwise equal to the call-by-call one."""
    g, dt, nit = 5.0 / 3.0, 0.001, 100
    ps = frhip.FRPSpace2D(0.0, 2.0, 100, 0.0, 1.0, 50, 2, 1, 1)
    u0 = ics.shock_vortex_script(ps, g)
    S0 = -3.0 * np.log10(2)
    s = frhip.FREuler2D(ps, g, "hll", "ghost")
    s.set_state(u0)
    s.positivity()
    nf = 0
    for n in range(nit):
        s.step(dt, 1, "midpoint")
        nf += s.sensor_filter("modal", 5e-4, S0, 9.0)
        s.fill_ghost("inflow_left")
        if n + 1 < nit:
            s.positivity()
    got = s.get_state()
    orc = O.Fr2d(100, 50, 2, 0.02, 0.02, g, "hll", "ghost")
    ref = u0.copy(order="F")
    bad = orc.positivity(ref)
    nref = 0
    for n in range(nit):
        orc.midpoint(ref, dt, 1)
        nref += orc.sensor_filter(ref, 5e-4, S0, 9.0)[0]
        ref[...] = _ghost_rule_numpy(ref, "inflow_left")
        if n + 1 < nit:
            bad += orc.positivity(ref)
    assert bad == 0 and nf == nref and nf > 0
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= STATE_TOL
    s.set_state(u0)
    s.run_host_loop(dt, nit, "modal", 5e-4, S0, 9.0, "inflow_left", scheme="midpoint")
    assert np.array_equal(s.get_state(), got)
    s.close()
