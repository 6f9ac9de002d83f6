"""GPU parity of the 1D discrete-velocity BGK model (fr_bgk1d_*, example/bgk_wave.jl:69-129 as the
script writes it) against the CPU oracle on the same inputs.

Tolerances: residual ||du_gpu - du_cpu||_inf / ||du_cpu||_inf <= 1e-12; states after stepping
||u_gpu - u_cpu||_2 / ||u_cpu||_2 <= 1e-10 (FP64)."""
import numpy as np
import pytest

import frhip
from frhip import bgk
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
STATE_TOL = 1e-10


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def script_case(nx=20, deg=2, nu=100):
    ps = frhip.FRPSpace1D(0.0, 1.0, nx, deg)
    vs = bgk.VSpace1D(-5.0, 5.0, nu)
    return ps, vs, bgk.bgk_wave_initial(ps, vs)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5, 8])
def test_residual_parity(deg):
    ps, vs, f0 = script_case(nx=37, deg=deg, nu=64)
    f = frhip.ics.perturbed(f0, seed=deg, amplitude=0.05)
    g = bgk.BGK1D(ps, vs)
    orc = O.Bgk1d(37, deg, vs.u, vs.weights, 1.0 / 37)
    assert rel_inf(g.rhs(f), orc.rhs(f)) <= RES_TOL
    g.close()


def test_bgk_wave_to_t_end():
    """example/bgk_wave.jl as written: 20 cells, degree 2, 100 velocities on [-5, 5], tau = 1e-2,
    Midpoint with dt = cfl dx / u1 = 1e-3 to t = 1 (1000 steps), against the oracle's run."""
    ps, vs, f0 = script_case()
    dt = 0.1 * (1.0 / 20) / 5.0
    g = bgk.BGK1D(ps, vs)
    g.set_state(f0)
    g.step(dt, 1000, "midpoint")
    got = g.get_state()
    orc = O.Bgk1d(20, 2, vs.u, vs.weights, 1.0 / 20)
    ref = orc.step(f0, dt, 1000, "midpoint")
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= STATE_TOL
    # the script's post-processing: density of every point against its own t = 0 profile
    w = g.moments()[0]
    wo = orc.moments(ref)
    assert rel_inf(w[0], wo[..., 0]) <= 1e-12
    g.close()


@pytest.mark.parametrize("scheme", ["ssprk3", "euler", "midpoint"])
def test_schemes(scheme):
    ps, vs, f0 = script_case(nx=16, deg=3, nu=48)
    dt = 2e-4
    g = bgk.BGK1D(ps, vs)
    g.set_state(f0)
    g.step(dt, 7, scheme)
    ref = O.Bgk1d(16, 3, vs.u, vs.weights, 1.0 / 16).step(f0, dt, 7, scheme)
    assert rel_l2(g.get_state(), ref) <= STATE_TOL
    g.close()


def test_tau_ensemble_equals_single_problems():
    ps, vs, f0 = script_case(nx=24, deg=2, nu=40)
    taus = [1e-2, 3e-3, 1e-1]
    ens = bgk.BGK1D(ps, vs, tau=taus)
    ens.set_state(np.concatenate([f0.ravel(order="F")] * 3))
    ens.step(5e-4, 6, "ssprk3")
    got = ens.get_state()
    for m, t in enumerate(taus):
        one = bgk.BGK1D(ps, vs, tau=t)
        one.set_state(f0)
        one.step(5e-4, 6, "ssprk3")
        assert np.array_equal(got[..., m], one.get_state())
        one.close()
    ens.close()


def test_moments_launch_shapes_agree_bitwise():
    """The moments kernel sums the velocities in eight interleaved slices added in slice order, with
    one wavefront per slice on states below 262,144 points and one thread per point above: a
    43-member ensemble of 6,144-point members (264,192 points, one thread per point) gives every
    member the bits of the single problem (one wavefront per slice), residual and one SSP-RK3 step."""
    ps, vs, f0 = script_case(nx=2048, deg=2, nu=40)
    f = frhip.ics.perturbed(f0, seed=11, amplitude=0.05)
    nm = 43
    assert ps.nx * 3 * nm >= 262144 > ps.nx * 3
    ens = bgk.BGK1D(ps, vs, tau=[1e-2] * nm)
    one = bgk.BGK1D(ps, vs)
    flat = np.tile(f.ravel(order="F"), nm)
    de = ens.rhs(flat)
    d1 = one.rhs(f)
    for m in (0, 17, nm - 1):
        assert np.array_equal(de[..., m], d1)
    ens.set_state(flat)
    one.set_state(f)
    ens.step(2e-5, 1, "ssprk3")
    one.step(2e-5, 1, "ssprk3")
    assert np.array_equal(ens.get_state()[..., nm - 1], one.get_state())
    ref = O.Bgk1d(2048, 2, vs.u, vs.weights, 1.0 / 2048).rhs(f)
    assert rel_inf(d1, ref) <= RES_TOL
    ens.close()
    one.close()


def test_residual_parity_random_shapes():
    """Seeded sweep: 1-300 cells, degrees 1-8, 2-70 velocities (fewer than the moments kernel's
    eight slices too), one to five members with their own tau; residual against the oracle."""
    rng = np.random.default_rng(2030)
    for draw in range(16):
        nx, deg, nu = int(rng.integers(1, 301)), int(rng.integers(1, 9)), int(rng.integers(2, 71))
        nm = int(rng.integers(1, 6))
        ps = frhip.FRPSpace1D(0.0, 1.0, nx, deg)
        vs = bgk.VSpace1D(-5.0, 5.0, nu)
        f0 = frhip.ics.perturbed(bgk.bgk_wave_initial(ps, vs), seed=draw, amplitude=0.05)
        taus = list(10.0 ** rng.uniform(-3, -1, size=nm))
        g = bgk.BGK1D(ps, vs, tau=taus if nm > 1 else taus[0])
        flat = np.concatenate([frhip.ics.perturbed(f0, seed=100 * draw + m, amplitude=0.02).ravel(order="F")
                               for m in range(nm)])
        du = g.rhs(flat)
        du = du.reshape(f0.shape + ((nm,) if nm > 1 else ()), order="F")
        n1 = f0.size
        for m in range(nm):
            fm = flat[m * n1:(m + 1) * n1].reshape(f0.shape, order="F")
            ref = O.Bgk1d(nx, deg, vs.u, vs.weights, 1.0 / nx, tau=taus[m]).rhs(fm)
            got = du[..., m] if nm > 1 else du
            assert rel_inf(got, ref) <= RES_TOL, (nx, deg, nu, nm, m)
        g.close()


@pytest.mark.parametrize("scheme,chunk", [("midpoint", 4), ("ssprk3", 5), ("euler", 3), ("euler", 4)])
def test_step_graph_equals_eager_steps(scheme, chunk):
    """Graph replays give the bits of eager steps: 13 steps as replays of a captured graph of
    `chunk` steps plus eager remainder steps (Euler with an odd chunk ends a replay in the second
    buffer and copies back inside the graph); the kept graph is replayed again after eager work
    and checked against the eager run, and a new dt is captured afresh."""
    ps, vs, f0 = script_case(nx=12, deg=2, nu=40)
    f = frhip.ics.perturbed(f0, seed=5, amplitude=0.05)
    dt = 1e-3
    a, b = bgk.BGK1D(ps, vs), bgk.BGK1D(ps, vs)
    a.set_state(f)
    b.set_state(f)
    a.step(dt, 13, scheme)
    b.step_graph(dt, 13, scheme, chunk=chunk)
    assert np.array_equal(a.get_state(), b.get_state())
    for rep in range(3):  # the kept graph, with eager steps of both contexts in between
        a.step(dt, 1, scheme)
        b.step(dt, 1, scheme)
        a.step(dt, 2 * chunk, scheme)
        b.step_graph(dt, 2 * chunk, scheme, chunk=chunk)
        assert np.array_equal(a.get_state(), b.get_state()), rep
    a.step(0.5 * dt, 7, scheme)
    b.step_graph(0.5 * dt, 7, scheme, chunk=chunk)
    assert np.array_equal(a.get_state(), b.get_state())
    ref = O.Bgk1d(12, 2, vs.u, vs.weights, 1.0 / 12).step(f, dt, 13, scheme)
    b.set_state(f)
    b.step_graph(dt, 13, scheme, chunk=chunk)
    assert rel_l2(b.get_state(), ref) <= STATE_TOL
    a.close()
    b.close()


def test_bgk_wave_to_t_end_graph_replayed():
    """The script's run (20 cells, 100 velocities, Midpoint, 1000 steps) as ten replays of a
    100-step graph: the bits of the eager run."""
    ps, vs, f0 = script_case()
    dt = 0.1 * (1.0 / 20) / 5.0
    a, b = bgk.BGK1D(ps, vs), bgk.BGK1D(ps, vs)
    a.set_state(f0)
    b.set_state(f0)
    a.step(dt, 1000, "midpoint")
    b.step_graph(dt, 1000, "midpoint", chunk=100)
    assert np.array_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


RESIDENT_CASES = [(20, 2, 100, 1), (7, 5, 33, 1), (2, 1, 5, 1), (5, 8, 20, 1), (13, 3, 17, 3)]


@pytest.mark.parametrize("scheme", ["ssprk3", "midpoint", "euler"])
@pytest.mark.parametrize("nx,deg,nu,nm", RESIDENT_CASES)
def test_resident_path_equals_streaming_path(nx, deg, nu, nm, scheme):
    """The LDS-resident path (a workgroup per member, up to 256 steps per launch) gives the bits of
    the streaming path over 300 steps (two resident launches), for one member and a 3-member tau
    ensemble, degrees 1-8, fewer velocities than the eight velocity slices too."""
    ps = frhip.FRPSpace1D(0.0, 1.0, nx, deg)
    vs = bgk.VSpace1D(-5.0, 5.0, nu)
    f0 = frhip.ics.perturbed(bgk.bgk_wave_initial(ps, vs), seed=nx + deg, amplitude=0.05)
    taus = [1e-2, 3e-2, 5e-3][:nm]
    flat = np.concatenate([frhip.ics.perturbed(f0, seed=m, amplitude=0.02).ravel(order="F") for m in range(nm)])
    res, st = bgk.BGK1D(ps, vs, tau=taus if nm > 1 else taus[0]), bgk.BGK1D(ps, vs, tau=taus if nm > 1 else taus[0])
    res.set_path("resident")
    st.set_path("stream")
    assert res.path(scheme) == "resident" and st.path(scheme) == "stream"
    dt = 0.1 * (1.0 / nx) / (5.0 * (deg + 1) ** 2)
    for g in (res, st):
        g.set_state(flat)
        g.step(dt, 300, scheme)
    a, b = res.get_state(), st.get_state()
    assert np.all(np.isfinite(b))
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    res.close()
    st.close()


def test_resident_path_ensemble_of_script_members():
    """64 members of the script's case with their own tau (one resident workgroup each) over 40
    Midpoint steps: the bits of the streaming path, every member finite."""
    ps, vs, f0 = script_case()
    taus = list(np.exp(np.random.default_rng(8).uniform(np.log(3e-3), np.log(1e-1), 64)))
    flat = np.concatenate([frhip.ics.perturbed(f0, seed=m, amplitude=0.02).ravel(order="F") for m in range(64)])
    got = []
    for path in ("stream", "resident"):
        g = bgk.BGK1D(ps, vs, tau=taus)
        g.set_path(path)
        g.set_state(flat)
        g.step(1e-3, 40, "midpoint")
        got.append(g.get_state())
        g.close()
    assert np.all(np.isfinite(got[0]))
    assert np.array_equal(got[0], got[1])


def test_resident_path_selection():
    """auto takes the resident path when a member's arrays fit the workgroup's LDS (the script's
    case under every scheme) and the streaming path otherwise; forcing the resident path on a member
    that does not fit is refused at step time."""
    ps, vs, f0 = script_case()
    g = bgk.BGK1D(ps, vs)
    assert [g.path(s) for s in ("ssprk3", "midpoint", "euler")] == ["resident"] * 3
    g.close()
    big_ps = frhip.FRPSpace1D(0.0, 1.0, 2048, 2)
    big = bgk.BGK1D(big_ps, bgk.VSpace1D(-5.0, 5.0, 40))
    assert big.path("midpoint") == "stream"
    big.set_path("resident")
    big.set_state(bgk.bgk_wave_initial(big_ps, bgk.VSpace1D(-5.0, 5.0, 40)))
    with pytest.raises(frhip.FRError):
        big.step(1e-5, 1, "midpoint")
    big.set_path("auto")
    big.step(1e-5, 1, "midpoint")
    big.close()


def test_mol_on_device_arrays_side_stream():
    import torch

    ps, vs, f0 = script_case(nx=33, deg=2, nu=50)
    f = frhip.ics.perturbed(f0, seed=3, amplitude=0.05)
    g = bgk.BGK1D(ps, vs)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        fd = torch.from_numpy(f.ravel(order="F").copy()).cuda()
        dd = torch.full_like(fd, float("nan"))
        g.mol(dd, fd)
        out = dd.clone()
    torch.cuda.synchronize()
    ref = O.Bgk1d(33, 2, vs.u, vs.weights, 1.0 / 33).rhs(f)
    assert rel_inf(out.cpu().numpy().reshape(f.shape, order="F"), ref) <= RES_TOL
    # the context's own stream is ordered after that call: a step right after sees its state
    g.set_state(f)
    g.step(1e-4, 1, "euler")
    assert rel_l2(g.get_state(), f + 1e-4 * ref) <= STATE_TOL
    g.close()


def test_create_rejects_bad_input():
    ps, vs, _ = script_case()
    with pytest.raises(frhip.FRError):
        bgk.BGK1D(ps, vs, tau=0.0)
    with pytest.raises(frhip.FRError):
        bgk.BGK1D(ps, vs, gamma=1.0)
    g = bgk.BGK1D(ps, vs)
    with pytest.raises(frhip.FRError):
        g.step(1e-3, 1, "rk4")
    with pytest.raises(frhip.FRError):
        g.step_graph(1e-3, 4, "rk4")
    with pytest.raises(ValueError):
        g.set_state(np.zeros(7))
    g.close()
