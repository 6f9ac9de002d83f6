"""Pins the initial states of the workloads (frhip.ics) by what the gas dynamics fixes:
- the Rankine-Hugoniot states (in place of KitBase's ib_rh, shock-vortex.jl:162,198) carry the
  same mass, momentum and energy flux across the shock, reduce to no jump at Mach 1, give the
  textbook downstream Mach number and raise the entropy;
- prim_conserve / conserve_prim (KitBase's lambda = rho / (2 p) convention) invert each other
  and give p = 1, E = 1 / (gamma - 1) for rho = 1 at rest with lambda = 1/2;
- the isentropic vortex is isentropic, has the closed-form temperature dip at its centre and
  the free stream far from it;
- the Lamb-Oseen vortex ahead of the normal shock is isentropic, in radial balance
  dp/dr = rho v_theta^2 / r, and peaks at the requested swirl Mach number.
A planted mistake in the Rankine-Hugoniot pressure ratio is rejected."""
import numpy as np
import pytest

from frhip import ics


class Points:
    """Stand-in for a solution space: only the point coordinates xpg[..., 0:2] are read."""

    def __init__(self, x, y):
        self.xpg = np.stack([x, y], axis=-1)


def rh_fluxes(state, gamma):
    r, u, p = state
    E = p / (gamma - 1.0) + 0.5 * r * u * u
    return np.array([r * u, r * u * u + p, (E + p) * u])


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0, 1.2])
@pytest.mark.parametrize("mach", [1.0, 1.1, 1.12, 2.0, 5.0])
def test_rankine_hugoniot_conserves_fluxes(mach, gamma):
    s1, s2 = ics.rankine_hugoniot(mach, gamma)
    f1, f2 = rh_fluxes(s1, gamma), rh_fluxes(s2, gamma)
    assert np.allclose(f1, f2, rtol=1e-14, atol=0)
    r1, u1, p1 = s1
    r2, u2, p2 = s2
    assert (r1, p1) == (1.0, 1.0) and np.isclose(u1, mach * np.sqrt(gamma), rtol=1e-15)
    m2 = u2 / np.sqrt(gamma * p2 / r2)
    m2_ref = np.sqrt(((gamma - 1.0) * mach ** 2 + 2.0) / (2.0 * gamma * mach ** 2 - (gamma - 1.0)))
    assert np.isclose(m2, m2_ref, rtol=1e-14)
    if mach == 1.0:
        assert np.allclose(s2, s1, rtol=1e-15)
    else:
        assert m2 < 1.0 and r2 > 1.0 and p2 > 1.0
        assert p2 / r2 ** gamma > p1 / r1 ** gamma  # entropy rises across the shock


def test_rankine_hugoniot_rejects_a_planted_pressure_mistake():
    gamma, mach = 1.4, 2.0
    (r1, u1, p1), (r2, u2, _) = ics.rankine_hugoniot(mach, gamma)
    m2 = mach * mach
    wrong = 1.0 + 2.0 * gamma / (gamma - 1.0) * (m2 - 1.0)  # (gamma - 1) in place of (gamma + 1)
    f1, f2 = rh_fluxes((r1, u1, p1), gamma), rh_fluxes((r2, u2, wrong), gamma)
    assert not np.allclose(f1, f2, rtol=1e-6)


def test_prim_conserve_round_trip_and_known_values():
    gamma = 1.4
    rng = np.random.default_rng(3)
    prim = np.stack([rng.uniform(0.5, 2.0, 50), rng.normal(size=50), rng.normal(size=50),
                     rng.uniform(0.2, 3.0, 50)], axis=-1)
    back = ics.conserve_prim(ics.prim_conserve(prim, gamma), gamma)
    assert np.allclose(back, prim, rtol=1e-13, atol=1e-15)
    w = ics.prim_conserve(np.array([1.0, 0.0, 0.0, 0.5]), gamma)
    assert np.allclose(w, [1.0, 0.0, 0.0, 1.0 / (gamma - 1.0)], rtol=1e-15)
    w1 = ics.prim_conserve1(np.array([2.0, 3.0, 0.25]), gamma)  # p = rho / (2 lambda) = 4
    assert np.allclose(w1, [2.0, 6.0, 4.0 / (gamma - 1.0) + 9.0], rtol=1e-15)


def pressure(w, gamma):
    return (gamma - 1.0) * (w[..., 3] - 0.5 * (w[..., 1] ** 2 + w[..., 2] ** 2) / w[..., 0])


def test_isentropic_vortex_is_isentropic_with_the_closed_form_dip():
    gamma, eps = 1.4, 5.0
    x, y = np.meshgrid(np.linspace(0.0, 10.0, 41), np.linspace(0.0, 10.0, 41), indexing="ij")
    pts = Points(x, y)
    pts.x0, pts.x1, pts.y0, pts.y1 = 0.0, 10.0, 0.0, 10.0
    w = ics.isentropic_vortex(pts, gamma, strength=eps)
    rho, p = w[..., 0], pressure(w, gamma)
    assert np.allclose(p / rho ** gamma, 1.0, rtol=1e-13)
    T0 = 1.0 - (gamma - 1.0) * eps ** 2 / (8.0 * gamma * np.pi ** 2) * np.e
    c = (20, 20)  # the centre (5, 5)
    assert np.isclose(p[c] / rho[c], T0, rtol=1e-14)
    assert np.isclose(w[c][1] / rho[c], 1.0, rtol=1e-15) and np.isclose(w[c][2] / rho[c], 1.0, rtol=1e-15)
    corner = (0, 0)  # r^2 = 50: the perturbation is below 1e-20
    assert np.allclose(w[corner], ics.prim_conserve(np.array([1.0, 1.0, 1.0, 0.5]), gamma), rtol=1e-15)


def test_lamb_oseen_vortex_ahead_of_the_shock():
    gamma, mach, vm, rc = 1.4, 1.1, 0.9, 0.075
    xc, yc = 0.25, 0.5
    r = np.linspace(0.2 * rc, 6.0 * rc, 3001)
    pts = Points(xc + r, np.full_like(r, yc))  # a ray along +x from the centre, ahead of x = 0.9
    w = ics.shock_vortex(pts, gamma, shock_mach=mach, shock_x=0.9, vortex_mach=vm, vortex_r=rc, vortex_x=xc,
                         vortex_y=yc)
    rho, p = w[..., 0], pressure(w, gamma)
    assert np.allclose(p / rho ** gamma, 1.0, rtol=1e-12)  # upstream state rho = p = 1, isentropic
    # on this ray the swirl is V; radial balance by centred differences
    vt = w[..., 2] / rho
    dpdr = np.gradient(p, r)
    bal = rho * vt ** 2 / r
    assert np.max(np.abs(dpdr[1:-1] - bal[1:-1])) <= 2e-4 * np.max(bal)
    # peak swirl Mach against the upstream sound speed sqrt(gamma)
    assert np.isclose(np.max(np.abs(vt)) / np.sqrt(gamma), vm, rtol=1e-6)
    # downstream of the shock: the uniform Rankine-Hugoniot state
    (_, _, _), (r2, u2, p2) = ics.rankine_hugoniot(mach, gamma)
    wd = ics.shock_vortex(Points(np.array([0.9]), np.array([0.1])), gamma, shock_mach=mach)
    assert np.allclose(wd[0], ics.prim_conserve(np.array([r2, u2, 0.0, r2 / (2.0 * p2)]), gamma), rtol=1e-15)
