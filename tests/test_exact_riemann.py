"""Pins the exact 1D Euler Riemann solver of tests/np_reference.py (used as the exact solution of
the reference's Sod runs) against the published star states of Toro, Riemann Solvers and
Numerical Methods for Fluid Dynamics, Table 4.2 (gamma = 1.4, five significant digits), and
checks its sampled solution: the far-field states, the constant star states and the continuity of
the rarefaction fan at its head and tail."""
import numpy as np
import pytest

from np_reference import exact_riemann

G = 1.4
TORO = [  # (rho, u, p) left, right -> p*, u*, rho*L, rho*R as printed
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), ("0.30313", "0.92745", "0.42632", "0.26557")),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), ("0.00189", "0.00000", "0.02185", "0.02185")),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), ("460.894", "19.5975", "0.57506", "5.99924")),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), ("46.0950", "-6.19633", "5.99242", "0.57511")),
]


def close_to_printed(x, s):
    """x rounds to the printed value s: within half a unit of its last digit."""
    dec = len(s.split(".")[1])
    return abs(x - float(s)) <= 0.5 * 10.0**-dec + 1e-12


@pytest.mark.parametrize("case", range(len(TORO)))
def test_star_states_of_toro_table_4_2(case):
    wl, wr, (pst, ust, rl, rr) = TORO[case]
    _, _, _, ps, us = exact_riemann(wl, wr, G, np.array([0.0]))
    assert close_to_printed(ps, pst) and close_to_printed(us, ust), (ps, us)
    # the star densities, sampled just left and right of the contact
    rho, u, p, _, _ = exact_riemann(wl, wr, G, np.array([us - 1e-9, us + 1e-9]))
    assert close_to_printed(rho[0], rl) and close_to_printed(rho[1], rr), rho
    assert np.allclose(p, ps, rtol=1e-12) and np.allclose(u, us, rtol=1e-12, atol=1e-15)


def test_far_field_and_fan_continuity():
    wl, wr = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)  # Sod: rarefaction left, shock right
    rho, u, p, ps, us = exact_riemann(wl, wr, G, np.array([-10.0, 10.0]))
    assert (rho[0], u[0], p[0]) == wl and (rho[1], u[1], p[1]) == wr
    cl = np.sqrt(G * wl[2] / wl[0])
    head, tail = -cl, us - cl * (ps / wl[2]) ** ((G - 1) / (2 * G))
    e = 1e-10
    r, uu, pp, _, _ = exact_riemann(wl, wr, G, np.array([head - e, head + e, tail - e, tail + e]))
    assert abs(r[0] - r[1]) < 1e-8 and abs(pp[0] - pp[1]) < 1e-8
    assert abs(r[2] - r[3]) < 1e-8 and abs(uu[2] - uu[3]) < 1e-8
