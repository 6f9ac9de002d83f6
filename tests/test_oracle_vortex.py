"""Whole-scheme pins of the 2D oracle by the isentropic vortex, whose exact solution is the initial
state carried with the free stream (u, v) = (1, 1) over the periodic square: at t = 2 it is the
initial vortex moved by (2, 2), each point taking the nearest periodic image of the centre.  The
order of accuracy of the p = 3 scheme (Radau correction, i.e. nodal DG, on the reference's dudt!
of example/shock-vortex.jl:26-118) must be p + 1 = 4 with the Rusanov flux of the benchmark, and
p + 1 = 3 at p = 2 with HLL.  A dropped or mis-signed correction term, a wrong Riemann flux or a
transposed derivative operator leaves at most first order or no convergence at all (one such
mistake is planted below).  With the Rusanov flux p = 2 converges at about 2.4 only (measured
2.63 / 2.35 at 16 / 32 / 64 cells, against 3.2 for HLL and Roe): the flux's extra dissipation, not
a defect of one degree, since the same operators give order 3.2 with HLL."""
import numpy as np
import pytest

import frhip
from frhip import ics
from oracle import oracle as O

GAMMA = 1.4


def vortex_error(n, deg, L, flux="rusanov", t_end=2.0, cfl=0.05):
    """Density L2 error (Gauss quadrature over the domain) of the oracle's SSP-RK3 run at t_end
    against the translated exact vortex."""
    ps = frhip.FRPSpace2D(0.0, L, n, 0.0, L, n, deg, 1, 1)
    c = (0.5 * L, 0.5 * L)
    u0 = ics.isentropic_vortex(ps, GAMMA, strength=5.0, center=c)
    dt, ns = ics.cfl_steps(u0, GAMMA, ps.dx0, cfl, t_end)
    orc = O.Fr2d(n, n, deg, L / n, L / n, GAMMA, flux, "periodic")
    u = u0.copy(order="F")
    orc.ssprk3(u, dt, ns)
    ue = ics.isentropic_vortex(ps, GAMMA, strength=5.0, center=c, t=t_end)
    d = (u - ue)[1:-1, 1:-1, :, :, 0]
    jac = (L / n) ** 2 / 4.0
    return float(np.sqrt(np.einsum("ijkl,kl->", d * d, ps.wp) * jac)), u, ue


def test_vortex_order_p3():
    # the benchmark square [0,10]^2 and flux; measured 8.51e-3, 5.07e-4, 2.19e-5 (orders 4.07, 4.53)
    e = [vortex_error(n, 3, 10.0)[0] for n in (8, 16, 32)]
    orders = [np.log2(e[i] / e[i + 1]) for i in range(2)]
    assert all(3.6 <= q <= 5.5 for q in orders), (e, orders)
    assert e[-1] <= 3e-5, e


def test_vortex_order_p2_hll():
    # measured 3.69e-3, 4.10e-4 (order 3.17)
    e = [vortex_error(n, 2, 10.0, flux="hll")[0] for n in (16, 32)]
    assert 2.8 <= np.log2(e[0] / e[1]) <= 4.5, e


def test_exact_solution_is_periodic_in_time():
    # one period (t = 10 moves the vortex by the side of the square) gives the initial state back
    # on the interior (ghost points lie outside the square, where t = 0 uses the raw distance)
    ps = frhip.FRPSpace2D(0.0, 10.0, 6, 0.0, 10.0, 6, 3, 1, 1)
    u0 = ics.isentropic_vortex(ps, GAMMA, strength=5.0)
    u1 = ics.isentropic_vortex(ps, GAMMA, strength=5.0, t=10.0)
    assert np.max(np.abs(u1 - u0)[1:-1, 1:-1]) <= 1e-12
    # and t = 2 is the initial state moved by two cells of width 10/5 on a 5 x 5 grid
    ps5 = frhip.FRPSpace2D(0.0, 10.0, 5, 0.0, 10.0, 5, 3, 1, 1)
    a = ics.isentropic_vortex(ps5, GAMMA, strength=5.0)
    b = ics.isentropic_vortex(ps5, GAMMA, strength=5.0, t=2.0)
    assert np.max(np.abs(b[2:-1, 2:-1] - a[1:-2, 1:-2])) <= 1e-12


def test_vortex_order_detects_a_wrong_correction():
    """The pin must fail for a plausible mistake: the left correction derivative taken as the
    right one reversed without its sign change (correct: dhl(x_k) = -dhr(x_{p+2-k})).  The mistake
    is planted in the oracle's operators, on the CPU."""
    L, deg = 10.0, 3
    errs = []
    for n in (8, 16):
        ps = frhip.FRPSpace2D(0.0, L, n, 0.0, L, n, deg, 1, 1)
        ops = O.Ops(deg)
        ops.dhl = ops.dhr[::-1].copy()
        orc = O.Fr2d(n, n, deg, L / n, L / n, GAMMA, "rusanov", "periodic", ops=ops)
        u0 = ics.isentropic_vortex(ps, GAMMA, strength=5.0)
        dt, ns = ics.cfl_steps(u0, GAMMA, ps.dx0, 0.05, 0.5)
        u = u0.copy(order="F")
        with np.errstate(all="ignore"):
            orc.ssprk3(u, dt, ns)
        ue = ics.isentropic_vortex(ps, GAMMA, strength=5.0, t=0.5)
        d = (u - ue)[1:-1, 1:-1, :, :, 0]
        errs.append(float(np.sqrt(np.mean(d * d))) if np.all(np.isfinite(d)) else np.inf)
    assert not (np.isfinite(errs[1]) and np.log2(errs[0] / errs[1]) > 3.0), errs
