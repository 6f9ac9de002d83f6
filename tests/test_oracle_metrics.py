"""Cells with point-varying metrics (an FRPSpace2D over a curvilinear mesh, src/struct.jl:130-142).

Pins, on the CPU:
- the host metrics: the batched Jacobians equal the pinned single-point rs_jacobi
  (src/Geometry/geo_jacobi.jl:77-88) bitwise, iJ J = I, and the quadrature of det J over each
  polar cell equals the cell's shoelace area (det J of a bilinear map is linear in (r, s), so the
  Gauss rule integrates it exactly); the polar mesh's vertices lie on their circles;
- the oracle's per-point residual (example/shock-vortex.jl:26-118 reading iJ[i, j][k, l]):
  * a point array that holds the rectangle metric gives the rectangle oracle bitwise;
  * metrics constant per cell give, cell by cell, the rectangle oracle run with that cell's
    metric (du of a cell reads only its own metric and the neighbours' traces of u);
  * du is linear in the metric;
  * an independent numpy restatement agrees on a polar mesh, and the same restatement with the
    point indices (k, l) of the metric swapped does not (so a transposed index is caught).
"""
import numpy as np
import pytest

import np_reference as NR
from oracle import oracle as O

G = 1.4


def smooth_state(ps, seed=1, amp=0.01):
    """A positive smooth state on the physical solution points, then perturbed per value."""
    from frhip import ics

    x, y = ps.xpg[..., 0], ps.xpg[..., 1]
    rho = 1.0 + 0.2 * np.sin(x) * np.cos(y)
    u = 0.3 * np.cos(y)
    v = -0.2 * np.sin(x)
    p = 1.0 + 0.1 * np.cos(x + y)
    w = np.stack([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)], axis=-1)
    return ics.perturbed(np.asfortranarray(w), seed=seed, amplitude=amp)


def polar_space(nr, nth, deg):
    import frhip

    return frhip.FRPSpace2D(frhip.CSpace2D(1.0, 6.0, nr, 0.0, np.pi, nth, 1, 1), deg)


def oracle_for(ps, flux="hll", bc="ghost"):
    return O.Fr2d(ps.nx, ps.ny, ps.deg, 1.0, 1.0, G, flux, bc, iJp=ps.iJ)


def test_batched_jacobians_match_the_pinned_single_point_routine():
    import frhip

    ps = polar_space(6, 5, 2)
    assert ps.J.shape == (8, 7, 3, 3, 2, 2) and not ps.uniform
    r = ps.xpl
    for (i, j, k, l) in [(0, 0, 0, 0), (3, 2, 1, 2), (7, 6, 2, 1), (5, 0, 0, 2)]:
        J = frhip.rs_jacobi(r[k], r[l], ps.vertices[i, j])
        assert np.array_equal(ps.J[i, j, k, l], J)
    prod = np.einsum("...ab,...bc->...ac", ps.iJ, ps.J)
    assert np.max(np.abs(prod - np.eye(2))) <= 1e-14
    assert np.all(np.linalg.det(ps.J) > 0)


def test_polar_mesh_geometry_and_exact_cell_areas():
    import frhip

    ps = polar_space(7, 9, 2)
    b = ps.base
    rad = np.hypot(b.vertices[..., 0], b.vertices[..., 1])
    assert np.allclose(rad[..., 0], b.r - b.dr / 2, rtol=1e-14)
    assert np.allclose(rad[..., 2], b.r + b.dr / 2, rtol=1e-14)
    assert np.allclose(np.hypot(b.x, b.y), b.r, rtol=1e-14)
    assert ps.nx == 7 and ps.ny == 9 and ps.ngx == 1 and ps.nθ == 9
    # area of each straight-sided cell: the Gauss quadrature of det J over [-1, 1]^2
    quad = np.einsum("kl,ijkl->ij", ps.wp, np.linalg.det(ps.J))
    v = b.vertices
    xs, ys = v[..., 0], v[..., 1]
    shoelace = 0.5 * np.abs(np.sum(xs * np.roll(ys, -1, axis=-1) - np.roll(xs, -1, axis=-1) * ys, axis=-1))
    assert np.max(np.abs(quad - shoelace) / shoelace) <= 1e-13
    # interface Jacobians: the reference's points (r_1, 0), (1, r_1), (r_n, 1), (0, r_n)
    r = ps.xpl
    for f, (rr, ss) in enumerate(((r[0], 0.0), (1.0, r[0]), (r[-1], 1.0), (0.0, r[-1]))):
        for q in range(3):
            assert np.array_equal(ps.Ji[2, 3, f, q], frhip.rs_jacobi(rr, ss, b.vertices[2, 3]))


def test_point_metric_holding_the_rectangle_metric_is_the_rectangle_oracle():
    import frhip

    nx, ny, deg = 6, 5, 3
    ps = frhip.FRPSpace2D(0.0, 10.0, nx, 0.0, 10.0, ny, deg, 1, 1)
    u = smooth_state(ps, seed=4)
    iJp = np.zeros((nx + 2, ny + 2, deg + 1, deg + 1, 2, 2), order="F")
    iJp[..., :, :] = ps.iJ_elem
    for flux in ("hll", "rusanov", "roe"):
        for bc in ("ghost", "periodic"):
            rect = O.Fr2d(nx, ny, deg, 10 / nx, 10 / ny, G, flux, bc).rhs(u)
            pt = O.Fr2d(nx, ny, deg, 1.0, 1.0, G, flux, bc, iJp=iJp).rhs(u)
            assert np.array_equal(rect, pt)


def test_cellwise_constant_metrics_are_rectangle_runs_cell_by_cell():
    import frhip

    nx, ny, deg = 7, 6, 2
    ps = frhip.FRPSpace2D(0.0, 7.0, nx, 0.0, 6.0, ny, deg, 1, 1)
    u = smooth_state(ps, seed=9)
    values = [(2.0, 2.0), (3.5, 1.25), (0.8, 4.0)]
    pick = np.add.outer(np.arange(nx + 2), 2 * np.arange(ny + 2)) % 3
    iJp = np.zeros((nx + 2, ny + 2, deg + 1, deg + 1, 2, 2), order="F")
    for c, (a, d) in enumerate(values):
        iJp[pick == c, :, :, 0, 0] = a
        iJp[pick == c, :, :, 1, 1] = d
    for bc in ("ghost", "periodic"):
        du = O.Fr2d(nx, ny, deg, 1.0, 1.0, G, "hll", bc, iJp=iJp).rhs(u)
        for c, (a, d) in enumerate(values):
            ref = O.Fr2d(nx, ny, deg, 2.0 / a, 2.0 / d, G, "hll", bc).rhs(u)
            cells = pick[1:-1, 1:-1] == c
            assert np.array_equal(du[1:-1, 1:-1][cells], ref[1:-1, 1:-1][cells])


def test_residual_is_linear_in_the_metric():
    ps = polar_space(5, 6, 2)
    u = smooth_state(ps, seed=2)
    rng = np.random.default_rng(3)
    M1 = np.asfortranarray(ps.iJ)
    M2 = np.asfortranarray(rng.uniform(-1.0, 1.0, ps.iJ.shape))
    R = lambda M: O.Fr2d(ps.nx, ps.ny, ps.deg, 1.0, 1.0, G, "rusanov", "ghost", iJp=M).rhs(u)  # noqa: E731
    lhs = R(np.asfortranarray(M1 + 2.0 * M2))
    rhs = R(M1) + 2.0 * R(M2)
    assert np.max(np.abs(lhs - rhs)) <= 1e-13 * np.max(np.abs(rhs))


@pytest.mark.parametrize("deg", [1, 2, 3])
@pytest.mark.parametrize("flux", ["hll", "rusanov"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_polar_mesh_residual_vs_numpy_restatement(deg, flux, bc):
    ps = polar_space(5, 4, deg)
    u = smooth_state(ps, seed=deg)
    du = oracle_for(ps, flux, bc).rhs(u)
    ref = NR.fr2d_rhs_metric(u, NR.operators(deg), G, np.asarray(ps.iJ), flux, bc == "periodic")
    assert np.max(np.abs(du - ref)) <= 5e-14 * np.max(np.abs(ref))
    # the metric with its point indices (k, l) swapped gives a different residual
    swapped = np.swapaxes(np.asarray(ps.iJ), 2, 3)
    bad = NR.fr2d_rhs_metric(u, NR.operators(deg), G, swapped, flux, bc == "periodic")
    assert np.max(np.abs(bad - ref)) > 1e-6 * np.max(np.abs(ref))
