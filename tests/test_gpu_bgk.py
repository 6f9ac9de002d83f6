"""GPU parity of the 2D discrete-velocity BGK residual and stepping (and its batched ensemble)
against the CPU oracle (tensorised example/bgk_wave.jl:69-129)."""
import numpy as np
import pytest

from frhip import bgk
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("deg", [1, 2, 3])
@pytest.mark.parametrize("nvel", [6, 8])
def test_bgk_residual_parity(deg, nvel):
    nx, ny = 19, 13  # partial 16 x 16 tiles
    u, wu = bgk.velocity_grid(nvel)
    v, wv = bgk.velocity_grid(nvel + 2)
    f = bgk.maxwellian_state(nx, ny, deg, u, v, amp=0.1)
    rng = np.random.default_rng(deg * 10 + nvel)
    f *= 1.0 + 0.02 * rng.uniform(-1, 1, size=f.size)
    g = bgk.BGK2D(nx, ny, deg, u, wu, v, wv, 1e-2)
    df = g.rhs(f)
    ref = O.bgk2d_rhs(f, nx, ny, deg, u, wu, v, wv, 1e-2)
    assert np.max(np.abs(df - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_bgk_ensemble_stepping_parity():
    nx, ny, deg, nvel = 16, 16, 2, 8
    u, wu = bgk.velocity_grid(nvel)
    taus = np.array([1e-3, 1e-2, 1e-1])
    f0 = bgk.maxwellian_state(nx, ny, deg, u, u, amp=0.1, nmem=len(taus))
    g = bgk.BGK2D(nx, ny, deg, u, wu, u, wu, taus)
    dt = 2e-4
    g.set_state(f0)
    g.step(dt, 5)
    out = g.get_state().reshape(len(taus), -1)
    # steps of one call alternate the state between two arrays: a split call gives the same state
    g.set_state(f0)
    g.step(dt, 2)
    g.step(dt, 3)
    assert np.array_equal(g.get_state().reshape(len(taus), -1), out)
    m = f0.size // len(taus)
    for i, tau in enumerate(taus):
        fc = f0[i * m:(i + 1) * m].copy()
        L = lambda x: O.bgk2d_rhs(x, nx, ny, deg, u, wu, u, wu, tau)  # noqa: E731
        for _ in range(5):
            a = (0.0 + 1.0 * fc) + dt * L(fc)
            b = (0.75 * fc + 0.25 * a) + (0.25 * dt) * L(a)
            fc = ((1 / 3) * fc + (2 / 3) * b) + ((2 / 3) * dt) * L(b)
        assert np.linalg.norm(out[i] - fc) <= 1e-10 * np.linalg.norm(fc)
    # mass is conserved by transport (periodic) and relaxation (moments of M equal those of f)
    w = g.moments()
    npts = nx * ny * (deg + 1) ** 2
    wq = np.polynomial.legendre.leggauss(deg + 1)[1]
    wpt = np.einsum("k,l->kl", wq, wq).ravel(order="F")
    for i in range(len(taus)):
        rho = w[i, 0].reshape((nx * ny, -1), order="F")
        mass = float((rho * wpt[None, :]).sum())
        assert abs(mass - 4.0 * nx * ny) <= 1e-10 * 4.0 * nx * ny
    assert w.shape == (len(taus), 4, npts)


def test_bgk_residual_parity_random_shapes():
    """Seeded sweep: grids 1-40 x 1-40 (single and partial 16 x 16 tiles), degrees 1-4, 2-10
    velocities per direction (unequal counts), one or three members; residual against the oracle."""
    rng = np.random.default_rng(2029)
    for draw in range(14):
        nx, ny = int(rng.integers(1, 41)), int(rng.integers(1, 41))
        deg = int(rng.integers(1, 5))
        nu, nv = int(rng.integers(2, 11)), int(rng.integers(2, 11))
        nmem = 1 if draw % 2 == 0 else 3
        u, wu = bgk.velocity_grid(nu)
        v, wv = bgk.velocity_grid(nv)
        taus = [1e-2] if nmem == 1 else [1e-3, 1e-2, 1e-1]
        f = bgk.maxwellian_state(nx, ny, deg, u, v, amp=0.1, nmem=nmem)
        f = f * (1.0 + 0.02 * rng.uniform(-1, 1, size=f.size))
        g = bgk.BGK2D(nx, ny, deg, u, wu, v, wv, taus if nmem > 1 else taus[0])
        df = g.rhs(f).reshape(nmem, -1)
        m = f.size // nmem
        for i, tau in enumerate(taus):
            ref = O.bgk2d_rhs(f[i * m:(i + 1) * m], nx, ny, deg, u, wu, v, wv, tau)
            case = (nx, ny, deg, nu, nv, nmem, i)
            assert np.max(np.abs(df[i] - ref.ravel())) <= 1e-12 * np.max(np.abs(ref)), case
