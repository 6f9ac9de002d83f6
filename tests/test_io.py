"""Solution output (example/shock-vortex.jl:333-352 post-processing, VTK and state files): the
oracle's point-field restatement pinned by closed forms, and the host-side writers/readers."""
import numpy as np

from frhip import io
from oracle import oracle as O


def test_solution_field_oracle_closed_form():
    """Uniform state (rho, U, V, p) = (1.3, 0.4, -0.7, 2.1): every point of sol is
    (rho, U, V, 1/lambda) = (rho, U, V, 2p/rho); a second, random state goes back to its
    conserved variables through the independent prim_conserve2 (KitBase lambda = rho/(2p))."""
    g = 1.4
    nx, ny, n = 3, 2, 3
    rho, U, V, p = 1.3, 0.4, -0.7, 2.1
    w = O.prim_conserve2([rho, U, V, rho / (2 * p)], g)
    u = np.zeros((nx + 2, ny + 2, n, n, 4), order="F")
    u[...] = w
    xpg = np.random.default_rng(1).random((nx + 2, ny + 2, n, n, 2))
    x, y, sol = O.solution_field2d(u, xpg, g)
    assert np.allclose(sol[..., 0], rho, rtol=1e-15)
    assert np.allclose(sol[..., 1], U, rtol=1e-14)
    assert np.allclose(sol[..., 2], V, rtol=1e-14)
    assert np.allclose(sol[..., 3], 2 * p / rho, rtol=1e-14)
    # placement: element (i, j), point (k, l) -> ((i-1)*n + k, (j-1)*n + l)
    assert x[1 * n + 2, 0 * n + 1] == xpg[2, 1, 2, 1, 0] and y[2 * n + 0, 1 * n + 2] == xpg[3, 2, 0, 2, 1]
    rng = np.random.default_rng(7)
    prim = np.array([0.5 + rng.random(), rng.random() - 0.5, rng.random() - 0.5, 0.2 + rng.random()])
    u[2, 1, 1, 2] = O.prim_conserve2(prim, g)
    _, _, sol = O.solution_field2d(u, xpg, g)
    back = sol[1 * n + 1, 0 * n + 2].copy()
    back[3] = 1.0 / back[3]
    assert np.allclose(O.prim_conserve2(back, g), u[2, 1, 1, 2], rtol=1e-14)


def test_vtk_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    x, y = np.meshgrid(np.linspace(0, 1, 7), np.linspace(0, 2, 5), indexing="ij")
    f = {"rho": rng.random((7, 5)), "U": rng.random((7, 5)) - 0.5}
    path = str(tmp_path / "s.vtk")
    io.write_vtk(path, x, y, f)
    with open(path, "rb") as fh:
        head = fh.read(120).split(b"\n")
    assert head[0].startswith(b"# vtk DataFile Version 3.0") and head[2] == b"BINARY"
    assert head[4] == b"DIMENSIONS 7 5 1"
    x2, y2, f2 = io.read_vtk(path)
    assert np.array_equal(x2, x) and np.array_equal(y2, y)
    assert set(f2) == {"rho", "U"} and all(np.array_equal(f2[k], f[k]) for k in f)


def test_state_file_roundtrip(tmp_path):
    rng = np.random.default_rng(5)
    u = np.asfortranarray(rng.random((6, 4, 3, 3, 4)))
    path = str(tmp_path / "state.npz")
    io.save_state(path, u, nx=4, ny=2, deg=2, gamma=1.4, t=0.125, case="vortex")
    v, meta = io.load_state(path)
    assert v.flags["F_CONTIGUOUS"] and np.array_equal(v, u)
    assert meta == {"nx": 4, "ny": 2, "deg": 2, "gamma": 1.4, "t": 0.125, "case": "vortex"}
