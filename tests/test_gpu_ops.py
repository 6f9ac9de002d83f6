"""GPU parity of the library-level contractions (fr_interp_face_dev = interp_face!,
src/interpolate.jl:1-10; fr_poly_derivative_dev = poly_derivative!, src/derivative.jl:1-10) and of
the nodal exponential filter applied on the device (filter_exp, src/Polynomial/poly_filter.jl:11-24)
against numpy contractions of the oracle's operators: rel inf-norm <= 1e-14.  Exactness on
polynomials (traces at +-1 and derivatives of x^q at the solution points) pins the operators."""
import numpy as np
import pytest

import frhip
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_reference_test_call_interp_face():
    # test/runtests.jl:39-41: f = randn(5, deg + 1), fδ = randn(5, 2) with FRPSpace1D(0, 1, 20, 5)
    ops = O.Ops(5)
    rng = np.random.default_rng(0)
    f = np.asfortranarray(rng.standard_normal((5, 6)))
    ff = np.asfortranarray(rng.standard_normal((5, 2)))
    frhip.interp_face(ff, f, ops.ll, ops.lr)
    assert rel_inf(ff[:, 0], f @ ops.ll) <= 1e-14 and rel_inf(ff[:, 1], f @ ops.lr) <= 1e-14
    # the vector method: one row
    v = np.zeros(2)
    frhip.interp_face(v, f[2], ops.ll, ops.lr)
    assert rel_inf(v, np.array([f[2] @ ops.ll, f[2] @ ops.lr])) <= 1e-14


@pytest.mark.parametrize("n", [1, 2, 4, 5, 9, 16])
def test_interp_face_batches_and_polynomials(n):
    r = np.cos(np.pi * (np.arange(n) + 0.5) / n)  # any distinct points
    ll = np.array([np.prod([(-1 - r[m]) / (r[q] - r[m]) for m in range(n) if m != q]) for q in range(n)])
    lr = np.array([np.prod([(1 - r[m]) / (r[q] - r[m]) for m in range(n) if m != q]) for q in range(n)])
    rng = np.random.default_rng(n)
    nrow, nb = 20011, 3
    f = np.asfortranarray(rng.standard_normal((nrow, n, nb)))
    ff = np.zeros((nrow, 2, nb), order="F")
    frhip.interp_face(ff, f, ll, lr)
    ref = np.stack([np.einsum("iqb,q->ib", f, ll), np.einsum("iqb,q->ib", f, lr)], axis=1)
    assert rel_inf(ff, ref) <= 1e-13
    # rows that sample a polynomial of degree < n give its exact end values
    c = rng.standard_normal((nrow, n))
    g = np.asfortranarray((c[:, None, :] * r[None, :, None] ** np.arange(n)[None, None, :]).sum(-1))
    gg = np.zeros((nrow, 2), order="F")
    frhip.interp_face(gg, g, ll, lr)
    ends = np.stack([(c * (-1.0) ** np.arange(n)).sum(1), c.sum(1)], 1)
    assert np.max(np.abs(gg - ends)) <= 1e-9 * np.max(np.abs(ends))


@pytest.mark.parametrize("n", [3, 4, 9])
def test_row_pairs_and_single_rows_agree(n):
    """An even row count on 16-byte aligned device arrays takes the two-rows-per-thread kernel
    (16-byte loads and stores); the same data one double off alignment takes the one-row kernel.
    Both give the same bits, and the numpy contraction to round-off."""
    import torch

    rng = np.random.default_rng(40 + n)
    nrow, nb = 20012, 5
    f = np.asfortranarray(rng.standard_normal((nrow, n, nb)))
    pdm = np.asfortranarray(rng.standard_normal((n, n)))
    flat = torch.from_numpy(f.ravel(order="F").copy()).cuda()
    buf = torch.zeros(flat.numel() + 1, dtype=torch.float64, device="cuda")
    buf[1:] = flat
    shifted = buf[1:]  # 8 bytes past a 16-byte boundary
    assert flat.data_ptr() % 16 == 0 and shifted.data_ptr() % 16 == 8
    outs = []
    for src in (flat, shifted):
        d = torch.empty(nrow * n * nb, dtype=torch.float64, device="cuda")
        frhip.poly_derivative(d.view(nb, n, nrow), src.view(nb, n, nrow), pdm)
        ff = torch.empty(nrow * 2 * nb, dtype=torch.float64, device="cuda")
        frhip.interp_face(ff.view(nb, 2, nrow), src.view(nb, n, nrow), pdm[0], pdm[1])
        torch.cuda.synchronize()
        outs.append((d.cpu().numpy(), ff.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    ref = np.einsum("iqb,pq->ipb", f, pdm).ravel(order="F")
    assert rel_inf(outs[0][0], ref) <= 1e-13
    reff = np.stack([np.einsum("iqb,q->ib", f, pdm[0]), np.einsum("iqb,q->ib", f, pdm[1])], axis=1).ravel(order="F")
    assert rel_inf(outs[0][1], reff) <= 1e-13


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 6])
def test_poly_derivative_state_and_exactness(deg):
    ops = O.Ops(deg)
    n = deg + 1
    rng = np.random.default_rng(deg)
    # a 1D Euler state u(i, p, k): rows = cells, q = solution point, batch = variable
    u = np.asfortranarray(rng.standard_normal((3001, n, 3)))
    du = np.zeros_like(u, order="F")
    frhip.poly_derivative(du, u, ops.dl)
    assert rel_inf(du, np.einsum("ipk,mp->imk", u, ops.dl)) <= 1e-13
    # derivatives of x^q (q <= p) are exact at the solution points
    for q in range(n):
        x = np.asfortranarray(np.tile(ops.r ** q, (7, 1)))
        d = np.zeros_like(x, order="F")
        frhip.poly_derivative(d, x, ops.dl)
        exact = q * ops.r ** (q - 1) if q else np.zeros(n)
        assert np.max(np.abs(d - exact[None, :])) <= 1e-12
    # the reference's vector method
    v = np.zeros(n)
    frhip.poly_derivative(v, u[5, :, 1].copy(), ops.dl)
    assert rel_inf(v, ops.dl @ u[5, :, 1]) <= 1e-14
    with pytest.raises(ValueError):
        frhip.poly_derivative(v, u[5, :, 1].copy(), np.zeros((n + 1, n + 1)))


def test_derivative_along_a_2d_state_axis_on_torch_tensors():
    """The r-derivative of a 2D state u(i, j, k, l, m) on the device: rows = (i, j), q = k,
    batch = (l, m); torch tensors in reversed shape on a side stream, ordered by the stream."""
    import torch

    nx, ny, deg = 37, 21, 3
    n = deg + 1
    ops = O.Ops(deg)
    rng = np.random.default_rng(4)
    u = np.asfortranarray(rng.standard_normal((nx + 2, ny + 2, n, n, 4)))
    rows = (nx + 2) * (ny + 2)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ud = torch.from_numpy(u.ravel(order="F").copy()).cuda().view(n * 4, n, rows)
        dd = torch.full_like(ud, float("nan"))
        frhip.poly_derivative(dd, ud, ops.dl)
        out = dd.clone()
    torch.cuda.synchronize()
    got = out.cpu().numpy().ravel().reshape(u.shape, order="F")
    assert rel_inf(got, np.einsum("ijqlm,kq->ijklm", u, ops.dl)) <= 1e-13


@pytest.mark.parametrize("deg", [2, 4])
def test_filter_exp_applied_on_the_device(deg):
    """F = filter_exp(N, s, V) applied to every cell of a 1D state through the row contraction
    keeps cell means, removes the highest mode, and equals the oracle's F times the state."""
    ops = O.Ops(deg)
    n = deg + 1
    F = frhip.filter_exp(deg, 8, ops.V1)
    Fo = O.filter_exp(deg, 8, ops.V1)
    rng = np.random.default_rng(deg)
    u = np.asfortranarray(rng.standard_normal((4099, n, 3)))
    out = np.zeros_like(u, order="F")
    frhip.poly_derivative(out, u, F)
    assert rel_inf(out, np.einsum("ipk,mp->imk", u, Fo)) <= 1e-13
    w = ops.w
    assert np.allclose(np.einsum("ipk,p->ik", out, w), np.einsum("ipk,p->ik", u, w), atol=1e-12)
    modal = np.einsum("mp,ipk->imk", np.linalg.inv(ops.V1), out)
    assert np.max(np.abs(modal[:, deg, :])) <= 1e-12
