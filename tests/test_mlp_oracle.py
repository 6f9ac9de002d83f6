"""CPU pins of the learned-sensor oracle (oracle/mlp_oracle.py): the float64 gradient against
central finite differences of the loss and against torch's CPU autograd of the same network, the
forward pass against torch, and a planted mistake (W2 used transposed in the backward pass) that
the finite-difference check must reject.  The network has no counterpart in the reference."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

NPAR = M.NH * M.NIN + M.NH + M.NH * M.NH + M.NH + M.NH + 1


def _params(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((M.NH, M.NIN)) * np.sqrt(2.0 / M.NIN)
    b1 = 0.1 * rng.standard_normal(M.NH)
    W2 = rng.standard_normal((M.NH, M.NH)) * np.sqrt(2.0 / M.NH)
    b2 = 0.1 * rng.standard_normal(M.NH)
    W3 = rng.standard_normal(M.NH) * np.sqrt(1.0 / M.NH)
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [0.05]])


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, M.NIN))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = (rng.uniform(size=n) > 0.7).astype(np.float64)
    return x, y


# parameter indices spread over every block: W1, b1, W2, b2, W3, b3
_BLOCKS = np.cumsum([0, M.NH * M.NIN, M.NH, M.NH * M.NH, M.NH, M.NH, 1])


def _indices(seed, per_block=6):
    rng = np.random.default_rng(seed)
    idx = []
    for a, b in zip(_BLOCKS[:-1], _BLOCKS[1:]):
        idx.extend(rng.choice(np.arange(a, b), size=min(per_block, b - a), replace=False).tolist())
    return idx


def test_parameter_count():
    assert _BLOCKS[-1] == NPAR == 1633


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_finite_differences(seed):
    p = _params(seed)
    x, y = _data(512, seed + 10)
    g, _ = M.grad(p, x, y, rounding=False)
    idx = _indices(seed + 20)
    fd = M.finite_difference(p, x, y, idx, h=1e-6)
    scale = np.max(np.abs(g))
    # central differences of a piecewise-quadratic loss: error O(h^2) away from ReLU kinks, plus
    # round-off of ~1e-16 / h in the difference quotient
    assert np.max(np.abs(g[idx] - fd)) <= 1e-7 * scale, (g[idx], fd)


def test_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    p = _params(7)
    x, y = _data(1000, 8)
    g, loss = M.grad(p, x, y, rounding=False)
    W1, b1, W2, b2, W3, b3 = (torch.tensor(np.array(a), dtype=torch.float64, requires_grad=True)
                              for a in M.unpack(p))
    X, Y = torch.tensor(x), torch.tensor(y)
    o = torch.relu(torch.relu(X @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    L = torch.mean((o - Y) ** 2)
    L.backward()
    gt = np.concatenate([W1.grad.numpy().ravel(), b1.grad.numpy(), W2.grad.numpy().ravel(), b2.grad.numpy(),
                         W3.grad.numpy(), [b3.grad.item()]])
    assert abs(loss - L.item()) <= 1e-14 * abs(loss)
    assert np.max(np.abs(g - gt)) <= 1e-13 * np.max(np.abs(gt))
    of, _ = M.forward(p, x, rounding=False)
    assert np.max(np.abs(of - o.detach().numpy())) <= 1e-13 * np.max(np.abs(of))


def test_rounded_gradient_is_a_bf16_approximation():
    # the BF16 rounding points of the device kernels move the gradient by BF16 round-off only
    p = _params(3)
    x, y = _data(2048, 4)
    gr, lr_ = M.grad(p, x, y, rounding=True)
    g, l64 = M.grad(p, x, y, rounding=False)
    assert abs(lr_ - l64) <= 2e-2 * l64
    assert np.linalg.norm(gr - g) <= 2e-2 * np.linalg.norm(g)


def test_finite_differences_reject_a_transposed_w2():
    # a plausible backward-pass mistake: delta1 = (z1 > 0) * (delta2 @ W2.T) instead of @ W2
    p = _params(5)
    x, y = _data(512, 6)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    o, a = M.forward(p, x, rounding=False)
    dout = 2.0 * (o - y) / x.shape[0]
    d2 = (a["z2"] > 0) * (dout[:, None] * W3[None, :])
    d1_wrong = (a["z1"] > 0) * (d2 @ W2.T)
    gW1_wrong = d1_wrong.T @ a["x"]
    idx = list(range(0, M.NH * M.NIN, 37))
    fd = M.finite_difference(p, x, y, idx, h=1e-6)
    g, _ = M.grad(p, x, y, rounding=False)
    scale = np.max(np.abs(g))
    assert np.max(np.abs(g[idx] - fd)) <= 1e-7 * scale
    assert np.max(np.abs(gW1_wrong.ravel()[idx] - fd)) > 1e-3 * scale
