"""The surrogate's FP64 oracle: MLP gradients by finite differences and
torch autograd, Adam against its closed form, and the straight-line estimate
against exact cases and the RK4 oracle."""
import numpy as np
import pytest
import torch

from oracle import field as F
from oracle import surrogate as S
from oracle import tracer as T


def rand_params(dims, seed=0):
    rs = np.random.RandomState(seed)
    return [(rs.randn(dims[i + 1], dims[i]) / np.sqrt(dims[i]), rs.randn(dims[i + 1]) * 0.1)
            for i in range(len(dims) - 1)]


def test_mlp_backward_finite_differences():
    dims = (6, 9, 7, 3)
    params = rand_params(dims)
    rs = np.random.RandomState(1)
    X = rs.randn(5, 6)
    G = rs.randn(5, 3)   # loss = sum(G * Y)
    Y, cache = S.mlp_forward(params, X)
    grads = S.mlp_backward(params, cache, G)
    eps = 1e-6
    for li in range(len(params)):
        W, b = params[li]
        for (i, j) in [(0, 0), (2, 3), (W.shape[0] - 1, W.shape[1] - 1)]:
            Wp, Wm = W.copy(), W.copy()
            Wp[i, j] += eps
            Wm[i, j] -= eps
            pp = [(Wp, b) if k == li else params[k] for k in range(len(params))]
            pm = [(Wm, b) if k == li else params[k] for k in range(len(params))]
            fd = (np.sum(G * S.mlp_forward(pp, X)[0]) - np.sum(G * S.mlp_forward(pm, X)[0])) / (2 * eps)
            assert abs(fd - grads[li][0][i, j]) < 1e-7 * max(1, abs(fd))
        bp, bm = b.copy(), b.copy()
        bp[1] += eps
        bm[1] -= eps
        pp = [(W, bp) if k == li else params[k] for k in range(len(params))]
        pm = [(W, bm) if k == li else params[k] for k in range(len(params))]
        fd = (np.sum(G * S.mlp_forward(pp, X)[0]) - np.sum(G * S.mlp_forward(pm, X)[0])) / (2 * eps)
        assert abs(fd - grads[li][1][1]) < 1e-7 * max(1, abs(fd))


def test_mlp_matches_torch_autograd():
    dims = (8, 16, 16, 4)
    params = rand_params(dims, 3)
    X = np.random.RandomState(4).randn(10, 8)
    G = np.random.RandomState(5).randn(10, 4)
    Y, cache = S.mlp_forward(params, X)
    grads = S.mlp_backward(params, cache, G)
    tp = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)) for W, b in params]
    h = torch.tensor(X)
    for li, (W, b) in enumerate(tp):
        h = h @ W.T + b
        if li < len(tp) - 1:
            h = torch.nn.functional.gelu(h)
    np.testing.assert_allclose(h.detach().numpy(), Y, rtol=1e-12, atol=1e-12)
    (h * torch.tensor(G)).sum().backward()
    for (W, b), (gW, gb) in zip(tp, grads):
        np.testing.assert_allclose(W.grad.numpy(), gW, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(b.grad.numpy(), gb, rtol=1e-10, atol=1e-12)


def test_adam_closed_form_first_step():
    # first step of Adam moves every weight by lr * sign(g) (up to eps)
    w = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -0.1, 0.0])
    w1, m, v = S.adam_step(w, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(w1[:2], w[:2] - 0.01 * np.sign(g[:2]), rtol=1e-6)
    assert w1[2] == w[2]


def cube_of(ne, n, lwl=1064e-9):
    x = np.linspace(-5e-3, 5e-3, n)
    c = T.Cube(x, x.copy(), x.copy())
    c.external_ne(ne)
    c.calc_dndr(lwl=lwl)
    c.extent = c.extent_z
    return c


def test_straight_null_and_uniform_exact():
    n = 9
    c = cube_of(np.zeros((n, n, n)), n)
    s0 = c.init_beam(30, 2e-3, 1e-3, rng=np.random.RandomState(0))
    base, feat = S.straight_estimate(T.NodeGrid(c), s0)
    rf, ph, _, _ = T.zmarch_rk4(T.NodeGrid(c), s0)
    np.testing.assert_allclose(base[:4], rf, atol=1e-15)
    c = cube_of(np.full((n, n, n), 3e25), n)
    s0 = c.init_beam(30, 2e-3, 0.0, rng=np.random.RandomState(0))
    base, feat = S.straight_estimate(T.NodeGrid(c), s0)
    r = 3e25 / T.critical_density(1064e-9)
    np.testing.assert_allclose(base[4], T.wavenumber(1064e-9) * (np.sqrt(1 - r) - 1) * 10e-3, rtol=1e-12)


def test_straight_linear_slab_first_order_exact():
    """A constant gradient has no second-order term: the estimate's angle is
    the exact deflection of the parabolic ray."""
    n = 11
    x = np.linspace(-5e-3, 5e-3, n)
    XX = np.meshgrid(x, x, x, indexing="ij")[0]
    c = cube_of(1e25 * (1 + 2 * XX / 5e-3), n, lwl=1053e-9)
    s0 = c.init_beam(20, 1e-3, 0.0, rng=np.random.RandomState(1))
    base, feat = S.straight_estimate(T.NodeGrid(c), s0)
    rf, ph, _, _ = T.zmarch_rk4(T.NodeGrid(c), s0)
    np.testing.assert_allclose(base[1], rf[1], rtol=1e-9)
    np.testing.assert_allclose(base[0], rf[0], rtol=0, atol=1e-12)
    assert np.all(feat[:, 36:40] == 0)  # T and R vanish


def test_straight_second_order_beats_first_order():
    n = 33
    d = F.powerlaw_field((n, n, n), -11 / 3, seed=3)
    d /= np.sqrt((d ** 2).mean())
    c = cube_of(1e24 * (1 + 0.1 * d), n)
    s0 = c.init_beam(500, 2e-3, 50e-6, rng=np.random.RandomState(1))
    g = T.NodeGrid(c)
    rf, ph, _, _ = T.zmarch_rk4(g, s0)
    base, feat = S.straight_estimate(g, s0)
    U0, W0 = s0[3] / T.C, s0[5] / T.C
    Sx, Sz = feat[:, 32], feat[:, 40]
    born1 = np.arctan((U0 + Sx / W0) / (W0 + Sz / W0))
    e2 = np.sqrt(np.mean((base[1] - rf[1]) ** 2))
    e1 = np.sqrt(np.mean((born1 - rf[1]) ** 2))
    assert e2 < 0.2 * e1
    assert np.abs(base[[0, 2]] - rf[[0, 2]]).max() < 1e-3 * (x_step := 10e-3 / (n - 1))
