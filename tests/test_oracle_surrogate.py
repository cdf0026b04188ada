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
            h = torch.nn.functional.gelu(h, approximate="tanh")
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


def test_path_null_and_uniform_exact():
    n = 9
    c = cube_of(np.zeros((n, n, n)), n)
    s0 = c.init_beam(30, 2e-3, 1e-3, rng=np.random.RandomState(0))
    base, feat = S.path_estimate(T.NodeGrid(c), s0)
    rf, ph, _, _ = T.zmarch_rk4(T.NodeGrid(c), s0)
    np.testing.assert_allclose(base[:4], rf, atol=1e-15)
    c = cube_of(np.full((n, n, n), 3e25), n)
    s0 = c.init_beam(30, 2e-3, 0.0, rng=np.random.RandomState(0))
    base, feat = S.path_estimate(T.NodeGrid(c), s0)
    r = 3e25 / T.critical_density(1064e-9)
    np.testing.assert_allclose(base[4], T.wavenumber(1064e-9) * (np.sqrt(1 - r) - 1) * 10e-3, rtol=1e-12)


def test_path_hands_over_the_rays_the_zmarch_hands_over():
    """On a null field the estimate's path is the z-march's straight line, so
    the rays it gives a NaN estimate (for the surrogate's fallback) are
    exactly those the oracle z-march hands over (status 1): steep at launch
    (W <= 0.97), or crossing a side face of the grid; the others are
    finite and equal the z-march's exits."""
    n = 9
    c = cube_of(np.zeros((n, n, n)), n)
    C = T.C
    th = np.array([0.0, 0.1, 0.2, 0.23, 0.26, 0.3, 0.05, 0.05, -0.05, 0.2])  # slope angle in x
    x0 = np.array([0.0, 0.0, -1e-3, -2e-3, 0.0, 0.0, 4.9e-3, 4.0e-3, -4.9e-3, -4.9e-3])
    s0 = np.zeros((6, th.size))
    s0[0] = x0
    s0[2] = -5e-3
    s0[3] = C * np.sin(th)
    s0[5] = C * np.cos(th)
    base, _ = S.path_estimate(T.NodeGrid(c), s0)
    rf, _, _, st = T.zmarch_rk4(T.NodeGrid(c), s0)
    nan = ~np.isfinite(base).all(axis=0)
    # W = cos(th) <= 0.97 for |th| >= 0.2443; 4.9 mm + 10 mm * tan(0.05) leaves the +-5 mm grid
    np.testing.assert_array_equal(nan, [False, False, False, False, True, True, True, False, True, False])
    np.testing.assert_array_equal(nan, st == 1)
    np.testing.assert_allclose(base[:4, ~nan], rf[:, ~nan], atol=1e-15)


def test_path_constant_gradient_matches_parabola():
    """A constant transverse gradient: the mid-plane kick integrates the
    parabolic ray exactly (positions), and the angle to the order of W's
    change."""
    n = 11
    x = np.linspace(-5e-3, 5e-3, n)
    XX = np.meshgrid(x, x, x, indexing="ij")[0]
    c = cube_of(1e25 * (1 + 2 * XX / 5e-3), n, lwl=1053e-9)
    s0 = c.init_beam(20, 1e-3, 0.0, rng=np.random.RandomState(1))
    base, feat = S.path_estimate(T.NodeGrid(c), s0)
    rf, ph, _, _ = T.zmarch_rk4(T.NodeGrid(c), s0)
    np.testing.assert_allclose(base[1], rf[1], rtol=1e-9)
    np.testing.assert_allclose(base[0], rf[0], rtol=0, atol=1e-12)


def beam_at_angles(n, extent, theta, seed):
    """n rays from z = -extent across the central 4 mm, all at angles
    (theta, -theta/2) (s0 layout 6 x n)."""
    rs = np.random.RandomState(seed)
    c = T.C
    s0 = np.zeros((6, n))
    s0[0] = rs.uniform(-2e-3, 2e-3, n)
    s0[1] = rs.uniform(-2e-3, 2e-3, n)
    s0[2] = -extent
    v = np.array([np.sin(theta), -0.5 * np.sin(theta), 0.0])
    v[2] = np.sqrt(1 - v[0] ** 2 - v[1] ** 2)
    s0[3:] = c * v[:, None]
    return s0


@pytest.mark.parametrize("k", [8, 16])
def test_block_position_moment_exact_for_z_linear_kick(k):
    """dn/dx growing linearly in z and uniform in x, y: the z-block kick at
    the block centre plus the -s M_x / W position term integrates the cubic
    path exactly, so the block estimate matches the fine RK4 for any block
    size; without the moment terms it is off by the first moment."""
    n = 33
    x = np.linspace(-5e-3, 5e-3, n)
    XX, _, ZZ = np.meshgrid(x, x, x, indexing="ij")
    c = cube_of(1e24 * (1 + 0.5 * XX * (ZZ + 5e-3) / 25e-6), n)
    # a narrow beam about x = 0, where dn/dz (proportional to x) and so the
    # change of W are negligible
    s0 = c.init_beam(40, 1e-4, 0.0, rng=np.random.RandomState(2))
    g = T.NodeGrid(c)
    rf, ph, _, _ = T.zmarch_rk4(g, s0, h=0.125)
    base, _ = S.path_estimate(g, s0, k=k)
    plain, _ = S.path_estimate(g, s0, k=k, moments=False)
    disp = np.abs(rf[0] - s0[0]).max()
    err = np.abs(base[0] - rf[0]).max()
    assert err < 1e-4 * disp
    assert np.abs(plain[0] - rf[0]).max() > 30 * err


def test_block_kick_moment_exact_for_tilted_rays():
    """dn/dx = a x z: along a tilted ray the field is bilinear in the ray's
    x and z, and the centre value G plus v . dM/dx is its exact block
    integral.  What remains at 16-cell blocks is second order (the ray's
    bending and W's change inside a block, which here are comparable with
    the small deflection): within 1.2e-2 of the deflection, against about
    25% without the moment term."""
    n = 33
    x = np.linspace(-5e-3, 5e-3, n)
    XX, _, ZZ = np.meshgrid(x, x, x, indexing="ij")
    c = cube_of(1e24 * (1 + 0.5 * XX ** 2 * ZZ / 1.25e-7), n)
    s0 = beam_at_angles(40, c.extent, 2e-3, 3)
    g = T.NodeGrid(c)
    rf, ph, _, _ = T.zmarch_rk4(g, s0, h=0.125)
    base, _ = S.path_estimate(g, s0, k=16)
    plain, _ = S.path_estimate(g, s0, k=16, moments=False)
    defl = np.abs(rf[1] - np.arctan(s0[3] / s0[5])).max()
    err = np.abs(base[1] - rf[1]).max()
    assert err < 1.2e-2 * defl
    assert np.abs(plain[1] - rf[1]).max() > 10 * err


@pytest.mark.parametrize("k", [1, 4, 8])
def test_path_estimate_close_to_rk4_turbulent(k):
    n = 33
    d = F.powerlaw_field((n, n, n), -11 / 3, seed=3)
    d /= np.sqrt((d ** 2).mean())
    c = cube_of(1e24 * (1 + 0.1 * d), n)
    s0 = c.init_beam(500, 2e-3, 50e-6, rng=np.random.RandomState(1))
    g = T.NodeGrid(c)
    rf, ph, _, _ = T.zmarch_rk4(g, s0)
    base, feat = S.path_estimate(g, s0, k=k)
    defl = np.sqrt(np.mean(rf[[1, 3]] ** 2))
    # one evaluation per k z-cells against eight per cell: the error grows
    # about as k^2 from 1e-2 of the deflection at k = 1
    assert np.sqrt(np.mean((base[[1, 3]] - rf[[1, 3]]) ** 2)) < 1e-2 * max(1, k * k / 4) * defl
    dx = 10e-3 / (n - 1)
    assert np.abs(base[[0, 2]] - rf[[0, 2]]).max() < 1e-3 * max(1, k * k / 4) * dx
    assert np.abs(base[4] - ph).max() < 1e-4 * max(1, k * k / 4) * np.abs(ph).max()


@pytest.mark.parametrize("k", [4, 16])
def test_block_kick_uses_mid_block_w(k):
    """Constant gradients in x and in z: W changes linearly in z (dW/dZ =
    g_z / W), so the exact kick is g_x (W_out - W_in) / g_z, which the block
    kick G_x / W_mid reproduces to second order in W's change when W_mid =
    W + G_z / 2 W (G_z being the whole block's change): within 1e-4 of the
    deflection here.  Taking W's change over half a block as k / 2 times the
    block's change is off by 4e-3 of it at k = 4 and 9e-2 at k = 16."""
    n = 33
    x = np.linspace(-5e-3, 5e-3, n)
    XX, _, ZZ = np.meshgrid(x, x, x, indexing="ij")
    c = cube_of(1e25 * (1 + 0.5 * XX / 5e-3 + 2.0 * ZZ / 5e-3), n)
    s0 = c.init_beam(20, 1e-3, 0.0, rng=np.random.RandomState(4))
    g = T.NodeGrid(c)
    rf, ph, _, _ = T.zmarch_rk4(g, s0, h=0.125)
    base, _ = S.path_estimate(g, s0, k=k)
    defl = np.abs(rf[1]).max()
    err = np.abs(base[1] - rf[1]).max()
    assert err < 1e-4 * defl, err / defl
