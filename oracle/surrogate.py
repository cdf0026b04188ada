"""FP64 oracle of the surrogate path: straight-line (Born) estimate and
features, and the MLP forward/backward/Adam.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).  Nothing here follows a
reference function: the reference has no surrogate.  Pins:
  * ``straight_estimate`` -- the same midpoint-per-z-cell sums the GPU kernel
    forms; pinned by exactness cases (null field; uniform slab: first order
    is exact for the phase; linear slab: Born-1 deflection is exact) and
    checked against the FP64 RK4 oracle for its order of accuracy.
  * ``mlp_forward`` / ``mlp_backward`` -- pinned by central finite
    differences (tests/test_oracle_surrogate.py) and by torch autograd in
    float64 on the CPU.
"""
import numpy as np
from scipy.special import erf

from .tracer import C, trilinear

NSEG = 16


def _eval_d(grid, px, py, iz, tz):
    """Value of the 4 channels and d/dx, d/dy of channels x, y, w (per cell)."""
    nx, ny, nz = grid.shape
    inside = (px >= 0) & (px <= nx - 1) & (py >= 0) & (py <= ny - 1)
    fx = np.where(inside, px, 0.0)
    fy = np.where(inside, py, 0.0)
    ix = np.minimum(np.floor(fx).astype(np.int64), nx - 2)
    iy = np.minimum(np.floor(fy).astype(np.int64), ny - 2)
    tx, ty = fx - ix, fy - iy
    out = []
    dgx, dgy = [], []
    for ch, v in enumerate(grid.fields):
        c = {}
        for a in (0, 1):
            for b in (0, 1):
                q0 = v[ix + a, iy + b, iz]
                q1 = v[ix + a, iy + b, iz + 1]
                c[a, b] = q0 + tz * (q1 - q0)
        r0 = c[0, 0] + tx * (c[1, 0] - c[0, 0])
        r1 = c[0, 1] + tx * (c[1, 1] - c[0, 1])
        out.append(np.where(inside, r0 + ty * (r1 - r0), 0.0))
        if ch in (0, 1, 3):
            d_x = (c[1, 0] - c[0, 0]) + ty * ((c[1, 1] - c[0, 1]) - (c[1, 0] - c[0, 0]))
            dgx.append(np.where(inside, d_x, 0.0))
            dgy.append(np.where(inside, r1 - r0, 0.0))
    return out, dgx, dgy


def straight_estimate(grid, s0):
    """Second-order straight-line estimate and the 48 features, FP64.

    ``grid`` is an oracle.tracer.NodeGrid (fields in the oracle's units:
    gamma in 1/m and eta); the per-z-cell sums use G = gamma dz and
    Phi = k0 dz eta like the GPU node field.  Returns base (5 x N: x, theta,
    y, phi, phase) and feat (N x 48)."""
    nx, ny, nz = grid.shape
    x0, y0, z0 = grid.origin
    dx, dy, dz = grid.spacing
    k0 = grid.k0
    s0 = np.asarray(s0, dtype=np.float64)
    N = s0.shape[1]
    U0, V0, W0 = s0[3] / C, s0[4] / C, s0[5] / C
    sx, sy = dz / dx, dz / dy
    slx, sly = sx * U0 / W0, sy * V0 / W0
    px0 = (s0[0] - x0) / dx
    py0 = (s0[1] - y0) / dy
    Zs = (s0[2] - z0) / dz
    pre = Zs < 0
    px0 = np.where(pre, px0 + slx * -Zs, px0)
    py0 = np.where(pre, py0 + sly * -Zs, py0)
    Zs = np.where(pre, 0.0, Zs)
    Zend = float(nz - 1)
    iW0 = 1.0 / W0
    sp = np.sqrt(U0 ** 2 + V0 ** 2 + W0 ** 2) / W0
    fsx, fsy = sx * iW0 ** 2, sy * iW0 ** 2
    scale = np.array([dz, dz, dz, k0 * dz])
    Sx, Sy, Sz, Mx, My, P = (np.zeros(N) for _ in range(6))
    Tx, Ty, Rx, Ry, Dx, Dy = (np.zeros(N) for _ in range(6))
    seg = np.zeros((N, 2 * NSEG))
    ncell = nz - 1
    for iz in range(ncell):
        act = Zs < iz + 1
        zl = np.maximum(Zs, iz)
        w = np.where(act, (iz + 1) - zl, 0.0)
        zm = zl + 0.5 * w
        px = px0 + slx * (zm - Zs)
        py = py0 + sly * (zm - Zs)
        vals, dgx, dgy = _eval_d(grid, px, py, iz, zm - iz)
        g = [v * s for v, s in zip(vals, scale)]
        dgx = [d * s for d, s in zip(dgx, (dz, dz, k0 * dz))]
        dgy = [d * s for d, s in zip(dgy, (dz, dz, k0 * dz))]
        h = 0.5 * w
        d1x = fsx * (Dx + Sx * h + 0.5 * g[0] * h * h)
        d1y = fsy * (Dy + Sy * h + 0.5 * g[1] * h * h)
        f2x = dgx[0] * d1x + dgy[0] * d1y
        f2y = dgx[1] * d1x + dgy[1] * d1y
        lever = Zend - zm
        Tx += w * f2x
        Ty += w * f2y
        Rx += w * lever * f2x
        Ry += w * lever * f2y
        Mx += w * lever * g[0]
        My += w * lever * g[1]
        Dx += Sx * w + 0.5 * g[0] * w * w
        Dy += Sy * w + 0.5 * g[1] * w * w
        Sx += w * g[0]
        Sy += w * g[1]
        Sz += w * g[2]
        P += w * sp * (g[3] + dgx[2] * d1x + dgy[2] * d1y)
        q = (iz * NSEG) // ncell
        seg[:, q] += w * g[0]
        seg[:, NSEG + q] += w * g[1]
    Ue = U0 + (Sx + Tx) * iW0
    Ve = V0 + (Sy + Ty) * iW0
    We = W0 + Sz * iW0
    L = Zend - Zs
    pxe = px0 + slx * L + fsx * (Mx + Rx)
    pye = py0 + sly * L + fsy * (My + Ry)
    base = np.stack([x0 + pxe * dx, np.arctan(Ue / We), y0 + pye * dy, np.arctan(Ve / We), P])
    glob = np.stack([Sx, Sy, Mx, My, Tx, Ty, Rx, Ry, Sz, P, U0, V0, W0 - 1.0,
                     px0 - np.floor(px0), py0 - np.floor(py0), Zend - Zs], axis=1)
    return base, np.concatenate([seg, glob], axis=1)


# --------------------------------------------------------------------- MLP
def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def mlp_forward(params, X):
    """params: list of (W [out, in], b [out]); GELU after every layer but the
    last.  Returns the output and the cache for backward."""
    H = X
    cache = []
    for li, (W, b) in enumerate(params):
        Z = H @ W.T + b
        cache.append((H, Z))
        H = gelu(Z) if li < len(params) - 1 else Z
    return H, cache


def mlp_backward(params, cache, dY):
    """Gradients (dW, db) per layer for the upstream gradient dY."""
    grads = [None] * len(params)
    d = dY
    for li in range(len(params) - 1, -1, -1):
        W, b = params[li]
        H, Z = cache[li]
        dZ = d if li == len(params) - 1 else d * gelu_grad(Z)
        grads[li] = (dZ.T @ H, dZ.sum(axis=0))
        d = dZ @ W
    return grads


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """Kingma & Ba (2015), Algorithm 1, with L2 weight decay added to g."""
    g = g + wd * w
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
