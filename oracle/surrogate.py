"""FP64 oracle of the surrogate path: the cheap path estimate and its
features, and the MLP forward/backward/Adam.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).  Nothing here follows a
reference function: the reference has no surrogate.  Pins:
  * ``path_estimate`` -- the one-evaluation-per-z-cell drift-kick-drift
    estimate the GPU kernel forms; pinned by exact cases (null field;
    uniform slab phase; constant gradient: the mid-plane kick integrates the
    parabola exactly) and by its agreement with the FP64 RK4 oracle.
  * ``mlp_forward`` / ``mlp_backward`` -- pinned by central finite
    differences (tests/test_oracle_surrogate.py) and by torch autograd in
    float64 on the CPU.
"""
import numpy as np

from .tracer import C, W_HANDOVER, trilinear

NSEG = 16


def block_fields(grid, k):
    """Straight-line z-block integrals of the node field, FP64, in the GPU's
    per-cell units (G = gamma dz, Phi = k0 dz eta).  Block b covers cells
    [b k, min((b + 1) k, ncell)); with m_c the mid-plane value of cell c (the
    mean of its two node planes, i.e. the exact z-integral of the trilinear
    field over the cell), returns G[b] = sum_c m_c and the exact first moment
    M[b] = int (z - zmid_b) f dz = sum_c (c + 1/2 - zmid_b) m_c
    + (f[b1] - f[b0]) / 12, zmid_b the block's centre, as arrays
    [4, nx, ny, nb]."""
    nx, ny, nz = grid.shape
    dz = grid.spacing[2]
    scale = np.array([dz, dz, dz, grid.k0 * dz])
    f = np.stack([v * s for v, s in zip(grid.fields, scale)])
    mid = 0.5 * (f[..., :-1] + f[..., 1:])
    ncell = nz - 1
    nb = -(-ncell // k)
    G = np.zeros((4, nx, ny, nb))
    M = np.zeros((4, nx, ny, nb))
    for bb in range(nb):
        b0, b1 = bb * k, min(ncell, (bb + 1) * k)
        zc = np.arange(b0, b1) + 0.5 - 0.5 * (b0 + b1)
        G[..., bb] = mid[..., b0:b1].sum(-1)
        # plus each cell's own moment: its field is linear in z, slope
        # f[c + 1] - f[c], so int (z - z_c) f dz = (f[c + 1] - f[c]) / 12,
        # which telescopes over the block
        M[..., bb] = (mid[..., b0:b1] * zc).sum(-1) + (f[..., b1] - f[..., b0]) / 12.0
    return G, M


def _bilinear(A, mx, my, inside):
    """Bilinear value and x / y slopes (per cell) of A [4, nx, ny] at cell
    positions (mx, my); zero outside."""
    nx, ny = A.shape[1:3]
    ix = np.where(inside, np.minimum(np.floor(np.where(inside, mx, 0)).astype(np.int64), nx - 2), 0)
    iy = np.where(inside, np.minimum(np.floor(np.where(inside, my, 0)).astype(np.int64), ny - 2), 0)
    tx, ty = mx - ix, my - iy
    a00, a10, a01, a11 = A[:, ix, iy], A[:, ix + 1, iy], A[:, ix, iy + 1], A[:, ix + 1, iy + 1]
    v = (a00 * (1 - tx) + a10 * tx) * (1 - ty) + (a01 * (1 - tx) + a11 * tx) * ty
    ddx = (a10 - a00) * (1 - ty) + (a11 - a01) * ty
    ddy = (a01 - a00) * (1 - tx) + (a11 - a10) * tx
    z = np.where(inside, 1.0, 0.0)
    return v * z, ddx * z, ddy * z, ix, iy


def path_estimate(grid, s0, k=1, moments=True):
    """Cheap path estimate and the 48 features, FP64.

    The ray is advanced by drift-kick-drift steps over z-blocks of k cells:
    drift half a block along the current direction, kick by the block's
    straight-line field integral G at that position, drift the second half.
    Two first-moment corrections make the straight-line integral follow the
    ray: the field integral gets v . grad(M) (the ray's transverse drift
    inside the block, v in cells per cell) and the position gets
    -s M_x / W (a kick spread over the block rather than at its centre).
    A ray entering inside the grid first takes single-cell steps from the node
    field (the first cell partial) up to the next block boundary.

    ``grid`` is an oracle.tracer.NodeGrid.  Returns base (5 x N: x, theta, y,
    phi, phase) and feat (N x 48): 16 z-segment sums of the x and of the y
    kicks (a step belongs to the segment of its first cell, q = 16 c / ncell),
    then U - U0, V - V0, W - W0, the exit displacement from the entry line
    (x, y, cells), the phase, U0, V0, W0 - 1, the entry-cell fractions x and
    y, the x and y column changes between steps, the summed second-order
    proxies q_x = (dGx/dx Gx + dGx/dy Gy) / W^2 and q_y likewise, and the
    path length in cells.  ``moments=False`` drops both first-moment
    corrections (a check that they matter, for the tests).

    A ray the full trace would hand over to the time integrator, judged on
    the estimate's own path (W <= W_HANDOVER at launch or at the exit, or x / y
    inside the grid at the exit but not at the entry or the reverse), gets a
    NaN base, so the surrogate flags it for the fallback (path.hip
    path_kernel)."""
    nx, ny, nz = grid.shape
    x0, y0, z0 = grid.origin
    dx, dy, dz = grid.spacing
    k0 = grid.k0
    s0 = np.asarray(s0, dtype=np.float64)
    N = s0.shape[1]
    U0, V0, W0 = s0[3] / C, s0[4] / C, s0[5] / C
    sx, sy = dz / dx, dz / dy
    px0 = (s0[0] - x0) / dx
    py0 = (s0[1] - y0) / dy
    Zs = (s0[2] - z0) / dz
    pre = Zs < 0
    px0 = np.where(pre, px0 + sx * U0 / W0 * -Zs, px0)
    py0 = np.where(pre, py0 + sy * V0 / W0 * -Zs, py0)
    Zs = np.where(pre, 0.0, Zs)
    ncell = nz - 1
    G, M = block_fields(grid, k)
    scale = np.array([dz, dz, dz, k0 * dz])
    st = dict(px=px0.copy(), py=py0.copy(), U=U0.copy(), V=V0.copy(), W=W0.copy(), ph=np.zeros(N),
              cx=np.zeros(N), cy=np.zeros(N), qx=np.zeros(N), qy=np.zeros(N), cix=np.full(N, -1),
              ciy=np.full(N, -1))
    seg = np.zeros((N, 2 * NSEG))
    rows = np.arange(N)

    def step(act, c0, w, sample):
        """One drift-kick-drift step of width w (cells) starting at cell c0
        for the rays in act; sample(mx, my, inside) -> (g, mval, q)."""
        U, V, W = st["U"], st["V"], st["W"]
        hm = 0.5 * w
        mx = st["px"] + hm * sx * U / W
        my = st["py"] + hm * sy * V / W
        inside = (mx >= 0) & (mx <= nx - 1) & (my >= 0) & (my <= ny - 1)
        g, mval, qx, qy, ix, iy = sample(mx, my, inside)
        sp = np.sqrt(U * U + V * V + W * W) / W
        iWm = 1.0 / (W + 0.5 * g[2] / W)  # W at the middle: g[2] is the whole step's change
        Un, Vn, Wn = U + g[0] * iWm, V + g[1] * iWm, W + g[2] * iWm
        pxn = mx + hm * sx * Un / Wn - sx * mval[0] / Wn
        pyn = my + hm * sy * Vn / Wn - sy * mval[1] / Wn
        upd = lambda key, val: st.__setitem__(key, np.where(act, val, st[key]))
        first = st["cix"] < 0
        upd("cx", st["cx"] + np.where(~first & (ix != st["cix"]), 1.0, 0.0))
        upd("cy", st["cy"] + np.where(~first & (iy != st["ciy"]), 1.0, 0.0))
        upd("qx", st["qx"] + qx)
        upd("qy", st["qy"] + qy)
        upd("cix", ix)
        upd("ciy", iy)
        upd("ph", st["ph"] + sp * g[3])
        upd("U", Un)
        upd("V", Vn)
        upd("W", Wn)
        upd("px", pxn)
        upd("py", pyn)
        q = (c0 * NSEG) // ncell
        q = np.minimum(np.where(act, q, 0), NSEG - 1)
        np.add.at(seg, (rows[act], q[act]), g[0][act])
        np.add.at(seg, (rows[act], NSEG + q[act]), g[1][act])

    # single-cell steps from the entry cell to the next block boundary
    iz0 = np.where(Zs >= ncell, ncell, np.floor(Zs)).astype(np.int64)
    on_boundary = (Zs == iz0) & (iz0 % k == 0)
    bstart = np.where(on_boundary, iz0, np.minimum(ncell, (iz0 // k + 1) * k))
    for c in range(ncell):
        act = (c >= iz0) & (c < bstart)
        if not act.any():
            continue
        w = np.where(c == iz0, (c + 1) - Zs, 1.0)
        tz = 1.0 - 0.5 * w

        def cell_sample(mx, my, inside, c=c, tz=tz):
            vals = trilinear(grid.fields, grid.origin, grid.spacing, grid.shape, np.where(inside, mx, -1.0),
                             np.where(inside, my, -1.0), c + tz)
            g = np.stack([v * s for v, s in zip(vals, scale)]) * w
            ix = np.where(inside, np.minimum(np.floor(np.where(inside, mx, 0)).astype(np.int64), nx - 2), 0)
            iy = np.where(inside, np.minimum(np.floor(np.where(inside, my, 0)).astype(np.int64), ny - 2), 0)
            return g, np.zeros_like(g), 0.0, 0.0, ix, iy
        step(act, c, w, cell_sample)
    # blocks
    nb = G.shape[-1]
    for bb in range(nb):
        b0, b1 = bb * k, min(ncell, (bb + 1) * k)
        act = bstart <= b0
        if not act.any():
            continue
        w = float(b1 - b0)

        def block_sample(mx, my, inside, bb=bb):
            Gv, Gdx, Gdy, ix, iy = _bilinear(G[..., bb], mx, my, inside)
            Mv, Mdx, Mdy, _, _ = _bilinear(M[..., bb], mx, my, inside)
            U, V, W = st["U"], st["V"], st["W"]
            vx, vy = sx * U / W, sy * V / W
            g = Gv + vx * Mdx + vy * Mdy if moments else Gv
            iW2 = 1.0 / (W * W)
            qx = (Gdx[0] * Gv[0] + Gdy[0] * Gv[1]) * iW2
            qy = (Gdx[1] * Gv[0] + Gdy[1] * Gv[1]) * iW2
            return g, (Mv if moments else 0 * Mv), qx, qy, ix, iy
        step(act, b0, w, block_sample)
    px, py, U, V, W, ph = st["px"], st["py"], st["U"], st["V"], st["W"], st["ph"]
    L = ncell - Zs
    base = np.stack([x0 + px * dx, np.arctan(U / W), y0 + py * dy, np.arctan(V / W), ph])
    fx1, fy1 = nx - 1.0, ny - 1.0
    in0 = (px0 >= 0) & (px0 <= fx1) & (py0 >= 0) & (py0 <= fy1)
    in1 = (px >= 0) & (px <= fx1) & (py >= 0) & (py <= fy1)
    handover = ~(W0 - 1.0 > W_HANDOVER - 1.0) | ~(W - 1.0 > W_HANDOVER - 1.0) | (in0 != in1)
    base[:, handover] = np.nan
    glob = np.stack([U - U0, V - V0, W - W0, px - (px0 + sx * U0 / W0 * L), py - (py0 + sy * V0 / W0 * L), ph,
                     U0, V0, W0 - 1.0, px0 - np.floor(px0), py0 - np.floor(py0), st["cx"], st["cy"], st["qx"],
                     st["qy"], L], axis=1)
    return base, np.concatenate([seg, glob], axis=1)


# --------------------------------------------------------------------- MLP
GELU_K0 = np.sqrt(2.0 / np.pi)
GELU_K1 = 0.044715


def gelu(x):
    """GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(GELU_K0 * (x + GELU_K1 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_K0 * (x + GELU_K1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_K0 * (1.0 + 3.0 * GELU_K1 * x * x)


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
