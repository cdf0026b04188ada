"""FP64 restatement of the ray tracer (ElectronCube) and its integrators.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

Reference: particle_tracking/particle_tracker.py (PT below).
  * ``Cube``                 PT:121-145, 212-218 (construction, external_ne)
  * ``Cube.test_*``          PT:147-210 (density builders)
  * ``Cube.calc_dndr``       PT:220-241 (omega, n_c, clip, np.gradient, RGIs)
  * ``Cube.dndr``            PT:243-256
  * ``Cube.init_beam``       PT:258-310 (draw order t, u, u, phi, chi)
  * ``Cube.solve``           PT:312-331 (one global RK45 solve_ivp, as shipped)
  * ``ray_at_exit``          PT:333-380
  * ``dsdt``                 PT:398-419
Integrators that are not in the reference (they define the accuracy bar):
  * ``solve_tight``          per-ray RK45 at tight tolerance with a 7th
                             component, the phase phi = k0 * int (n - 1) ds.
  * ``zmarch_rk4``           the fixed-step RK4 the GPU tracer runs, in FP64:
                             z is the independent variable, steps of h cells
                             land exactly on grid planes and on the exit plane.
  * ``FieldCube``            a Cube over given node fields (a window of a
                             device node field) for ``solve_tight``.
Pins: tests/golden/reference_vectors.npz and builder_vectors.npz (written
from the reference's own functions), the null / slab / uniform-slab closed
forms, and solve_tight against the z-march (tests/test_oracle_golden.py).
"""
import numpy as np
import scipy.constants as sc
from scipy.integrate import solve_ivp
from scipy.interpolate import RegularGridInterpolator

C = sc.c                      # PT:119
NC_COEFF = 3.14207787e-4      # PT:228, = eps0 m_e / e^2
# The z-march hands a ray over to the time integrator below this W = vz / c
# (a slope above ~0.25 or a ray slowing towards a turning point): beyond it
# the fixed half-cell z-step misses 1e-3 cells (the GPU compares W - 1 with
# -0.03 in float32).
W_HANDOVER = 0.97
# largest stage-to-stage relative change of W that the GPU's cell-cached step
# takes (beyond it the step is redone by gathers, as a face crossing)
NEWTON_TOL = 2.0 ** -12


def critical_density(lwl):
    """n_c for vacuum wavelength lwl (PT:227-228)."""
    omega = 2 * np.pi * (C / lwl)
    return NC_COEFF * omega ** 2


def wavenumber(lwl):
    return 2 * np.pi / lwl


class Cube:
    """Restatement of ElectronCube (PT:121)."""

    def __init__(self, x, y, z, probing_direction="z"):
        self.z, self.y, self.x = z, y, x
        self.extent_x = x.max()
        self.extent_y = y.max()
        self.extent_z = z.max()
        self.probing_direction = probing_direction

    def external_ne(self, ne):
        self.ne = ne

    # density builders, PT:147-210 (meshgrid 'ij', extents from the axes)
    def _mesh(self):
        return np.meshgrid(self.x, self.y, self.z, indexing="ij")

    def test_null(self):
        XX, _, _ = self._mesh()
        self.ne = np.zeros_like(XX)

    def test_slab(self, s=1, n_e0=2e23):
        XX, _, _ = self._mesh()
        self.ne = n_e0 * (1.0 + s * XX / self.extent_x)

    def test_linear_cos(self, s1=0.1, s2=0.1, n_e0=2e23, Ly=1):
        XX, YY, _ = self._mesh()
        self.ne = n_e0 * (1.0 + s1 * XX / self.extent_x) * (1 + s2 * np.cos(2 * np.pi * YY / Ly))

    def test_exponential_cos(self, n_e0=1e24, Ly=1e-3, s=2e-3):
        XX, YY, _ = self._mesh()
        self.ne = n_e0 * 10 ** (XX / s) * (1 + np.cos(2 * np.pi * YY / Ly))

    def test_lens(self, n_e0=1e24, LR=1e-3):
        XX, YY, _ = self._mesh()
        RR = np.sqrt(XX ** 2 + YY ** 2)
        self.ne = n_e0 * np.exp(-RR ** 2 / LR ** 2)

    def test_liner(self, n_e0=1e24, LR=1e-3):
        XX, _, ZZ = self._mesh()
        RR = np.sqrt(XX ** 2 + ZZ ** 2)
        self.ne = n_e0 * np.exp(-RR ** 2 / LR ** 2)

    def calc_dndr(self, lwl=1053e-9, ne_max=1):
        self.lwl = lwl
        self.omega = 2 * np.pi * (C / lwl)
        nc = NC_COEFF * self.omega ** 2
        ne_nc = self.ne / nc
        ne_nc[ne_nc > ne_max] = ne_max
        self.ne_nc = ne_nc
        self.dndx = -0.5 * C ** 2 * np.gradient(self.ne_nc, self.x, axis=0)
        self.dndy = -0.5 * C ** 2 * np.gradient(self.ne_nc, self.y, axis=1)
        self.dndz = -0.5 * C ** 2 * np.gradient(self.ne_nc, self.z, axis=2)
        g = (self.x, self.y, self.z)
        self.dndx_interp = RegularGridInterpolator(g, self.dndx, bounds_error=False, fill_value=0.0)
        self.dndy_interp = RegularGridInterpolator(g, self.dndy, bounds_error=False, fill_value=0.0)
        self.dndz_interp = RegularGridInterpolator(g, self.dndz, bounds_error=False, fill_value=0.0)

    def dndr(self, x):
        grad = np.zeros_like(x)
        grad[0, :] = self.dndx_interp(x.T)
        grad[1, :] = self.dndy_interp(x.T)
        grad[2, :] = self.dndz_interp(x.T)
        return grad

    def init_beam(self, Np, beam_size, divergence, rng=None):
        rnd = np.random if rng is None else rng
        s0 = np.zeros((6, Np))
        t = 2 * np.pi * rnd.rand(Np)
        u = rnd.rand(Np) + rnd.rand(Np)
        u[u > 1] = 2 - u[u > 1]
        phi = np.pi * rnd.rand(Np)
        chi = divergence * rnd.randn(Np)
        self.s0 = beam_from_draws(t, u, phi, chi, beam_size, self._extent_for_beam(), self.probing_direction)
        return self.s0

    def _extent_for_beam(self):
        d = self.probing_direction
        self.extent = {"x": self.extent_x, "y": self.extent_y, "z": self.extent_z}[d]
        return self.extent

    def solve(self, method="RK45", **kw):
        """As shipped: one solve_ivp over the flattened 6N system (PT:312-331)."""
        t = np.linspace(0.0, np.sqrt(8.0) * self.extent / C, 2)
        s0 = self.s0.flatten()
        sol = solve_ivp(lambda t, y: dsdt(t, y, self), [0, t[-1]], s0, t_eval=t, method=method, **kw)
        Np = s0.size // 6
        self.sol = sol
        self.sf = sol.y[:, -1].reshape(6, Np)
        self.rf = ray_at_exit(self.sf, self.extent, self.probing_direction)
        return self.rf

    # --- node fields used by the fixed-step integrators ---------------------
    def node_fields(self):
        """gamma = -0.5 grad(ne/nc) (1/m, i.e. dndr / c^2) and eta = n - 1 on
        the grid nodes.  eta uses the clipped ne/nc of calc_dndr."""
        gx = self.dndx / C ** 2
        gy = self.dndy / C ** 2
        gz = self.dndz / C ** 2
        r = self.ne_nc
        eta = -r / (1.0 + np.sqrt(np.maximum(1.0 - r, 0.0)))
        return gx, gy, gz, eta


def beam_from_draws(t, u, phi, chi, beam_size, extent, probing_direction="z"):
    """PT:280-309: the ray state from the four draws."""
    Np = t.size
    s0 = np.zeros((6, Np))
    if probing_direction == "x":
        s0[3] = C * np.cos(chi)
        s0[4] = C * np.sin(chi) * np.cos(phi)
        s0[5] = C * np.sin(chi) * np.sin(phi)
        s0[0] = extent
        s0[1] = beam_size * u * np.cos(t)
        s0[2] = beam_size * u * np.sin(t)
    elif probing_direction == "y":
        s0[4] = C * np.cos(chi)
        s0[3] = C * np.sin(chi) * np.cos(phi)
        s0[5] = C * np.sin(chi) * np.sin(phi)
        s0[0] = beam_size * u * np.cos(t)
        s0[1] = -extent
        s0[2] = beam_size * u * np.sin(t)
    elif probing_direction == "z":
        s0[3] = C * np.sin(chi) * np.cos(phi)
        s0[4] = C * np.sin(chi) * np.sin(phi)
        s0[5] = C * np.cos(chi)
        s0[0] = beam_size * u * np.cos(t)
        s0[1] = beam_size * u * np.sin(t)
        s0[2] = -extent
    return s0


def ray_at_exit(sf, extent, probing_direction="z"):
    """PT:333-380."""
    Np = sf.shape[1]
    ray_p = np.zeros((4, Np))
    x, y, z, vx, vy, vz = sf
    with np.errstate(divide="ignore", invalid="ignore"):
        if probing_direction == "x":
            t_bp = (x - extent) / vx
            ray_p[0] = y - vy * t_bp
            ray_p[2] = z - vz * t_bp
            ray_p[1] = np.arctan(vy / vx)
            ray_p[3] = np.arctan(vz / vx)
        elif probing_direction == "y":
            t_bp = (y - extent) / vy
            ray_p[0] = x - vx * t_bp
            ray_p[2] = z - vz * t_bp
            ray_p[1] = np.arctan(vx / vy)
            ray_p[3] = np.arctan(vz / vy)
        elif probing_direction == "z":
            t_bp = (z - extent) / vz
            ray_p[0] = x - vx * t_bp
            ray_p[2] = y - vy * t_bp
            ray_p[1] = np.arctan(vx / vz)
            ray_p[3] = np.arctan(vy / vz)
    return ray_p


def dsdt(t, s, cube):
    """PT:398-419."""
    Np = s.size // 6
    s = s.reshape(6, Np)
    sprime = np.zeros_like(s)
    v = s[3:, :]
    x = s[:3, :]
    sprime[3:6, :] = cube.dndr(x)
    sprime[:3, :] = v
    return sprime.flatten()


# ----------------------------------------------------------------------------
# Trilinear interpolation with RegularGridInterpolator semantics on a uniform
# grid (linear method, fill_value 0 strictly outside, NaN in -> NaN out).
# ----------------------------------------------------------------------------
def trilinear(values, origin, spacing, shape, px, py, pz):
    """``values`` is a tuple of (nx,ny,nz) arrays; (px,py,pz) are positions in
    cell units (0 .. n-1).  Returns a tuple of interpolated arrays."""
    nx, ny, nz = shape
    out_of_bounds = (px < 0) | (px > nx - 1) | (py < 0) | (py > ny - 1) | (pz < 0) | (pz > nz - 1)
    nan = np.isnan(px) | np.isnan(py) | np.isnan(pz)
    fx = np.where(nan | out_of_bounds, 0.0, px)
    fy = np.where(nan | out_of_bounds, 0.0, py)
    fz = np.where(nan | out_of_bounds, 0.0, pz)
    ix = np.clip(np.floor(fx).astype(np.int64), 0, nx - 2)
    iy = np.clip(np.floor(fy).astype(np.int64), 0, ny - 2)
    iz = np.clip(np.floor(fz).astype(np.int64), 0, nz - 2)
    tx, ty, tz = fx - ix, fy - iy, fz - iz
    res = []
    for v in values:
        acc = 0.0
        for ax in (0, 1):
            wx = tx if ax else 1 - tx
            for ay in (0, 1):
                wy = ty if ay else 1 - ty
                for az in (0, 1):
                    wz = tz if az else 1 - tz
                    acc = acc + v[ix + ax, iy + ay, iz + az] * (wx * wy * wz)
        acc = np.where(out_of_bounds, 0.0, acc)
        acc = np.where(nan, np.nan, acc)
        res.append(acc)
    return res


class NodeGrid:
    """Uniform-grid description plus the 4 node fields the GPU stores
    (gamma_x, gamma_y, gamma_z, eta)."""

    def __init__(self, cube, k0=None):
        self.x, self.y, self.z = cube.x, cube.y, cube.z
        self.shape = (len(cube.x), len(cube.y), len(cube.z))
        self.origin = (cube.x[0], cube.y[0], cube.z[0])
        self.spacing = tuple((a[-1] - a[0]) / (len(a) - 1) for a in (cube.x, cube.y, cube.z))
        self.fields = cube.node_fields()
        self.k0 = wavenumber(cube.lwl) if k0 is None else k0
        self.extent = cube.extent_z

    def sample(self, px, py, pz):
        return trilinear(self.fields, self.origin, self.spacing, self.shape, px, py, pz)


def zmarch_rk4(grid, s0, h=0.5, cT=None):
    """Fixed-step RK4 with z (in cells) as the independent variable.

    State per ray: px, py (cells), U, V, W (= v / c), tau (= c t, metres) and
    the phase phi = k0 int eta ds with ds/dZ = dz |v| / vz.  Steps of ``h``
    cells (h = 1/m) run from the entry plane to the last grid plane
    Z = nz - 1, so every step stays inside one z-cell and the exit lands
    exactly on the plane that ray_at_exit back-projects to.  A ray that
    starts before the grid is carried straight to Z = 0 (the field is zero
    outside, PT:239-241); a ray starting between lattice points takes one
    partial first step.

    The phase is summed with the RK4 weights of each step, except for
    h = 1/2, where it follows the GPU's cell-cached march: a cell [c, c + 1]
    whose two steps each keep all four stage points inside the x / y column
    cell they start in (and W within NEWTON_TOL from stage to stage) takes
    Simpson's rule on the ray's points, (f(c) + 4 f(c + 1/2) + f(c + 1)) / 6;
    a step whose stage points cross a face (or whose W changes more) takes the
    RK4 weights, and the other step of such a cell Simpson's
    rule on its half cell with the midpoint integrand taken a quarter cell
    back along the ray from that step's end (or, for the cell's first step,
    from its second step's start).  A partial first step and a leading odd
    step keep the RK4 weights.

    Hand-over (status 1, outputs NaN, as the GPU's; the GPU then traces such
    rays in time): W = vz / c - 1 <= W_HANDOVER - 1 at launch or at the end
    of any step (steep rays, and rays slowing towards a turning point), or a
    change of the ray's x / y inside-the-grid state between the lattice start
    and any step end (a side face crossed: the field jumps to zero there).

    Returns rf (4 x N: x, theta, y, phi_angle), phase (N), sf (6 x N state at
    c t = cT, straight continuation past the exit plane), status (N; 0 ok,
    1 handed over, 2 NaN input)."""
    nx, ny, nz = grid.shape
    x0, y0, z0 = grid.origin
    dx, dy, dz = grid.spacing
    k0 = grid.k0
    s0 = np.asarray(s0, dtype=np.float64)
    N = s0.shape[1]
    m = int(round(1.0 / h))
    assert abs(m * h - 1.0) < 1e-12, "h must be 1/m"
    sx, sy = dz / dx, dz / dy

    px = (s0[0] - x0) / dx
    py = (s0[1] - y0) / dy
    Zs = (s0[2] - z0) / dz
    U, V, W = s0[3] / C, s0[4] / C, s0[5] / C
    tau = np.zeros(N)
    phase = np.zeros(N)
    status = np.zeros(N, dtype=np.int32)
    status[np.isnan(s0).any(axis=0)] = 2
    status[(status == 0) & ~(W - 1.0 > W_HANDOVER - 1.0)] = 1
    ok = status == 0

    # straight carry to the first plane
    pre = ok & (Zs < 0)
    dZ = np.where(pre, -Zs, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        px = np.where(pre, px + sx * U / W * dZ, px)
        py = np.where(pre, py + sy * V / W * dZ, py)
        tau = np.where(pre, tau + dz * dZ / W, tau)
    Zs = np.where(pre, 0.0, Zs)

    Zend = float(nz - 1)
    K = (nz - 1) * m
    y = [px, py, U, V, W, tau]

    def rhs(Z, y):
        px, py, U, V, W, _ = y
        gx, gy, gz, e = grid.sample(px, py, Z)
        iW = 1.0 / W
        sp = np.sqrt(U * U + V * V + W * W) * iW
        return [sx * U * iW, sy * V * iW, dz * gx * iW, dz * gy * iW, dz * gz * iW, dz * iW], k0 * dz * e * sp

    def rk4(Z, y, hh, act):
        k1, p1 = rhs(Z, y)
        y2 = [a + 0.5 * hh * b for a, b in zip(y, k1)]
        k2, p2 = rhs(Z + 0.5 * hh, y2)
        y3 = [a + 0.5 * hh * b for a, b in zip(y, k2)]
        k3, p3 = rhs(Z + 0.5 * hh, y3)
        y4 = [a + hh * b for a, b in zip(y, k3)]
        k4, p4 = rhs(Z + hh, y4)
        yn = [a + hh / 6.0 * (b + 2 * c + 2 * d + e) for a, b, c, d, e in zip(y, k1, k2, k3, k4)]
        dp = hh / 6.0 * (p1 + 2 * p2 + 2 * p3 + p4)
        yn = [np.where(act, a, b) for a, b in zip(yn, y)]
        return yn, np.where(act, dp, 0.0), np.where(act, p1, 0.0)

    with np.errstate(divide="ignore", invalid="ignore"):
        # lattice start index per ray; a partial first step if off-lattice
        kf = Zs * m
        kstart = np.ceil(kf - 1e-9).astype(np.int64)
        part = ok & (kstart * h - Zs > 1e-12) & (Zs < Zend)
        if part.any():
            hh = np.where(part, kstart * h - Zs, 0.0)
            y, dp, _ = rk4(Zs, y, hh, part)
            phase += dp
        def lateral(yy):
            return (yy[0] >= 0) & (yy[0] <= nx - 1) & (yy[1] >= 0) & (yy[1] <= ny - 1)

        in0 = lateral(y)
        cellwise = m == 2
        kfirst = kstart + (kstart % 2) if cellwise else np.full(N, K + 1)
        pend = np.zeros(N)   # weight owed to the next step's start-point integrand
        phi0 = np.zeros(N)   # the cell's first step's start-point integrand
        s1 = np.zeros(N, dtype=bool)

        def column_cell(yy):
            return np.floor(yy[0]), np.floor(yy[1])

        def phi_back(Z, yy, back):
            """integrand at plane Z, a quarter cell back along the ray from yy"""
            px, py, U, V, W, tau = yy
            return rhs(Z, [px - back * sx * U / W, py - back * sy * V / W, U, V, W, tau])[1]

        for k in range(K):
            act = ok & (kstart <= k)
            if not act.any():
                continue
            ys0 = y
            y, dp, p1 = rk4(k * h, ys0, h, act)
            # hand-over: W at or below W_HANDOVER, or a side face crossed
            turned = act & (~(y[4] - 1.0 > W_HANDOVER - 1.0) | (lateral(y) != in0))
            if turned.any():
                status[turned] = 1
                ok = ok & ~turned
                act = act & ~turned
                y = [np.where(turned, a_, b_) for a_, b_ in zip(ys0, y)]
            if not cellwise:
                phase += dp
                continue
            # stage points of this step, for the smooth (no face crossed) test
            k1, _ = rhs(k * h, ys0)
            y2 = [a_ + 0.5 * h * b_ for a_, b_ in zip(ys0, k1)]
            k2, _ = rhs(k * h + 0.5 * h, y2)
            y3 = [a_ + 0.5 * h * b_ for a_, b_ in zip(ys0, k2)]
            k3, _ = rhs(k * h + 0.5 * h, y3)
            y4 = [a_ + h * b_ for a_, b_ in zip(ys0, k3)]
            cx, cy = column_cell(ys0)
            smooth = np.ones(N, dtype=bool)
            for yy in (y2, y3, y4):
                ex, ey = column_cell(yy)
                smooth &= (ex == cx) & (ey == cy)
            # the GPU refines 1/W from stage to stage by one Newton step and
            # redoes a step exactly (as a face crossing) when W changed by
            # more than 2^-12 between two stages
            for ya, yb in ((ys0, y2), (y2, y3), (y3, y4)):
                smooth &= np.abs(1.0 - yb[4] / ya[4]) <= NEWTON_TOL
            lead = act & (k < kfirst)          # a leading odd step: RK4
            phase += np.where(lead, dp, 0.0)
            cw = act & (k >= kfirst)
            if k % 2 == 0:
                add = pend * p1 + np.where(smooth, 0.0, dp)
                phase += np.where(cw, add, 0.0)
                pend = np.where(cw, 0.0, pend)
                s1 = np.where(cw, smooth, s1)
                phi0 = np.where(cw, p1, phi0)
            else:
                Zh = k * h
                pa = phi_back(Zh - 0.25, ys0, 0.25)   # the first step's midpoint
                pb = phi_back(Zh + 0.25, y, 0.25)    # this step's midpoint
                add = np.where(s1 & smooth, (phi0 + 4 * p1) / 6.0,
                               np.where(s1, (phi0 + 4 * pa + p1) / 12.0 + dp,
                                        np.where(smooth, (p1 + 4 * pb) / 12.0, dp)))
                phase += np.where(cw, add, 0.0)
                pend = np.where(cw, np.where(s1 & smooth, 1.0 / 6.0, np.where(~s1 & smooth, 1.0 / 12.0, 0.0)), pend)
        if cellwise:  # the exit point's integrand, owed by the last step
            _, p_end = rhs(Zend, y)
            phase += np.where(ok & (kfirst < K), pend * p_end, 0.0)

    px, py, U, V, W, tau = y
    xe = x0 + px * dx
    ye = y0 + py * dy
    ze = np.full(N, z0 + Zend * dz)
    theta = np.arctan(U / W)
    phia = np.arctan(V / W)
    rf = np.stack([xe, theta, ye, phia])
    if cT is None:
        cT = np.sqrt(8.0) * grid.extent
    rem = cT - tau
    sf = np.stack([xe + U * rem, ye + V * rem, ze + W * rem, C * U, C * V, C * W])
    bad = status != 0
    rf[:, bad] = np.nan
    sf[:, bad] = np.nan
    phase[bad] = np.nan
    return rf, phase, sf, status


class FieldCube(Cube):
    """A Cube whose node fields are given directly (gamma_x, gamma_y, gamma_z
    in 1/m, i.e. dndr / c^2, and eta = n - 1 on the nodes of the x, y, z
    axes), e.g. a window of a device node field; calc_dndr is not run.  Its
    dndr and interpolators are the RegularGridInterpolators of PT:239-241 over
    c^2 gamma."""

    def __init__(self, x, y, z, gx, gy, gz, eta, lwl, probing_direction="z"):
        super().__init__(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64),
                         np.asarray(z, dtype=np.float64), probing_direction)
        self.lwl = lwl
        self.dndx, self.dndy, self.dndz = gx * C ** 2, gy * C ** 2, gz * C ** 2
        self._eta = eta
        g = (self.x, self.y, self.z)
        self.dndx_interp = RegularGridInterpolator(g, self.dndx, bounds_error=False, fill_value=0.0)
        self.dndy_interp = RegularGridInterpolator(g, self.dndy, bounds_error=False, fill_value=0.0)
        self.dndz_interp = RegularGridInterpolator(g, self.dndz, bounds_error=False, fill_value=0.0)
        self.extent = self._extent_for_beam()

    def node_fields(self):
        return self.dndx / C ** 2, self.dndy / C ** 2, self.dndz / C ** 2, self._eta


def solve_tight(cube, s0, k0=None, rtol=1e-11, atol_pos=1e-14, atol_vel=1e-6, atol_phase=1e-10,
                max_step=np.inf, bundle=False):
    """RK45 at tight tolerance on the reference's own RHS (dsdt: the
    gradient RegularGridInterpolators of calc_dndr, PT:239-241, 398-419),
    with the phase as a 7th component d phi/dt = k0 * eta(x) * |v|.
    Integrates to t = sqrt(8) extent / c as PT:317 does, then ray_at_exit.

    bundle=False: one solve_ivp per ray.  bundle=True: one solve_ivp over
    all rays together, as the reference's solve does (PT:323), but at the
    tight tolerance; the error norm then holds every ray to it, and since
    all rays cross the z-planes (where the trilinear field has its kinks) at
    nearly the same times, the bundle costs about as much as a few single
    rays.  Returns rf (4 x N), phase (N), sf (6 x N), nfev."""
    if k0 is None:
        k0 = wavenumber(cube.lwl)
    gx, gy, gz, eta = cube.node_fields()
    eta_i = RegularGridInterpolator((cube.x, cube.y, cube.z), eta, bounds_error=False, fill_value=0.0)
    T = np.sqrt(8.0) * cube.extent / C
    N = s0.shape[1]
    sf = np.zeros((6, N))
    phase = np.zeros(N)
    nfev = 0

    def f(t, y):
        n = y.size // 7
        out = np.empty(7 * n)
        p = y[:3 * n].reshape(3, n)
        v = y[3 * n:6 * n].reshape(3, n)
        out[:3 * n] = y[3 * n:6 * n]
        out[3 * n:6 * n] = cube.dndr(p).ravel()
        out[6 * n:] = k0 * eta_i(p.T) * np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
        return out

    groups = [np.arange(N)] if bundle else [np.array([i]) for i in range(N)]
    for idx in groups:
        n = idx.size
        atol = np.repeat(np.array([atol_pos] * 3 + [atol_vel] * 3 + [atol_phase]), n)
        y0 = np.concatenate([s0[:, idx].ravel(), np.zeros(n)])
        sol = solve_ivp(f, [0, T], y0, method="RK45", rtol=rtol, atol=atol, max_step=max_step)
        yT = sol.y[:, -1]
        sf[:, idx] = yT[:6 * n].reshape(6, n)
        phase[idx] = yT[6 * n:]
        nfev += sol.nfev
    rf = ray_at_exit(sf, cube.extent, cube.probing_direction)
    return rf, phase, sf, nfev
