"""Drop-in for particle_tracking/particle_tracker.py on MI355X.

``ElectronCube`` keeps the reference's constructor, density builders,
``external_ne``, ``calc_dndr``, ``dndr``, ``init_beam``, ``solve``,
``ray_at_exit`` and ``save_output_rays`` (PT:121-395) and the module-level
``dsdt`` (PT:398-419).  Arrays live on the GPU as torch tensors; numpy inputs
are accepted and moved there.

What changes underneath:
  * calc_dndr runs tt_grad_stencil: one float4 node field {-0.5 dz grad(ne/nc),
    k0 dz (n - 1)} replaces the three float64 gradient arrays and
    RegularGridInterpolators (dndx/dndy/dndz/ne_nc stay available as
    attributes, computed on demand).
  * solve runs tt_trace_rk4: fixed-step RK4 with z as the independent
    variable (h = 0.5 cells by default) instead of one adaptive RK45 over all
    rays.  It lands every ray on the exit plane that ray_at_exit projects to,
    and also returns the phase k0 int (n - 1) ds per ray (``self.phase``).
    Rays the z-march hands over (launched backwards or steeper than ~0.25,
    slowed towards a turning point by an overdense region, or crossing a
    side face of the grid) are then traced in time by tt_trace_time with
    the reference's RK45 pair at a tight per-ray tolerance
    (``self.status`` 3).  The ``method`` argument is accepted for signature
    compatibility.
  * init_beam draws from numpy's global RNG exactly as the reference does
    (so seeded runs give the same rays) unless ``seed=`` is given, in which
    case the GPU Philox sampler draws the same distribution on the device.
  * probing_direction 'x' and 'y' are traced in an axis-permuted frame whose
    march axis is the probing axis (the node field of that frame is built by
    the stencil from a strided view of the density), so the exit rays are
    the reference's ray_at_exit rows for that direction (PT:353-368).
"""
import time
from datetime import datetime

import numpy as np
import torch

from . import _lib

c = 299792458.0           # scipy.constants.c (PT:119)
NC_COEFF = 3.14207787e-4  # PT:228


def _dev(a, dtype=torch.float64, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=device)


def _uniform_axis(a, name):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 1 or a.size < 2:
        raise ValueError(f"{name} must be a 1-D coordinate array with >= 2 points")
    d = np.diff(a)
    step = (a[-1] - a[0]) / (a.size - 1)
    if not (step > 0 and np.allclose(d, step, rtol=1e-9, atol=0)):
        raise ValueError(f"{name} must be uniformly spaced and increasing")
    return a[0], step


class ElectronCube:
    """A class to hold and generate electron density cubes (PT:121)."""

    def __init__(self, x, y, z, probing_direction="z", device="cuda"):
        self.z, self.y, self.x = np.asarray(z), np.asarray(y), np.asarray(x)
        self.extent_x = self.x.max()
        self.extent_y = self.y.max()
        self.extent_z = self.z.max()
        self.probing_direction = probing_direction
        self.device = device
        x0, dx = _uniform_axis(self.x, "x")
        y0, dy = _uniform_axis(self.y, "y")
        z0, dz = _uniform_axis(self.z, "z")
        self.grid = _lib.Grid(self.x.size, self.y.size, self.z.size, x0, y0, z0, dx, dy, dz)
        self.lib = _lib.load()
        self.field4 = None

    # ----------------------------------------------------- density builders
    def _mesh(self):
        X = torch.as_tensor(self.x, dtype=torch.float64, device=self.device)
        Y = torch.as_tensor(self.y, dtype=torch.float64, device=self.device)
        Z = torch.as_tensor(self.z, dtype=torch.float64, device=self.device)
        return torch.meshgrid(X, Y, Z, indexing="ij")

    def test_null(self):
        """Null test, an empty cube (PT:147-152)."""
        XX, _, _ = self._mesh()
        self.ne = torch.zeros_like(XX)

    def test_slab(self, s=1, n_e0=2e23):
        """n_e = n_e0 (1 + s x / extent) (PT:154-165)."""
        XX, _, _ = self._mesh()
        self.ne = n_e0 * (1.0 + s * XX / self.extent_x)

    def test_linear_cos(self, s1=0.1, s2=0.1, n_e0=2e23, Ly=1):
        """PT:167-177."""
        XX, YY, _ = self._mesh()
        self.ne = n_e0 * (1.0 + s1 * XX / self.extent_x) * (1 + s2 * torch.cos(2 * np.pi * YY / Ly))

    def test_exponential_cos(self, n_e0=1e24, Ly=1e-3, s=2e-3):
        """PT:179-188."""
        XX, YY, _ = self._mesh()
        self.ne = n_e0 * 10 ** (XX / s) * (1 + torch.cos(2 * np.pi * YY / Ly))

    def test_lens(self, n_e0=1e24, LR=1e-3):
        """PT:190-199."""
        XX, YY, _ = self._mesh()
        RR = torch.sqrt(XX ** 2 + YY ** 2)
        self.ne = n_e0 * torch.exp(-RR ** 2 / LR ** 2)

    def test_liner(self, n_e0=1e24, LR=1e-3):
        """PT:201-210."""
        XX, _, ZZ = self._mesh()
        RR = torch.sqrt(XX ** 2 + ZZ ** 2)
        self.ne = n_e0 * torch.exp(-RR ** 2 / LR ** 2)

    def external_ne(self, ne):
        """Load an externally generated (Mx, My, Mz) density grid, m^-3 (PT:212-218)."""
        if isinstance(ne, torch.Tensor):
            self.ne = ne if ne.device.type == "cuda" else ne.to(self.device)
        else:
            self.ne = torch.as_tensor(np.asarray(ne), device=self.device)

    # ------------------------------------------------------------ gradients
    def calc_dndr(self, lwl=1053e-9, ne_max=1):
        """Gradient field of ne/nc (PT:220-241) as one float4 node field."""
        self.lwl = lwl
        self.omega = 2 * np.pi * (c / lwl)
        self.nc = NC_COEFF * self.omega ** 2
        self.ne_max = ne_max
        self.k0 = 2 * np.pi / lwl
        ne = self.ne
        if tuple(ne.shape) != (self.grid.nx, self.grid.ny, self.grid.nz):
            raise ValueError(f"ne has shape {tuple(ne.shape)}, grid is {(self.grid.nx, self.grid.ny, self.grid.nz)}")
        if ne.dtype not in (torch.float32, torch.float64):
            ne = ne.double()
        dtype = 0 if ne.dtype == torch.float32 else 1
        sx, sy, sz = ne.stride()
        self.field4 = torch.empty((self.grid.nz, self.grid.ny, self.grid.nx, 4), dtype=torch.float32,
                                  device=self.device)
        self.lib.tt_grad_stencil(_lib.ptr(ne), dtype, sx, sy, sz, _lib.C.byref(self.grid), 0.0, 1.0 / self.nc,
                                 float(ne_max), self.k0, None, 0, _lib.ptr(self.field4), _lib.stream_ptr())
        self._ne_ref = ne  # keep alive until the kernel has run

    def _channel_ij(self, k):
        return self.field4[..., k].permute(2, 1, 0)

    @property
    def dndx(self):
        return self._channel_ij(0).double() * (c ** 2 / self.grid.dz)

    @property
    def dndy(self):
        return self._channel_ij(1).double() * (c ** 2 / self.grid.dz)

    @property
    def dndz(self):
        return self._channel_ij(2).double() * (c ** 2 / self.grid.dz)

    @property
    def dndx_interp(self):
        """Callable like the RegularGridInterpolator of PT:239: values of
        dndx at points xi of shape (..., 3), zero outside the grid."""
        return _ChannelInterp(self, 0)

    @property
    def dndy_interp(self):
        """As dndx_interp, for dndy (PT:240)."""
        return _ChannelInterp(self, 1)

    @property
    def dndz_interp(self):
        """As dndx_interp, for dndz (PT:241)."""
        return _ChannelInterp(self, 2)

    @property
    def ne_nc(self):
        r = self.ne.double() / self.nc
        return torch.clamp(r, max=self.ne_max)

    def sample_field(self, pts):
        """Raw field channels at 3 x N points (m): 4 x N float64."""
        pts = _dev(pts, device=self.device).contiguous()
        n = pts.shape[1]
        out = torch.empty((4, n), dtype=torch.float64, device=self.device)
        self.lib.tt_sample_field(_lib.ptr(self.field4), _lib.C.byref(self.grid), _lib.ptr(pts), n, n,
                                 _lib.ptr(out), n, _lib.stream_ptr())
        return out

    def dndr(self, x):
        """Gradient -(c^2/2) grad(ne/nc) at 3 x N locations (PT:243-256)."""
        g = self.sample_field(x)[:3]
        return g * (c ** 2 / self.grid.dz)

    # ----------------------------------------------------------------- rays
    def _extent(self):
        d = self.probing_direction
        if d not in ("x", "y", "z"):
            raise ValueError(f"probing_direction must be 'x', 'y' or 'z', got {d!r}")
        self.extent = {"x": self.extent_x, "y": self.extent_y, "z": self.extent_z}[d]
        return self.extent

    def init_beam(self, Np, beam_size, divergence, seed=None):
        """Rays in a disc of radius beam_size with polar-angle spread
        ``divergence`` (PT:258-310).  seed=None: numpy global RNG, the
        reference's draws; seed=int: the GPU Philox sampler."""
        extent = self._extent()
        axis = {"x": 0, "y": 1, "z": 2}[self.probing_direction]
        Np = int(Np)
        if seed is None:
            t = 2 * np.pi * np.random.rand(Np)
            u = np.random.rand(Np) + np.random.rand(Np)
            u[u > 1] = 2 - u[u > 1]
            phi = np.pi * np.random.rand(Np)
            chi = divergence * np.random.randn(Np)
            s0 = np.zeros((6, Np))
            sc, cc, sp, cp = np.sin(chi), np.cos(chi), np.sin(phi), np.cos(phi)
            pa, pb = beam_size * u * np.cos(t), beam_size * u * np.sin(t)
            if axis == 0:
                s0[3], s0[4], s0[5] = c * cc, c * sc * cp, c * sc * sp
                s0[0], s0[1], s0[2] = extent, pa, pb
            elif axis == 1:
                s0[4], s0[3], s0[5] = c * cc, c * sc * cp, c * sc * sp
                s0[0], s0[1], s0[2] = pa, -extent, pb
            else:
                s0[3], s0[4], s0[5] = c * sc * cp, c * sc * sp, c * cc
                s0[0], s0[1], s0[2] = pa, pb, -extent
            self.s0 = torch.as_tensor(s0, device=self.device)
        else:
            self.s0 = torch.empty((6, Np), dtype=torch.float64, device=self.device)
            self.lib.tt_beam_sample(Np, float(beam_size), float(divergence), float(extent), axis, int(seed), 0,
                                    _lib.ptr(self.s0), Np, _lib.stream_ptr())
        return self.s0

    def _sort_order(self, s0, grid=None):
        grid = grid or self.grid
        n = s0.shape[1]
        nbytes = _lib.C.c_size_t()
        self.lib.tt_ray_sort_scratch_bytes(_lib.C.byref(grid), 0, _lib.C.byref(nbytes))
        scratch = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        order = torch.empty(n, dtype=torch.int32, device=self.device)
        self.lib.tt_ray_sort(_lib.ptr(s0), s0.stride(0), n, _lib.C.byref(grid), 0, _lib.ptr(scratch),
                             _lib.ptr(order), _lib.stream_ptr())
        return order

    # tracer frame per probing direction: tracer axis a' is cube axis PERM[a'],
    # so the probing axis becomes the march axis z' and the tracer's exit
    # rays [x', theta', y', phi'] are the reference's ray_at_exit layout
    # (PT:353-378): 'x' -> (y, z, x), 'y' -> (x, z, y)
    PERM = {"x": (1, 2, 0), "y": (0, 2, 1), "z": (0, 1, 2)}

    def _frame(self):
        """(field4, grid) of the tracer frame of the probing direction; for
        'x' / 'y' the node field is rebuilt from the axis-permuted density
        (a strided view, no copy) and kept until calc_dndr runs again."""
        d = self.probing_direction
        if d == "z":
            return self.field4, self.grid
        cached = getattr(self, "_probe_frame", None)
        if cached is not None and cached[0] == d and cached[1] is self.field4:
            return cached[2], cached[3]
        p = self.PERM[d]
        g = self.grid
        n = (g.nx, g.ny, g.nz)
        o = (g.x0, g.y0, g.z0)
        sp = (g.dx, g.dy, g.dz)
        grid = _lib.Grid(n[p[0]], n[p[1]], n[p[2]], o[p[0]], o[p[1]], o[p[2]], sp[p[0]], sp[p[1]], sp[p[2]])
        ne = self._ne_ref.permute(*p)
        field4 = torch.empty((grid.nz, grid.ny, grid.nx, 4), dtype=torch.float32, device=self.device)
        sx, sy, sz = ne.stride()
        self.lib.tt_grad_stencil(_lib.ptr(ne), 0 if ne.dtype == torch.float32 else 1, sx, sy, sz, _lib.C.byref(grid),
                                 0.0, 1.0 / self.nc, float(self.ne_max), self.k0, None, 0, _lib.ptr(field4),
                                 _lib.stream_ptr())
        self._probe_frame = (d, self.field4, field4, grid)
        return field4, grid

    def solve(self, method="RK45", h=0.5, sort=True, verbose=True, variant=0):
        """Trace all rays through the cube and project them to the exit plane
        (PT:312-331).  Returns rf (4 x N: the reference's ray_at_exit rows).
        Probing along 'x' or 'y' traces in an axis-permuted frame whose march
        axis is the probing axis; as in the reference, 'x' rays start at
        x = +extent (PT:280-289) and so leave the cube at once."""
        if self.field4 is None:
            raise RuntimeError("call calc_dndr() before solve()")
        extent = self._extent()
        field4, grid = self._frame()
        p = self.PERM[self.probing_direction]
        s0 = _dev(self.s0, device=self.device)
        if self.probing_direction != "z":
            s0 = s0[[p[0], p[1], p[2], 3 + p[0], 3 + p[1], 3 + p[2]]]
        s0 = s0.contiguous()
        n = s0.shape[1]
        self.rf = torch.empty((4, n), dtype=torch.float64, device=self.device)
        sf = torch.empty((6, n), dtype=torch.float64, device=self.device)
        self.phase = torch.empty(n, dtype=torch.float64, device=self.device)
        self.status = torch.empty(n, dtype=torch.int32, device=self.device)
        opts = _lib.TraceOpts(float(h), float(np.sqrt(8.0) * extent), int(variant))
        start = time.time()
        order = self._sort_order(s0, grid) if (sort and n > 0) else None
        self.lib.tt_trace_rk4(_lib.ptr(field4), _lib.C.byref(grid), _lib.C.byref(opts), _lib.ptr(s0), n, n,
                              _lib.ptr(order), _lib.ptr(self.rf), n, _lib.ptr(self.phase), _lib.ptr(sf), n,
                              _lib.ptr(self.status), _lib.stream_ptr())
        # rays the z-march hands over (status 1) are traced in time, as the
        # reference integrates every ray (status 3)
        self.lib.tt_trace_time(_lib.ptr(field4), _lib.C.byref(grid), _lib.C.byref(opts), _lib.ptr(s0), n, n,
                               _lib.ptr(order), _lib.ptr(self.rf), n, _lib.ptr(self.phase), _lib.ptr(sf), n,
                               _lib.ptr(self.status), 1, _lib.stream_ptr())
        if self.probing_direction != "z":  # back to the cube's axes
            inv = [p.index(a) for a in range(3)]
            sf = sf[[inv[0], inv[1], inv[2], 3 + inv[0], 3 + inv[1], 3 + inv[2]]]
        self.sf = sf
        torch.cuda.current_stream().synchronize()
        if verbose:
            print("Ray trace completed in:\t", time.time() - start, "s")
        return self.rf

    def ray_at_exit(self):
        """4D rays from the 6D state sf (PT:333-380)."""
        x, y, z, vx, vy, vz = self.sf
        ray_p = torch.zeros((4, self.sf.shape[1]), dtype=torch.float64, device=self.sf.device)
        d = self.probing_direction
        if d == "x":
            t_bp = (x - self.extent) / vx
            ray_p[0], ray_p[2] = y - vy * t_bp, z - vz * t_bp
            ray_p[1], ray_p[3] = torch.atan(vy / vx), torch.atan(vz / vx)
        elif d == "y":
            t_bp = (y - self.extent) / vy
            ray_p[0], ray_p[2] = x - vx * t_bp, z - vz * t_bp
            ray_p[1], ray_p[3] = torch.atan(vx / vy), torch.atan(vz / vy)
        elif d == "z":
            t_bp = (z - self.extent) / vz
            ray_p[0], ray_p[2] = x - vx * t_bp, y - vy * t_bp
            ray_p[1], ray_p[3] = torch.atan(vx / vz), torch.atan(vy / vz)
        return ray_p

    def save_output_rays(self, fn=None):
        """rf as .npy, named by date and time unless fn is given (PT:382-395)."""
        if fn is None:
            fn = "{} rays.npy".format(datetime.now().strftime("%Y-%m-%d_%H-%M-%S"))
        else:
            fn = "{}.npy".format(fn)
        with open(fn, "wb") as f:
            np.save(f, self.rf.cpu().numpy())
        return fn


class _ChannelInterp:
    """One gradient channel of the node field as a callable with the
    RegularGridInterpolator call convention of PT:239-241 and PT:253-255:
    interp(xi) with xi of shape (..., 3) (metres) returns an array of shape
    (...) of dndx / dndy / dndz (m/s^2), linear inside the grid, zero
    strictly outside it and NaN at NaN points (tt_sample_field)."""

    def __init__(self, cube, channel):
        self.cube = cube
        self.channel = channel

    def __call__(self, xi):
        xi = _dev(xi, device=self.cube.device)
        if xi.shape[-1] != 3:
            raise ValueError(f"points must have shape (..., 3), got {tuple(xi.shape)}")
        lead = xi.shape[:-1]
        pts = xi.reshape(-1, 3).T.contiguous()
        g = self.cube.sample_field(pts)[self.channel]
        return (g * (c ** 2 / self.cube.grid.dz)).reshape(lead)


def dsdt(t, s, ElectronCube):
    """ODE right-hand side for a flattened 6N state (PT:398-419)."""
    s = _dev(s, device=ElectronCube.device)
    Np = s.numel() // 6
    s = s.reshape(6, Np)
    sprime = torch.zeros_like(s)
    sprime[3:6, :] = ElectronCube.dndr(s[:3, :])
    sprime[:3, :] = s[3:, :]
    return sprime.flatten()
