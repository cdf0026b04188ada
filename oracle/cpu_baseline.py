"""CPU baseline harness: the reference's CPU path, timed on host cores.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``): used by bench.py's
cpu_baseline leg.

Pattern of particle_tracking/example_multiprocess.py:41-62 (a Pool of P
workers, each solving its own bundle of rays) against the real API of the
library: each worker builds the three RegularGridInterpolators of calc_dndr
(PT:239-241) over the shared float64 gradient arrays and integrates the
reference's dsdt (PT:398-419), then ray_at_exit (PT:333-380).  Two modes:

  * "as_shipped": one solve_ivp RK45 over the whole bundle at scipy's default
    tolerances (PT:312-331, rtol 1e-3 on one error norm over all rays);
  * "stated_tolerance": one solve_ivp RK45 per ray at rtol = RTOL_STATED,
    a per-ray tolerance whose exit angles stay within 1e-3 of the RMS
    deflection of a tight (rtol 1e-11) solve at 512^3: rtol 1e-5 leaves
    8.7e-3 of it and 1e-6 leaves 1.0-1.3e-3 (on two samples), so 5e-7.

The gradient arrays live in POSIX shared memory, written by the parent, so
workers (forked before the parent touches the GPU) attach by name.
"""
import os
import time
from multiprocessing import get_context, shared_memory

import numpy as np

RTOL_STATED = 5e-7
MAX_PROCESSES = 16  # the host-core share of one GPU on the measuring machine

_G = {}


def _worker_init():
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[v] = "1"


def _attach(name, shape):
    if _G.get("name") == name:
        return _G["cube"]
    from . import tracer as T
    shm = shared_memory.SharedMemory(name=name)
    try:  # the parent owns the segment; keep this process's tracker off it
        from multiprocessing import resource_tracker
        resource_tracker.unregister(shm._name, "shared_memory")
    except Exception:
        pass
    arr = np.ndarray((3,) + tuple(shape[:3]), dtype=np.float64, buffer=shm.buf)
    axes = shape[3]
    cube = T.Cube(axes[0], axes[1], axes[2])
    from scipy.interpolate import RegularGridInterpolator
    g = (cube.x, cube.y, cube.z)
    cube.dndx_interp = RegularGridInterpolator(g, arr[0], bounds_error=False, fill_value=0.0)
    cube.dndy_interp = RegularGridInterpolator(g, arr[1], bounds_error=False, fill_value=0.0)
    cube.dndz_interp = RegularGridInterpolator(g, arr[2], bounds_error=False, fill_value=0.0)
    cube.extent = cube.extent_z
    _G.update(name=name, shm=shm, cube=cube)
    return cube


def _solve_bundle(args):
    """As shipped: the reference's solve over the bundle."""
    name, shape, s0 = args
    cube = _attach(name, shape)
    cube.s0 = s0
    t = time.perf_counter()
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        rf = cube.solve()
    return rf, time.perf_counter() - t, int(cube.sol.nfev) * s0.shape[1]


def _solve_per_ray(args):
    """One RK45 per ray at RTOL_STATED on the reference's dsdt."""
    from scipy.integrate import solve_ivp

    from . import tracer as T
    name, shape, s0 = args
    cube = _attach(name, shape)
    Tend = np.sqrt(8.0) * cube.extent / T.C
    sf = np.zeros_like(s0)
    nfev = 0
    t = time.perf_counter()
    for i in range(s0.shape[1]):
        sol = solve_ivp(lambda tt, y: T.dsdt(tt, y, cube), [0, Tend], s0[:, i], method="RK45", rtol=RTOL_STATED)
        sf[:, i] = sol.y[:, -1]
        nfev += sol.nfev
    rf = T.ray_at_exit(sf, cube.extent, "z")
    return rf, time.perf_counter() - t, nfev


_MODES = {"as_shipped": _solve_bundle, "stated_tolerance": _solve_per_ray}


class CPUBaseline:
    """Create before the process initialises the GPU (the pool forks)."""

    def __init__(self, processes=None):
        self.cpu_count = os.cpu_count() or 1
        try:
            self.affinity = len(os.sched_getaffinity(0))
        except AttributeError:
            self.affinity = self.cpu_count
        self.P = processes or max(1, min(MAX_PROCESSES, self.affinity))
        self.pool = get_context("fork").Pool(self.P, initializer=_worker_init)
        self.shm = None

    def load_gradients(self, dndx, dndy, dndz, x, y, z):
        """float64 (nx, ny, nz) gradient arrays in m/s^2 (PT:235-237)."""
        shape = dndx.shape
        nbytes = 3 * dndx.size * 8
        self.shm = shared_memory.SharedMemory(create=True, size=nbytes)
        arr = np.ndarray((3,) + shape, dtype=np.float64, buffer=self.shm.buf)
        arr[0], arr[1], arr[2] = dndx, dndy, dndz
        self.shape = tuple(shape) + ((np.asarray(x), np.asarray(y), np.asarray(z)),)

    def run(self, s0, bundles, mode="as_shipped"):
        """Solve the rays s0 (6 x N) split into ``bundles`` bundles over the
        pool.  Returns rf (4 x N), wall seconds, total right-hand-side
        evaluations per ray summed over the rays (ray-RHS evaluations)."""
        fn = _MODES[mode]
        parts = np.array_split(s0, bundles, axis=1)
        t = time.perf_counter()
        out = self.pool.map(fn, [(self.shm.name, self.shape, p) for p in parts], chunksize=1)
        wall = time.perf_counter() - t
        rf = np.concatenate([o[0] for o in out], axis=1)
        return rf, wall, sum(o[2] for o in out)

    def run_timed(self, s0, bundle, budget_s, mode="as_shipped", log=None):
        """Rounds of P bundles of ``bundle`` rays (one per process) until the
        wall time reaches ``budget_s`` or the rays run out.  Returns
        (rf of the rays done, rays done, wall seconds, ray-RHS evaluations)."""
        done, nfev, wall = 0, 0, 0.0
        N = s0.shape[1]
        rfs = []
        r = 0
        while done + self.P * bundle <= N and wall < budget_s:
            chunk = s0[:, done:done + self.P * bundle]
            rf, w, nf = self.run(chunk, self.P, mode)
            rfs.append(rf)
            done += chunk.shape[1]
            wall += w
            nfev += nf
            r += 1
            if log:
                log(f"cpu baseline ({mode}) round {r}: {chunk.shape[1]} rays in {w:.2f} s")
        rf = np.concatenate(rfs, axis=1) if rfs else np.zeros((4, 0))
        return rf, done, wall, nfev

    def close(self):
        self.pool.close()
        self.pool.join()
        if self.shm is not None:
            self.shm.close()
            self.shm.unlink()
            self.shm = None
