"""Write the small golden fixtures in tests/golden/ from the reference itself.

Run once, in a container where the reference repository is mounted at
/root/reference (it never travels to a GPU machine):

    python tests/golden/make_golden.py

It imports the reference's library modules (turboGen, particle_tracker,
ray_transfer_matrix), calls them on tiny seeded inputs and stores inputs and
outputs as .npz data.  These vectors pin what the closed-form tests cannot:
the reference's draw order, 'xy' meshgrid, flip/shift/zero-mode composition,
np.gradient/RegularGridInterpolator composition, the as-shipped global RK45
trajectory and the histogram transpose and edges.
"""
import contextlib
import io
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference():
    sys.path.insert(0, os.path.join(REF, "gaussian_fields"))
    sys.path.insert(0, os.path.join(REF, "particle_tracking"))
    import matplotlib
    matplotlib.use("Agg")
    import turboGen
    import particle_tracker
    import ray_transfer_matrix
    return turboGen, particle_tracker, ray_transfer_matrix


def kfunc_band(M):
    def k_func(K):
        with np.errstate(divide="ignore"):
            P = K ** (-11.0 / 3.0)
        P[(K < 1.0 / M) | (K > 0.5)] = 0.0
        return P
    return k_func


def main():
    tg, pt, rtm = _import_reference()
    out = {}

    # 1. gaussian3D_FFT at N = 2, 4 (M = 5, 9), seeded global RNG
    for N, seed in [(2, 11), (4, 12)]:
        np.random.seed(seed)
        sig = tg.gaussian3D_FFT(N, kfunc_band(2 * N + 1))
        out[f"field_N{N}_seed"] = np.array(seed)
        out[f"field_N{N}"] = sig
    np.random.seed(13)
    out["field_N3_pow53"] = tg.gaussian3D_FFT(3, lambda K: K ** (-5.0 / 3.0))
    out["field_N3_pow53_seed"] = np.array(13)

    # 2. a small cube built from a seeded turbulent field; calc_dndr, dndr,
    #    init_beam, solve (as shipped), ray_at_exit
    n = 11
    ext = 5e-3
    x = np.linspace(-ext, ext, n)
    rs = np.random.RandomState(5)
    ne = 1e24 * (1.0 + 0.1 * rs.randn(n, n, n))
    cube = pt.ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    out["cube_x"] = x
    out["cube_ne"] = ne
    out["cube_dndx"] = cube.dndx
    out["cube_dndy"] = cube.dndy
    out["cube_dndz"] = cube.dndz
    out["cube_ne_nc"] = cube.ne_nc
    pts = np.stack([rs.uniform(-1.2 * ext, 1.2 * ext, 64) for _ in range(3)])
    pts[:, 0] = [ext, ext, ext]             # upper corner (inside, inclusive)
    pts[:, 1] = [-ext, -ext, -ext]          # lower corner
    pts[:, 2] = [ext * (1 + 1e-12), 0, 0]   # just outside in x
    out["dndr_pts"] = pts
    out["dndr_vals"] = cube.dndr(pts)
    np.random.seed(21)
    cube.init_beam(Np=64, beam_size=2e-3, divergence=5e-3)
    out["beam_seed"] = np.array(21)
    out["beam_s0"] = cube.s0
    with contextlib.redirect_stdout(io.StringIO()):
        rf = cube.solve()
    out["solve_sf"] = cube.sf
    out["solve_rf"] = rf

    # 3. slab known answer (NB:163): test_slab(s=8, n_e0=1e25) on a coarse grid
    ns = 21
    xs = np.linspace(-ext, ext, ns)
    slab = pt.ElectronCube(xs, xs.copy(), xs.copy())
    slab.test_slab(s=8, n_e0=1e25)
    slab.calc_dndr()
    np.random.seed(22)
    slab.init_beam(Np=16, beam_size=1e-3, divergence=0.0)
    with contextlib.redirect_stdout(io.StringIO()):
        out["slab_rf"] = slab.solve()
    out["slab_s0"] = slab.s0
    out["slab_x"] = xs

    # 4. ray-transfer optics and histograms on deterministic rays (metres/rad)
    rs = np.random.RandomState(7)
    Nr = 1000
    r0 = np.stack([rs.uniform(-4e-3, 4e-3, Nr), rs.normal(0, 4e-3, Nr),
                   rs.uniform(-4e-3, 4e-3, Nr), rs.normal(0, 4e-3, Nr)])
    r0[:, 0] = [3e-3, 0.0, -3e-3, 0.0]     # lands exactly on the last x edge
    r0[:, 1] = [np.nan, 0, 0, 0]           # NaN ray is dropped
    out["rtm_r0"] = r0
    for name, cls, kw in [("sh", rtm.Shadowgraphy, {}), ("df", rtm.Schlieren_DF, {"R": 1}),
                          ("lf", rtm.Schlieren_LF, {"R": 1}),
                          ("afr", rtm.AFR, {"Rs": np.array([0.5, 1.0, 2.0, 3.0])})]:
        d = cls(r0.copy(), focal_plane=1.0, L=400, R=25, Lx=6, Ly=6)
        d.solve(**kw)
        d.histogram(bin_scale=1, pix_x=64, pix_y=48)
        out[f"rtm_{name}_rf"] = d.rf
        out[f"rtm_{name}_H"] = d.H
        out[f"rtm_{name}_xedges"] = d.xedges
        out[f"rtm_{name}_yedges"] = d.yedges
    ke = rtm.knife_edge(rtm.m_to_mm(r0.copy()), 0.5, "x", 1)
    out["rtm_knife_x_pos"] = ke
    ke = rtm.knife_edge(rtm.m_to_mm(r0.copy()), -0.25, "y", -1)
    out["rtm_knife_y_neg"] = ke

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
