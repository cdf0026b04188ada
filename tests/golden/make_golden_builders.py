"""Write tests/golden/builder_vectors.npz from the reference itself.

Run once, in a container where the reference repository is mounted at
/root/reference (it never travels to a GPU machine):

    python tests/golden/make_golden_builders.py

It calls the reference's ElectronCube density builders (test_null, test_slab,
test_linear_cos, test_exponential_cos, test_lens, test_liner;
particle_tracking/particle_tracker.py:147-210) on a small non-cubic grid with
non-default parameters, rect_aperture (particle_tracking/
ray_transfer_matrix.py:128-136) on seeded rays, and init_beam + the
as-shipped solve + ray_at_exit for probing along 'y' and 'x'
(PT:258-331, 333-380) on a small seeded cube, and stores inputs and outputs as
.npz data.  These pin the builders' formulas and axis conventions and the
'x' / 'y' branches of init_beam and ray_at_exit, which reference_vectors.npz
(written by make_golden.py) does not cover.
"""
import contextlib
import io
import os

import numpy as np

from make_golden import HERE, _import_reference


def main():
    _, pt, rtm = _import_reference()
    out = {}

    # 1. density builders on a non-cubic grid
    x = np.linspace(-5e-3, 5e-3, 7)
    y = np.linspace(-4e-3, 4e-3, 6)
    z = np.linspace(-3e-3, 3e-3, 5)
    out["b_x"], out["b_y"], out["b_z"] = x, y, z
    cube = pt.ElectronCube(x, y, z)
    calls = {
        "null": ("test_null", {}),
        "slab": ("test_slab", {"s": 3, "n_e0": 5e23}),
        "linear_cos": ("test_linear_cos", {"s1": 0.3, "s2": 0.2, "n_e0": 4e23, "Ly": 3e-3}),
        "exponential_cos": ("test_exponential_cos", {"n_e0": 2e24, "Ly": 2e-3, "s": 4e-3}),
        "lens": ("test_lens", {"n_e0": 3e24, "LR": 2e-3}),
        "liner": ("test_liner", {"n_e0": 3e24, "LR": 2e-3}),
    }
    for key, (fn, kw) in calls.items():
        getattr(cube, fn)(**kw)
        out[f"b_{key}"] = np.asarray(cube.ne, dtype=np.float64)

    # 2. rect_aperture on rays in mm, some outside in x only, y only, both
    rs = np.random.RandomState(41)
    r = np.stack([rs.uniform(-3, 3, 200), rs.normal(0, 1e-3, 200), rs.uniform(-3, 3, 200),
                  rs.normal(0, 1e-3, 200)])
    out["rect_r"] = r.copy()
    out["rect_out"] = rtm.rect_aperture(r.copy(), 2.0, 1.5)

    # 3. probing along y and x: init_beam draws, as-shipped solve, ray_at_exit
    nx, ny, nz = 9, 11, 9
    px = np.linspace(-4e-3, 4e-3, nx)
    py = np.linspace(-5e-3, 5e-3, ny)
    pz = np.linspace(-4e-3, 4e-3, nz)
    ne = 1e24 * (1.0 + 0.1 * rs.randn(nx, ny, nz))
    out["p_x"], out["p_y"], out["p_z"], out["p_ne"] = px, py, pz, ne
    for d, seed in (("y", 31), ("x", 32)):
        c = pt.ElectronCube(px, py, pz, probing_direction=d)
        c.external_ne(ne.copy())
        c.calc_dndr(lwl=1064e-9)
        np.random.seed(seed)
        c.init_beam(Np=32, beam_size=1.5e-3, divergence=5e-3)
        out[f"p_{d}_seed"] = np.array(seed)
        out[f"p_{d}_s0"] = c.s0.copy()
        with contextlib.redirect_stdout(io.StringIO()):
            rf = c.solve()
        out[f"p_{d}_sf"] = c.sf
        out[f"p_{d}_rf"] = rf

    path = os.path.join(HERE, "builder_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
