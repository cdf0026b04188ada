"""The FP64 oracle against the vectors written from the reference itself
(tests/golden/make_golden.py) and against closed forms."""
import numpy as np
import pytest

from oracle import field as F
from oracle import philox
from oracle import rtm as R
from oracle import tracer as T


# ---------------------------------------------------------------- philox
def test_philox_known_answers():
    # Salmon et al. SC'11 / Random123 known-answer vectors for philox4x32-10
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in cases:
        got = tuple(int(v) for v in philox.philox4x32(*c, *k))
        assert got == want


def test_philox_normals_statistics():
    wr, wi = philox.field_noise(np.arange(200000, dtype=np.uint64), seed=1234)
    for w in (wr, wi):
        assert abs(w.mean()) < 0.01 and abs(w.std() - 1) < 0.01
    assert abs(np.corrcoef(wr, wi)[0, 1]) < 0.01


# ---------------------------------------------------------------- field
@pytest.mark.parametrize("N", [2, 4])
def test_field_matches_reference_golden(golden, N):
    M = 2 * N + 1

    def k_func(K):
        with np.errstate(divide="ignore"):
            P = K ** (-11.0 / 3.0)
        P[(K < 1.0 / M) | (K > 0.5)] = 0.0
        return P

    rng = np.random.RandomState(int(golden[f"field_N{N}_seed"]))
    got = F.gaussian3D_FFT(N, k_func, rng=rng)
    np.testing.assert_allclose(got, golden[f"field_N{N}"], rtol=0, atol=1e-13 * np.abs(got).max())


def test_field_pow53_golden(golden):
    rng = np.random.RandomState(int(golden["field_N3_pow53_seed"]))
    with np.errstate(divide="ignore"):
        got = F.gaussian3D_FFT(3, lambda K: K ** (-5.0 / 3.0), rng=rng)
    np.testing.assert_allclose(got, golden["field_N3_pow53"], rtol=0, atol=1e-13 * np.abs(got).max())


def test_general_form_reduces_to_flip_for_odd():
    for n in (5, 9, 33):
        np.testing.assert_array_equal(F.partner_index(n), np.arange(n)[::-1])


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 10, 4), (9, 8, 7)])
def test_general_field_is_hermitian_real(shape):
    rs = np.random.RandomState(0)
    Wr, Wi = rs.randn(*shape), rs.randn(*shape)
    S = F.sqrt_powerlaw(shape, -11 / 3)
    spec = F.shaped_spectrum(Wr, Wi, S)
    full = np.fft.ifftn(spec)
    assert np.abs(full.imag).max() < 1e-14 * np.abs(full.real).max()
    assert abs(full.real.mean()) < 1e-14 * np.abs(full.real).max()


def test_single_mode_closed_form():
    # P = delta at frequency (1, 2, 3)/n: the field is a pure cosine whose
    # amplitude and phase are fixed by the noise at that mode.
    n = 8
    shape = (n, n, n)
    rs = np.random.RandomState(3)
    Wr, Wi = rs.randn(*shape), rs.randn(*shape)
    j = (1, 2, 3)
    c = tuple((jj + n // 2) % n for jj in j)          # centred index
    p = tuple((2 * (n // 2) - cc) % n for cc in c)    # partner
    S = np.zeros(shape)
    S[c] = 1.0
    S[p] = 1.0
    f = F.gaussian_field(Wr, Wi, S)
    a = Wr[c] + Wr[p]
    b = Wi[c] - Wi[p]
    ii = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    arg = 2 * np.pi * sum(jj * i for jj, i in zip(j, ii)) / n
    want = 2.0 / n ** 3 * (a * np.cos(arg) - b * np.sin(arg))
    np.testing.assert_allclose(f, want, atol=1e-15)


def test_radial_psd_slope():
    f = F.powerlaw_field((64, 64, 64), -11 / 3, seed=5)
    k, P, cnt = F.radial_psd(f, nbins=32)
    sel = (k > 0.05) & (k < 0.4)
    slope = np.polyfit(np.log(k[sel]), np.log(P[sel]), 1)[0]
    assert abs(slope - (-11 / 3)) < 0.15


# ---------------------------------------------------------------- tracer
@pytest.fixture(scope="module")
def ref_cube(golden):
    x = golden["cube_x"]
    cube = T.Cube(x, x.copy(), x.copy())
    cube.external_ne(golden["cube_ne"].copy())
    cube.calc_dndr(lwl=1064e-9)
    return cube


def test_calc_dndr_golden(golden, ref_cube):
    for k in ("dndx", "dndy", "dndz", "ne_nc"):
        np.testing.assert_allclose(getattr(ref_cube, k), golden["cube_" + k], rtol=1e-14, atol=0)


def test_dndr_golden(golden, ref_cube):
    got = ref_cube.dndr(golden["dndr_pts"])
    np.testing.assert_allclose(got, golden["dndr_vals"], rtol=1e-13, atol=0)
    assert np.all(got[:, 2] == 0)  # just outside in x


def test_own_trilinear_matches_rgi(golden, ref_cube):
    g = T.NodeGrid(ref_cube)
    pts = golden["dndr_pts"]
    px = (pts[0] - g.origin[0]) / g.spacing[0]
    py = (pts[1] - g.origin[1]) / g.spacing[1]
    pz = (pts[2] - g.origin[2]) / g.spacing[2]
    gx, gy, gz, _ = g.sample(px, py, pz)
    want = golden["dndr_vals"] / T.C ** 2
    np.testing.assert_allclose(np.stack([gx, gy, gz]), want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_init_beam_golden(golden, ref_cube):
    rng = np.random.RandomState(int(golden["beam_seed"]))
    s0 = ref_cube.init_beam(64, 2e-3, 5e-3, rng=rng)
    np.testing.assert_array_equal(s0, golden["beam_s0"])


def test_solve_as_shipped_golden(golden, ref_cube):
    ref_cube.s0 = golden["beam_s0"].copy()
    ref_cube.extent = ref_cube.extent_z
    rf = ref_cube.solve()
    np.testing.assert_allclose(ref_cube.sf, golden["solve_sf"], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(rf, golden["solve_rf"], rtol=1e-10, atol=1e-15)


def test_zmarch_agrees_with_reference_solve(golden, ref_cube):
    # the as-shipped global RK45 is only accurate to its rtol=1e-3; the
    # z-march RK4 must sit inside that band around it
    g = T.NodeGrid(ref_cube)
    rf, ph, sf, st = T.zmarch_rk4(g, golden["beam_s0"], h=0.5)
    assert (st == 0).all()
    ref = golden["solve_rf"]
    dx = g.spacing[0]
    assert np.abs(rf[[0, 2]] - ref[[0, 2]]).max() < 1e-2 * dx
    ang = np.abs(ref[[1, 3]]).max()
    assert np.abs(rf[[1, 3]] - ref[[1, 3]]).max() < 1e-2 * ang


def test_slab_golden_and_closed_form(golden):
    xs = golden["slab_x"]
    cube = T.Cube(xs, xs.copy(), xs.copy())
    XX = np.meshgrid(xs, xs, xs, indexing="ij")[0]
    cube.external_ne(1e25 * (1.0 + 8 * XX / xs.max()))  # test_slab, PT:154-165
    cube.calc_dndr()
    g = T.NodeGrid(cube)
    rf, ph, sf, st = T.zmarch_rk4(g, golden["slab_s0"], h=0.5)
    # closed form: theta_x = -1/2 d(ne/nc)/dx * L_cube for a ray along z
    nc = T.critical_density(1053e-9)
    grad = 1e25 * 8 / xs.max() / nc
    L = xs[-1] - xs[0]
    slope = -0.5 * grad * L        # vx / vz at exit (vz stays c)
    theta = np.arctan(slope)
    np.testing.assert_allclose(rf[1], theta, rtol=1e-12)
    # x_exit = x0 + 1/2 slope L  (parabola in z)
    np.testing.assert_allclose(rf[0] - golden["slab_s0"][0], 0.5 * slope * L, rtol=1e-10)
    # the as-shipped global RK45 restatement reproduces the reference's own
    # trace of the same rays (which is 5% off the closed form in angle: its
    # default rtol steps across the exit-plane discontinuity)
    cube.s0 = golden["slab_s0"].copy()
    cube.extent = cube.extent_z
    np.testing.assert_allclose(cube.solve(), golden["slab_rf"], rtol=1e-9, atol=1e-15)


def test_null_straight_lines():
    x = np.linspace(-5e-3, 5e-3, 9)
    cube = T.Cube(x, x.copy(), x.copy())
    cube.external_ne(np.zeros((9, 9, 9)))
    cube.calc_dndr()
    s0 = cube.init_beam(50, 2e-3, 5e-3, rng=np.random.RandomState(0))
    rf, ph, sf, st = T.zmarch_rk4(T.NodeGrid(cube), s0)
    tx = s0[3] / s0[5]
    ty = s0[4] / s0[5]
    np.testing.assert_allclose(rf[0], s0[0] + tx * 10e-3, atol=1e-15)
    np.testing.assert_allclose(rf[1], np.arctan(tx), atol=1e-15)
    np.testing.assert_allclose(rf[2], s0[1] + ty * 10e-3, atol=1e-15)
    assert np.all(ph == 0)


def test_uniform_slab_phase_closed_form():
    x = np.linspace(-5e-3, 5e-3, 9)
    cube = T.Cube(x, x.copy(), x.copy())
    cube.external_ne(np.full((9, 9, 9), 3e25))
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(20, 2e-3, 0.0, rng=np.random.RandomState(0))
    rf, ph, sf, st = T.zmarch_rk4(T.NodeGrid(cube), s0)
    r = 3e25 / T.critical_density(1064e-9)
    want = T.wavenumber(1064e-9) * (np.sqrt(1 - r) - 1) * 10e-3
    np.testing.assert_allclose(ph, want, rtol=1e-12)


def test_zmarch_matches_tight_solver_turbulent():
    n = 13
    x = np.linspace(-5e-3, 5e-3, n)
    d = F.powerlaw_field((n, n, n), -11 / 3, seed=2)
    d /= np.sqrt((d ** 2).mean())
    cube = T.Cube(x, x.copy(), x.copy())
    cube.external_ne(1e24 * (1 + 0.1 * d))
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(4, 2e-3, 50e-6, rng=np.random.RandomState(1))
    rf, ph, sf, st = T.zmarch_rk4(T.NodeGrid(cube), s0)
    rt, pt_, sft, _ = T.solve_tight(cube, s0)
    dx = x[1] - x[0]
    assert np.abs(rf[[0, 2]] - rt[[0, 2]]).max() < 1e-8 * dx
    assert np.abs(rf[[1, 3]] - rt[[1, 3]]).max() < 1e-9
    assert np.abs(ph - pt_).max() < 1e-6
    np.testing.assert_allclose(sf[:2], sft[:2], atol=1e-12)


def tilted_beam(n_rays, lo, hi, theta, seed):
    """Rays from z = -extent with launch angles theta (rad, x and y) plus a
    small jitter, positions uniform in [lo, hi]^2."""
    rs = np.random.RandomState(seed)
    s0 = np.zeros((6, n_rays))
    s0[0] = rs.uniform(lo, hi, n_rays)
    s0[1] = rs.uniform(lo, hi, n_rays)
    s0[2] = -5e-3
    tx = theta[0] + rs.normal(0, 0.01, n_rays)
    ty = theta[1] + rs.normal(0, 0.01, n_rays)
    vz = 1.0 / np.sqrt(1 + np.tan(tx) ** 2 + np.tan(ty) ** 2)
    s0[3], s0[4], s0[5] = T.C * vz * np.tan(tx), T.C * vz * np.tan(ty), T.C * vz
    return s0


def turbulent_cube(n, eps, seed=2):
    x = np.linspace(-5e-3, 5e-3, n)
    d = F.powerlaw_field((n, n, n), -11 / 3, seed=seed)
    d /= np.sqrt((d ** 2).mean())
    cube = T.Cube(x, x.copy(), x.copy())
    cube.external_ne(1e24 * (1 + eps * d))
    cube.calc_dndr(lwl=1064e-9)
    cube.extent = cube.extent_z
    return x, cube


def test_tight_bundle_equals_per_ray():
    """One tight solve_ivp over a bundle of rays gives every ray the
    per-ray tight answer."""
    x, cube = turbulent_cube(13, 0.1)
    s0 = tilted_beam(5, -3e-3, 3e-3, (0.05, -0.04), seed=3)
    r1, p1, s1, _ = T.solve_tight(cube, s0)
    r2, p2, s2, _ = T.solve_tight(cube, s0, bundle=True)
    dx = x[1] - x[0]
    assert np.abs(r1[[0, 2]] - r2[[0, 2]]).max() < 1e-9 * dx
    # both are tight to ~1e-10 rad in angle (rtol 1e-11 on |v| = c per step,
    # summed over the steps) and ~1e-8 relative in phase, far inside the
    # 1e-3 of the deflection and 1e-5 of the phase the z-march is held to
    assert np.abs(r1[[1, 3]] - r2[[1, 3]]).max() < 2e-9
    assert np.abs(p1 - p2).max() < 1e-7 * np.abs(p1).max()


def test_zmarch_matches_tight_solver_across_cell_faces():
    """The fixed-step z-march (h = 1/2 cell, what the GPU runs) against tight
    solve_ivp on a 33^3 turbulent cube (20% RMS) with rays tilted ~0.15 rad
    so that every ray crosses several x and y cell faces, where the
    trilinear field has kinks and RK4 loses its order.  The bar of the path:
    exit positions within 1e-3 cells, exit angles within 1e-3 of the RMS
    deflection (exit minus launch angle) and the phase within 1e-5 of its
    magnitude."""
    n = 33
    x, cube = turbulent_cube(n, 0.2)
    dx = x[1] - x[0]
    s0 = tilted_beam(8, -3e-3, 3e-3, (0.15, -0.12), seed=4)
    rf, ph, sf, st = T.zmarch_rk4(T.NodeGrid(cube), s0)
    rt, pt_, sft, _ = T.solve_tight(cube, s0, bundle=True)
    assert (st == 0).all()
    # every ray crossed at least 3 faces in x and 3 in y
    cx = np.abs(np.floor((rt[0] - x[0]) / dx) - np.floor((s0[0] - x[0]) / dx))
    cy = np.abs(np.floor((rt[2] - x[0]) / dx) - np.floor((s0[1] - x[0]) / dx))
    assert cx.min() >= 3 and cy.min() >= 3
    launch = np.arctan(s0[[3, 4]] / s0[5])
    defl = np.sqrt(np.mean((rt[[1, 3]] - launch) ** 2))
    assert defl > 1e-4
    assert np.abs(rf[[0, 2]] - rt[[0, 2]]).max() <= 1e-3 * dx
    assert np.abs(rf[[1, 3]] - rt[[1, 3]]).max() <= 1e-3 * defl
    assert np.abs(ph - pt_).max() <= 1e-5 * np.abs(pt_).max()


def test_zmarch_hands_over_steep_and_side_crossing_rays():
    """The z-march's hand-over rule on the turbulent cube: rays launched at
    0.3 rad (W = 0.955, where the half-cell z-step leaves 1e-3 of the
    deflection) and rays that cross a side face of the grid (the field jumps
    to zero there) get status 1; the rays it keeps match tight solve_ivp to
    the bar of the path."""
    n = 33
    x, cube = turbulent_cube(n, 0.2)
    dx = x[1] - x[0]
    g = T.NodeGrid(cube)
    steep = tilted_beam(6, -2e-3, 0.0, (0.3, 0.0), seed=5)
    _, _, _, st = T.zmarch_rk4(g, steep)
    assert (st == 1).all()
    # 0.2 rad from x in [3, 4.5] mm: every ray leaves through x = 5 mm
    side = tilted_beam(6, 3e-3, 4.5e-3, (0.2, 0.0), seed=6)
    side[1] = np.linspace(-2e-3, 2e-3, 6)
    _, _, _, st = T.zmarch_rk4(g, side)
    assert (st == 1).all()
    # kept: inside and at most 0.2 rad
    keep = tilted_beam(8, -3e-3, 0.0, (0.2, 0.1), seed=7)
    rf, ph, _, st = T.zmarch_rk4(g, keep)
    assert (st == 0).all()
    rt, pt_, _, _ = T.solve_tight(cube, keep, bundle=True)
    launch = np.arctan(keep[[3, 4]] / keep[5])
    defl = np.sqrt(np.mean((rt[[1, 3]] - launch) ** 2))
    assert np.abs(rf[[0, 2]] - rt[[0, 2]]).max() <= 1e-3 * dx
    assert np.abs(rf[[1, 3]] - rt[[1, 3]]).max() <= 1e-3 * defl


def test_zmarch_keeps_paraxial_rays_of_an_overdense_liner():
    """An overdense liner (test_liner at twice the critical density, clipped
    at ne_max = 1; PT:201-210, 230-231) refracts by up to ~0.15 of a unit
    velocity per cell: rays that come near turning are handed over, and the
    ones the z-march keeps (W > 0.97 all the way) match tight solve_ivp
    within 1e-3 cells and 1e-3 of the RMS deflection."""
    n = 41
    x = np.linspace(-5e-3, 5e-3, n)
    dx = x[1] - x[0]
    cube = T.Cube(x, x.copy(), x.copy())
    cube.test_liner(n_e0=2e27, LR=1e-3)
    cube.calc_dndr()
    cube.extent = cube.extent_z
    s0 = cube.init_beam(60, 3e-3, 1e-3, rng=np.random.RandomState(1))
    rf, ph, _, st = T.zmarch_rk4(T.NodeGrid(cube), s0)
    kept = np.flatnonzero(st == 0)
    assert 10 <= kept.size <= 50      # both kinds present
    rt, pt_, _, _ = T.solve_tight(cube, s0[:, kept], bundle=True)
    defl = np.sqrt(np.mean(rt[[1, 3]] ** 2))
    assert defl > 1e-2
    assert np.abs(rf[[0, 2]][:, kept] - rt[[0, 2]]).max() <= 1e-3 * dx
    assert np.abs(rf[[1, 3]][:, kept] - rt[[1, 3]]).max() <= 1e-3 * defl
    # phase: 5.6e-5 of its magnitude on the most deflected kept ray (0.13 rad)
    assert np.abs(ph[kept] - pt_).max() <= 1e-4 * np.abs(pt_).max()


def test_zmarch_partial_first_step_and_outside_start():
    n = 9
    x = np.linspace(-5e-3, 5e-3, n)
    rs = np.random.RandomState(4)
    cube = T.Cube(x, x.copy(), x.copy())
    cube.external_ne(1e24 * (1 + 0.1 * rs.randn(n, n, n)))
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(6, 1e-3, 1e-4, rng=np.random.RandomState(2))
    s0[2, :3] = -6e-3          # before the grid: straight carry
    s0[2, 3:] = -3.3e-3        # off-lattice: partial first step
    rf, ph, sf, st = T.zmarch_rk4(T.NodeGrid(cube), s0, h=0.5)
    cube.extent = cube.extent_z
    rt, pt_, _, _ = T.solve_tight(cube, s0)
    dx = x[1] - x[0]
    assert np.abs(rf[[0, 2]] - rt[[0, 2]]).max() < 1e-6 * dx
    assert np.abs(rf[[1, 3]] - rt[[1, 3]]).max() < 1e-8


# ---------------------------------------------------------------- optics
@pytest.mark.parametrize("name", ["sh", "df", "lf", "afr"])
def test_rtm_golden(golden, name):
    r0 = golden["rtm_r0"]
    kw = dict(focal_plane=1.0, L=400, R=25)
    rf = {"sh": lambda: R.shadowgraphy(r0, **kw),
          "df": lambda: R.schlieren_df(r0, 1, **kw),
          "lf": lambda: R.schlieren_lf(r0, 1, **kw),
          "afr": lambda: R.afr(r0, np.array([0.5, 1.0, 2.0, 3.0]), **kw)}[name]()
    np.testing.assert_allclose(rf, golden[f"rtm_{name}_rf"], rtol=1e-14, atol=1e-14, equal_nan=True)
    H, xe, ye = R.histogram(rf, 6, 6, bin_scale=1, pix_x=64, pix_y=48)
    np.testing.assert_array_equal(H, golden[f"rtm_{name}_H"])
    np.testing.assert_array_equal(xe, golden[f"rtm_{name}_xedges"])
    np.testing.assert_array_equal(ye, golden[f"rtm_{name}_yedges"])


def test_knife_edge_golden(golden):
    r0 = golden["rtm_r0"]
    np.testing.assert_array_equal(R.knife_edge(R.m_to_mm(r0), 0.5, "x", 1), golden["rtm_knife_x_pos"])
    np.testing.assert_array_equal(R.knife_edge(R.m_to_mm(r0), -0.25, "y", -1), golden["rtm_knife_y_neg"])


def test_shadowgraph_closed_form():
    rs = np.random.RandomState(1)
    r0 = np.stack([rs.uniform(-2e-3, 2e-3, 100), rs.normal(0, 1e-3, 100),
                   rs.uniform(-2e-3, 2e-3, 100), rs.normal(0, 1e-3, 100)])
    fp = 3.0
    rf = R.shadowgraphy(r0, focal_plane=fp, L=400, R=1e9)
    # 4f relay with the object plane shifted: x = -x0 + fp * theta0 (mm, rad)
    np.testing.assert_allclose(rf[0], -r0[0] * 1e3 + fp * r0[1], atol=1e-12)
    np.testing.assert_allclose(rf[1], -r0[1], atol=1e-15)


def test_df_plus_lf_is_shadowgraph():
    rs = np.random.RandomState(2)
    r0 = np.stack([rs.uniform(-2e-3, 2e-3, 5000), rs.normal(0, 4e-3, 5000),
                   rs.uniform(-2e-3, 2e-3, 5000), rs.normal(0, 4e-3, 5000)])
    Hs = R.histogram(R.shadowgraphy(r0), 6, 6, 1, 64, 64)[0]
    Hd = R.histogram(R.schlieren_df(r0, 1), 6, 6, 1, 64, 64)[0]
    Hl = R.histogram(R.schlieren_lf(r0, 1), 6, 6, 1, 64, 64)[0]
    # the Schlieren_DF/LF trains have the same 4f matrix as the shadowgraph;
    # rays exactly at |x_F| = 1 are rejected by neither stop
    np.testing.assert_array_equal(Hd + Hl, Hs)


def test_knife_edge_fourier_plane_closed_form():
    # at the Fourier plane x_F = L * theta0 exactly (independent of x0, fp)
    rs = np.random.RandomState(3)
    r0 = np.stack([rs.uniform(-2e-3, 2e-3, 2000), rs.normal(0, 2e-3, 2000),
                   rs.uniform(-2e-3, 2e-3, 2000), rs.normal(0, 2e-3, 2000)])
    rf = R.schlieren_knife(r0, offset=0.3, axis="x", direction=1, L=400, R=1e9, focal_plane=0.7)
    passed = ~np.isnan(rf[0])
    np.testing.assert_array_equal(passed, ~(400 * r0[1] > 0.3))


# ------------------------------------------- builders, probing axes, rect
@pytest.fixture(scope="module")
def builders():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "builder_vectors.npz"))


@pytest.mark.parametrize("key,fn,kw", [
    ("null", "test_null", {}),
    ("slab", "test_slab", {"s": 3, "n_e0": 5e23}),
    ("linear_cos", "test_linear_cos", {"s1": 0.3, "s2": 0.2, "n_e0": 4e23, "Ly": 3e-3}),
    ("exponential_cos", "test_exponential_cos", {"n_e0": 2e24, "Ly": 2e-3, "s": 4e-3}),
    ("lens", "test_lens", {"n_e0": 3e24, "LR": 2e-3}),
    ("liner", "test_liner", {"n_e0": 3e24, "LR": 2e-3})])
def test_density_builders_golden(builders, key, fn, kw):
    cube = T.Cube(builders["b_x"], builders["b_y"], builders["b_z"])
    getattr(cube, fn)(**kw)
    np.testing.assert_array_equal(cube.ne, builders[f"b_{key}"])


def test_rect_aperture_golden(builders):
    got = R.rect_aperture(builders["rect_r"].copy(), 2.0, 1.5)
    np.testing.assert_array_equal(got, builders["rect_out"])
    # the reference rejects a ray only when it is outside in x AND in y
    r = builders["rect_r"]
    np.testing.assert_array_equal(np.isnan(got[0]), (np.abs(r[0]) > 2.0) & (np.abs(r[2]) > 1.5))


@pytest.mark.parametrize("d", ["y", "x"])
def test_probing_axis_beam_and_as_shipped_solve_golden(builders, d):
    cube = T.Cube(builders["p_x"], builders["p_y"], builders["p_z"], probing_direction=d)
    cube.external_ne(builders["p_ne"].copy())
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(32, 1.5e-3, 5e-3, rng=np.random.RandomState(int(builders[f"p_{d}_seed"])))
    np.testing.assert_array_equal(s0, builders[f"p_{d}_s0"])
    rf = cube.solve()
    np.testing.assert_allclose(cube.sf, builders[f"p_{d}_sf"], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(rf, builders[f"p_{d}_rf"], rtol=1e-10, atol=1e-15)


def test_probing_y_tight_solve_near_as_shipped(builders):
    """The tight per-ray solve with probing along y sits within the
    as-shipped RK45's own tolerance band of the reference's trace."""
    cube = T.Cube(builders["p_x"], builders["p_y"], builders["p_z"], probing_direction="y")
    cube.external_ne(builders["p_ne"].copy())
    cube.calc_dndr(lwl=1064e-9)
    cube.extent = cube.extent_y
    s0 = builders["p_y_s0"]
    rt, _, _, _ = T.solve_tight(cube, s0[:, :8], bundle=True)
    ref = builders["p_y_rf"][:, :8]
    dx = builders["p_x"][1] - builders["p_x"][0]
    assert np.abs(rt[[0, 2]] - ref[[0, 2]]).max() < 1e-2 * dx
    assert np.abs(rt[[1, 3]] - ref[[1, 3]]).max() < 1e-2 * np.abs(ref[[1, 3]]).max()
