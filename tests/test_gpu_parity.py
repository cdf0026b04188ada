"""HIP kernels (through the C ABI) against the FP64 oracle on the same inputs.

Tolerances (stated by the parity bar of this path):
  field         <= 1e-5 relative RMS against numpy FP64 on identical noise
  ray exit x,y  <= 1e-3 grid cells against the FP64 integrator
  exit angles   <= 1e-3 of the RMS deflection
  histograms    exact counts on identical rays (integer work); Poisson
                agreement for traced rays
"""
import numpy as np
import pytest
import torch

from oracle import field as OF
from oracle import philox as OP
from oracle import rtm as OR
from oracle import tracer as OT

pytestmark = pytest.mark.gpu

DEV = "cuda"


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


# ------------------------------------------------------------------ field
@pytest.mark.parametrize("N", [2, 4])
def test_field_dropin_matches_reference_golden(lib, golden, N):
    from turbulence_tracing_amd import turboGen
    M = 2 * N + 1

    def k_func(K):
        with np.errstate(divide="ignore"):
            P = K ** (-11.0 / 3.0)
        P[(K < 1.0 / M) | (K > 0.5)] = 0.0
        return P

    rng = np.random.RandomState(int(golden[f"field_N{N}_seed"]))
    got = turboGen.gaussian3D_FFT(N, k_func, rng=rng, as_numpy=True)
    assert got.shape == (M, M, M)
    assert rel_rms(got, golden[f"field_N{N}"]) <= 1e-5


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (9, 15, 7), (64, 64, 64), (30, 21, 14), (37, 11, 6),
                                   (128, 96, 80),
                                   # the fixed-length passes along each axis
                                   (10, 6, 256), (12, 256, 8), (512, 6, 10), (6, 4, 1024),
                                   (6, 8, 192), (384, 4, 6), (10, 128, 6),
                                   # reference sizes 2N + 1 with large prime factors (1025 = 5^2 41,
                                   # 2049 = 3 683): direct-DFT passes
                                   (1025, 3, 4), (3, 4, 2049)])
def test_field_from_noise_any_shape(lib, shape):
    from turbulence_tracing_amd import turboGen
    rs = np.random.RandomState(sum(shape))
    Wr, Wi = rs.randn(*shape), rs.randn(*shape)
    S = OF.sqrt_powerlaw(shape, -11 / 3)
    want = OF.gaussian_field(Wr, Wi, S)
    got = turboGen.gaussian_field_from_noise(Wr, Wi, S).double().cpu().numpy()
    assert rel_rms(got, want) <= 1e-5


@pytest.mark.parametrize("shape,beta,aniso", [((32, 32, 32), -11 / 3, 1.0), ((24, 20, 16), -3.0, 0.5),
                                              ((17, 19, 23), -3.5, 2.0), ((96, 64, 48), -11 / 3, 1.3),
                                              ((256, 12, 10), -11 / 3, 1.0), ((10, 8, 512), -3.0, 1.5),
                                              ((192, 10, 8), -11 / 3, 0.8), ((8, 6, 384), -3.2, 1.0),
                                              ((6, 128, 10), -11 / 3, 1.0)])
def test_native_field_matches_philox_twin(lib, shape, beta, aniso):
    from turbulence_tracing_amd import turboGen
    f, rms = turboGen.gaussian_field(shape, beta=beta, aniso=aniso, seed=77)
    want = OF.powerlaw_field(shape, beta, seed=77, aniso=aniso)
    got = f.double().cpu().numpy()
    assert rel_rms(got, want) <= 1e-5
    assert abs(float(rms) - np.sqrt(np.mean(want ** 2))) <= 1e-5 * np.sqrt(np.mean(want ** 2))


def test_field_512_properties(lib):
    """Full-size field: unit RMS after normalisation, zero mean, the radial
    spectrum slope, and agreement with the FP64 twin on a sub-block of the
    spectrum-free statistics."""
    from turbulence_tracing_amd import turboGen
    f, rms = turboGen.gaussian_field((512, 512, 512), seed=3, normalise=True)
    assert abs(float(f.double().pow(2).mean()) - 1.0) < 1e-4
    assert abs(float(f.double().mean())) < 1e-4
    # spectrum slope on a 128^3 field from the same generator
    g, _ = turboGen.gaussian_field((128, 128, 128), seed=4)
    k, P, _ = OF.radial_psd(g.double().cpu().numpy(), nbins=64)
    sel = (k > 0.05) & (k < 0.4)
    slope = np.polyfit(np.log(k[sel]), np.log(P[sel]), 1)[0]
    assert abs(slope + 11 / 3) < 0.15


# --------------------------------------------------------------- gradient
def small_cube(n=33, seed=2, eps=0.1, ne0=1e24, L=10e-3):
    x = np.linspace(-L / 2, L / 2, n)
    d = OF.powerlaw_field((n, n, n), -11 / 3, seed=seed)
    d /= np.sqrt((d ** 2).mean())
    return x, ne0 * (1 + eps * d)


def test_grad_stencil_matches_calc_dndr(lib, golden):
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x = golden["cube_x"]
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(golden["cube_ne"])
    cube.calc_dndr(lwl=1064e-9)
    for k in ("dndx", "dndy", "dndz"):
        got = getattr(cube, k).cpu().numpy()
        want = golden["cube_" + k]
        assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()
    # float32 canonical-layout input through the drop-in's calc_dndr (no
    # normalisation: the density is passed as it is); the Pipeline's fused
    # unit-RMS normalisation and density map are pinned in test_gpu_pins.py
    n = 24
    xx, ne = small_cube(n)
    ref = OT.Cube(xx, xx.copy(), xx.copy())
    ref.external_ne(ne.copy())
    ref.calc_dndr(lwl=1064e-9)
    gx, gy, gz, eta = ref.node_fields()
    cube2 = ElectronCube(xx, xx.copy(), xx.copy())
    cube2.external_ne(torch.as_tensor(ne, dtype=torch.float32, device=DEV))
    cube2.calc_dndr(lwl=1064e-9)
    f4 = cube2.field4.double().cpu().numpy()  # [z][y][x][4]
    dz = xx[1] - xx[0]
    k0 = 2 * np.pi / 1064e-9
    for ch, want in enumerate([gx * dz, gy * dz, gz * dz, eta * k0 * dz]):
        got = f4[..., ch].transpose(2, 1, 0)
        assert np.abs(got - want).max() <= 1e-3 * np.abs(want).max(), ch


def test_grad_stencil_clip_and_linear_exact(lib):
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 17
    x = np.linspace(-5e-3, 5e-3, n)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.test_slab(s=8, n_e0=1e25)
    cube.calc_dndr()
    ref = OT.Cube(x, x.copy(), x.copy())
    XX = np.meshgrid(x, x, x, indexing="ij")[0]
    ref.external_ne(1e25 * (1 + 8 * XX / x.max()))
    ref.calc_dndr()
    assert np.abs(cube.dndx.cpu().numpy() - ref.dndx).max() <= 1e-6 * np.abs(ref.dndx).max()
    # clipping above ne_max (PT:231): liner beyond critical density
    cube.test_liner(n_e0=2e27, LR=1e-3)
    cube.calc_dndr(ne_max=1)
    ref2 = OT.Cube(x, x.copy(), x.copy())
    ref2.external_ne(cube.ne.cpu().numpy())
    ref2.calc_dndr(ne_max=1)
    assert np.abs(cube.dndz.cpu().numpy() - ref2.dndz).max() <= 1e-6 * np.abs(ref2.dndz).max()


# ----------------------------------------------------------------- tracer
def trace_both(n=33, n_rays=4096, seed=2, eps=0.1, divergence=50e-6, beam=2e-3, h=0.5, sort=True, variant=0):
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x, ne = small_cube(n, seed=seed, eps=eps)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(n_rays, beam, divergence, seed=seed + 100)
    rf = cube.solve(h=h, sort=sort, verbose=False, variant=variant).cpu().numpy()
    ph = cube.phase.cpu().numpy()
    sf = cube.sf.cpu().numpy()
    ref = OT.Cube(x, x.copy(), x.copy())
    ref.external_ne(ne)
    ref.calc_dndr(lwl=1064e-9)
    ref.extent = ref.extent_z
    g = OT.NodeGrid(ref)
    rf_o, ph_o, sf_o, st = OT.zmarch_rk4(g, s0.cpu().numpy(), h=h)
    return x, rf, ph, sf, rf_o, ph_o, sf_o, cube


def check_tolerance(x, rf, ph, rf_o, ph_o, s0_theta_rms=None):
    dx = x[1] - x[0]
    pos_err = np.abs(rf[[0, 2]] - rf_o[[0, 2]]).max() / dx
    defl = np.sqrt(np.mean(rf_o[[1, 3]] ** 2))
    ang_err = np.abs(rf[[1, 3]] - rf_o[[1, 3]]).max()
    ph_err = np.abs(ph - ph_o).max()
    assert pos_err <= 1e-3, pos_err
    assert ang_err <= 1e-3 * defl, (ang_err, defl)
    assert ph_err <= 1e-5 * np.abs(ph_o).max(), ph_err
    return pos_err, ang_err / defl, ph_err


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n,h,divergence", [(33, 0.5, 50e-6), (20, 0.25, 2e-3), (48, 1.0, 5e-3)])
def test_trace_matches_oracle_turbulent(lib, variant, n, h, divergence):
    x, rf, ph, sf, rf_o, ph_o, sf_o, _ = trace_both(n=n, h=h, divergence=divergence, variant=variant)
    check_tolerance(x, rf, ph, rf_o, ph_o)
    # sf: positions at c t = cT to the position bar, velocities to the angle bar
    dx = x[1] - x[0]
    defl = np.sqrt(np.mean(rf_o[[1, 3]] ** 2))
    assert np.abs(sf[:3] - sf_o[:3]).max() <= 1e-3 * dx
    assert np.abs(sf[3:] - sf_o[3:]).max() / OT.C <= 1e-3 * defl


def test_trace_unsorted_equals_sorted(lib):
    a = trace_both(n=17, n_rays=2000, sort=True)
    b = trace_both(n=17, n_rays=2000, sort=False)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])


def test_plane_forms_match_numpy(lib):
    """tt_plane_forms writes, per node, the bilinear form of its z-plane
    cell, A = f00, B = f10 - f00, C = f01 - f00, E = (f11 - f01) - B, with
    the same float32 subtractions as numpy (bit for bit), and zero B, C, E on
    the last column and row."""
    from turbulence_tracing_amd import _lib
    nx, ny, nz = 13, 9, 5
    rs = np.random.RandomState(3)
    f4 = rs.randn(nz, ny, nx, 4).astype(np.float32)
    grid = _lib.Grid(nx, ny, nz, 0, 0, 0, 1, 1, 1)
    dev = torch.as_tensor(f4, device=DEV)
    out = torch.full((nz, ny, nx, 16), 7.0, dtype=torch.float32, device=DEV)
    lib.tt_plane_forms(dev.data_ptr(), __import__("ctypes").byref(grid), out.data_ptr(), _lib.stream_ptr())
    got = out.cpu().numpy().reshape(nz, ny, nx, 4, 4)
    A = f4
    B = np.zeros_like(f4)
    Cc = np.zeros_like(f4)
    E = np.zeros_like(f4)
    B[:, :-1, :-1] = f4[:, :-1, 1:] - f4[:, :-1, :-1]
    Cc[:, :-1, :-1] = f4[:, 1:, :-1] - f4[:, :-1, :-1]
    E[:, :-1, :-1] = (f4[:, 1:, 1:] - f4[:, 1:, :-1]) - B[:, :-1, :-1]
    for i, want in enumerate((A, B, Cc, E)):
        np.testing.assert_array_equal(got[..., i, :], want)


@pytest.mark.parametrize("n", [64, 97])
def test_trace_with_plane_forms_is_bit_identical(lib, n):
    """The cell cache loading the plane forms (tt_plane_forms) instead of
    forming them from the node field: the same exit rays and phase, bit for
    bit, through the Pipeline."""
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    res = []
    for forms in (True, False):
        p = Pipeline(Config(n=n, n_rays=30000, pix=64, plane_forms=forms))
        p.run(seed=4)
        torch.cuda.synchronize()
        res.append((p.rf.cpu().numpy(), p.phase.cpu().numpy(), p.order.cpu().numpy()))
    (rf_a, ph_a, o_a), (rf_b, ph_b, o_b) = res
    # per ray in beam order (the order within a column is not fixed)
    A, Bv = np.empty_like(rf_a), np.empty_like(rf_b)
    A[:, o_a] = rf_a
    Bv[:, o_b] = rf_b
    np.testing.assert_array_equal(A, Bv)
    pa, pb = np.empty_like(ph_a), np.empty_like(ph_b)
    pa[o_a] = ph_a
    pb[o_b] = ph_b
    np.testing.assert_array_equal(pa, pb)


def test_trace_golden_cube_vs_reference_solve(lib, golden):
    """The reference's own as-shipped trace (golden) sits within its RK45
    tolerance of the GPU trace of the same rays."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x = golden["cube_x"]
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(golden["cube_ne"])
    cube.calc_dndr(lwl=1064e-9)
    cube.s0 = torch.as_tensor(golden["beam_s0"], device=DEV)
    cube.extent = cube.extent_z
    rf = cube.solve(verbose=False).cpu().numpy()
    ref = golden["solve_rf"]
    dx = x[1] - x[0]
    assert np.abs(rf[[0, 2]] - ref[[0, 2]]).max() < 1e-2 * dx
    assert np.abs(rf[[1, 3]] - ref[[1, 3]]).max() < 1e-2 * np.abs(ref[[1, 3]]).max()
    # and sf / ray_at_exit consistency: ray_at_exit(sf) reproduces rf
    np.testing.assert_allclose(cube.ray_at_exit().cpu().numpy(), rf, rtol=1e-9, atol=1e-12)


def test_trace_null_and_slab_closed_forms(lib):
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 21
    x = np.linspace(-5e-3, 5e-3, n)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.test_null()
    cube.calc_dndr()
    s0 = cube.init_beam(1000, 2e-3, 5e-3, seed=1)
    rf = cube.solve(verbose=False).cpu().numpy()
    s = s0.cpu().numpy()
    tx, ty = s[3] / s[5], s[4] / s[5]
    dx = x[1] - x[0]
    # straight lines to FP32 state precision (1e-6 cells, 1e-6 relative angle)
    np.testing.assert_allclose(rf[0], s[0] + tx * 10e-3, rtol=0, atol=1e-6 * dx)
    np.testing.assert_allclose(rf[2], s[1] + ty * 10e-3, rtol=0, atol=1e-6 * dx)
    np.testing.assert_allclose(rf[1], np.arctan(tx), rtol=1e-6, atol=1e-15)
    assert np.all(cube.phase.cpu().numpy() == 0)
    cube.test_slab(s=8, n_e0=1e25)
    cube.calc_dndr()
    cube.init_beam(1000, 1e-3, 0.0, seed=2)
    rf = cube.solve(verbose=False).cpu().numpy()
    nc = OT.critical_density(1053e-9)
    slope = -0.5 * (1e25 * 8 / x.max() / nc) * 10e-3
    np.testing.assert_allclose(rf[1], np.arctan(slope), rtol=1e-5)


def test_trace_nan_and_backward_rays_flagged(lib):
    """A NaN ray stays NaN (status 2); a ray launched backwards is traced in
    time (status 3) and, in the empty cube, reported by ray_at_exit's
    back-projection of its straight path (PT:371-378)."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 9
    x = np.linspace(-5e-3, 5e-3, n)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.test_null()
    cube.calc_dndr()
    s0 = cube.init_beam(8, 1e-3, 1e-3, seed=3)
    s0[0, 1] = float("nan")
    s0[5, 2] = -s0[5, 2]
    cube.s0 = s0
    rf = cube.solve(verbose=False).cpu().numpy()
    st = cube.status.cpu().numpy()
    assert st[1] == 2 and st[2] == 3 and (st[[0, 3, 4, 5, 6, 7]] == 0).all()
    assert np.isnan(rf[:, 1]).all()
    s = s0.cpu().numpy()[:, 2]
    t_bp = (s[2] - 5e-3) / s[5]  # straight path: back-projected from its start
    np.testing.assert_allclose(rf[:, 2], [s[0] - s[3] * t_bp, np.arctan(s[3] / s[5]), s[1] - s[4] * t_bp,
                                          np.arctan(s[4] / s[5])], rtol=1e-12, atol=1e-15)
    assert cube.phase.cpu().numpy()[2] == 0.0


def test_turning_rays_traced_in_time_match_tight_solve(lib):
    """An overdense liner (test_liner beyond critical density, clipped at
    ne_max = 1: PT:201-210, 230-231; the reference's notebook shows ray
    angles reaching pi/2) turns rays back.  The z-march hands over the rays
    that turn or come near turning (W <= 0.97) and tt_trace_time traces them
    in time; both kinds, and the rays the z-march keeps, agree with tight
    solve_ivp on the same cube: exit positions within 1e-3 cells, exit angles
    within 1e-3 of the RMS deflection, phase within 1e-4 of its magnitude,
    and the state at c t = cT (the reference's sol.y[:, -1]) within 1e-3
    cells."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 41
    x = np.linspace(-5e-3, 5e-3, n)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.test_liner(n_e0=2e27, LR=1e-3)
    cube.calc_dndr()
    s0 = cube.init_beam(400, 3e-3, 1e-3, seed=8)
    rf = cube.solve(verbose=False).cpu().numpy()
    st = cube.status.cpu().numpy()
    ph = cube.phase.cpu().numpy()
    sf = cube.sf.cpu().numpy()
    timed = np.flatnonzero(st == 3)
    kept = np.flatnonzero(st == 0)
    assert np.isin(st, (0, 3)).all()
    assert np.isfinite(rf[:, timed]).all() and np.isfinite(ph[timed]).all()
    back = timed[sf[5, timed] < 0]    # left through the entry side
    fwd = timed[sf[5, timed] > 0]
    assert back.size >= 10 and fwd.size >= 4 and kept.size >= 10
    ref = OT.Cube(x, x.copy(), x.copy())
    ref.test_liner(n_e0=2e27, LR=1e-3)
    ref.calc_dndr()
    ref.extent = ref.extent_z
    sel = np.concatenate([back[:8], fwd[:4], kept[:4]])
    rt, pt_, sft, _ = OT.solve_tight(ref, s0.cpu().numpy()[:, sel], bundle=False)
    assert (sft[5, :8] < 0).all() and (sft[5, 8:] > 0).all()
    dx = x[1] - x[0]
    defl = np.sqrt(np.mean(rt[[1, 3]] ** 2))
    pos = np.abs(rf[[0, 2]][:, sel] - rt[[0, 2]]) / dx
    ang = np.abs(rf[[1, 3]][:, sel] - rt[[1, 3]])
    phe = np.abs(ph[sel] - pt_) / np.abs(pt_).max()
    print(f"liner: pos {pos.max():.2e} cells, angle {ang.max():.2e} rad of RMS {defl:.2e}, "
          f"phase {phe.max():.2e}; time-traced back {back.size}, forward {fwd.size}; kept {kept.size}")
    assert pos.max() <= 1e-3 and ang.max() <= 1e-3 * defl and phe.max() <= 1e-4
    assert np.abs(sf[:3, sel] - sft[:3]).max() <= 1e-3 * dx


def test_trace_where_w_changes_fast_between_stages(lib):
    """A steep z-ramp of the density (about four cells wide, up to ne/nc =
    0.05, so every ray keeps W > 0.97): W changes by ~2e-3 between RK4 stages
    there, more than the 2^-12 within which the cached step's Newton-refined
    1/W is exact, so those steps are redone by gathers (and summed by the RK4
    weights).  The GPU agrees with the FP64 z-march that follows the same
    rule, to the bar and to 1e-5 of the phase, and with tight solve_ivp."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 41
    x = np.linspace(-5e-3, 5e-3, n)
    dx = x[1] - x[0]
    XX, _, ZZ = np.meshgrid(x, x, x, indexing="ij")
    nc = OT.critical_density(1053e-9)
    ne = nc * 0.05 * 0.5 * (1 + np.tanh(ZZ / (2 * dx))) * (1 + 0.3 * XX / 5e-3)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr()
    s0 = cube.init_beam(2000, 2e-3, 1e-3, seed=21)
    rf = cube.solve(verbose=False).cpu().numpy()
    ph = cube.phase.cpu().numpy()
    assert (cube.status.cpu().numpy() == 0).all()
    ref = OT.Cube(x, x.copy(), x.copy())
    ref.external_ne(ne)
    ref.calc_dndr()
    ref.extent = ref.extent_z
    s = s0.cpu().numpy()
    rf_o, ph_o, _, st = OT.zmarch_rk4(OT.NodeGrid(ref), s)
    assert (st == 0).all()
    launch = np.arctan(s[[3, 4]] / s[5])
    defl = np.sqrt(np.mean((rf_o[[1, 3]] - launch) ** 2))
    assert defl > 1e-3
    assert np.abs(rf[[0, 2]] - rf_o[[0, 2]]).max() <= 1e-3 * dx
    assert np.abs(rf[[1, 3]] - rf_o[[1, 3]]).max() <= 1e-3 * defl
    assert np.abs(ph - ph_o).max() <= 1e-5 * np.abs(ph_o).max()
    sel = np.arange(0, 2000, 250)
    rt, pt_, _, _ = OT.solve_tight(ref, s[:, sel], bundle=True)
    assert np.abs(rf[[0, 2]][:, sel] - rt[[0, 2]]).max() <= 1e-3 * dx
    assert np.abs(rf[[1, 3]][:, sel] - rt[[1, 3]]).max() <= 1e-3 * defl
    assert np.abs(ph[sel] - pt_).max() <= 1e-5 * np.abs(pt_).max()


def test_dndr_and_dsdt_match_reference_interpolator(lib, golden):
    from turbulence_tracing_amd import particle_tracker as pt
    x = golden["cube_x"]
    cube = pt.ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(golden["cube_ne"])
    cube.calc_dndr(lwl=1064e-9)
    got = cube.dndr(golden["dndr_pts"]).cpu().numpy()
    want = golden["dndr_vals"]
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()
    assert np.all(got[:, 2] == 0)
    s = np.concatenate([golden["dndr_pts"], np.ones((3, 64))]).flatten()
    d = pt.dsdt(0.0, s, cube).cpu().numpy().reshape(6, 64)
    np.testing.assert_array_equal(d[:3], np.ones((3, 64)))


# ------------------------------------------------------------------- beam
def test_beam_sampler_matches_philox_twin(lib):
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x = np.linspace(-5e-3, 5e-3, 9)
    cube = ElectronCube(x, x.copy(), x.copy())
    s0 = cube.init_beam(5000, 2e-3, 50e-6, seed=1234).cpu().numpy()
    ut, ua, ub, uphi, nchi = OP.beam_uniforms(np.arange(5000), 1234)
    u = ua + ub
    u[u > 1] = 2 - u[u > 1]
    want = OT.beam_from_draws(2 * np.pi * ut, u, np.pi * uphi, 50e-6 * nchi, 2e-3, 5e-3, "z")
    np.testing.assert_allclose(s0, want, rtol=1e-12, atol=1e-18)
    r = np.hypot(s0[0], s0[1])
    assert r.max() <= 2e-3 and abs(np.mean(r ** 2) / 2e-3 ** 2 - 0.5) < 0.02   # uniform in the disc
    th = s0[3] / s0[5]
    assert abs(np.var(th) / (50e-6 ** 2 / 2) - 1) < 0.1                        # Var(theta_x) = sigma^2 / 2


def test_ray_sort_is_permutation_by_column(lib):
    from turbulence_tracing_amd import _lib
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 65
    x = np.linspace(-5e-3, 5e-3, n)
    cube = ElectronCube(x, x.copy(), x.copy())
    s0 = cube.init_beam(100000, 3e-3, 1e-3, seed=9)
    order = cube._sort_order(s0).cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(100000))
    s = s0.cpu().numpy()[:, order]
    ix = np.floor((s[0] + 5e-3) / (x[1] - x[0])).astype(int)
    iy = np.floor((s[1] + 5e-3) / (x[1] - x[0])).astype(int)

    def morton(a, b):
        k = np.zeros_like(a)
        for bit in range(8):
            k |= ((a >> bit) & 1) << (2 * bit) | ((b >> bit) & 1) << (2 * bit + 1)
        return k
    keys = morton(ix, iy)
    assert np.all(np.diff(keys) >= 0)


@pytest.mark.parametrize("n,N", [(100000, 100000), (777, 1000), (1, 1)])
def test_ray_sort_rays_moves_rays_exactly(lib, n, N):
    """tt_ray_sort_rays writes s0's columns in the sort order (bit for bit),
    for a full and a partial batch (out rows of length N > n)."""
    import ctypes as C
    import torch
    from turbulence_tracing_amd import _lib
    x = np.linspace(-5e-3, 5e-3, 65)
    grid = _lib.Grid(65, 65, 65, -5e-3, -5e-3, -5e-3, x[1] - x[0], x[1] - x[0], x[1] - x[0])
    rs = np.random.RandomState(n)
    s0 = torch.as_tensor(np.concatenate([rs.uniform(-3e-3, 3e-3, (2, N)), rs.randn(4, N)]), device="cuda")
    nbytes = C.c_size_t()
    lib.tt_ray_sort_scratch_bytes(C.byref(grid), 0, C.byref(nbytes))
    scratch = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    order = torch.full((N,), -1, dtype=torch.int32, device="cuda")
    staging = torch.empty((N, 8), dtype=torch.float64, device="cuda")
    out = torch.full((6, N), 7.0, dtype=torch.float64, device="cuda")
    lib.tt_ray_sort_rays(_lib.ptr(s0), N, n, C.byref(grid), 0, _lib.ptr(scratch), _lib.ptr(order), _lib.ptr(staging),
                         _lib.ptr(out), N, _lib.stream_ptr())
    o = order[:n].long().cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(n))
    np.testing.assert_array_equal(out[:, :n].cpu().numpy(), s0[:, :n].cpu().numpy()[:, o])
    assert bool((out[:, n:] == 7.0).all())
    # same order as tt_ray_sort's key order: entry-column keys non-decreasing
    ix = np.floor((out[0, :n].cpu().numpy() + 5e-3) / (x[1] - x[0])).astype(np.int64)
    iy = np.floor((out[1, :n].cpu().numpy() + 5e-3) / (x[1] - x[0])).astype(np.int64)
    k = np.zeros_like(ix)
    for bit in range(8):
        k |= ((ix >> bit) & 1) << (2 * bit) | ((iy >> bit) & 1) << (2 * bit + 1)
    assert np.all(np.diff(k) >= 0)


# -------------------------------------------------------------- detectors
@pytest.mark.parametrize("name", ["sh", "df", "lf", "afr"])
def test_detectors_match_reference_golden(lib, golden, name):
    from turbulence_tracing_amd import ray_transfer_matrix as rtm
    r0 = golden["rtm_r0"]
    cls, kw = {"sh": (rtm.Shadowgraphy, {}), "df": (rtm.Schlieren_DF, {"R": 1}), "lf": (rtm.Schlieren_LF, {"R": 1}),
               "afr": (rtm.AFR, {"Rs": np.array([0.5, 1.0, 2.0, 3.0])})}[name]
    d = cls(r0, focal_plane=1.0, L=400, R=25, Lx=6, Ly=6)
    d.solve(**kw)
    d.histogram(bin_scale=1, pix_x=64, pix_y=48)
    np.testing.assert_allclose(d.rf.cpu().numpy(), golden[f"rtm_{name}_rf"], rtol=1e-12, atol=1e-12,
                               equal_nan=True)
    np.testing.assert_array_equal(d.H.cpu().numpy(), golden[f"rtm_{name}_H"])
    np.testing.assert_array_equal(d.xedges, golden[f"rtm_{name}_xedges"])


def test_histogram_edges_exact(lib):
    """Rays exactly on bin edges, on the last edge, outside and NaN."""
    from turbulence_tracing_amd import ray_transfer_matrix as rtm
    nb = 10
    edges = np.linspace(-3, 3, nb + 1)
    xs = np.concatenate([edges, [-3.0000001, 3.0000001, np.nan, 0.123, -2.9999999]])
    xs_m = xs / 1e3
    r0 = np.zeros((4, xs.size))
    r0[0] = xs_m
    r0[2] = 0.0
    d = rtm.Rays(r0, Lx=6, Ly=6)
    d.rf = d.r0
    d.histogram(bin_scale=1, pix_x=nb, pix_y=4)
    H_want, _, _ = OR.histogram(rtm.m_to_mm(r0).cpu().numpy(), 6, 6, 1, nb, 4)
    np.testing.assert_array_equal(d.H.cpu().numpy(), H_want)


def test_fused_detectors_match_oracle_on_traced_rays(lib):
    from turbulence_tracing_amd import ray_transfer_matrix as rtm
    x, rf, ph, sf, rf_o, ph_o, sf_o, _ = trace_both(n=17, n_rays=20000, divergence=2e-3)
    tilt = 2 * np.pi * 20 / 6
    Hs, Hk, Hi = rtm.detect_fused(torch.as_tensor(rf, device=DEV), torch.as_tensor(ph, device=DEV), 6, 6, 64, 64,
                                  knife=(0.0, "x", 1), tilt_k=tilt)
    # identical rays -> identical integer histograms
    Ws = OR.histogram(OR.shadowgraphy(rf), 6, 6, 1, 64, 64)[0]
    Wk = OR.histogram(OR.schlieren_knife(rf, 0.0, "x", 1), 6, 6, 1, 64, 64)[0]
    np.testing.assert_array_equal(Hs.cpu().numpy(), Ws)
    np.testing.assert_array_equal(Hk.cpu().numpy(), Wk)
    sh = OR.shadowgraphy(rf)
    Wi = OR.histogram(sh, 6, 6, 1, 64, 64, weights=OR.interferogram_weights(sh, ph, tilt))[0]
    np.testing.assert_allclose(Hi.cpu().numpy(), Wi, rtol=1e-5, atol=1e-3)
    # the oracle's own rays (same beam) -> counts agree within Poisson noise
    Wo = OR.histogram(OR.shadowgraphy(rf_o), 6, 6, 1, 64, 64)[0]
    H = Hs.cpu().numpy().astype(float)
    m = (H + Wo) > 0
    chi2 = np.sum((H[m] - Wo[m]) ** 2 / (H[m] + Wo[m]))
    assert chi2 <= m.sum() * 0.05


def test_interferogram_uniform_slab_fringes(lib):
    """Uniform slab: phase is the same for every ray, so the fringes are
    straight with period 2 pi / tilt_k."""
    from turbulence_tracing_amd import ray_transfer_matrix as rtm
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    n = 9
    x = np.linspace(-5e-3, 5e-3, n)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(np.full((n, n, n), 3e25))
    cube.calc_dndr(lwl=1064e-9)
    cube.init_beam(200000, 3e-3, 0.0, seed=5)
    rf = cube.solve(verbose=False)
    ph = cube.phase
    r = 3e25 / OT.critical_density(1064e-9)
    want = 2 * np.pi / 1064e-9 * (np.sqrt(1 - r) - 1) * 10e-3
    np.testing.assert_allclose(ph.cpu().numpy(), want, rtol=1e-6)
    tilt = 2 * np.pi / 1.0   # 1 mm period
    Hs, Hk, Hi = rtm.detect_fused(rf, ph, 6, 6, 60, 1, tilt_k=tilt)
    w = Hi.cpu().numpy()[0] / np.maximum(Hs.cpu().numpy()[0], 1)
    xc = (np.arange(60) + 0.5) * 0.1 - 3
    pred = 1 + np.cos(want - tilt * xc)
    ok = Hs.cpu().numpy()[0] > 50
    assert np.abs(w[ok] - pred[ok]).max() < 0.2


# --------------------------------------------------------------- pipeline
def test_pipeline_small_end_to_end(lib):
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    cfg = Config(n=64, n_rays=20000, pix=128)
    p = Pipeline(cfg)
    p.run(seed=5)
    torch.cuda.synchronize()
    # field is the Philox twin, normalised in the stencil
    want = OF.powerlaw_field((64, 64, 64), cfg.beta, seed=5)
    got = p.field.double().cpu().numpy().transpose(2, 1, 0)
    assert rel_rms(got, want) <= 1e-5
    # trace of the sorted rays against the FP64 integrator on the same field
    f4 = p.field4.double().cpu().numpy()
    ref = OT.Cube.__new__(OT.Cube)
    g = OT.NodeGrid.__new__(OT.NodeGrid)
    dz = p.grid.dz
    g.shape = (64, 64, 64)
    g.origin = (p.grid.x0, p.grid.y0, p.grid.z0)
    g.spacing = (p.grid.dx, p.grid.dy, p.grid.dz)
    k0 = 2 * np.pi / cfg.lwl
    g.fields = tuple(f4[..., ch].transpose(2, 1, 0) / (dz if ch < 3 else k0 * dz) for ch in range(4))
    g.k0 = k0
    g.extent = cfg.L / 2
    s0 = p.s0.cpu().numpy()
    rf_o, ph_o, _, _ = OT.zmarch_rk4(g, s0)
    x = np.linspace(-cfg.L / 2, cfg.L / 2, 64)
    check_tolerance(x, p.rf.cpu().numpy(), p.phase.cpu().numpy(), rf_o, ph_o)
    assert int(p.H_shadow.sum()) > 0.9 * cfg.n_rays


def test_pipeline_full_size_subset_matches_oracle(lib):
    """The bench workload at full size (512^3 field, 1e7 rays, 2048^2
    detectors): every ray leaves finite, 512 rays spread over the
    entry-column order match the FP64 integrator on the same node field
    (cropped in x and y around their paths), the shadowgraph holds nearly
    every ray, and both count histograms equal the FP64 optics and
    np.histogram2d on all 1e7 exit rays exactly."""
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    cfg = Config()
    assert (cfg.n, cfg.n_rays, cfg.pix) == (512, 10_000_000, 2048)
    p = Pipeline(cfg)
    p.run(seed=5)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(p.rf).all()) and bool(torch.isfinite(p.phase).all())
    idx = torch.linspace(0, cfg.n_rays - 1, 512, device=p.s0.device).long()
    s0 = p.s0[:, idx].cpu().numpy()
    rf = p.rf[:, idx].cpu().numpy()
    ph = p.phase[idx].cpu().numpy()
    gr = p.grid
    # x / y window around the rays (paths drift far less than a cell here)
    ix = (s0[0] - gr.x0) / gr.dx
    iy = (s0[1] - gr.y0) / gr.dy
    x_lo, x_hi = max(int(ix.min()) - 4, 0), min(int(ix.max()) + 5, gr.nx - 1)
    y_lo, y_hi = max(int(iy.min()) - 4, 0), min(int(iy.max()) + 5, gr.ny - 1)
    f4 = p.field4[:, y_lo:y_hi + 1, x_lo:x_hi + 1].double().cpu().numpy()
    dz = gr.dz
    k0 = 2 * np.pi / cfg.lwl
    g = OT.NodeGrid.__new__(OT.NodeGrid)
    g.shape = (x_hi - x_lo + 1, y_hi - y_lo + 1, gr.nz)
    g.origin = (gr.x0 + x_lo * gr.dx, gr.y0 + y_lo * gr.dy, gr.z0)
    g.spacing = (gr.dx, gr.dy, gr.dz)
    g.fields = tuple(np.ascontiguousarray(f4[..., ch].transpose(2, 1, 0)) / (dz if ch < 3 else k0 * dz)
                     for ch in range(4))
    g.k0 = k0
    g.extent = cfg.L / 2
    rf_o, ph_o, _, st = OT.zmarch_rk4(g, s0)
    assert (st == 0).all()
    # the window held every oracle path
    assert ((rf_o[0] - g.origin[0]) / gr.dx > 1).all() and ((rf_o[0] - g.origin[0]) / gr.dx < g.shape[0] - 2).all()
    assert ((rf_o[2] - g.origin[1]) / gr.dy > 1).all() and ((rf_o[2] - g.origin[1]) / gr.dy < g.shape[1] - 2).all()
    x = np.linspace(-cfg.L / 2, cfg.L / 2, cfg.n)
    check_tolerance(x, rf, ph, rf_o, ph_o)
    assert int(p.H_shadow.sum()) > 0.99 * cfg.n_rays
    # the 2048^2 detectors over all 1e7 exit rays: counts exactly those of the
    # FP64 optics and np.histogram2d on the same rays
    rf_all = p.rf.cpu().numpy()
    Ws = OR.histogram(OR.shadowgraphy(rf_all, L=cfg.L_optics, R=cfg.R_optics), cfg.det_size, cfg.det_size, 1,
                      cfg.pix, cfg.pix)[0]
    np.testing.assert_array_equal(p.H_shadow.cpu().numpy(), Ws)
    Wk = OR.histogram(OR.schlieren_knife(rf_all, 0.0, "x", 1, L=cfg.L_optics, R=cfg.R_optics), cfg.det_size,
                      cfg.det_size, 1, cfg.pix, cfg.pix)[0]
    np.testing.assert_array_equal(p.H_knife.cpu().numpy(), Wk)


# ------------------------------------------------------- probing directions
def test_probing_y_equals_z_through_swapped_cube(lib):
    """Probing along y (PT:290-299, 361-368) is probing along z through the
    cube with its y and z axes swapped: same exit rays, phase and final state
    bit for bit, on a non-cubic turbulent grid."""
    from oracle import field as OF
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x = np.linspace(-5e-3, 5e-3, 33)
    y = np.linspace(-4e-3, 4e-3, 29)
    z = np.linspace(-5e-3, 5e-3, 41)
    d = OF.powerlaw_field((33, 29, 41), -11 / 3, seed=4)
    ne = 1e24 * (1 + 0.1 * d / np.sqrt((d ** 2).mean()))
    cy = ElectronCube(x, y, z, probing_direction="y")
    cy.external_ne(ne)
    cy.calc_dndr(lwl=1064e-9)
    s0 = cy.init_beam(3000, 2e-3, 1e-3, seed=5)
    rf_y = cy.solve(verbose=False).cpu().numpy()
    cz = ElectronCube(x, z, y, probing_direction="z")
    cz.external_ne(np.ascontiguousarray(ne.transpose(0, 2, 1)))
    cz.calc_dndr(lwl=1064e-9)
    cz.init_beam(1, 2e-3, 1e-3, seed=5)
    cz.s0 = s0[[0, 2, 1, 3, 5, 4]].contiguous()
    rf_z = cz.solve(verbose=False).cpu().numpy()
    np.testing.assert_array_equal(rf_y, rf_z)
    np.testing.assert_array_equal(cy.phase.cpu().numpy(), cz.phase.cpu().numpy())
    np.testing.assert_array_equal(cy.sf.cpu().numpy(), cz.sf.cpu().numpy()[[0, 2, 1, 3, 5, 4]])
    assert np.abs(rf_y[1] - np.arctan(s0[3].cpu().numpy() / s0[4].cpu().numpy())).max() > 1e-7  # deflected
    # ray_at_exit of the 6-D state gives the same rays
    np.testing.assert_allclose(cy.ray_at_exit().cpu().numpy(), rf_y, rtol=0, atol=1e-12)


def test_probing_x_rays_start_past_the_cube(lib):
    """The reference launches x-probing rays at x = +extent moving +x
    (PT:280-289): they never cross the cube, and ray_at_exit (PT:353-360)
    returns their launch state."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x = np.linspace(-5e-3, 5e-3, 33)
    cube = ElectronCube(x, x.copy(), x.copy(), probing_direction="x")
    cube.test_lens()
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(2000, 2e-3, 1e-3, seed=6).cpu().numpy()
    assert np.all(s0[0] == 5e-3)
    rf = cube.solve(verbose=False).cpu().numpy()
    np.testing.assert_allclose(rf[0], s0[1], rtol=0, atol=1e-15)
    np.testing.assert_allclose(rf[2], s0[2], rtol=0, atol=1e-15)
    # the ray state carries U, V, W - 1 in float32: angles to float rounding
    np.testing.assert_allclose(rf[1], np.arctan(s0[4] / s0[3]), rtol=2e-7, atol=0)
    np.testing.assert_allclose(rf[3], np.arctan(s0[5] / s0[3]), rtol=2e-7, atol=0)


def test_detectors_pickle_like_the_reference(lib, tmp_path):
    """Diagnostics pickle (example_MPI.py:154-166) to numpy-backed objects
    that load without touching the GPU and keep H, edges and rays."""
    import pickle
    from turbulence_tracing_amd import ray_transfer_matrix as rtm
    rs = np.random.RandomState(3)
    r0 = np.stack([rs.randn(5000) * 1e-3, rs.randn(5000) * 1e-4, rs.randn(5000) * 1e-3, rs.randn(5000) * 1e-4])
    sh = rtm.Shadowgraphy(r0, L=400, R=25, Lx=18, Ly=13.5)
    sh.solve()
    sh.histogram(bin_scale=10)
    path = tmp_path / "Shadowgraphy.pkl"
    with open(path, "wb") as f:
        pickle.dump(sh, f)
    with open(path, "rb") as f:
        back = pickle.load(f)  # written by this test
    assert isinstance(back.H, np.ndarray) and isinstance(back.rf, np.ndarray)
    np.testing.assert_array_equal(back.H, sh.H.cpu().numpy())
    np.testing.assert_array_equal(back.xedges, sh.xedges)
    assert back.H.sum() == np.isfinite(sh.rf.cpu().numpy()).all(0).sum() - (
        (np.abs(sh.rf.cpu().numpy()[0]) > 9) | (np.abs(sh.rf.cpu().numpy()[2]) > 6.75)).sum()


def test_pipeline_back_to_back_steps_match_isolated_runs(lib):
    """Pipeline.step runs its stages on side streams; results taken on the
    caller's stream after each of several back-to-back steps (no host
    synchronisation) equal those of the same seeds run one at a time with a
    full synchronisation in between."""
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    cfg = Config(n=64, n_rays=20000, pix=128)
    seeds = [3, 8, 21]

    def grab(p):
        # clones on the caller's stream: ordered after this step's detectors
        return [h.clone() for h in (p.H_shadow, p.H_knife, p.H_interf)] + [p.field4.clone()]

    p = Pipeline(cfg)
    back = []
    for sd in seeds:
        p.step(sd)
        back.append(grab(p))
    torch.cuda.synchronize()
    q = Pipeline(cfg)
    for sd, got in zip(seeds, back):
        q.run(sd)
        torch.cuda.synchronize()
        want = grab(q)
        torch.cuda.synchronize()
        for i in range(2):  # counts are order-independent
            assert torch.equal(got[i], want[i])
        torch.testing.assert_close(got[2], want[2], rtol=1e-5, atol=1e-3)
        assert torch.equal(got[3], want[3])
        assert int(got[0].sum()) > 0


def test_pipeline_steps_overlapped_last_equals_isolated(lib):
    """Pipeline.steps overlaps consecutive realisations (the next field and
    beam during the detectors, gradient and ordering after the previous
    trace); its last realisation equals the same seed run alone."""
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    cfg = Config(n=64, n_rays=20000, pix=128)
    p = Pipeline(cfg)
    p.steps([3, 8, 21, 5])
    got = [h.clone() for h in (p.H_shadow, p.H_knife, p.H_interf)] + [p.field4.clone(), p.s0.clone()]
    q = Pipeline(cfg)
    q.run(5)
    want = [h.clone() for h in (q.H_shadow, q.H_knife, q.H_interf)] + [q.field4.clone(), q.s0.clone()]
    torch.cuda.synchronize()
    for i in (0, 1, 3):
        assert torch.equal(got[i], want[i])
    torch.testing.assert_close(got[2], want[2], rtol=1e-5, atol=1e-3)
    # the ordered rays: same multiset (order within a column is not fixed)
    a = got[4].cpu().numpy()
    b = want[4].cpu().numpy()
    np.testing.assert_array_equal(np.sort(a, axis=1), np.sort(b, axis=1))
