"""Surrogate kernels against the FP64 oracle: the MFMA GEMM in every stride
mode the MLP uses, MLP forward/backward/Adam, the straight-line estimate,
the trained network's gain over the estimate alone, and the fallback."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import field as OF
from oracle import surrogate as OS
from oracle import tracer as OT

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bf16(a):
    return torch.as_tensor(np.asarray(a), dtype=torch.float32).to(torch.bfloat16)


def f64(t):
    return t.float().cpu().numpy().astype(np.float64)


def gemm(lib, A, sam, sak, B, sbk, sbn, M, N, K, **kw):
    from turbulence_tracing_amd import _lib
    Cf = kw.pop("Cf")
    a = _lib.GemmArgs(A.data_ptr(), sam, sak, B.data_ptr(), sbk, sbn, M, N, K, None, 0, Cf.data_ptr(), None,
                      kw.pop("ldc", N), None, kw.pop("ksplit", 0))
    lib.tt_gemm_bf16(C.byref(a), _lib.stream_ptr())
    torch.cuda.synchronize()


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 72, 48), (1000, 256, 256), (37, 8, 256), (257, 48, 8)])
def test_gemm_all_stride_modes(lib, M, N, K):
    rs = np.random.RandomState(M + N + K)
    A = rs.randn(M, K)
    B = rs.randn(K, N)
    Ab, Bb = bf16(A), bf16(B)
    want = f64(Ab) @ f64(Bb)
    scale = np.abs(want).max()
    # A row-major (k contiguous), B given as [N][K] (k contiguous): forward layout
    A1 = Ab.contiguous().to(DEV)
    B1 = Bb.T.contiguous().to(DEV)
    out = torch.zeros((M, N), dtype=torch.float32, device=DEV)
    gemm(lib, A1, K, 1, B1, 1, K, M, N, K, Cf=out)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=0, atol=1e-5 * scale * np.sqrt(K))
    # A given as [K][M] (m contiguous), B row-major [K][N] (n contiguous): dW layout, split K
    A2 = Ab.T.contiguous().to(DEV)
    B2 = Bb.contiguous().to(DEV)
    out.zero_()
    gemm(lib, A2, 1, M, B2, N, 1, M, N, K, Cf=out, ksplit=32)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=0, atol=1e-5 * scale * np.sqrt(K))
    # element-wise fallback (odd strides): A with a padded leading dimension
    A3 = torch.zeros((M, K + 3), dtype=torch.bfloat16)
    A3[:, :K] = Ab
    A3 = A3.to(DEV)
    out.zero_()
    gemm(lib, A3, K + 3, 1, B2, N, 1, M, N, K, Cf=out)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=0, atol=1e-5 * scale * np.sqrt(K))


def emulate_forward(Ws, bs, X):
    """The GPU's numerics: bf16 weights/inputs, fp32-exact products, bf16
    hidden activations."""
    H = f64(bf16(X))
    cache = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Z = H @ f64(bf16(W)).T + b
        cache.append((H, Z))
        if l < len(Ws) - 1:
            H = f64(bf16(OS.gelu(Z)))
        else:
            H = Z
    return H, cache


def test_mlp_forward_backward_adam_match_oracle(lib):
    from turbulence_tracing_amd.surrogate import MLP
    dims = (48, 256, 256, 256, 8)
    mlp = MLP(dims, seed=3)
    rs = np.random.RandomState(0)
    N = 512
    X = rs.randn(N, 48)
    Ws = [w.cpu().numpy().astype(np.float64) for w in mlp.W]
    bs = [b.cpu().numpy().astype(np.float64) for b in mlp.b]
    Y = mlp.forward(bf16(X).to(DEV), keep=True).cpu().numpy()
    Ye, cache = emulate_forward(Ws, bs, X)
    assert np.sqrt(np.mean((Y - Ye) ** 2) / np.mean(Ye ** 2)) < 2e-3
    G = rs.randn(N, 8) / N
    mlp.backward(bf16(G).to(DEV))
    torch.cuda.synchronize()
    # oracle backward with the GPU's bf16 upstream gradients
    params = [(f64(bf16(W)), b) for W, b in zip(Ws, bs)]
    grads = OS.mlp_backward(params, cache, f64(bf16(G)))
    for l in range(len(Ws)):
        gW = mlp.dW[l].cpu().numpy()
        rel = np.sqrt(np.mean((gW - grads[l][0]) ** 2) / np.mean(grads[l][0] ** 2))
        assert rel < 2e-2, (l, rel)
        gb = mlp.db[l].cpu().numpy()
        relb = np.sqrt(np.mean((gb - grads[l][1]) ** 2) / np.mean(grads[l][1] ** 2))
        assert relb < 2e-2, (l, relb)
    # Adam: one step against the oracle on the GPU's own gradients
    w_before = [w.cpu().numpy().astype(np.float64) for w in mlp.W]
    g_now = [g.cpu().numpy().astype(np.float64) for g in mlp.dW]
    mlp.adam(1e-3)
    for l in range(len(Ws)):
        w1, _, _ = OS.adam_step(w_before[l], g_now[l], 0 * w_before[l], 0 * w_before[l], 1, 1e-3)
        np.testing.assert_allclose(mlp.W[l].cpu().numpy(), w1, rtol=1e-5, atol=1e-7)
        np.testing.assert_array_equal(mlp.Wb[l].float().cpu().numpy(), mlp.W[l].cpu().to(torch.bfloat16).float().numpy())


@pytest.mark.parametrize("n", [1, 127, 128, 1000, 4099])
def test_fused_inference_matches_layerwise_forward(lib, n):
    """tt_mlp_infer (one kernel, transposed MFMA chain in registers) against
    the layer-by-layer GEMM forward on the same standardised bf16 inputs, and
    against the FP64 emulation of the bf16 network."""
    from turbulence_tracing_amd.surrogate import MLP, Norm, Surrogate
    mlp = MLP(seed=5)
    rs = np.random.RandomState(n)
    feat = torch.as_tensor(rs.randn(n, 48) * 2 + 0.5, dtype=torch.float32).to(DEV)
    mu = torch.as_tensor(rs.randn(48) * 0.1, dtype=torch.float32).to(DEV)
    isd = torch.as_tensor(0.3 + rs.rand(48), dtype=torch.float32).to(DEV)
    sur = Surrogate(Norm(mu.cpu().numpy(), isd.cpu().numpy(), 1.0, 1.0, 1.0), mlp=mlp)
    Y = mlp.infer(feat, mu, isd)
    Xs = sur.standardize(feat)
    Yl = mlp.forward(Xs)
    torch.cuda.synchronize()
    scale = float(Yl.pow(2).mean().sqrt())
    assert float((Y - Yl).pow(2).mean().sqrt()) < 2e-3 * scale
    Ws = [w.cpu().numpy().astype(np.float64) for w in mlp.W]
    bs = [b.cpu().numpy().astype(np.float64) for b in mlp.b]
    Ye, _ = emulate_forward(Ws, bs, f64(Xs))
    assert np.sqrt(np.mean((Y.cpu().numpy() - Ye) ** 2)) < 2e-3 * np.sqrt(np.mean(Ye ** 2))
    # rows past n untouched: a padded output keeps its sentinel
    out = torch.full((n + 5, 8), 7.0, device=DEV)
    mlp.infer(feat, mu, isd, out=out[:n])
    assert bool((out[n:] == 7.0).all())


def small_cube(n=33, seed=3):
    x = np.linspace(-5e-3, 5e-3, n)
    d = OF.powerlaw_field((n, n, n), -11 / 3, seed=seed)
    d /= np.sqrt((d ** 2).mean())
    return x, 1e24 * (1 + 0.1 * d)


@pytest.mark.parametrize("k,inside", [(1, False), (4, False), (16, False), (1, True), (4, True), (16, True)])
def test_path_features_match_oracle(lib, k, inside):
    """Block path estimate and features against the FP64 oracle on a
    turbulent 33^3 cube; ``inside`` starts the rays inside the grid off a
    block boundary (single-cell steps from the node field first)."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    from turbulence_tracing_amd.surrogate import path_estimate
    x, ne = small_cube()
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(2000, 2e-3, 50e-6, seed=9)
    if inside:
        s0[2] = -5e-3 + 5.37 * (x[1] - x[0])
    base, feat = path_estimate(cube.field4, cube.grid, s0, k=k)
    ref = OT.Cube(x, x.copy(), x.copy())
    ref.external_ne(ne)
    ref.calc_dndr(lwl=1064e-9)
    ref.extent = ref.extent_z
    bo, fo = OS.path_estimate(OT.NodeGrid(ref), s0.cpu().numpy(), k=k)
    b = base.cpu().numpy()
    dx = x[1] - x[0]
    assert np.abs(b[[0, 2]] - bo[[0, 2]]).max() < 1e-5 * dx
    assert np.abs(b[[1, 3]] - bo[[1, 3]]).max() < 1e-5 * np.abs(bo[[1, 3]]).max()
    assert np.abs(b[4] - bo[4]).max() < 1e-5 * np.abs(bo[4]).max()
    f = feat.cpu().numpy()
    scale = np.abs(fo).max(axis=0) + 1e-30
    assert (np.abs(f - fo).max(axis=0) / scale).max() < 1e-3


@pytest.mark.parametrize("inside", [False, True])
def test_path_k1_node_field_matches_block_fields(lib, inside):
    """At k = 1 the estimate reads the node field (no block fields) and
    evaluates each node plane's bilinear form; it must agree with the k = 1
    block-field path to float rounding, on a wide beam
    whose rays change column many times (a ray within rounding of a column
    boundary may count one column change more or less).  The beam's steep
    rays and those that leave through a side face are handed over (NaN
    estimate) by both paths alike."""
    import torch
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    from turbulence_tracing_amd.surrogate import block_fields, path_estimate
    x, ne = small_cube(n=41, seed=5)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne * 20)
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(20000, 4e-3, 0.1, seed=11)
    if inside:
        s0[2] = -5e-3 + 7.61 * (x[1] - x[0])
    blocks = block_fields(cube.field4, cube.grid, 1)
    b_blk, f_blk = path_estimate(cube.field4, cube.grid, s0, k=1, blocks=blocks)
    b_node, f_node = path_estimate(cube.field4, cube.grid, s0, k=1)
    bb, bn = b_blk.cpu().numpy(), b_node.cpu().numpy()
    fb_ok, fn_ok = np.isfinite(bb).all(axis=0), np.isfinite(bn).all(axis=0)
    assert (fb_ok != fn_ok).sum() <= 10  # a ray within rounding of a side face
    fin = fb_ok & fn_ok
    assert 0.5 < fin.mean() < 1.0
    bb, bn = bb[:, fin], bn[:, fin]
    dx = x[1] - x[0]
    assert np.abs(bn[[0, 2]] - bb[[0, 2]]).max() < 1e-5 * dx
    assert np.abs(bn[[1, 3]] - bb[[1, 3]]).max() < 1e-5 * np.abs(bb[[1, 3]]).max()
    assert np.abs(bn[4] - bb[4]).max() < 1e-5 * np.abs(bb[4]).max()
    fb, fn = f_blk.cpu().numpy(), f_node.cpu().numpy()
    cross = [2 * 16 + 11, 2 * 16 + 12]
    other = [q for q in range(fb.shape[1]) if q not in cross]
    scale = np.abs(fb[:, other]).max(axis=0) + 1e-30
    assert (np.abs(fn[:, other] - fb[:, other]).max(axis=0) / scale).max() < 1e-4
    dc = np.abs(fn[:, cross] - fb[:, cross])
    assert dc.max() <= 1 and (dc > 0).sum() <= 10
    assert fb[:, cross[0]].max() > 3  # column changes along the path
    from turbulence_tracing_amd import _lib
    with pytest.raises(_lib.TTError, match="k > 1"):
        lib.tt_path_features(cube.field4.data_ptr(), None, 4, C.byref(cube.grid), s0.data_ptr(), s0.stride(0), 10,
                             None, f_node.data_ptr(), f_node.stride(1), b_node.data_ptr(), b_node.shape[1], None)


def test_empty_ray_set_through_the_surrogate(lib):
    """An empty ray set goes through the estimate, the network, the
    correction and the surrogate trace as empty results."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    from turbulence_tracing_amd.surrogate import MLP, Norm, Surrogate, path_estimate
    x, ne = small_cube(n=17)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(16, 2e-3, 50e-6, seed=3)[:, :0]
    for k in (1, 4):
        base, feat = path_estimate(cube.field4, cube.grid, s0, k=k)
        assert tuple(base.shape) == (5, 0) and tuple(feat.shape) == (0, 48)
    sur = Surrogate(Norm(np.zeros(48, np.float32), np.ones(48, np.float32), 1e-6, 1e-6, 1e-3), mlp=MLP(seed=2), k=4)
    Y = sur.predict(feat)
    assert tuple(Y.shape) == (0, 8)
    rf, ph, flag = sur.trace(cube.field4, cube.grid, s0)
    torch.cuda.synchronize()
    assert tuple(rf.shape) == (4, 0) and ph.numel() == 0 and flag.numel() == 0


def _train_128(k, seed=100, fields=32, rays=50000, tau=0.9):
    from turbulence_tracing_amd.pipeline import Config
    from turbulence_tracing_amd.surrogate import Surrogate, build_dataset, make_norm, normalise_targets
    cfg = Config(n=128, pix=64)
    feat, resid, _ = build_dataset(fields, rays, cfg, seed=seed, k=k)
    norm = make_norm(feat, resid)
    sur = Surrogate(norm, seed=1, k=k, tau=tau)
    sur.train(feat, normalise_targets(resid, norm), epochs=6, batch=4096, lr=5e-4)
    return sur


def test_surrogate_gain_and_fallback(lib):
    """Trained on rays the tracer generates through 32 seeded 128^3
    realisations, the network corrects the block path estimate (k = 8) on a
    held-out realisation: it cuts the RMS error of positions, angles and
    phase (measured on MI355X: x0.80, x0.70, x0.22).  Flagged rays are
    re-traced by the full tracer, bit for bit; with zero thresholds every ray
    falls back and the result is the tracer's."""
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    from turbulence_tracing_amd.surrogate import errors, path_estimate
    sur = _train_128(8)
    norm = sur.norm
    p = Pipeline(Config(n=128, pix=64, n_rays=200000))
    p.run(seed=999)
    base, f2 = path_estimate(p.field4, p.grid, p.s0, k=8)
    Y = sur.predict(f2)
    rf, ph, flag, lst, cnt = sur.apply(base, Y, t_pos=1.0, t_ang=1.0)
    e_base = errors(torch.stack([base[0], base[1], base[2], base[3]]), base[4], p.rf, p.phase)
    e_mlp = errors(rf, ph, p.rf, p.phase)
    rms = lambda e: float(e.double().pow(2).mean().sqrt())
    gain = [rms(e_mlp[i]) / rms(e_base[i]) for i in range(3)]
    print(f"network / estimate error RMS: position {gain[0]:.3f} angle {gain[1]:.3f} phase {gain[2]:.3f}")
    assert gain[0] < 0.9 and gain[1] < 0.85 and gain[2] < 0.4, gain
    # an angle threshold at the 90th percentile of the network's own error
    # bound flags about a tenth of the rays; those are exactly the tracer's
    t_ang = float(torch.quantile((Y[:100000, 6].abs() * norm.e_ang).double(), 0.9))
    rf3, ph3, flag3 = sur.trace(p.field4, p.grid, p.s0, t_pos=1.0, t_ang=t_ang)
    fl = flag3.bool()
    assert 0 < int(fl.sum()) < p.s0.shape[1]
    assert int(sur.n_fallback) == int(fl.sum())
    np.testing.assert_array_equal(rf3[:, fl].cpu().numpy(), p.rf[:, fl].cpu().numpy())
    np.testing.assert_array_equal(ph3[fl].cpu().numpy(), p.phase[fl].cpu().numpy())
    # the rays cut into batches on two streams: every ray's result is the same
    rf4, ph4, flag4 = sur.trace(p.field4, p.grid, p.s0, t_pos=1.0, t_ang=t_ang, chunks=4)
    assert torch.equal(flag4, flag3)
    assert torch.equal(rf4.nan_to_num(7.0), rf3.nan_to_num(7.0)) and torch.equal(ph4.nan_to_num(7.0), ph3.nan_to_num(7.0))
    # fallback forced from the host: every ray re-traced -> exactly the tracer
    rf2, ph2, flag2 = sur.trace(p.field4, p.grid, p.s0, t_pos=0.0, t_ang=0.0)
    assert int(flag2.sum()) == p.s0.shape[1]
    np.testing.assert_array_equal(rf2.cpu().numpy(), p.rf.cpu().numpy())
    np.testing.assert_array_equal(ph2.cpu().numpy(), p.phase.cpu().numpy())


def test_surrogate_calibrated_fallback_on_harder_realisation(lib):
    """A harder setting, blocks of 32 cells at 128^3, where the bare
    estimate's angle error passes 1 urad on some rays.  The error heads
    (pinball loss, quantile 0.9) are calibrated on one held-out realisation
    so that every ray there whose corrected error exceeds 0.8 urad is
    flagged.  On another held-out realisation every ray whose corrected error
    exceeds the 1 um / 1 urad bar is flagged and re-traced, the surrogate's
    result is then within the bar everywhere, and the flagged rays carry
    larger errors than the rest."""
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    from turbulence_tracing_amd.surrogate import errors, path_estimate, residual
    k = 32
    sur = _train_128(k)
    pc = Pipeline(Config(n=128, pix=64, n_rays=200000))
    pc.run(seed=998)
    bc, fc = path_estimate(pc.field4, pc.grid, pc.s0, k=k)
    cal = sur.calibrate(fc, residual(pc.rf, pc.phase, bc), 1e-6, 1e-6, margin=0.8)
    print("calibration:", cal)
    assert cal["ang"]["n_over_margin"] > 0 and cal["ang"]["max_unflagged"] <= 0.8e-6
    p = Pipeline(Config(n=128, pix=64, n_rays=200000))
    p.run(seed=999)
    base, f2 = path_estimate(p.field4, p.grid, p.s0, k=k)
    eb = errors(torch.stack([base[0], base[1], base[2], base[3]]), base[4], p.rf, p.phase)
    assert int((eb[1] > 1e-6).sum()) > 0  # the bare estimate breaks the bar here
    Y = sur.predict(f2)
    rfn, phn, _, _, _ = sur.apply(base, Y, 1.0, 1.0)
    en = errors(rfn, phn, p.rf, p.phase)
    rf, ph, flag = sur.trace(p.field4, p.grid, p.s0, t_pos=1e-6, t_ang=1e-6)
    fl = flag.bool()
    over = (en[0] > 1e-6) | (en[1] > 1e-6)
    print(f"bare over bar {int((eb[1] > 1e-6).sum())}, corrected over bar {int(over.sum())}, "
          f"flagged {float(fl.double().mean()):.3f}")
    assert bool(fl[over].all())
    ea = errors(rf, ph, p.rf, p.phase)
    assert float(ea[0].max()) <= 1e-6 and float(ea[1].max()) <= 1e-6
    np.testing.assert_array_equal(rf[:, fl].cpu().numpy(), p.rf[:, fl].cpu().numpy())
    if bool((~fl).any()):
        assert float(en[1][fl].double().mean()) > float(en[1][~fl].double().mean())


def test_surrogate_loss_pinball_matches_numpy(lib):
    """tt_surrogate_loss: weighted squared error on the residual columns and
    the pinball loss at quantile tau on the error heads, whose target is max
    |T - Y| over each head's columns; gradients / rows in bf16."""
    from turbulence_tracing_amd import _lib
    rs = np.random.RandomState(8)
    rows, n_reg, n_pad, tau = 1000, 5, 8, 0.9
    Y = rs.randn(rows, n_pad).astype(np.float32)
    T = rs.randn(rows, n_reg).astype(np.float32)
    w = np.array([1, 1, 1, 1, 0.5], np.float32)
    lo, hi = np.array([0, 2], np.int32), np.array([2, 4], np.int32)
    Yd, Td, wd = (torch.as_tensor(a, device=DEV) for a in (Y, T, w))
    lod, hid = torch.as_tensor(lo, device=DEV), torch.as_tensor(hi, device=DEV)
    dY = torch.empty((rows, n_pad), dtype=torch.bfloat16, device=DEV)
    loss = torch.zeros(2, dtype=torch.float64, device=DEV)
    lib.tt_surrogate_loss(Yd.data_ptr(), n_pad, Td.data_ptr(), n_reg, dY.data_ptr(), n_pad, rows, n_reg, wd.data_ptr(),
                          2, lod.data_ptr(), hid.data_ptr(), tau, n_pad, loss.data_ptr(), _lib.stream_ptr())
    torch.cuda.synchronize()
    e = Y[:, :n_reg].astype(np.float64) - T
    want_d = np.zeros((rows, n_pad))
    want_d[:, :n_reg] = 2 * w * e / rows
    pin = 0.0
    for h in range(2):
        tgt = np.abs(T[:, lo[h]:hi[h]] - Y[:, lo[h]:hi[h]]).max(1)
        d = Y[:, n_reg + h].astype(np.float64) - tgt
        pin += np.sum(np.where(d > 0, (1 - tau) * d, -tau * d))
        want_d[:, n_reg + h] = np.where(d > 0, 1 - tau, -tau) / rows
    got = dY.float().cpu().numpy()
    np.testing.assert_allclose(got, want_d, rtol=1e-2, atol=1e-2 / rows)
    l = loss.cpu().numpy()
    np.testing.assert_allclose(l[0], np.sum(w * e ** 2), rtol=1e-5)
    np.testing.assert_allclose(l[1], pin, rtol=1e-5)


def test_counted_trace_and_wave_ordered_list(lib):
    """tt_surrogate_apply lists the flagged rays with one slot range per
    wave (ray order inside each wave); tt_trace_rk4_counted traces exactly
    the first *n_dev listed rays, bit for bit as tt_trace_rk4 on that list,
    and leaves the other rays' outputs untouched."""
    import ctypes
    from turbulence_tracing_amd import _lib
    from turbulence_tracing_amd.surrogate import MLP, Norm, Surrogate, path_estimate
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    x, ne = small_cube(n=33)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(20000, 2e-3, 50e-6, seed=7)
    n = s0.shape[1]
    sur = Surrogate(Norm(np.zeros(48, np.float32), np.ones(48, np.float32), 1e-6, 1e-6, 1e-3), mlp=MLP(seed=2), k=4)
    base, feat = path_estimate(cube.field4, cube.grid, s0, k=4)
    Y = sur.predict(feat)
    t_ang = float(torch.quantile((Y[:, 6].abs() * sur.norm.e_ang).double(), 0.7))
    rf, ph, flag, lst, cnt = sur.apply(base, Y, 1.0, t_ang)
    torch.cuda.synchronize()
    m = int(cnt.item())
    fl = flag.bool().cpu().numpy()
    L = lst[:m].cpu().numpy()
    assert m == fl.sum() and np.array_equal(np.sort(L), np.flatnonzero(fl))
    # each wave's flagged rays sit in consecutive slots, in ray order
    waves = L // 64
    for wv in np.unique(waves):
        pos = np.flatnonzero(waves == wv)
        assert np.array_equal(pos, np.arange(pos[0], pos[0] + pos.size))
        assert np.all(np.diff(L[pos]) > 0)
    opts = _lib.TraceOpts(0.5, 0.0, 0)
    want_rf = torch.full((4, n), 7.0, dtype=torch.float64, device=DEV)
    want_ph = torch.full((n,), 7.0, dtype=torch.float64, device=DEV)
    lib.tt_trace_rk4(cube.field4.data_ptr(), ctypes.byref(cube.grid), ctypes.byref(opts), s0.data_ptr(), n, m,
                     lst.data_ptr(), want_rf.data_ptr(), n, want_ph.data_ptr(), None, n, None, _lib.stream_ptr())
    got_rf = torch.full((4, n), 7.0, dtype=torch.float64, device=DEV)
    got_ph = torch.full((n,), 7.0, dtype=torch.float64, device=DEV)
    lib.tt_trace_rk4_counted(cube.field4.data_ptr(), ctypes.byref(cube.grid), ctypes.byref(opts), s0.data_ptr(), n,
                             n, lst.data_ptr(), cnt.data_ptr(), got_rf.data_ptr(), n, got_ph.data_ptr(), None,
                             _lib.stream_ptr())
    torch.cuda.synchronize()
    assert torch.equal(got_rf, want_rf) and torch.equal(got_ph, want_ph)
    assert bool((got_ph[torch.as_tensor(~fl, device=DEV)] == 7.0).all())


def test_surrogate_hands_over_like_the_plain_trace(lib):
    """Rays the full trace hands over to the time integrator (every 5th ray
    launched at 0.3 rad; every 7th other one started half a cell inside the
    +x face and heading out) get a NaN estimate, so the surrogate flags them
    whatever its network says, and its fallback re-traces them as the plain
    trace does (z-march, then the time integrator): rf and phase bit for bit,
    status 3."""
    from turbulence_tracing_amd import _lib
    from turbulence_tracing_amd.particle_tracker import ElectronCube, c
    from turbulence_tracing_amd.surrogate import MLP, Norm, Surrogate
    x, ne = small_cube(n=33)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    n = 3000
    s0 = cube.init_beam(n, 2e-3, 50e-6, seed=13).clone()
    steep = torch.zeros(n, dtype=torch.bool, device=DEV)
    steep[::5] = True
    side = torch.zeros(n, dtype=torch.bool, device=DEV)
    side[3::7] = True
    side &= ~steep
    dx = float(x[1] - x[0])
    s0[3, steep] = 0.3 * c
    s0[4, steep] = 0.0
    s0[5, steep] = float(np.sqrt(1.0 - 0.09)) * c
    s0[0, side] = float(x.max()) - 0.5 * dx
    s0[3, side] = 0.05 * c
    s0[4, side] = 0.0
    s0[5, side] = float(np.sqrt(1.0 - 0.0025)) * c
    opts = _lib.TraceOpts(0.5, float(np.sqrt(8.0) * x.max()), 0)
    want_rf = torch.empty((4, n), dtype=torch.float64, device=DEV)
    want_ph = torch.empty(n, dtype=torch.float64, device=DEV)
    want_st = torch.empty(n, dtype=torch.int32, device=DEV)
    args = (cube.field4.data_ptr(), C.byref(cube.grid), C.byref(opts), s0.data_ptr(), n, n, None,
            want_rf.data_ptr(), n, want_ph.data_ptr(), None, n, want_st.data_ptr())
    assert lib.tt_trace_rk4(*args, _lib.stream_ptr()) == 0
    assert lib.tt_trace_time(*args, 1, _lib.stream_ptr()) == 0
    sur = Surrogate(Norm(np.zeros(48, np.float32), np.ones(48, np.float32), 1e-6, 1e-6, 1e-3), mlp=MLP(seed=2), k=4)
    rf, ph, fl = sur.trace(cube.field4, cube.grid, s0, t_pos=1.0, t_ang=1.0)  # bounds flag nothing by themselves
    torch.cuda.synchronize()
    sel = steep | side
    assert bool((want_st[sel] == 3).all()), "the plain trace hands these rays over"
    assert bool(fl[sel].bool().all())
    assert torch.equal(rf[:, sel], want_rf[:, sel]) and torch.equal(ph[sel], want_ph[sel])
    assert bool((sur.status[sel] == 3).all())
    assert bool((sur.status[~fl.bool()] == 0).all())


def test_surrogate_trace_graph_capture_replays(lib):
    """Surrogate.trace waits on nothing from the host, so it can be
    captured in a graph: three replays, with eager work and
    synchronisations between them, each equal the eager result."""
    from turbulence_tracing_amd.particle_tracker import ElectronCube
    from turbulence_tracing_amd.surrogate import MLP, Norm, Surrogate, block_fields, path_estimate
    x, ne = small_cube(n=33)
    cube = ElectronCube(x, x.copy(), x.copy())
    cube.external_ne(ne)
    cube.calc_dndr(lwl=1064e-9)
    s0 = cube.init_beam(30000, 2e-3, 50e-6, seed=11)
    sur = Surrogate(Norm(np.zeros(48, np.float32), np.ones(48, np.float32), 1e-6, 1e-6, 1e-3), mlp=MLP(seed=5), k=4)
    blocks = block_fields(cube.field4, cube.grid, 4)
    # a threshold that flags about 30% of the rays of this untrained network
    _, feat = path_estimate(cube.field4, cube.grid, s0, k=4, blocks=blocks)
    t_ang = float(torch.quantile((sur.predict(feat)[:, 6].abs() * sur.norm.e_ang).double(), 0.7))
    rf0, ph0, fl0 = sur.trace(cube.field4, cube.grid, s0, t_pos=1.0, t_ang=t_ang, blocks=blocks)
    torch.cuda.synchronize()
    assert 0 < int(fl0.sum()) < s0.shape[1]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up on a side stream, as capture requires
        sur.trace(cube.field4, cube.grid, s0, t_pos=1.0, t_ang=t_ang, blocks=blocks)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        rf_g, ph_g, fl_g = sur.trace(cube.field4, cube.grid, s0, t_pos=1.0, t_ang=t_ang, blocks=blocks)
    for _ in range(3):
        junk = torch.randn(1 << 20, device=DEV).sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(fl_g, fl0)
        assert torch.equal(rf_g.nan_to_num(7.0), rf0.nan_to_num(7.0)) and torch.equal(ph_g, ph0)
        del junk
