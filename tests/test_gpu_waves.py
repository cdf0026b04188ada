"""Streaming a large volume's beam in waves: the waves together give exactly
the rays, exit states and histogram counts of one single-batch run through the
same field (the beam is counter-based, so wave k holds rays [k w, (k+1) w) of
the same draw); the surrogate form with every ray flagged gives the plain
tracer's result."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def unsort(p, n):
    o = p.order[:n].long()
    rf = torch.empty((4, n), dtype=torch.float64, device=p.rf.device)
    rf[:, o] = p.rf[:, :n]
    ph = torch.empty(n, dtype=torch.float64, device=p.rf.device)
    ph[o] = p.phase[:n]
    return rf.cpu().numpy(), ph.cpu().numpy()


def collect(run):
    parts = []

    def on_wave(w, n):
        parts.append(unsort(run.p, n))
    return parts, on_wave


@pytest.mark.parametrize("total,wave", [(25000, 10000), (7, 3), (4096, 4096)])
def test_waves_equal_single_batch(lib, total, wave):
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    from turbulence_tracing_amd.waves import WaveRun
    cfg = Config(n=48, pix=64, L=20e-3, ne0=5e24, eps=0.15, beam_size=4e-3, det_size=12.0)
    run = WaveRun(cfg, total, wave)
    run.field(11)
    parts, on_wave = collect(run)
    run.run(12, on_wave=on_wave)
    torch.cuda.synchronize()
    rf_w = np.concatenate([a for a, _ in parts], axis=1)
    ph_w = np.concatenate([b for _, b in parts])
    one = Pipeline(Config(**{**cfg.__dict__, "n_rays": total}))
    one.generate_field(11)
    one.gradient()
    one.beam(12)
    one.sort()
    one.trace()
    one.detect()
    rf1, ph1 = unsort(one, total)
    np.testing.assert_array_equal(rf_w, rf1)
    np.testing.assert_array_equal(ph_w, ph1)
    np.testing.assert_array_equal(run.p.H_shadow.cpu().numpy(), one.H_shadow.cpu().numpy())
    np.testing.assert_array_equal(run.p.H_knife.cpu().numpy(), one.H_knife.cpu().numpy())
    np.testing.assert_allclose(run.p.H_interf.cpu().numpy(), one.H_interf.cpu().numpy(), rtol=1e-5, atol=1e-3)
    assert int(run.p.H_shadow.sum()) > 0


def test_waves_surrogate_all_flagged_is_plain(lib):
    from turbulence_tracing_amd.pipeline import Config
    from turbulence_tracing_amd.surrogate import MLP, Norm, Surrogate
    from turbulence_tracing_amd.waves import WaveRun
    cfg = Config(n=48, pix=64)
    run = WaveRun(cfg, 9000, 4000)
    run.field(3)
    parts, on_wave = collect(run)
    run.run(4, on_wave=on_wave)
    plain = [(a.copy(), b.copy()) for a, b in parts]
    H = run.p.H_shadow.cpu().numpy().copy()
    sur = Surrogate(Norm(np.zeros(48, np.float32), np.ones(48, np.float32), 1e-6, 1e-6, 1e-3), MLP(seed=0), k=4)
    parts2, on_wave2 = collect(run)
    nfb = run.run(4, surrogate=sur, t_pos=0.0, t_ang=0.0, on_wave=on_wave2)
    assert nfb == 9000
    for (a, b), (c, d) in zip(plain, parts2):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, d)
    np.testing.assert_array_equal(H, run.p.H_shadow.cpu().numpy())
