"""Batched ensembles: the persistent multi-realisation trace against one
tt_trace_rk4 call per member (bit for bit, every launch shape, rays handed
over to the time integrator included), the captured
ensemble step against the eager one over several replays, and the autotuner's
choice and cache."""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def small_members():
    from turbulence_tracing_amd.ensemble import Member
    sizes = [33, 48, 64, 40, 33]
    rays = [3000, 20000, 7001, 1, 12345]
    rs = np.random.RandomState(4)
    return [Member(n=n, beta=float(rs.uniform(-11 / 3, -3)), aniso=float(rs.uniform(0.5, 2)),
                   rms=float(rs.uniform(0.02, 0.2)), ne0=float(rs.uniform(1e23, 1e25)), n_rays=r, seed=100 + i)
            for i, (n, r) in enumerate(zip(sizes, rays))]


def snapshot(ens):
    """Per member: rf and phase in the beam's own ray order (the ray order
    the counting sort gives within one entry column is not fixed between
    runs), the two count histograms, and the interferogram (a float sum in
    atomic order)."""
    torch.cuda.synchronize()
    out = []
    for p in ens.pipes:
        o = p.order.long()
        rf = torch.empty_like(p.rf)
        rf[:, o] = p.rf
        ph = torch.empty_like(p.phase)
        ph[o] = p.phase
        out.append((rf.cpu().numpy(), ph.cpu().numpy(), p.H_shadow.cpu().numpy().copy(),
                    p.H_knife.cpu().numpy().copy(), p.H_interf.cpu().numpy().copy()))
    return out


def assert_same(a, b):
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_allclose(a[4], b[4], rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("variant", [0, 1])
def test_batch_trace_matches_per_member(lib, variant):
    from turbulence_tracing_amd.ensemble import Ensemble
    ens = Ensemble(small_members(), pix=64)
    ens.generate()
    ens.order()
    ref = []
    for p in ens.pipes:
        p.opts.variant = variant
        p.trace()
        ref.append((p.rf.cpu().numpy().copy(), p.phase.cpu().numpy().copy()))
    for block, bpc, chunk in [(256, 0, 4096), (128, 2, 64), (64, 1, 1000)]:
        for p in ens.pipes:
            p.rf.fill_(7.0)
            p.phase.fill_(7.0)
        ens.tuning = {"block": block, "blocks_per_cu": bpc, "variant": variant, "chunk": chunk, "brick_shift": 0}
        ens.trace()
        torch.cuda.synchronize()
        for (rf, ph), p in zip(ref, ens.pipes):
            np.testing.assert_array_equal(p.rf.cpu().numpy(), rf)
            np.testing.assert_array_equal(p.phase.cpu().numpy(), ph)


def test_batch_trace_hands_over_like_per_member(lib):
    """Every 7th ray of every member launched at 0.3 rad (W = 0.954): the
    z-march hands them over (status 1) and the batched call traces them in
    time, as tt_trace_rk4 + tt_trace_time do per member: rf, phase and status
    bit for bit, over two launch shapes."""
    from turbulence_tracing_amd.ensemble import Ensemble
    from turbulence_tracing_amd.particle_tracker import c
    ens = Ensemble(small_members(), pix=64)
    ens.generate()
    ens.order()
    for p in ens.pipes:
        p.s0[3, ::7] = 0.3 * c
        p.s0[4, ::7] = 0.0
        p.s0[5, ::7] = float(np.sqrt(1.0 - 0.09)) * c
    ref = []
    for p in ens.pipes:
        p.trace()
        ref.append((p.rf.cpu().numpy().copy(), p.phase.cpu().numpy().copy(), p.status.cpu().numpy().copy()))
    for rf, ph, st in ref:
        assert (st[::7] == 3).all() and np.isfinite(rf[:, ::7]).all(), "steep rays traced in time"
        assert set(np.unique(st).tolist()) <= {0, 3}
    for block, bpc, chunk in [(256, 0, 4096), (128, 2, 64)]:
        for p in ens.pipes:
            p.rf.fill_(7.0)
            p.phase.fill_(7.0)
            p.status.fill_(-1)
        ens.tuning = {"block": block, "blocks_per_cu": bpc, "variant": 0, "chunk": chunk, "brick_shift": 0}
        ens.trace()
        torch.cuda.synchronize()
        for (rf, ph, st), p in zip(ref, ens.pipes):
            np.testing.assert_array_equal(p.rf.cpu().numpy(), rf)
            np.testing.assert_array_equal(p.phase.cpu().numpy(), ph)
            np.testing.assert_array_equal(p.status.cpu().numpy(), st)


def test_ensemble_graph_replays_equal_eager(lib):
    from turbulence_tracing_amd.ensemble import Ensemble
    ens = Ensemble(small_members(), pix=64)
    ens.step()
    eager = snapshot(ens)
    ens.capture()
    for rep in range(3):
        for p in ens.pipes:  # eager work between replays clobbers every output
            p.rf.fill_(0.0)
            p.H_shadow.fill_(-1)
        torch.cuda.synchronize()
        ens.replay()
        got = snapshot(ens)
        for a, b in zip(eager, got):
            assert_same(a, b)
    # the step conserves rays: every finite ray inside the detector is counted
    assert all(int(p.H_shadow.sum()) > 0 for p in ens.pipes if p.cfg.n_rays > 100)


def test_autotune_choice_and_cache(lib, tmp_path):
    from turbulence_tracing_amd.autotune import Autotuner
    from turbulence_tracing_amd.ensemble import Ensemble
    ens = Ensemble(small_members(), pix=64)
    ens.generate()
    cands = [{"block": 128, "blocks_per_cu": 0, "variant": 0, "chunk": 1024, "brick_shift": 0},
             {"block": 256, "blocks_per_cu": 0, "variant": 1, "chunk": 4096, "brick_shift": 1},
             {"block": 512, "blocks_per_cu": 0, "variant": 0, "chunk": 4096, "brick_shift": 0}]  # rejected
    path = str(tmp_path / "tune.json")
    tuner = Autotuner(path)
    best = ens.autotune(tuner, cands)
    assert best in cands[:2]
    rec = json.load(open(path))
    entry = list(rec.values())[0]
    assert len(entry["all"]) == 3 and entry["all"][2]["ms"] is None
    # a second tuner reads the cache and runs nothing
    calls = []
    t2 = Autotuner(path)
    got = t2.tune("ensemble_trace", ens.signature(), cands, lambda c: calls.append(c))
    assert got == best and not calls
    # the tuned ensemble traces what the default one does for the same variant and order
    ens.step()
    tuned = snapshot(ens)
    ens.tuning = {**best, "block": 256, "blocks_per_cu": 0, "chunk": 4096}
    ens.step()
    for a, b in zip(tuned, snapshot(ens)):
        assert_same(a, b)


def test_full_ensemble_batch_equals_per_member(lib):
    """The ensemble workload at its real size (the seeded draw of 48 members,
    128-384^3, 2e5-2e6 rays each, ~51 M rays resident): the one persistent
    batched launch gives every member's exit rays and phases bit for bit as
    one trace per member (the z-march, then the rays it hands over traced in
    time), and nearly every ray leaves finite."""
    from turbulence_tracing_amd.ensemble import Ensemble, draw_members
    members = draw_members(48, 0)
    assert len(members) == 48 and sum(m.n_rays for m in members) > 40_000_000
    ens = Ensemble(members, pix=512)
    ens.generate()
    ens.order()
    ens.trace()
    torch.cuda.synchronize()
    batched = [(p.rf.clone(), p.phase.clone()) for p in ens.pipes]
    for (rf, ph), p in zip(batched, ens.pipes):
        p.rf.fill_(7.0)
        p.phase.fill_(7.0)
        p.trace()
        torch.cuda.synchronize()
        assert torch.equal(p.rf.nan_to_num(3.0), rf.nan_to_num(3.0))
        assert torch.equal(p.phase.nan_to_num(3.0), ph.nan_to_num(3.0))
        assert float(torch.isfinite(ph).double().mean()) > 0.999
