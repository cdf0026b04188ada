"""Host-side logic that needs no GPU: the fused inference kernel's weight
packing (checked by emulating its register data flow in numpy), the ensemble
draw, and the autotuner's cache keys."""
import numpy as np

from oracle import surrogate as OS


def test_fused_inference_packing_emulated():
    """Emulate mlp_infer_kernel on one 16-ray tile: MFMA 16x16x32 fragments
    (lane l: A[l%16][8(l/16)+j], B[8(l/16)+j][l%16], D[4(l/16)+i][l%16]) with
    the packed weights, each layer's D tiles reused as the next layer's B
    operand in place.  Equals the plain forward pass."""
    from turbulence_tracing_amd import surrogate as S
    rs = np.random.RandomState(0)
    dims = (48, 256, 256, 256, 8)
    W = [rs.randn(dims[i + 1], dims[i]) / np.sqrt(dims[i]) for i in range(4)]
    b = [rs.randn(dims[i + 1]) * 0.1 for i in range(4)]
    X = rs.randn(16, 48)
    ref, _ = OS.mlp_forward(list(zip(W, b)), X)
    packs = []
    for l in range(4):
        n_idx, k_idx, nt, kc = S._frag_index(l, dims[l + 1])
        Wp = np.zeros((nt * 16, kc * 32))
        Wp[:dims[l + 1], :dims[l]] = W[l]
        packs.append((Wp[n_idx.numpy(), k_idx.numpy()].reshape(nt, kc, 64, 8), nt, kc))
    lane = np.arange(64)
    g, r = lane // 16, lane % 16
    B = np.zeros((2, 64, 8))
    for c in range(2):
        f = 32 * c + 8 * g[:, None] + np.arange(8)[None, :]
        B[c] = np.where(f < 48, X[r[:, None], np.minimum(f, 47)], 0.0)
    for l in range(4):
        P, nt, kc = packs[l]
        bp = np.zeros(nt * 16)
        bp[:dims[l + 1]] = b[l]
        Y = np.zeros((8, 64, 8))
        out = np.zeros((16, 16))
        for t in range(nt):
            D = np.zeros((16, 16))
            for c in range(kc):
                Am = np.zeros((16, 32))
                Bm = np.zeros((32, 16))
                Am[r[:, None], 8 * g[:, None] + np.arange(8)] = P[t, c]
                Bm[8 * g[:, None] + np.arange(8), r[:, None]] = B[c]
                D += Am @ Bm
            for i in range(4):
                v = D[4 * g + i, r] + bp[16 * t + 4 * g + i]
                if l < 3:
                    Y[t >> 1, lane, 4 * (t & 1) + i] = OS.gelu(v)
                else:
                    out[r, 4 * g + i] = v
        B = Y
    np.testing.assert_allclose(out[:, :8], ref, rtol=1e-12, atol=1e-12)


def test_frag_index_covers_every_weight_once():
    from turbulence_tracing_amd import surrogate as S
    for layer, fout, fin in [(0, 256, 48), (1, 256, 256), (3, 8, 256)]:
        n_idx, k_idx, nt, kc = S._frag_index(layer, fout)
        pairs = set(zip(n_idx.tolist(), k_idx.tolist()))
        assert len(pairs) == nt * kc * 512          # no slot written twice
        assert {(n, k) for n in range(fout) for k in range(fin)} <= pairs


def test_ensemble_draw_ranges_and_seed():
    from turbulence_tracing_amd.ensemble import draw_members
    ms = draw_members(200, seed=3)
    assert {m.n for m in ms} == {128, 192, 256, 384}
    assert all(-11 / 3 <= m.beta <= -3 and 0.5 <= m.aniso <= 2 for m in ms)
    assert all(0.02 <= m.rms <= 0.2 and 1e23 <= m.ne0 <= 1e25 for m in ms)
    assert all(200_000 <= m.n_rays <= 2_000_000 for m in ms)
    assert draw_members(48, seed=3) == draw_members(48, seed=3)
    assert draw_members(48, seed=3) != draw_members(48, seed=4)


def test_autotuner_key_separates_problems(tmp_path, monkeypatch):
    from turbulence_tracing_amd import autotune
    monkeypatch.setattr(autotune, "device_key", lambda device=None: "gfx950:256cu")
    t = autotune.Autotuner(str(tmp_path / "c.json"))
    k1 = t.key("ensemble_trace", {"members": [(128, 10)], "h": 0.5})
    k2 = t.key("ensemble_trace", {"h": 0.5, "members": [(128, 10)]})
    k3 = t.key("ensemble_trace", {"members": [(128, 11)], "h": 0.5})
    assert k1 == k2 and k1 != k3 and "gfx950" in k1
    t.cache[k1] = {"best": {"block": 128}, "ms": 1.0, "all": []}
    assert t.tune("ensemble_trace", {"members": [(128, 10)], "h": 0.5}, [{"block": 256}], lambda c: None) == \
        {"block": 128}
