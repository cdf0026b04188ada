"""The spectrum oracle pinned without the reference: Parseval's theorem,
single Fourier modes of known power in known bins / shells, and the
reference's loop form of the shell sum on tiny grids."""
import numpy as np
import pytest

from oracle import spectrum as S


def mode_field(shape, k, amp=1.0):
    idx = np.meshgrid(*[np.arange(m) for m in shape], indexing="ij")
    ph = sum(2 * np.pi * ki * ii / m for ki, ii, m in zip(k, idx, shape))
    return amp * np.cos(ph)


def test_parseval_all_power_in_bins():
    rs = np.random.RandomState(0)
    # band-limited: nothing beyond half the grid's largest |k|
    F = np.fft.fftn(rs.randn(16, 12, 10))
    K = S._k_grid((16, 12, 10), 1.0)
    F[K > 0.3] = 0
    f = np.fft.ifftn(F).real
    sums, counts = S.binned(f, 1.0, 50)
    assert np.isclose(sums.sum(), f.size * (f ** 2).sum(), rtol=1e-12)
    assert counts.sum() == np.count_nonzero(K < 49 * S.bin_width((16, 12, 10), 1.0, 50))


@pytest.mark.parametrize("shape,k", [((16, 16, 16), (3, 0, 0)), ((15, 12, 9), (2, 1, 1)), ((8, 10, 12), (0, 0, 4))])
def test_single_mode_bin_and_power(shape, k):
    dx = 0.5
    f = mode_field(shape, k, amp=2.0)
    sums, counts = S.binned(f, dx, 40)
    kk = np.sqrt(sum((ki / (m * dx)) ** 2 for ki, m in zip(k, shape)))
    b = int(np.floor(kk / S.bin_width(shape, dx, 40)))
    n = np.prod(shape)
    # power (amp n / 2)^2 at +k and at -k, all in bin b
    assert np.isclose(sums[b], 2 * (2.0 * n / 2) ** 2, rtol=1e-9)
    assert np.isclose(sums.sum(), sums[b], rtol=1e-9)
    kw, sp = S.spectrum_3D_scalar(f, dx, 40)
    assert np.isclose(sp[b], sums[b] / counts[b]) and sp[-1] == 0.0


def test_shell_sum_matches_loop_form_and_known_mode():
    rs = np.random.RandomState(1)
    for shape in [(6, 5, 7), (8, 8, 8), (5, 4, 6)]:
        r = rs.randn(*shape)
        nx, ny, nz = shape
        tkeh = np.abs(np.fft.fftn(r) / r.size) ** 2
        ref = np.zeros(nx)
        for kx in range(-nx // 2, nx // 2 - 1):
            for ky in range(-ny // 2, ny // 2 - 1):
                for kz in range(-nz // 2, nz // 2 - 1):
                    k = int(np.round(np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)))
                    if k < nx:
                        ref[k] += tkeh[kx, ky, kz]
        sums, _ = S.shell_sums(r)
        np.testing.assert_allclose(sums, ref, rtol=1e-12, atol=1e-15)
    # a cosine at k = (2, 0, 0) on 8^3: power 1/4 at +-k, both visited
    # (index 3 = 8 // 2 - 1 is the one never visited), shell 2
    sums, _ = S.shell_sums(mode_field((8, 8, 8), (2, 0, 0)))
    assert np.isclose(sums[2], 0.5) and np.isclose(sums.sum(), 0.5)
    # k = (3, 0, 0): +3 is the unvisited index, only -3 counts
    sums, _ = S.shell_sums(mode_field((8, 8, 8), (3, 0, 0)))
    assert np.isclose(sums[3], 0.25) and np.isclose(sums.sum(), 0.25)
