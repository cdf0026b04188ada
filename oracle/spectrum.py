"""FP64 restatement of the reference's radial power-spectrum estimators.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

  * ``spectrum_3D_scalar``  gaussian_fields/calculate_spectrum_3d.py:3-59
    (C3): |fftn|^2, K from fftfreq(M_i, dx), bins of width K.max()/k_bin_num,
    the mean of the points with (i-1) w <= K < i w for i = 1 .. k_bin_num-1
    (the last entry stays 0; an empty bin gives NaN, numpy's mean of nothing).
  * ``compute3Dspectrum``   gaussian_fields/cmpspec.py:147-194 (CS): sums of
    |fftn(r)/nt|^2 into shells round(|k|) over the loop values
    range(-M//2, M//2 - 1) of each axis, divided by knorm, optional moving
    average (CS:35-37).
  * ``binned``              the (sums, counts) form the GPU returns.
Pinned in tests/test_oracle_spectrum.py by Parseval's theorem, single-mode
fields with known power, and the loop form of the shell sum on tiny grids.
"""
import numpy as np


def _k_grid(shape, dx):
    ks = [np.fft.fftfreq(m, dx) for m in shape]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    return np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)


def bin_width(shape, dx, k_bin_num):
    """K.max() / k_bin_num with K as C3:31-36 builds it (the corner point)."""
    f = [(m // 2) * (1.0 / (m * dx)) for m in shape]
    return np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2) / k_bin_num


def binned(data, dx, k_bin_num):
    """(sums, counts) of |fftn|^2 over bins b <= k_bin_num - 2 (C3's loop)."""
    data = np.asarray(data, dtype=np.float64)
    p = np.abs(np.fft.fftn(data)) ** 2
    K = _k_grid(data.shape, dx)
    w = bin_width(data.shape, dx, k_bin_num)
    b = np.floor(K / w).astype(np.int64)
    b = np.where(K < b * w, b - 1, np.where(K >= (b + 1) * w, b + 1, b))
    ok = b <= k_bin_num - 2
    sums = np.bincount(b[ok], weights=p[ok], minlength=k_bin_num)[:k_bin_num]
    counts = np.bincount(b[ok], minlength=k_bin_num)[:k_bin_num].astype(np.float64)
    return sums, counts


def spectrum_3D_scalar(data, dx, k_bin_num=100):
    sums, counts = binned(data, dx, k_bin_num)
    w = bin_width(np.shape(data), dx, k_bin_num)
    k_bins = w * np.arange(0, k_bin_num + 1)
    k_bins_weighted = (0.5 * (k_bins[:-1] ** 3 + k_bins[1:] ** 3)) ** (1 / 3)
    spect = np.zeros(k_bin_num)
    with np.errstate(invalid="ignore", divide="ignore"):
        spect[:-1] = sums[:-1] / counts[:-1]
    return k_bins_weighted, spect


def _loop_values(m):
    """Array index -> the reference's loop value, or None (CS:175-177)."""
    out = {}
    for v in range(-m // 2, m // 2 - 1):
        out[v % m] = v
    return out


def shell_sums(r):
    """(sums, counts) of |fftn(r)/nt|^2 per shell round(|v|) (CS:171-180)."""
    r = np.asarray(r, dtype=np.float64)
    nx, ny, nz = r.shape
    nt = nx * ny * nz
    tkeh = np.abs(np.fft.fftn(r) / nt) ** 2
    lv = [np.array([_loop_values(m).get(j, np.nan) for j in range(m)]) for m in r.shape]
    V0, V1, V2 = np.meshgrid(*lv, indexing="ij")
    ok = ~(np.isnan(V0) | np.isnan(V1) | np.isnan(V2))
    sh = np.rint(np.sqrt(np.where(ok, V0 ** 2 + V1 ** 2 + V2 ** 2, 0.0))).astype(np.int64)
    ok &= sh < nx
    sums = np.bincount(sh[ok], weights=tkeh[ok], minlength=nx)[:nx]
    counts = np.bincount(sh[ok], minlength=nx)[:nx].astype(np.float64)
    return sums, counts


def movingaverage(interval, window_size):
    window = np.ones(int(window_size)) / float(window_size)
    return np.convolve(interval, window, "same")


def compute3Dspectrum(r, lx, ly, lz, smooth):
    nx, ny, nz = np.shape(r)
    sums, _ = shell_sums(r)
    knorm = (2.0 * np.pi / lx + 2.0 * np.pi / ly + 2.0 * np.pi / lz) / 3.0
    wave_numbers = knorm * np.arange(0, nx)
    tke = sums / knorm
    knyquist = knorm * min(nx, ny, nz) / 2
    if smooth:
        sm = movingaverage(tke, 5)
        sm[0:4] = tke[0:4]
        tke = sm
    return knyquist, wave_numbers, tke
