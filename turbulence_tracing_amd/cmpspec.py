"""Drop-in for the 3-D estimator of gaussian_fields/cmpspec.py.

``compute3Dspectrum`` (CS:147-194) with the shell sum on the GPU
(tt_power_spectrum mode 1); ``movingaverage`` (CS:35-37) on the host.
"""
import numpy as np

from .calculate_spectrum_3d import power_bins


def movingaverage(interval, window_size):
    window = np.ones(int(window_size)) / float(window_size)
    return np.convolve(interval, window, "same")


def compute3Dspectrum(r, lx, ly, lz, smooth):
    """Returns (knyquist, wave_numbers, tke_spectrum) as the reference:
    tke_spectrum[k] = sum of |fftn(r)/nt|^2 over the loop values of
    range(-M//2, M//2 - 1) per axis with round(|k|) = k, divided by knorm."""
    nx, ny, nz = (int(s) for s in np.shape(r))
    sums, _ = power_bins(r, 1.0, nx, 1)
    k0x, k0y, k0z = 2.0 * np.pi / lx, 2.0 * np.pi / ly, 2.0 * np.pi / lz
    knorm = (k0x + k0y + k0z) / 3.0
    wave_numbers = knorm * np.arange(0, nx)
    tke_spectrum = sums.cpu().numpy() / knorm
    knyquist = knorm * min(nx, ny, nz) / 2
    if smooth:
        tkespecsmooth = movingaverage(tke_spectrum, 5)
        tkespecsmooth[0:4] = tke_spectrum[0:4]
        tke_spectrum = tkespecsmooth
    return knyquist, wave_numbers, tke_spectrum
