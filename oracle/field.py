"""FP64 restatement of the FFT Gaussian random field generator.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

Follows gaussian_fields/turboGen.py:488-538 (``gaussian3D_FFT``):
  M = 2N+1 (TG:515); k = fftfreq(M) (TG:516); K = |k| on an 'xy' meshgrid,
  fftshift-ed (TG:518-520); Wr, Wi = randn(M,M,M) twice from the global
  legacy RNG (TG:522-523); Wr + flip(Wr), Wi - flip(Wi) (TG:525-526);
  F = W * sqrt(k_func(K)) (TG:528-530); ifftshift (TG:532); zero mode
  (TG:534); ifftn(...).real (TG:536-538).

Shape-general form (used for even sizes, which the reference cannot make):
the centred-grid partner of index c on an axis of length n is
p(c) = (2*(n//2) - c) mod n.  For odd n this is exactly n-1-c, i.e. np.flip,
so the general form reduces bit-for-bit to the reference; for even n it is
the index of the negated frequency, which keeps the spectrum Hermitian.

Array convention: axis 0, 1, 2 = x, y, z of the tracer ('ij' indexing, as
ElectronCube.external_ne expects, particle_tracking/particle_tracker.py:212).
"""
import numpy as np

from . import philox


def partner_index(n):
    """Centred-grid index of the negated frequency, per axis."""
    c = np.arange(n)
    return (2 * (n // 2) - c) % n


def centred_freqs(shape):
    """Per-axis frequency (cycles/sample) on the centred (fftshift-ed) grid."""
    return [np.fft.fftshift(np.fft.fftfreq(n)) for n in shape]


def shaped_spectrum(Wr, Wi, sqrt_psd):
    """Symmetrise the centred-grid noise, scale by sqrt(P), ifftshift, zero
    the mean.  Returns the FFT-ordered complex spectrum (TG:525-534)."""
    shape = Wr.shape
    pidx = np.ix_(*[partner_index(n) for n in shape])
    W = (Wr + Wr[pidx]) + 1j * (Wi - Wi[pidx])
    with np.errstate(invalid="ignore", over="ignore"):
        F = W * sqrt_psd
    F_shift = np.fft.ifftshift(F)
    F_shift[(0,) * len(shape)] = 0
    return F_shift


def gaussian_field(Wr, Wi, sqrt_psd):
    """Real field = ifftn(shaped spectrum).real (TG:536-538), FP64."""
    return np.fft.ifftn(shaped_spectrum(Wr, Wi, sqrt_psd)).real


def gaussian3D_FFT(N, k_func, rng=None):
    """Reference-exact generator (TG:488-538).  ``rng`` defaults to the
    global legacy numpy RNG, which is what the reference draws from."""
    rnd = np.random if rng is None else rng
    M = 2 * N + 1
    k = np.fft.fftfreq(M)
    KX, KY, KZ = np.meshgrid(k, k, k)
    K = np.fft.fftshift(np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2))
    Wr = rnd.randn(M, M, M)
    Wi = rnd.randn(M, M, M)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        S = np.sqrt(k_func(K))
    return gaussian_field(Wr, Wi, S)


def K_centred(shape, aniso=1.0):
    """|k| on the centred grid with 'ij' axes; kz divided by the anisotropy."""
    fx, fy, fz = centred_freqs(shape)
    KX, KY, KZ = np.meshgrid(fx, fy, fz, indexing="ij")
    return np.sqrt(KX ** 2 + KY ** 2 + (KZ / aniso) ** 2)


def sqrt_powerlaw(shape, beta, aniso=1.0, kmin=None, kmax=0.5, amp=1.0):
    """sqrt(P) for P = amp * Ka**beta on kmin <= Ka <= kmax, zero elsewhere,
    with Ka = sqrt(kx^2 + ky^2 + (kz/aniso)^2) in cycles/sample.  kmin defaults
    to 1/max(shape) (the box scale 2 pi / L).  Zero mode gives 0."""
    if kmin is None:
        kmin = 1.0 / max(shape)
    Ka = K_centred(shape, aniso)
    band = (Ka >= kmin) & (Ka <= kmax) & (Ka > 0)
    out = np.zeros(shape)
    out[band] = np.sqrt(amp * Ka[band] ** beta)
    return out


def philox_noise(shape, seed):
    """Centred-grid noise (Wr, Wi) exactly as the GPU generator draws it:
    voxel (cx, cy, cz) uses counter index (cx*ny + cy)*nz + cz."""
    nx, ny, nz = shape
    lin = np.arange(nx * ny * nz, dtype=np.uint64)
    wr, wi = philox.field_noise(lin, seed)
    return wr.reshape(shape), wi.reshape(shape)


def powerlaw_field(shape, beta, seed, aniso=1.0, kmin=None, kmax=0.5, amp=1.0):
    """FP64 twin of the GPU's native (Philox) field generator."""
    Wr, Wi = philox_noise(shape, seed)
    return gaussian_field(Wr, Wi, sqrt_powerlaw(shape, beta, aniso, kmin, kmax, amp))


def radial_psd(field, nbins=None):
    """Shell-averaged |FFT|^2 versus |k| (cycles/sample): a validation tool
    for the spectrum slope (the role of gaussian_fields/calculate_spectrum_3d.py:3-59)."""
    F = np.fft.fftn(field)
    P = np.abs(F) ** 2
    fx, fy, fz = [np.fft.fftfreq(n) for n in field.shape]
    KX, KY, KZ = np.meshgrid(fx, fy, fz, indexing="ij")
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2).ravel()
    if nbins is None:
        nbins = min(field.shape) // 2
    edges = np.linspace(0, 0.5, nbins + 1)
    which = np.digitize(K, edges) - 1
    ok = (which >= 0) & (which < nbins)
    s = np.bincount(which[ok], weights=P.ravel()[ok], minlength=nbins)
    c = np.bincount(which[ok], minlength=nbins)
    centres = 0.5 * (edges[1:] + edges[:-1])
    with np.errstate(invalid="ignore"):
        return centres, s / np.maximum(c, 1), c
