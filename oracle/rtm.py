"""FP64 restatement of the ray-transfer-matrix optics and detectors.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

Reference: particle_tracking/ray_transfer_matrix.py (RTM below).
  m_to_mm RTM:37-40, lens/sym_lens RTM:42-60, distance RTM:62-71,
  circular_aperture RTM:73-79, circular_stop RTM:81-87, annular_stop
  RTM:89-97, angular_filter RTM:99-111, rect_aperture RTM:128-136,
  knife_edge RTM:138-154, Rays RTM:156-206, Shadowgraphy RTM:208-227,
  Schlieren_DF RTM:229-250, Schlieren_LF RTM:252-273, AFR RTM:275-299.
Not in the reference: the knife-edge schlieren (the Schlieren_DF train with
knife_edge in place of circular_stop) and the interferogram (a weighted
histogram of 1 + cos(phase - k_t x_det) over the shadowgraphy train).
"""
import numpy as np


def m_to_mm(r):
    rr = np.ndarray.copy(r)
    rr[0::2, :] *= 1e3
    return rr


def lens(r, f1, f2):
    L = np.zeros((4, 4))
    L[:2, :2] = [[1, 0], [-1 / f1, 1]]
    L[2:, 2:] = [[1, 0], [-1 / f2, 1]]
    return np.matmul(L, r)


def sym_lens(r, f):
    return lens(r, f, f)


def distance(r, d):
    D = np.zeros((4, 4))
    D[:2, :2] = [[1, d], [0, 1]]
    D[2:, 2:] = [[1, d], [0, 1]]
    return np.matmul(D, r)


def circular_aperture(r, R):
    filt = r[0, :] ** 2 + r[2, :] ** 2 > R ** 2
    r[:, filt] = None
    return r


def circular_stop(r, R):
    filt = r[0, :] ** 2 + r[2, :] ** 2 < R ** 2
    r[:, filt] = None
    return r


def annular_stop(r, R1, R2):
    filt1 = r[0, :] ** 2 + r[2, :] ** 2 > R1 ** 2
    filt2 = r[0, :] ** 2 + r[2, :] ** 2 < R2 ** 2
    return filt1 & filt2


def angular_filter(r, Rs):
    f = np.zeros((r.shape[1]), dtype=bool)
    for i in range(0, len(Rs) // 2):
        f += annular_stop(r, Rs[2 * i], Rs[2 * i + 1])
    r[:, f] = None
    return r


def rect_aperture(r, Lx, Ly):
    filt = (r[0, :] ** 2 > Lx ** 2) * (r[2, :] ** 2 > Ly ** 2)
    r[:, filt] = None
    return r


def knife_edge(r, offset, axis, direction):
    a = {"y": 2, "x": 0}[axis]
    if direction > 0:
        filt = r[a, :] > offset
    elif direction < 0:
        filt = r[a, :] < offset
    else:
        raise ValueError("Direction must be <0 or >0")
    r[:, filt] = None
    return r


def histogram(rf, Lx, Ly, bin_scale=10, pix_x=3448, pix_y=2574, weights=None):
    """Rays.histogram (RTM:173-191): NaN rays dropped, np.histogram2d over
    [-Lx/2, Lx/2] x [-Ly/2, Ly/2], then transposed to H[iy, ix]."""
    x = rf[0, :]
    y = rf[2, :]
    keep = ~np.isnan(x) & ~np.isnan(y)
    w = None if weights is None else weights[keep]
    H, xe, ye = np.histogram2d(x[keep], y[keep], bins=[pix_x // bin_scale, pix_y // bin_scale],
                               range=[[-Lx / 2, Lx / 2], [-Ly / 2, Ly / 2]], weights=w)
    return H.T, xe, ye


def shadowgraphy(r0_m, focal_plane=0, L=400, R=25):
    r = m_to_mm(r0_m)
    r = distance(r, L - focal_plane)
    r = circular_aperture(r, R)
    r = sym_lens(r, L)
    r = distance(r, L * 2)
    r = circular_aperture(r, R)
    r = sym_lens(r, L)
    return distance(r, L)


def schlieren_train(r0_m, stop, focal_plane=0, L=400, R=25):
    """The Schlieren_DF / Schlieren_LF train with ``stop`` applied at the
    Fourier plane (RTM:238-250, 261-273)."""
    r = m_to_mm(r0_m)
    r = distance(r, L - focal_plane)
    r = circular_aperture(r, R)
    r = sym_lens(r, L)
    r = distance(r, L)
    r = stop(r)
    r = distance(r, L)
    r = circular_aperture(r, R)
    r = sym_lens(r, L)
    return distance(r, L)


def schlieren_df(r0_m, Rstop=1, **kw):
    return schlieren_train(r0_m, lambda r: circular_stop(r, Rstop), **kw)


def schlieren_lf(r0_m, Rstop=1, **kw):
    return schlieren_train(r0_m, lambda r: circular_aperture(r, Rstop), **kw)


def schlieren_knife(r0_m, offset=0.0, axis="x", direction=1, **kw):
    return schlieren_train(r0_m, lambda r: knife_edge(r, offset, axis, direction), **kw)


def afr(r0_m, Rs, focal_plane=0, L=400, R=25):
    r = m_to_mm(r0_m)
    r = distance(r, L / 2 - focal_plane)
    r = circular_aperture(r, R)
    r = sym_lens(r, L / 2)
    r = distance(r, L / 4)
    r = angular_filter(r, Rs)
    r = distance(r, L / 4)
    r = circular_aperture(r, R)
    r = sym_lens(r, L / 2)
    return distance(r, L / 2)


def interferogram_weights(rf_mm, phase, tilt_k):
    """Two-beam interferogram weight per ray, 1 + cos(phase - k_t x_det),
    with x_det in mm and k_t in rad/mm."""
    return 1.0 + np.cos(phase - tilt_k * rf_mm[0])
