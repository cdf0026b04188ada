"""Drop-in for gaussian_fields/calculate_spectrum_3d.py (3-D estimator).

``spectrum_3D_scalar`` (C3:3-59) on the GPU: the field's forward FFT, |F|^2
and the radial binning run in libtt (tt_power_spectrum) without the host
argsort of the reference; the bin edges and weighted centres are the
reference's.
"""
import numpy as np
import torch

from . import _lib
from .turboGen import FieldPlan


def _device_field(data, device="cuda"):
    t = data if isinstance(data, torch.Tensor) else torch.as_tensor(np.asarray(data))
    if t.dim() != 3:
        raise ValueError(f"expected a 3-D field, got shape {tuple(t.shape)}")
    return t.to(device=device, dtype=torch.float32).contiguous()


def power_bins(data, dx, nbins, mode, bin_w=0.0, device="cuda"):
    """(sums, counts) float64 device tensors of tt_power_spectrum."""
    f = _device_field(data, device)
    m0, m1, m2 = f.shape
    plan = FieldPlan((m2, m1, m0), device)
    sums = torch.zeros(nbins, dtype=torch.float64, device=device)
    counts = torch.zeros(nbins, dtype=torch.float64, device=device)
    plan.lib.tt_power_spectrum(plan.handle, _lib.ptr(f), int(mode), int(nbins), float(dx), float(bin_w),
                               _lib.ptr(plan.workspace), _lib.ptr(sums), _lib.ptr(counts), _lib.stream_ptr())
    torch.cuda.current_stream().synchronize()
    plan.close()
    return sums, counts


def spectrum_3D_scalar(data, dx, k_bin_num=100):
    """Isotropic 3-D spectrum (C3:3-59): returns (k_bins_weighted, spect3D)
    as numpy arrays; spect3D[i] is the mean |fftn|^2 over
    i w <= K < (i + 1) w, w = K.max() / k_bin_num, NaN for an empty bin, and
    the last entry 0, as in the reference."""
    shape = tuple(int(s) for s in np.shape(data))
    f = [(m // 2) * (1.0 / (m * dx)) for m in shape]
    k_bin_width = np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2) / k_bin_num
    sums, counts = power_bins(data, dx, k_bin_num, 0, k_bin_width)
    sums, counts = sums.cpu().numpy(), counts.cpu().numpy()
    k_bins = k_bin_width * np.arange(0, k_bin_num + 1)
    k_bins_weighted = (0.5 * (k_bins[:-1] ** 3 + k_bins[1:] ** 3)) ** (1 / 3)
    spect3D = np.zeros_like(k_bins_weighted)
    with np.errstate(invalid="ignore", divide="ignore"):
        spect3D[:-1] = sums[:-1] / counts[:-1]
    return k_bins_weighted, spect3D
