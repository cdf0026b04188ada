"""Drop-in for gaussian_fields/turboGen.py's FFT generator on the GPU.

``gaussian3D_FFT(N, k_func)`` keeps the reference's signature and draws
(TG:488-538): the noise comes from numpy's global legacy RNG in the same
order and the spectrum from the caller's ``k_func`` evaluated on the host,
exactly as the reference builds them; symmetrisation, shaping, the zero mode
and the inverse FFT run in libtt (tt_field_from_noise).  The result is a
float32 CUDA tensor of shape (M, M, M) indexed like the reference's array
(``as_numpy=True`` returns float64 numpy).

``gaussian_field(shape, ...)`` is the native path: Philox noise and a
parameterised power-law spectrum generated inside the first FFT pass, for
any (even or odd) shape.
"""
import numpy as np
import torch

from . import _lib


class FieldPlan:
    """An FFT plan and its half-spectrum workspace for one grid shape."""

    def __init__(self, shape, device="cuda"):
        self.shape = tuple(int(s) for s in shape)
        nx, ny, nz = self.shape
        self.lib = _lib.load()
        h = _lib.P()
        self.lib.tt_fft_plan_create(nx, ny, nz, _lib.C.byref(h))
        self.handle = h
        nbytes = _lib.C.c_size_t()
        self.lib.tt_fft_workspace_bytes(h, _lib.C.byref(nbytes))
        self.workspace = torch.empty(nbytes.value, dtype=torch.uint8, device=device)
        self.device = device

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.tt_fft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def empty_field(self):
        nx, ny, nz = self.shape
        return torch.empty((nz, ny, nx), dtype=torch.float32, device=self.device)

    def generate(self, beta=-11.0 / 3.0, aniso=1.0, kmin=None, kmax=0.5, amp=1.0, seed=0, out=None, sumsq=None):
        """Native field into ``out`` ([nz][ny][nx] float32); adds the sum of
        squares to ``sumsq`` (1-element float64 tensor) when given."""
        if kmin is None:
            kmin = 1.0 / max(self.shape)
        out = self.empty_field() if out is None else out
        spec = _lib.Spectrum(beta, aniso, kmin, kmax, amp, int(seed) & 0xFFFFFFFFFFFFFFFF)
        self.lib.tt_field_generate(self.handle, _lib.C.byref(spec), _lib.ptr(self.workspace), _lib.ptr(out),
                                   _lib.ptr(sumsq), _lib.stream_ptr())
        return out

    def from_noise(self, wr, wi, sqrt_psd, out=None, sumsq=None):
        """Field from caller noise/spectrum on the centred grid (float32
        CUDA tensors of shape (nx, ny, nz), C order)."""
        out = self.empty_field() if out is None else out
        for t in (wr, wi, sqrt_psd):
            assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and tuple(t.shape) == self.shape
        self.lib.tt_field_from_noise(self.handle, _lib.ptr(wr), _lib.ptr(wi), _lib.ptr(sqrt_psd),
                                     _lib.ptr(self.workspace), _lib.ptr(out), _lib.ptr(sumsq), _lib.stream_ptr())
        return out


_PLANS = {}


def get_plan(shape, device="cuda"):
    key = (tuple(int(s) for s in shape), str(device))
    if key not in _PLANS:
        _PLANS[key] = FieldPlan(shape, device)
    return _PLANS[key]


def ij_view(field_zyx):
    """The [ix, iy, iz] view (reference / ElectronCube indexing) of a
    [z][y][x] device array."""
    return field_zyx.permute(2, 1, 0)


def gaussian_field_from_noise(Wr, Wi, sqrt_psd, device="cuda"):
    """Field from host noise and sqrt(P) on the centred grid (TG:525-538)."""
    shape = Wr.shape
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)
    plan = get_plan(shape, device)
    out = plan.from_noise(dev(Wr), dev(Wi), dev(sqrt_psd))
    return ij_view(out)


def gaussian3D_FFT(N, k_func, rng=None, device="cuda", as_numpy=False):
    """Drop-in for turboGen.gaussian3D_FFT (TG:488-538)."""
    rnd = np.random if rng is None else rng
    M = 2 * N + 1
    k = np.fft.fftfreq(M)
    KX, KY, KZ = np.meshgrid(k, k, k)
    K = np.fft.fftshift(np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2))
    Wr = rnd.randn(M, M, M)
    Wi = rnd.randn(M, M, M)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        S = np.sqrt(k_func(K))
    f = gaussian_field_from_noise(Wr, Wi, S, device=device)
    if as_numpy:
        return f.double().cpu().numpy()
    return f


def gaussian_field(shape, beta=-11.0 / 3.0, aniso=1.0, kmin=None, kmax=0.5, amp=1.0, seed=0, normalise=False,
                   device="cuda"):
    """Native GPU generator for any shape.  Returns the field as an
    [ix, iy, iz] view of a [z][y][x] float32 array and its RMS (a 1-element
    float64 tensor).  ``normalise`` divides the field by its RMS."""
    plan = get_plan(shape, device)
    sumsq = torch.zeros(1, dtype=torch.float64, device=device)
    out = plan.generate(beta, aniso, kmin, kmax, amp, seed, sumsq=sumsq)
    rms = torch.sqrt(sumsq / out.numel())
    if normalise:
        out.div_(rms.float())
    return ij_view(out), rms
