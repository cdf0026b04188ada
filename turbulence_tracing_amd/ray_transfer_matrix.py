"""Drop-in for particle_tracking/ray_transfer_matrix.py on MI355X.

The optics helpers (RTM:37-154) and the detector classes (RTM:156-299) keep
their names and arguments; rays are 4 x N float64 CUDA tensors
[x, theta, y, phi] (numpy accepted).  Each class builds its optical train as
a list of tt_optic ops, ``solve`` applies it in one kernel
(tt_optics_apply) and ``histogram`` bins with np.histogram2d semantics in one
kernel (tt_detect), returning H[iy, ix] as float64 like the reference.

Not in the reference: ``KnifeEdgeSchlieren`` (the Schlieren_DF train with
knife_edge at the Fourier plane) and ``Interferometry`` (a two-beam
interferogram, sum over rays of 1 + cos(phase - tilt_k x_det) imaged through
the shadowgraphy train), plus ``detect_fused``, which runs up to four
detectors over the same rays in a single pass.
"""
import numpy as np
import torch

from . import _lib


def _rays(r, device="cuda"):
    if isinstance(r, torch.Tensor):
        return r.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.asarray(r, dtype=np.float64), device=device).contiguous()


# ------------------------------------------------------------ optical trains
def op_distance(d):
    return (_lib.OP_DISTANCE, 0, float(d), 0.0)


def op_lens(f1, f2=None):
    return (_lib.OP_LENS, 0, float(f1), float(f1 if f2 is None else f2))


def op_aperture(R):
    return (_lib.OP_APERTURE, 0, float(R), 0.0)


def op_stop(R):
    return (_lib.OP_STOP, 0, float(R), 0.0)


def op_knife(offset, axis, direction):
    if direction == 0:
        raise ValueError("Direction must be <0 or >0")  # RTM:151-152 prints and then fails
    a = {"x": 0, "y": 2}[axis]
    return (_lib.OP_KNIFE, a, float(offset), float(np.sign(direction)))


def op_rect(Lx, Ly):
    return (_lib.OP_RECT, 0, float(Lx), float(Ly))


def op_annulus(R1, R2):
    return (_lib.OP_ANNULUS, 0, float(R1), float(R2))


def make_train(ops, to_mm=False):
    if len(ops) > _lib.TT_MAX_OPS:
        raise ValueError(f"optical train has {len(ops)} ops (max {_lib.TT_MAX_OPS})")
    t = _lib.Train()
    t.n_ops = len(ops)
    t.to_mm = 1 if to_mm else 0
    for i, (op, axis, a, b) in enumerate(ops):
        t.ops[i] = _lib.Optic(op, axis, a, b)
    return t


def apply_train(r, ops, to_mm=False, out=None):
    r = _rays(r)
    n = r.shape[1]
    out = torch.empty_like(r) if out is None else out
    t = make_train(ops, to_mm)
    _lib.load().tt_optics_apply(_lib.C.byref(t), _lib.ptr(r), r.stride(0), n, _lib.ptr(out), out.stride(0),
                                _lib.stream_ptr())
    return out


# ---------------------------------------------------- RTM:37-154 helpers
def m_to_mm(r):
    rr = _rays(r).clone()
    rr[0::2, :] *= 1e3
    return rr


def lens(r, f1, f2):
    return apply_train(r, [op_lens(f1, f2)])


def sym_lens(r, f):
    return lens(r, f, f)


def distance(r, d):
    return apply_train(r, [op_distance(d)])


def circular_aperture(r, R):
    return apply_train(r, [op_aperture(R)])


def circular_stop(r, R):
    return apply_train(r, [op_stop(R)])


def annular_stop(r, R1, R2):
    """The boolean filter of rays between R1 and R2 (RTM:89-97)."""
    r = _rays(r)
    r2 = r[0] ** 2 + r[2] ** 2
    return (r2 > R1 ** 2) & (r2 < R2 ** 2)


def angular_filter(r, Rs):
    Rs = np.asarray(Rs, dtype=np.float64)
    return apply_train(r, [op_annulus(Rs[2 * i], Rs[2 * i + 1]) for i in range(len(Rs) // 2)])


def rect_aperture(r, Lx, Ly):
    return apply_train(r, [op_rect(Lx, Ly)])


def knife_edge(r, offset, axis, direction):
    return apply_train(r, [op_knife(offset, axis, direction)])


# ------------------------------------------------------------ histograms
def hist_desc(Lx, Ly, nbx, nby):
    return _lib.Hist(int(nbx), int(nby), -Lx / 2, Lx / 2, -Ly / 2, Ly / 2)


def run_detectors(rf, dets, phase=None):
    """Run up to four (ops, to_mm, Lx, Ly, nbx, nby, mode, tilt_k) detectors
    over rf in one pass; returns a list of histograms H[iy, ix]
    (uint32 counts for mode 0, float32 weight sums for mode 1)."""
    rf = _rays(rf)
    n = rf.shape[1]
    if not 1 <= len(dets) <= _lib.TT_MAX_DETECTORS:
        raise ValueError(f"need 1..{_lib.TT_MAX_DETECTORS} detectors")
    arr = (_lib.Detector * len(dets))()
    outs = []
    for k, (ops, to_mm, Lx, Ly, nbx, nby, mode, tilt_k) in enumerate(dets):
        d = arr[k]
        d.train = make_train(ops, to_mm)
        d.hist = hist_desc(Lx, Ly, nbx, nby)
        d.mode = mode
        d.tilt_k = float(tilt_k)
        if mode == 0:
            H = torch.zeros((nby, nbx), dtype=torch.int32, device=rf.device)
            d.counts = H.data_ptr()
        else:
            H = torch.zeros((nby, nbx), dtype=torch.float32, device=rf.device)
            d.wsum = H.data_ptr()
        outs.append(H)
    ph = None
    if phase is not None:
        ph = phase.to(dtype=torch.float64).contiguous() if isinstance(phase, torch.Tensor) else \
            torch.as_tensor(np.asarray(phase, dtype=np.float64), device=rf.device)
    _lib.load().tt_detect(arr, len(dets), _lib.ptr(rf), rf.stride(0), _lib.ptr(ph), n, _lib.stream_ptr())
    return outs


# ------------------------------------------------------ detector classes
class Rays:
    """Inheritable class for ray diagnostics (RTM:156-206)."""

    def __init__(self, r0, focal_plane=0, L=400, R=25, Lx=18, Ly=13.5):
        self.focal_plane, self.L, self.R, self.Lx, self.Ly = focal_plane, L, R, Lx, Ly
        self.r0 = m_to_mm(r0)

    def _solve(self, ops):
        self.rf = apply_train(self.r0, ops)
        return self.rf

    def histogram(self, bin_scale=10, pix_x=3448, pix_y=2574, clear_mem=False):
        """np.histogram2d of the rays on the detector, H[iy, ix] (RTM:173-195)."""
        nbx, nby = pix_x // bin_scale, pix_y // bin_scale
        H = run_detectors(self.rf, [([], False, self.Lx, self.Ly, nbx, nby, 0, 0.0)])[0]
        self.H = H.double()
        self.xedges = np.linspace(-self.Lx / 2, self.Lx / 2, nbx + 1)
        self.yedges = np.linspace(-self.Ly / 2, self.Ly / 2, nby + 1)
        if clear_mem:
            self.clear_rays()

    def plot(self, ax, clim=None, cmap=None):
        ax.imshow(self.H.cpu().numpy(), interpolation="nearest", origin="lower", clim=clim, cmap=cmap,
                  extent=[self.xedges[0], self.xedges[-1], self.yedges[0], self.yedges[-1]])

    def clear_rays(self):
        self.r0 = None
        self.rf = None

    # The reference's drivers pickle whole diagnostics (example_MPI.py:154-166).
    # Device tensors are stored as numpy arrays, so a pickle loads on a host
    # without a GPU and reads like the reference's (H, rays as numpy).
    def __getstate__(self):
        st = dict(self.__dict__)
        for k, v in st.items():
            if isinstance(v, torch.Tensor):
                st[k] = v.detach().cpu().numpy()
        return st

    def __setstate__(self, st):
        self.__dict__.update(st)


def shadowgraphy_ops(L, R, focal_plane):
    """RTM:214-227."""
    return [op_distance(L - focal_plane), op_aperture(R), op_lens(L), op_distance(2 * L), op_aperture(R),
            op_lens(L), op_distance(L)]


def schlieren_ops(L, R, focal_plane, stop):
    """RTM:236-250 / 259-273 with ``stop`` at the Fourier plane."""
    return [op_distance(L - focal_plane), op_aperture(R), op_lens(L), op_distance(L)] + list(stop) + \
        [op_distance(L), op_aperture(R), op_lens(L), op_distance(L)]


def afr_ops(L, R, focal_plane, Rs):
    """RTM:285-299."""
    Rs = np.asarray(Rs, dtype=np.float64)
    ann = [op_annulus(Rs[2 * i], Rs[2 * i + 1]) for i in range(len(Rs) // 2)]
    return [op_distance(L / 2 - focal_plane), op_aperture(R), op_lens(L / 2), op_distance(L / 4)] + ann + \
        [op_distance(L / 4), op_aperture(R), op_lens(L / 2), op_distance(L / 2)]


class Shadowgraphy(Rays):
    def solve(self):
        return self._solve(shadowgraphy_ops(self.L, self.R, self.focal_plane))


class Schlieren_DF(Rays):
    def solve(self, R=1):
        return self._solve(schlieren_ops(self.L, self.R, self.focal_plane, [op_stop(R)]))


class Schlieren_LF(Rays):
    def solve(self, R=1):
        return self._solve(schlieren_ops(self.L, self.R, self.focal_plane, [op_aperture(R)]))


class AFR(Rays):
    def solve(self, Rs):
        return self._solve(afr_ops(self.L, self.R, self.focal_plane, Rs))


class KnifeEdgeSchlieren(Rays):
    """Schlieren with a knife edge at the Fourier plane (knife_edge, RTM:138-154,
    in the Schlieren_DF train).  Not a class of the reference."""

    def solve(self, offset=0.0, axis="x", direction=1):
        return self._solve(schlieren_ops(self.L, self.R, self.focal_plane, [op_knife(offset, axis, direction)]))


class Interferometry(Rays):
    """Two-beam interferogram imaged through the shadowgraphy train: each ray
    adds 1 + cos(phase - tilt_k x_det) to its pixel (x_det in mm, tilt_k in
    rad/mm).  Not a class of the reference, which tracks no phase."""

    def __init__(self, r0, phase, focal_plane=0, L=400, R=25, Lx=18, Ly=13.5, tilt_k=0.0):
        super().__init__(r0, focal_plane, L, R, Lx, Ly)
        self.phase = phase
        self.tilt_k = tilt_k

    def solve(self):
        return self._solve(shadowgraphy_ops(self.L, self.R, self.focal_plane))

    def histogram(self, bin_scale=10, pix_x=3448, pix_y=2574, clear_mem=False):
        nbx, nby = pix_x // bin_scale, pix_y // bin_scale
        H = run_detectors(self.rf, [([], False, self.Lx, self.Ly, nbx, nby, 1, self.tilt_k)], self.phase)[0]
        self.H = H.double()
        self.xedges = np.linspace(-self.Lx / 2, self.Lx / 2, nbx + 1)
        self.yedges = np.linspace(-self.Ly / 2, self.Ly / 2, nby + 1)
        if clear_mem:
            self.clear_rays()


def detect_fused(rf_m, phase, Lx, Ly, nbx, nby, L=400, R=25, focal_plane=0, knife=(0.0, "x", 1), tilt_k=0.0):
    """Shadowgraph, knife-edge schlieren and interferogram of rays rf (metres,
    radians) in ONE pass: returns (H_shadow uint32, H_knife uint32,
    H_interf float32), each H[iy, ix]."""
    sh = shadowgraphy_ops(L, R, focal_plane)
    ke = schlieren_ops(L, R, focal_plane, [op_knife(*knife)])
    dets = [(sh, True, Lx, Ly, nbx, nby, 0, 0.0), (ke, True, Lx, Ly, nbx, nby, 0, 0.0),
            (sh, True, Lx, Ly, nbx, nby, 1, tilt_k)]
    return run_detectors(rf_m, dets, phase)
