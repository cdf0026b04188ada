"""One realisation of the hot path on the GPU, with every buffer resident.

field (tt_field_generate: Philox noise x sqrt P(k), inverse 3-D FFT, sum of
squares) -> node field (tt_grad_stencil with the unit-RMS normalisation and
ne = ne0 (1 + eps delta) mapping applied on the device) and its plane forms
(tt_plane_forms) -> beam
(tt_beam_sample) -> ray order (tt_ray_sort, rays gathered into entry-column
order) -> trace (tt_trace_rk4, then tt_trace_time for the rays it hands
over) -> shadowgraph, knife-edge schlieren and
interferogram histograms in one pass (tt_detect).

This composes, on one device and without host round trips, what a user of
the reference composes by hand: gaussian3D_FFT (TG:488-538), normalisation
and the density map (caller code), ElectronCube.external_ne / calc_dndr /
init_beam / solve (PT:212-331) and the ray_transfer_matrix detectors
(RTM:156-250).
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .particle_tracker import NC_COEFF, c
from .ray_transfer_matrix import make_train, hist_desc, op_knife, schlieren_ops, shadowgraphy_ops
from .turboGen import FieldPlan


@dataclass
class Config:
    n: int = 512                 # grid nodes per axis
    L: float = 10e-3             # cube side, m (grid spans [-L/2, L/2])
    n_rays: int = 10_000_000
    ne0: float = 1e24            # mean electron density, m^-3
    eps: float = 0.1             # RMS fluctuation of ne / ne0
    lwl: float = 1064e-9         # probe wavelength, m
    beta: float = -11.0 / 3.0    # spectral index of P(k)
    aniso: float = 1.0
    beam_size: float = 2e-3      # beam radius, m
    divergence: float = 50e-6    # rad
    pix: int = 2048              # detector bins per axis
    det_size: float = 6.0        # detector side, mm
    L_optics: float = 400.0      # mm
    R_optics: float = 25.0       # mm
    fringes: float = 20.0        # interferogram fringes across the detector
    h: float = 0.5               # RK4 step, cells
    variant: int = 0             # tracer kernel: 0 cell-cached, 1 per-stage gathers
    brick_shift: int = 0         # ray order: entry columns grouped in bricks of 2^shift cells
    # build the plane-form field (64 B per node, tt_plane_forms) for the
    # tracer's cell cache: it saves ~10% of the trace's instructions (187
    # instead of 210 per wave-step) but costs a pass writing 4x the node field
    # and the trace's loads of it; measured slower overall both for a 1e7-ray
    # trace of a 512^3 field and for the 1024^3 waves, hence off by default
    plane_forms: bool = False

    @property
    def nc(self):
        return NC_COEFF * (2 * np.pi * c / self.lwl) ** 2

    @property
    def steps_per_ray(self):
        return int(round((self.n - 1) / self.h))


class Pipeline:
    def __init__(self, cfg: Config, device="cuda"):
        self.cfg = cfg
        self.device = device
        self.lib = _lib.load()
        n = cfg.n
        dx = cfg.L / (n - 1)
        self.grid = _lib.Grid(n, n, n, -cfg.L / 2, -cfg.L / 2, -cfg.L / 2, dx, dx, dx)
        self.plan = FieldPlan((n, n, n), device)
        f32, f64 = torch.float32, torch.float64
        self.field = torch.empty((n, n, n), dtype=f32, device=device)
        self.sumsq = torch.zeros(1, dtype=f64, device=device)
        self.field4 = torch.empty((n, n, n, 4), dtype=f32, device=device)
        # per node the bilinear form of its z-plane cell (tt_plane_forms): the
        # tracer loads two of these per cell instead of forming them
        self.forms = torch.empty((n, n, n, 16), dtype=f32, device=device) if cfg.plane_forms else None
        N = cfg.n_rays
        self.s0_raw = torch.empty((6, N), dtype=f64, device=device)
        self.s0 = torch.empty((6, N), dtype=f64, device=device)
        self.order = torch.empty(N, dtype=torch.int32, device=device)
        self.staging = torch.empty((N, 8), dtype=f64, device=device)  # one 64-byte record per ray
        nbytes = _lib.C.c_size_t()
        self.lib.tt_ray_sort_scratch_bytes(_lib.C.byref(self.grid), 0, _lib.C.byref(nbytes))
        self.sort_scratch = torch.empty(nbytes.value, dtype=torch.uint8, device=device)
        self.rf = torch.empty((4, N), dtype=f64, device=device)
        self.phase = torch.empty(N, dtype=f64, device=device)
        self.status = torch.empty(N, dtype=torch.int32, device=device)
        P = cfg.pix
        self.H_shadow = torch.zeros((P, P), dtype=torch.int32, device=device)
        self.H_knife = torch.zeros((P, P), dtype=torch.int32, device=device)
        self.H_interf = torch.zeros((P, P), dtype=f32, device=device)
        self.opts = _lib.TraceOpts(cfg.h, float(np.sqrt(8.0) * cfg.L / 2), cfg.variant,
                                   None if self.forms is None else self.forms.data_ptr())
        self._dets = self._detectors()
        self._side = None  # streams and events of step(), made on first use

    def _detectors(self):
        cfg = self.cfg
        sh = shadowgraphy_ops(cfg.L_optics, cfg.R_optics, 0.0)
        ke = schlieren_ops(cfg.L_optics, cfg.R_optics, 0.0, [op_knife(0.0, "x", 1)])
        tilt = 2 * np.pi * cfg.fringes / cfg.det_size
        arr = (_lib.Detector * 3)()
        for k, (ops, mode, tk, H) in enumerate([(sh, 0, 0.0, self.H_shadow), (ke, 0, 0.0, self.H_knife),
                                                (sh, 1, tilt, self.H_interf)]):
            arr[k].train = make_train(ops, to_mm=True)
            arr[k].hist = hist_desc(cfg.det_size, cfg.det_size, cfg.pix, cfg.pix)
            arr[k].mode = mode
            arr[k].tilt_k = tk
            if mode == 0:
                arr[k].counts = H.data_ptr()
            else:
                arr[k].wsum = H.data_ptr()
        return arr

    # individual stages (all asynchronous on the current stream) -------------
    def generate_field(self, seed):
        cfg = self.cfg
        self.sumsq.zero_()
        self.plan.generate(cfg.beta, cfg.aniso, 1.0 / cfg.n, 0.5, 1.0, seed, out=self.field, sumsq=self.sumsq)

    def gradient(self):
        cfg = self.cfg
        n = cfg.n
        alpha = cfg.ne0 / cfg.nc
        beta = cfg.ne0 * cfg.eps / cfg.nc
        self.lib.tt_grad_stencil(_lib.ptr(self.field), 0, 1, n, n * n, _lib.C.byref(self.grid), alpha, beta, 1.0,
                                 2 * np.pi / cfg.lwl, _lib.ptr(self.sumsq), n ** 3, _lib.ptr(self.field4),
                                 _lib.stream_ptr())
        if self.forms is not None:
            self.lib.tt_plane_forms(_lib.ptr(self.field4), _lib.C.byref(self.grid), _lib.ptr(self.forms),
                                    _lib.stream_ptr())

    def beam(self, seed, n=None, offset=0):
        """Rays [offset, offset + n) of the seeded beam (Philox counter =
        ray index), so waves of one beam are the rays of one large draw."""
        cfg = self.cfg
        n = cfg.n_rays if n is None else n
        self.lib.tt_beam_sample(n, cfg.beam_size, cfg.divergence, cfg.L / 2, 2, int(seed), int(offset),
                                _lib.ptr(self.s0_raw), cfg.n_rays, _lib.stream_ptr())

    def sort(self, n=None):
        N = self.cfg.n_rays
        n = N if n is None else n
        self.lib.tt_ray_sort_rays(_lib.ptr(self.s0_raw), N, n, _lib.C.byref(self.grid), self.cfg.brick_shift,
                                  _lib.ptr(self.sort_scratch), _lib.ptr(self.order), _lib.ptr(self.staging),
                                  _lib.ptr(self.s0), N, _lib.stream_ptr())

    def trace(self, n=None):
        """The z-march over rays [0, n), then the rays it hands over (status
        1: steep, near turning, or crossing a side face) traced in time."""
        N = self.cfg.n_rays
        n = N if n is None else n
        self.lib.tt_trace_rk4(_lib.ptr(self.field4), _lib.C.byref(self.grid), _lib.C.byref(self.opts),
                              _lib.ptr(self.s0), N, n, None, _lib.ptr(self.rf), N, _lib.ptr(self.phase), None, N,
                              _lib.ptr(self.status), _lib.stream_ptr())
        self.lib.tt_trace_time(_lib.ptr(self.field4), _lib.C.byref(self.grid), _lib.C.byref(self.opts),
                               _lib.ptr(self.s0), N, n, None, _lib.ptr(self.rf), N, _lib.ptr(self.phase), None, N,
                               _lib.ptr(self.status), 1, _lib.stream_ptr())

    def zero_histograms(self):
        self.H_shadow.zero_()
        self.H_knife.zero_()
        self.H_interf.zero_()

    def detect(self, n=None, accumulate=False):
        """Bin rays [0, n) into the three histograms (added to the current
        counts when accumulate is true)."""
        N = self.cfg.n_rays
        n = N if n is None else n
        if not accumulate:
            self.zero_histograms()
        self.lib.tt_detect(self._dets, 3, _lib.ptr(self.rf), N, _lib.ptr(self.phase), n, _lib.stream_ptr())

    def run(self, seed=0):
        """One realisation: field seed ``seed``, beam seed ``seed + 1``."""
        self.step(seed)

    def step(self, seed=0, trace_events=None):
        """One realisation with the independent stages on two side streams:
        field + gradient on one, beam + ordering on the other, then trace and
        detectors on the current stream, which waits for both.  The side
        streams first wait for everything already queued on the current
        stream (the previous step's detectors, and anything the caller queued
        to read this pipeline's buffers).  ``trace_events`` (two torch events)
        bracket the trace kernel."""
        self.steps([seed], None if trace_events is None else [trace_events])

    def steps(self, seeds, trace_events=None):
        """Realisations for ``seeds`` back to back, each as in step(), with
        consecutive ones overlapped: the next field and beam start while the
        previous detectors run, and only the gradient and the ordering (which
        overwrite the node field and the ordered rays) wait for the previous
        trace.  Only the last realisation's results remain; nothing else may
        read this pipeline's buffers until the call has been queued.
        ``trace_events`` holds one pair of torch events per seed."""
        import torch
        main = torch.cuda.current_stream()
        if self._side is None:
            self._side = (torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Event(), torch.cuda.Event(),
                          torch.cuda.Event())
        sa, sb, traced, grad_done, sorted_done = self._side
        traced.record(main)  # everything queued so far, before the first realisation's writes
        sa.wait_event(traced)
        sb.wait_event(traced)
        for i, seed in enumerate(seeds):
            with torch.cuda.stream(sa):
                self.generate_field(seed)
                sa.wait_event(traced)  # the node field is free once the previous trace is done
                self.gradient()
                grad_done.record(sa)
            with torch.cuda.stream(sb):
                self.beam(seed + 1)
                sb.wait_event(traced)  # likewise the ordered rays
                self.sort()
                sorted_done.record(sb)
            main.wait_event(grad_done)
            main.wait_event(sorted_done)
            if trace_events is not None:
                trace_events[i][0].record(main)
            self.trace()
            if trace_events is not None:
                trace_events[i][1].record(main)
            traced.record(main)
            self.detect()
        # later work on the caller's stream is ordered after the detectors;
        # the side streams wait for it at the next call

    @property
    def ray_steps(self):
        return self.cfg.n_rays * self.cfg.steps_per_ray
