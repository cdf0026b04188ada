"""Batched ensembles: many turbulence realisations of uneven size traced in
one persistent launch, the whole ensemble step captured in one HIP graph.

The reference runs realisations one after another (its drivers loop over
seeds or spread them over processes, particle_tracking/
example_multiprocess.py:41-62, example_MPI.py:82-149), each through
gaussian3D_FFT, calc_dndr, init_beam, solve and the detectors.  Here every
member keeps its own resident buffers (a Pipeline), the per-member stages
(field, gradient, beam, ray order, detectors) are stream-ordered launches, and
the rays of all members go through one persistent tracer launch
(tt_batch_trace) whose blocks take work chunks from a device counter, longest
rays first.  The launch shape, kernel variant and ray-ordering brick come from
the project's autotuner.  After one eager step the whole step is captured in
a torch CUDA graph (a hipGraph on ROCm) and replayed.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .autotune import Autotuner
from .pipeline import Config, Pipeline


@dataclass
class Member:
    n: int          # grid nodes per axis
    beta: float     # spectral index of P(k) = (kx^2 + ky^2 + (kz/a)^2)^(beta/2)
    aniso: float    # a
    rms: float      # RMS of ne / ne0 - 1
    ne0: float      # mean electron density, m^-3
    n_rays: int
    seed: int


def draw_members(count=48, seed=0, sizes=(128, 192, 256, 384), rays=(200_000, 2_000_000)):
    """Seeded draw of an ensemble: grid uniform over ``sizes``, beta uniform in
    [-11/3, -3], anisotropy uniform in [0.5, 2], RMS fluctuation uniform in
    [2%, 20%], mean ne uniform in [1e23, 1e25] m^-3, ray count uniform in
    ``rays``."""
    rs = np.random.RandomState(seed)
    out = []
    for i in range(count):
        out.append(Member(n=int(rs.choice(sizes)), beta=float(rs.uniform(-11 / 3, -3.0)),
                          aniso=float(rs.uniform(0.5, 2.0)), rms=float(rs.uniform(0.02, 0.2)),
                          ne0=float(rs.uniform(1e23, 1e25)), n_rays=int(rs.randint(rays[0], rays[1] + 1)),
                          seed=int(rs.randint(0, 2 ** 31 - 1))))
    return out


TRACE_CANDIDATES = [
    {"block": b, "blocks_per_cu": bpc, "variant": v, "chunk": ch, "brick_shift": bs}
    for v in (0, 1) for b in (128, 256) for bpc in (0, 2) for ch in (2048, 8192) for bs in (0, 2)
]


class Ensemble:
    def __init__(self, members, L=10e-3, lwl=1064e-9, beam_size=2e-3, divergence=50e-6, pix=512, h=0.5,
                 device="cuda"):
        self.members = list(members)
        self.device = device
        self.lib = _lib.load()
        self.h = h
        self.pipes = [Pipeline(Config(n=m.n, L=L, n_rays=m.n_rays, ne0=m.ne0, eps=m.rms, lwl=lwl, beta=m.beta,
                                      aniso=m.aniso, beam_size=beam_size, divergence=divergence, pix=pix, h=h),
                               device) for m in self.members]
        self.tuning = {"block": 256, "blocks_per_cu": 0, "variant": 0, "chunk": 4096, "brick_shift": 0}
        self.ws = None
        self.info = _lib.BatchInfo()
        self.graph = None
        self._prepared_chunk = None
        self.n_side_streams = 4
        self._sides = None

    # ------------------------------------------------------------- stages
    def _items(self):
        arr = (_lib.BatchItem * len(self.pipes))()
        for i, p in enumerate(self.pipes):
            N = p.cfg.n_rays
            # with the status array, the rays the z-march hands over are traced in time in the same call
            arr[i] = _lib.BatchItem(p.field4.data_ptr(), p.grid, p.s0.data_ptr(), N, N, None, p.rf.data_ptr(), N,
                                    p.phase.data_ptr(), p.status.data_ptr(), p.opts.cT)
        return arr

    def prepare(self, chunk=None):
        """Upload the member descriptors and the chunk table (synchronous)."""
        chunk = int(chunk or self.tuning["chunk"])
        items = self._items()
        nbytes = C.c_size_t()
        self.lib.tt_batch_workspace_bytes(items, len(self.pipes), chunk, C.byref(nbytes))
        if self.ws is None or self.ws.numel() < nbytes.value:
            self.ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        torch.cuda.synchronize()
        self.lib.tt_batch_prepare(items, len(self.pipes), self.h, chunk, self.ws.data_ptr(), self.ws.numel(),
                                  C.byref(self.info))
        self._prepared_chunk = chunk

    def generate(self):
        for m, p in zip(self.members, self.pipes):
            p.generate_field(m.seed)
            p.gradient()
            p.beam(m.seed + 1)

    def order(self):
        for p in self.pipes:
            p.cfg.brick_shift = self.tuning["brick_shift"]
            p.sort()

    def trace(self):
        if self._prepared_chunk != self.tuning["chunk"]:
            self.prepare()
        t = self.tuning
        tune = _lib.BatchTuning(t["block"], t["blocks_per_cu"], t["variant"])
        self.lib.tt_batch_trace(self.ws.data_ptr(), C.byref(self.info), C.byref(tune), _lib.stream_ptr())

    def detect(self):
        for p in self.pipes:
            p.detect()

    def _side_streams(self):
        if self._sides is None:
            self._sides = [torch.cuda.Stream() for _ in range(self.n_side_streams)]
        return self._sides

    def step(self):
        """One ensemble step: every member's field, gradient, beam and ray
        order, one batched trace, every member's detectors.  The per-member
        stages are small launches, so members are dealt round-robin to a few
        side streams forked from and joined back into the current stream (in
        a captured graph they become parallel branches); the trace runs on the
        current stream."""
        main = torch.cuda.current_stream()
        sides = self._side_streams()
        for sd in sides:
            sd.wait_stream(main)
        for i, (m, p) in enumerate(zip(self.members, self.pipes)):
            with torch.cuda.stream(sides[i % len(sides)]):
                p.generate_field(m.seed)
                p.gradient()
                p.beam(m.seed + 1)
                p.cfg.brick_shift = self.tuning["brick_shift"]
                p.sort()
        for sd in sides:
            main.wait_stream(sd)
        self.trace()
        for sd in sides:
            sd.wait_stream(main)
        for i, p in enumerate(self.pipes):
            with torch.cuda.stream(sides[i % len(sides)]):
                p.detect()
        for sd in sides:
            main.wait_stream(sd)

    # ------------------------------------------------------------ tuning
    def signature(self):
        return {"members": [(m.n, m.n_rays) for m in self.members], "h": self.h}

    def autotune(self, tuner=None, candidates=None, force=False):
        """Pick the trace launch shape, variant, chunk and ordering brick
        with the autotuner (fields and beams must be generated)."""
        tuner = tuner or Autotuner()

        def run(c):
            self.tuning = dict(c)
            self.order()
            self.trace()

        best = tuner.tune("ensemble_trace", self.signature(), candidates or TRACE_CANDIDATES, run, reps=2,
                          warmup=1, force=force)
        best.pop("rejected", None)
        self.tuning = best
        self.prepare()
        return best

    # ------------------------------------------------------------- graph
    def capture(self):
        """Capture step() in a HIP graph (after an eager warm-up step)."""
        if self._prepared_chunk != self.tuning["chunk"]:
            self.prepare()
        self.step()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError("Ensemble.replay: capture() first")
        self.graph.replay()

    @property
    def ray_steps(self):
        return sum(p.ray_steps for p in self.pipes)

    @property
    def n_rays(self):
        return sum(p.cfg.n_rays for p in self.pipes)
