"""One large volume, rays streamed in waves.

A realisation too large to trace in one batch of rays (1024^3 field, 1e8
rays) keeps its field resident in HBM and pushes the beam through it in waves
of a fixed number of rays: each wave draws its rays from the same seeded beam
(Philox counter = global ray index, so the waves together are exactly the rays
of one large draw), orders them by entry column, traces them with the plain
RK4 tracer or the surrogate + fallback form, and adds them into the detector
histograms.  The reference does the same work in one solve over all rays
(particle_tracking/particle_tracker.py:312-331) or split across processes
whose histograms are summed (particle_tracking/example_MPI.py:136-149).
"""
from .pipeline import Config, Pipeline


class WaveRun:
    def __init__(self, cfg: Config, total_rays, wave=10_000_000, device="cuda"):
        self.total = int(total_rays)
        self.wave = int(min(wave, total_rays))
        self.p = Pipeline(Config(**{**cfg.__dict__, "n_rays": self.wave}), device)

    @property
    def n_waves(self):
        return -(-self.total // self.wave)

    def field(self, seed):
        self.p.generate_field(seed)
        self.p.gradient()

    def waves(self):
        for w in range(self.n_waves):
            yield w, w * self.wave, min(self.wave, self.total - w * self.wave)

    def run(self, beam_seed, surrogate=None, t_pos=1e-6, t_ang=1e-6, blocks=None, on_wave=None):
        """Trace all rays in waves into freshly zeroed histograms.  With a
        ``surrogate`` (surrogate.Surrogate) each wave goes through the
        surrogate + fallback form instead of the plain tracer.  ``on_wave(w,
        n)`` is called after each wave is queued (e.g. to read rf / phase).
        Returns the number of rays re-traced by the fallback (0 for plain),
        summed on the device and read once, after the last wave is queued,
        so the waves themselves never wait on the host."""
        import torch
        p = self.p
        p.zero_histograms()
        fallback = torch.zeros((), dtype=torch.int64, device=p.rf.device)
        for w, off, n in self.waves():
            p.beam(beam_seed, n=n, offset=off)
            p.sort(n)
            if surrogate is None:
                p.trace(n)
            else:
                rf, ph, flag = surrogate.trace(p.field4, p.grid, p.s0[:, :n], t_pos=t_pos, t_ang=t_ang,
                                               blocks=blocks)
                p.rf[:, :n].copy_(rf)
                p.phase[:n].copy_(ph)
                fallback += flag.sum()
            p.detect(n, accumulate=True)
            if on_wave is not None:
                on_wave(w, n)
        return int(fallback)

    @property
    def ray_steps(self):
        return self.total * self.p.cfg.steps_per_ray


def large_volume_config(**kw):
    """The large single-volume workload: 1024^3 FP32 field, k^-11/3, 15% RMS,
    mean ne 5e24 m^-3, L = 20 mm, 1e8 rays in an 8 mm disc, dz = 0.5 dx,
    4096^2 detector histograms over 12 mm (the 8 mm beam fits)."""
    base = dict(n=1024, L=20e-3, ne0=5e24, eps=0.15, beam_size=4e-3, pix=4096, det_size=12.0, h=0.5)
    base.update(kw)
    return Config(**base)
