"""MI355X-native turbulence ray tracing: the data-parallel hot path of
jdhare/turbulence_tracing (FFT Gaussian field -> refractive-index gradients ->
ray integration -> shadowgraph / schlieren / interferogram detectors) as
hand-written HIP kernels for gfx950 behind a C ABI (include/tt.h), with the
reference's Python call surface in ``turboGen``, ``particle_tracker`` and
``ray_transfer_matrix``.
"""
__version__ = "0.1.0"
