"""CPU FP64 oracle for the turbulence-tracing hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference's algorithms
(jdhare/turbulence_tracing) in plain numpy/scipy so that the HIP kernels of
``turbulence_tracing_amd`` can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline; the product path
never calls into it and has no CPU fallback.

Modules and what pins each one (reference paths are relative to the
reference repository root):

* ``philox``  -- Philox4x32-10 counter RNG twin of the GPU sampler.  Pinned by
  the published known-answer vectors of Salmon et al. (SC'11).
* ``field``   -- ``gaussian3D_FFT`` (gaussian_fields/turboGen.py:488-538) and
  its shape-general form.  Pinned by golden vectors written from the
  reference (tests/golden), Hermitian => real output, zero mean, and a
  single-mode closed form.
* ``tracer``  -- ``ElectronCube`` calc_dndr / dndr / dsdt / init_beam / solve /
  ray_at_exit (particle_tracking/particle_tracker.py:125-419), the tight
  per-ray integrator, and the fixed-step z-march RK4 that the GPU tracer
  follows.  Pinned by golden vectors, null and slab closed forms, and the
  tight-tolerance solve_ivp.
* ``rtm``     -- ray-transfer optics and detectors
  (particle_tracking/ray_transfer_matrix.py:37-299) plus the interferogram.
  Pinned by golden vectors, the 4f closed forms and DF + LF = shadowgraph.
* ``mlp``     -- FP64 forward/backward/Adam of the surrogate MLP.  Pinned by
  finite differences and torch autograd on the CPU.
"""
