"""Philox4x32-10 counter-based RNG, numpy twin of the GPU sampler.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

Algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11, section 3 (Philox-4x32 with 10 rounds).  The round
multipliers and Weyl key increments are the published constants.  Pinned by
the paper's known-answer vectors (tests/test_oracle_philox.py).

Conversions used by both the GPU and this twin (bit-identical integer part):
  uniform(x) = ((x >> 8) + 0.5) * 2**-24          in (0, 1), exact in fp32
  box_muller(u1, u2) = sqrt(-2 ln u1) * (cos 2 pi u2, sin 2 pi u2)
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

# stream ids (counter word 2) -- shared with csrc/tt_rng.h
STREAM_FIELD = 0
STREAM_BEAM = 1


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32; all inputs are integer arrays/scalars < 2**32."""
    c0 = np.asarray(c0, dtype=np.uint64) & MASK
    c1 = np.asarray(c1, dtype=np.uint64) & MASK
    c2 = np.asarray(c2, dtype=np.uint64) & MASK
    c3 = np.asarray(c3, dtype=np.uint64) & MASK
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    c0, c1, c2, c3 = c0.copy(), c1.copy(), c2.copy(), c3.copy()
    k0 = np.uint64(int(k0) & 0xFFFFFFFF)
    k1 = np.uint64(int(k1) & 0xFFFFFFFF)
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
    return c0, c1, c2, c3


def uniform(x):
    x = np.asarray(x, dtype=np.uint64)
    return ((x >> np.uint64(8)).astype(np.float64) + 0.5) * 2.0 ** -24


def box_muller(u1, u2):
    r = np.sqrt(-2.0 * np.log(u1))
    a = 2.0 * np.pi * u2
    return r * np.cos(a), r * np.sin(a)


def split_seed(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def field_noise(lin_index, seed):
    """(Wr, Wi) for centred linear voxel indices, as the GPU field generator
    draws them: counter (idx_lo, idx_hi, STREAM_FIELD, 0), key = seed."""
    lin_index = np.asarray(lin_index, dtype=np.uint64)
    k0, k1 = split_seed(seed)
    r = philox4x32(lin_index & MASK, lin_index >> np.uint64(32), STREAM_FIELD, 0, k0, k1)
    return box_muller(uniform(r[0]), uniform(r[1]))


def beam_uniforms(ray_index, seed):
    """The five uniforms and one normal the GPU beam sampler draws per ray.

    Returns (ut, ua, ub, uphi, nchi): ut -> polar angle of the position,
    ua + ub -> radial coordinate (reflected), uphi -> azimuth of the velocity,
    nchi -> standard normal for the polar angle of the velocity."""
    ray_index = np.asarray(ray_index, dtype=np.uint64)
    k0, k1 = split_seed(seed)
    lo, hi = ray_index & MASK, ray_index >> np.uint64(32)
    a = philox4x32(lo, hi, STREAM_BEAM, 0, k0, k1)
    b = philox4x32(lo, hi, STREAM_BEAM, 1, k0, k1)
    n1, _ = box_muller(uniform(b[0]), uniform(b[1]))
    return uniform(a[0]), uniform(a[1]), uniform(a[2]), uniform(a[3]), n1
