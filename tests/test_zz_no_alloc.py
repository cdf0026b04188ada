"""tt_ray_sort_noalloc (the ray ordering that allocates nothing: its column
counts live in the library's static device scratch) against tt_ray_sort with
a caller-allocated scratch buffer, through the C ABI, on two calls in a row,
the second smaller."""
import ctypes as C

import numpy as np
import pytest
import torch

from turbulence_tracing_amd import _lib


def _noalloc(lib):
    f = lib.dll.tt_ray_sort_noalloc
    f.restype = C.c_int
    f.argtypes = [_lib.P, _lib.I64, _lib.I64, C.POINTER(_lib.Grid), _lib.I32, _lib.P, _lib.P]
    return f


def _keys(s0, grid, shift):
    """The entry-column Morton key of each ray (the sort's key)."""
    fx = np.floor((s0[0] - grid.x0) / grid.dx)
    fy = np.floor((s0[1] - grid.y0) / grid.dy)
    ix = np.clip(np.nan_to_num(fx, nan=0.0), 0, grid.nx - 1).astype(np.int64) >> shift
    iy = np.clip(np.nan_to_num(fy, nan=0.0), 0, grid.ny - 1).astype(np.int64) >> shift
    k = np.zeros_like(ix)
    for bit in range(16):
        k |= ((ix >> bit) & 1) << (2 * bit) | ((iy >> bit) & 1) << (2 * bit + 1)
    return k


def _by_key(order, keys):
    """(key, ray) pairs of an ordering: the order within one key (column) is
    not fixed between runs (atomic slots), so orderings are compared as the
    set of rays of each key, in key order."""
    k = keys[order]
    assert np.all(np.diff(k) >= 0)
    i = np.lexsort((order, k))
    return k[i], order[i]


@pytest.mark.gpu
def test_zz_noalloc_sort_equals_scratch_sort(lib):
    f = _noalloc(lib)
    cases = [(512, 10e-3, 300_000, 4e-3, 0), (65, 10e-3, 777, 3e-3, 1)]  # the second call is smaller
    for n_nodes, L, n_rays, spread, shift in cases:
        print(f"case: grid {n_nodes}^3, {n_rays} rays, brick_shift {shift}", flush=True)
        dx = L / (n_nodes - 1)
        grid = _lib.Grid(n_nodes, n_nodes, n_nodes, -L / 2, -L / 2, -L / 2, dx, dx, dx)
        rs = np.random.RandomState(n_rays)
        s0 = np.concatenate([rs.uniform(-spread, spread, (2, n_rays)), rs.randn(4, n_rays)])
        s0[0, :3] = [np.nan, -1.0, 1.0]  # NaN and off-grid rays key to the clamped columns
        t = torch.as_tensor(s0, device="cuda")
        nbytes = C.c_size_t()
        lib.tt_ray_sort_scratch_bytes(C.byref(grid), shift, C.byref(nbytes))
        scratch = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
        want = torch.full((n_rays,), -1, dtype=torch.int32, device="cuda")
        lib.tt_ray_sort(_lib.ptr(t), n_rays, n_rays, C.byref(grid), shift, _lib.ptr(scratch), _lib.ptr(want),
                        _lib.stream_ptr())
        got = torch.full((n_rays,), -1, dtype=torch.int32, device="cuda")
        rc = f(_lib.ptr(t), n_rays, n_rays, C.byref(grid), shift, _lib.ptr(got), _lib.stream_ptr())
        assert rc == 0, lib.dll.tt_last_error().decode()
        torch.cuda.synchronize()
        g = got.long().cpu().numpy()
        w = want.long().cpu().numpy()
        assert np.array_equal(np.sort(g), np.arange(n_rays))
        keys = _keys(s0, grid, shift)
        kg, og = _by_key(g, keys)
        kw, ow = _by_key(w, keys)
        np.testing.assert_array_equal(kg, kw)
        np.testing.assert_array_equal(og, ow)


def test_zz_noalloc_sort_rejects_grids_beyond_its_scratch():
    """More than 2^20 entry-column bins (a grid of 2049 nodes per axis at
    brick_shift 0) is an argument error raised before any device work; the
    same grid at brick_shift 2 (512 column bricks per axis) fits."""
    lib = _lib.load()
    f = _noalloc(lib)
    grid = _lib.Grid(2049, 2049, 8, 0, 0, 0, 1, 1, 1)
    rc = f(C.c_void_p(16), 4, 4, C.byref(grid), 0, C.c_void_p(16), None)
    assert rc == _lib.C.c_int(1).value
    assert "static scratch" in lib.dll.tt_last_error().decode()
    # an empty ray set at a size that fits touches no device memory
    assert f(None, 0, 0, C.byref(grid), 2, None, None) == 0
