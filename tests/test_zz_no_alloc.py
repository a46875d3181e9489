"""sph_sort_noalloc, the context-free sort that allocates nothing (its
working memory is a fixed device array of the library), against sph_sort
on a context: sorted (key, index) pairs and every cell's (first, last) slot
bit for bit, on two calls in a row into the same output buffers, the second
with fewer particles; sizes beyond the array's maximum are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return hip


def _alloc(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p.value


def _read(hip, ptr, nbytes, dtype):
    out = np.zeros(nbytes // np.dtype(dtype).itemsize, dtype)
    assert hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0  # device to host
    return out


def test_sort_noalloc_matches_sph_sort(sph, tables256, impact):
    hip = _hip()
    prm = sph.default_params()
    ncells = 1 << (3 * prm.cell_power)
    n_big, n_small = 100000, 20000
    pairs = _alloc(hip, 8 * n_big)
    cells = _alloc(hip, 8 * ncells)
    try:
        for n, seed in ((n_big, 41), (n_small, 42)):
            print(f"[no-alloc] n={n} seed={seed}", flush=True)
            ctx = sph.Context(prm, impact(n, seed), tables256)
            ctx.sort()
            ctx.sync()
            pos_ptr, nb = ctx.device_field("pos")
            assert nb == 16 * n
            sph.sort_noalloc(pos_ptr, n, prm.cell_width, prm.cell_power, pairs, cells)
            assert hip.hipDeviceSynchronize() == 0
            got_pairs = _read(hip, pairs, 8 * n, np.uint32)
            got_cells = _read(hip, cells, 8 * ncells, np.uint32).reshape(ncells, 2)
            assert np.array_equal(got_pairs, ctx.download("cell_array").reshape(-1))
            assert np.array_equal(got_cells[:, 0], ctx.download("cell_start"))
            assert np.array_equal(got_cells[:, 1], ctx.download("cell_end"))
            occupied = int((got_cells[:, 0] != 0xFFFFFFFF).sum())
            print(f"[no-alloc] n={n}: pairs and {occupied} occupied cells identical", flush=True)
            ctx.close()
    finally:
        hip.hipFree(pairs)
        hip.hipFree(cells)


def test_sort_noalloc_refuses_bad_sizes(sph):
    prm = sph.default_params()
    with pytest.raises(sph.SphError, match="exceeds the scratch maximum"):
        sph.sort_noalloc(16, sph.SORT_NOALLOC_MAX_N + 1, prm.cell_width, prm.cell_power, 16, 16)
    with pytest.raises(sph.SphError, match="at least 1"):
        sph.sort_noalloc(16, 0, prm.cell_width, prm.cell_power, 16, 16)
    with pytest.raises(sph.SphError, match="null pointer"):
        sph.sort_noalloc(0, 10, prm.cell_width, prm.cell_power, 16, 16)
    with pytest.raises(sph.SphError, match="cell_power"):
        sph.sort_noalloc(16, 10, prm.cell_width, 11, 16, 16)
