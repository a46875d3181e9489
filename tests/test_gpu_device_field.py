"""Zero-copy access to the per-particle buffers (sph_device_field), which
stands in for the reference's public positionBuffer / materialIdBuffer /
densityBuffer (Compute.hpp:55-57, bound by the renderer at Render.cpp:26):
a device-to-host copy from the returned address equals sph_download, after
steps, and Headless's Compute exposes the same three addresses."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PER_PARTICLE = ("pos", "vel", "rho", "u", "mass", "pressure", "h", "mat", "gradterm", "soundspeed", "du",
                "balsara", "dhdt", "active", "next_active", "gravnorm", "nr_iters", "potential", "dlnh")


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    return hip


def _read(hip, ptr, nbytes):
    out = np.zeros(nbytes, np.uint8)
    rc = hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, rc
    return out


def test_device_field_equals_download(sph, orc, tables256, impact):
    n = 20000
    prm = sph.default_params()
    ctx = sph.Context(prm, impact(n, 31), tables256)
    for _ in range(2):
        ctx.upload_tree(orc.build_octree(ctx.download("pos"), 4, 7))
        ctx.sort()
        ctx.gravity()
        ctx.density()
        ctx.acceleration()
        ctx.step()
    ctx.sync()
    hip = _hip()
    for f in PER_PARTICLE:
        ptr, nbytes = ctx.device_field(f)
        ref = ctx.download(f)
        assert ptr != 0 and nbytes == ref.nbytes, f
        assert np.array_equal(_read(hip, ptr, nbytes), ref.view(np.uint8).reshape(-1)), f
    # the address is stable: the next step updates it in place
    p0, nb = ctx.device_field("pos")
    ctx.upload_tree(orc.build_octree(ctx.download("pos"), 4, 7))
    ctx.sort()
    ctx.gravity()
    ctx.density()
    ctx.acceleration()
    ctx.step()
    ctx.sync()
    assert ctx.device_field("pos")[0] == p0
    assert np.array_equal(_read(hip, p0, nb), ctx.download("pos").view(np.uint8).reshape(-1))
    with pytest.raises(sph.SphError):
        ctx.device_field("acc")
    with pytest.raises(sph.SphError):
        ctx.device_field("cell_array")
    ctx.close()


def test_headless_render_buffers(sph):
    prm = sph.default_params()
    hl = sph.Headless(prm, sph.impact_config(20000, 32), eos_res=256, tree_mode=sph.TREE_EVERY, tree_threads=4)
    hl.step()
    pos, mat, rho = hl.render_buffers()
    ctx = hl.ctx
    assert (pos, mat, rho) == (ctx.device_field("pos")[0], ctx.device_field("mat")[0], ctx.device_field("rho")[0])
    hip = _hip()
    n = ctx.n
    assert np.array_equal(_read(hip, pos, 16 * n).view(np.float32).reshape(n, 4), ctx.download("pos"))
    assert np.array_equal(_read(hip, mat, 4 * n).view(np.int32), ctx.download("mat"))
    assert np.array_equal(_read(hip, rho, 4 * n).view(np.float32), ctx.download("rho"))
    hl.close()
