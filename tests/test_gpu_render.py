"""Snapshot raster on the GPU (drawToSnapshot, render.metal:49-66) against a
numpy FP32 restatement, bit-exact; and Headless snapshots: which steps write a
PNG (Compute.cpp:416-446: the time of the last completed step passes the next
multiple of the interval) and that each PNG shows the positions drawn one
step earlier (the reference encodes the raster with the step and writes the
file from the previous step's image)."""
import os

import numpy as np
import pytest

from test_render import colourise, decode_png

pytestmark = pytest.mark.gpu


def raster(pos, mat, res):
    f = np.float32
    scale = f(res) / (f(20.0) * f(6.3710))
    ux = (pos[:, 0].astype(f) - f(250.0)) * scale
    uy = (pos[:, 1].astype(f) - f(250.0)) * scale
    ok = (ux >= 0) & (uy >= 0) & (ux <= f(res)) & (uy <= f(res))
    # out-of-window values (NaN, -1e9) are masked before the cast
    px, py = np.where(ok, ux, 0).astype(np.int64), np.where(ok, uy, 0).astype(np.int64)
    ok &= (px < res) & (py < res)
    img = np.zeros(res * res, np.int64)
    np.maximum.at(img, (px + (res - 1 - py) * res)[ok], mat[ok].astype(np.int64))
    return img.reshape(res, res).astype(np.int32), int(ok.sum())


def test_draw_snapshot_matches_restatement(sph, tables256, demo_ic):
    ic = {k: v.copy() for k, v in demo_ic.items()}
    n = ic["mass"].shape[0]
    # particles at the window's edges and outside it, and pixel collisions
    extra = np.array([[250.0, 250.0], [249.999, 300.0], [300.0, 249.999], [250.0 + 20 * 6.371, 300.0],
                      [377.41, 377.41], [377.43, 300.0], [-1e9, 0.0], [np.nan, 300.0], [300.0, 300.0],
                      [300.0, 300.0], [300.001, 300.001]], np.float32)
    k = extra.shape[0]
    ic["pos"][:k, :2] = extra
    ic["mat"][:k] = [400, 402, 400, 402, 400, 402, 400, 402, 400, 402, 400]
    ctx = sph.Context(sph.default_params(), ic, tables256)
    for res in (1024, 333):
        img = ctx.draw_snapshot(res)
        want, inside = raster(ic["pos"], ic["mat"], res)
        print(f"[snapshot res {res}] {inside} of {n} particles in the window, "
              f"{int((want > 0).sum())} pixels set", flush=True)
        assert np.array_equal(img, want)
    ctx.close()


def test_headless_snapshots_timing_and_content(sph, tables256, demo_ic, tmp_path):
    prm = sph.default_params()
    path = str(tmp_path / "demo.hdf5")
    sph.write_hdf(path, demo_ic)
    hl = sph.Headless.from_file(prm, path, eos_res=256, tree_mode=sph.TREE_EVERY, tree_threads=4)
    outdir = tmp_path / "snaps"
    outdir.mkdir()
    res, steps = 256, 8
    # an interval that the first steps cross a few times
    hl.step()
    dt1 = float(hl.ctx.download("global_time")[0]) * 32.0 / 65536
    interval = max(1, int(2.5 * dt1))
    hl.set_snapshots(str(outdir), res, interval)
    times, positions = [], []
    for _ in range(steps):
        times.append(float(np.float32(hl.ctx.download("global_time")[0]) * np.float32(32.0 / 65536)))
        positions.append(hl.ctx.download("pos"))
        hl.step()
    mats = hl.ctx.download("mat")
    hl.close()
    written, nxt = {}, 0
    for call in range(steps):
        if call > 0 and times[call] > nxt:          # an image exists from the previous call
            written[nxt // interval] = call - 1     # drawn from the positions at that call
            nxt += interval
    files = sorted(os.listdir(outdir))
    print(f"[snapshots] interval {interval}, times {times}, files {files}", flush=True)
    assert files == sorted(f"snapshot_{k}.png" for k in written) and len(files) >= 2
    for k, drawn_at in written.items():
        img = decode_png((outdir / f"snapshot_{k}.png").read_bytes())
        want, _ = raster(positions[drawn_at], mats, res)
        assert np.array_equal(img, colourise(want)), k
