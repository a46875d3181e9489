"""Initial conditions from HDF5 files and checkpoint / resume on the GPU.

  Headless from a file: the demo IC (the reference's SPH/demo_impact_n40.hdf5,
      from the golden fixture) written with sph_write_hdf and read back by
      Headless::Headless(filepath); three steps against the oracle.
  Checkpoint / resume: a run checkpointed after three steps and resumed in a
      new Headless from that file continues bit-identically to the run that
      was not interrupted (tree rebuilt before every step), for the reference
      step as written and for individual timesteps with the inactive-drift
      correction.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_headless_from_file_demo_matches_oracle(sph, orc, tables256, demo_ic, tmp_path):
    prm = sph.default_params()
    path = str(tmp_path / "demo.hdf5")
    sph.write_hdf(path, demo_ic)
    hl = sph.Headless.from_file(prm, path, eos_res=256, tree_mode=sph.TREE_EVERY, tree_threads=4)
    st = orc.State(prm, demo_ic, tables256)
    for s in range(3):
        hl.step()
        st.full_step(orc.build_octree(st["pos"], prm.max_children_in_leaf, prm.max_tree_depth))
        d = np.abs(hl.ctx.download("pos")[:, :3] - st["pos"][:, :3]).max()
        dh = np.abs(hl.ctx.download("h") / st["h"] - 1).max()
        print(f"[demo from file] step {s} max |dx| {d:.3e} max |dh/h| {dh:.3e}", flush=True)
    assert int(hl.ctx.download("global_time")[0]) == int(st["global_time"][0])
    assert d < 1e-4 and dh < 1e-4
    hl.close()


FIELDS = ("pos", "vel", "acc", "rho", "u", "h", "pressure", "du", "dhdt", "gradterm", "soundspeed", "balsara",
          "gravnorm", "next_active", "active", "global_time", "dt", "dlnh")


def _state(ctx):
    out = {}
    for k in FIELDS:
        a = ctx.download(k)
        out[k] = a[:, :3] if a.ndim == 2 else a
    return out


@pytest.mark.parametrize("form,tree", [("as-written", "every"), ("rungs", "every"), ("as-written", "gpu")])
def test_checkpoint_resume_is_bit_identical(sph, tmp_path, form, tree):
    n = 20000
    if form == "rungs":
        prm = sph.default_params(activity_rule=sph.SPH_ACTIVE_README, max_rungs=4,
                                 fixes=sph.SPH_FIX_INACTIVE_DRIFT)
    else:
        prm = sph.default_params()
    mode = sph.TREE_EVERY if tree == "every" else sph.TREE_GPU
    cfg = sph.impact_config(n, 1)
    path = str(tmp_path / "ckpt.hdf5")
    hl = sph.Headless(prm, cfg, eos_res=256, tree_mode=mode, tree_threads=4)
    for _ in range(3):
        hl.step()
    hl.ctx.checkpoint_write(path)
    for _ in range(3):
        hl.step()
    want = _state(hl.ctx)
    hl.close()
    back = sph.Headless.from_file(prm, path, eos_res=256, tree_mode=mode, tree_threads=4)
    back.ctx.checkpoint_read(path)
    for _ in range(3):
        back.step()
    got = _state(back.ctx)
    back.close()
    diff = [k for k in FIELDS if not np.array_equal(got[k], want[k])]
    active = float(np.mean(want["active"]))
    print(f"[checkpoint {form} tree {tree}] active fraction {active:.3f}; fields that differ: {diff}", flush=True)
    assert not diff
