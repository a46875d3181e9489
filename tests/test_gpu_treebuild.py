"""Octree built on the GPU (sph_build_tree) against the oracle's restatement
of the same Morton-key build (oracle.build_octree_morton): the flat array,
the level lists and the node count bit-exact; gravity on it bit-identical to
gravity on the same tree uploaded from the host; Headless steps with the GPU
tree against the oracle stepping on the oracle's tree; the Morton tree's
gravity accuracy against an FP64 direct sum beside octree.cpp's mean-split
tree; and the error paths (non-finite position, depth above 10)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx_with_positions(sph, impact, tables256, pos3, max_leaf, max_depth):
    n = pos3.shape[0]
    base = impact(max(n, 2000), 5)
    ic = {k: np.ascontiguousarray(v[:n]) for k, v in base.items()}
    ic["pos"] = np.zeros((n, 4), np.float32)
    ic["pos"][:, :3] = pos3
    prm = sph.default_params(max_children_in_leaf=max_leaf, max_tree_depth=max_depth)
    return sph.Context(prm, ic, tables256), ic


def _cases():
    rng = np.random.default_rng(0)
    out = []
    for n, ml, md in [(1, 4, 7), (3, 4, 7), (5, 4, 7), (2000, 4, 7), (3000, 1, 10), (4000, 4, 0), (500, 1, 3)]:
        out.append((f"gauss-{n}-{ml}-{md}", rng.normal(size=(n, 3)).astype(np.float32), ml, md))
    c = rng.normal(size=(10, 3))
    p = c[rng.integers(0, 10, 30000)] + (rng.random((30000, 1)) < 0.5) * rng.normal(size=(30000, 3)) * 1e-3
    out.append(("coincident-clumps", p.astype(np.float32), 4, 10))
    p = np.zeros((1000, 3), np.float32)
    p[:, 0] = rng.random(1000)
    out.append(("line", p, 4, 7))
    out.append(("one-point", np.full((100, 3), 3.0, np.float32), 4, 7))
    p = (rng.random((50000, 3)) * 1e-4 - 7.5e6).astype(np.float32)
    out.append(("tiny-extent-far", p, 4, 7))
    return out


@pytest.mark.parametrize("name,pos3,max_leaf,max_depth", _cases(), ids=[c[0] for c in _cases()])
def test_gpu_tree_bit_exact(sph, orc, tables256, impact, name, pos3, max_leaf, max_depth):
    ctx, ic = _ctx_with_positions(sph, impact, tables256, pos3, max_leaf, max_depth)
    ctx.build_tree()
    g = ctx.download_tree()
    c = orc.build_octree_morton(ic["pos"], max_leaf, max_depth)
    print(f"[tree {name}] n={pos3.shape[0]} nodes {g['node_count']}/{c['node_count']} "
          f"levels {list(g['level_counts'])}", flush=True)
    assert g["node_count"] == c["node_count"]
    assert np.array_equal(g["level_counts"], c["level_counts"])
    assert np.array_equal(g["octree"], c["octree"])
    assert np.array_equal(g["level_nodes"], c["level_nodes"])
    ctx.close()


@pytest.mark.parametrize("n,seed,headon", [(100000, 3, False), (1000000, 3, False), (400000, 4, True)])
def test_gpu_tree_impact_bit_exact(sph, orc, tables256, impact, n, seed, headon):
    """The impact initial conditions, twice in a row on one context (the
    second build reuses the buffers and swaps the tree's double buffer)."""
    ic = impact(n, seed, headon=headon)
    prm = sph.default_params()
    ctx = sph.Context(prm, ic, tables256)
    c = orc.build_octree_morton(ic["pos"], prm.max_children_in_leaf, prm.max_tree_depth)
    for rep in range(2):
        ctx.build_tree()
        g = ctx.download_tree()
        print(f"[tree impact n={n} headon={headon} rep {rep}] nodes {g['node_count']} ints {g['octree'].size} "
              f"levels {list(g['level_counts'])}", flush=True)
        assert g["node_count"] == c["node_count"]
        assert np.array_equal(g["level_counts"], c["level_counts"])
        assert np.array_equal(g["octree"], c["octree"])
        assert np.array_equal(g["level_nodes"], c["level_nodes"])
    ctx.close()


@pytest.mark.parametrize("form", ["as-written", "near"])
def test_gravity_on_gpu_tree_matches_uploaded(sph, orc, tables256, impact, form):
    """Gravity on the GPU-built tree equals, bit for bit, gravity on the
    oracle's Morton tree uploaded from the host; and the oracle's gravity on
    that tree within the gravity parity bar (1e-5 at the median, 1e-3 at
    q99.9 of the relative field difference)."""
    n = 50000
    ic = impact(n, 6)
    prm = sph.default_params(fixes=sph.SPH_FIX_ALL if form == "near" else 0)
    tree = orc.build_octree_morton(ic["pos"], prm.max_children_in_leaf, prm.max_tree_depth)
    out = []
    for built in (True, False):
        ctx = sph.Context(prm, ic, tables256)
        if built:
            ctx.build_tree()
        else:
            ctx.upload_tree(tree)
        ctx.gravity()
        ctx.gravity()
        ctx.sync()
        out.append((ctx.download("acc"), ctx.download("multipoles")))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    st = orc.State(prm, ic, tables256)
    st.set_tree(tree)
    st.gravity()
    st.gravity()
    ag, ac = out[0][0][:, :3].astype(np.float64), st["acc"][:, :3].astype(np.float64)
    err = np.linalg.norm(ag - ac, axis=1) / np.maximum(np.linalg.norm(ac, axis=1), 1e-30)
    q = np.quantile(err, [0.5, 0.999])
    print(f"[gpu tree gravity {form}] vs oracle rel err q50 {q[0]:.2e} q99.9 {q[1]:.2e}", flush=True)
    assert q[0] < 1e-5 and q[1] < 1e-3


def test_headless_gpu_tree_steps_match_oracle(sph, orc, tables256, impact):
    """Five Headless steps with the tree rebuilt on the GPU before every step,
    against the oracle stepping on its restatement of the same tree."""
    n = 20000
    prm = sph.default_params()
    hl = sph.Headless(prm, sph.impact_config(n, 1), eos_res=256, tree_mode=sph.TREE_GPU)
    st = orc.State(prm, impact(n, 1), tables256)
    for s in range(5):
        hl.step()
        st.full_step(orc.build_octree_morton(st["pos"], prm.max_children_in_leaf, prm.max_tree_depth))
        d = np.abs(hl.ctx.download("pos")[:, :3] - st["pos"][:, :3]).max()
        print(f"[headless gpu tree] step {s} max |dx| = {d:.3e}", flush=True)
    assert int(hl.ctx.download("global_time")[0]) == int(st["global_time"][0])
    assert d < 1e-4
    hl.close()


def test_morton_tree_gravity_accuracy(sph, orc, tables256, impact):
    """Near-field form, 1e5 particles: gravity RMS against an FP64 direct sum
    (4096 sampled particles) on the GPU-built Morton tree and on octree.cpp's
    mean-split tree; the Morton tree is within 1.5x of the other."""
    n = 100000
    ic = impact(n, 10)
    prm = sph.default_params(fixes=sph.SPH_FIX_ALL, precision=1)
    targets = np.random.default_rng(0).choice(n, 4096, replace=False).astype(np.int32)
    ref, _ = orc.direct_sum(ic["pos"], ic["mass"], targets, prm.gravity_g, -prm.gravity_softening)
    rms = {}
    for name in ("morton", "mean-split"):
        ctx = sph.Context(prm, ic, tables256)
        if name == "morton":
            ctx.build_tree()
        else:
            ctx.upload_tree(orc.build_octree(ic["pos"], prm.max_children_in_leaf, prm.max_tree_depth))
        ctx.gravity()
        ctx.gravity()
        ctx.sync()
        a = ctx.download("acc")[targets, :3].astype(np.float64)
        rel = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
        rms[name] = float(np.sqrt(np.mean(rel ** 2)))
        print(f"[tree accuracy] {name}: RMS {rms[name]:.3e} max {rel.max():.3e}", flush=True)
        ctx.close()
    assert rms["morton"] < 1.5 * rms["mean-split"]


def test_gpu_tree_build_time(sph, tables256, impact):
    """Build time at 1e6 particles (reported; bounded loosely)."""
    ic = impact(1000000, 3)
    ctx = sph.Context(sph.default_params(), ic, tables256)
    ctx.build_tree()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(10):
        ctx.build_tree()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 100.0
    print(f"[tree build 1e6] {ms:.3f} ms per build", flush=True)
    assert ms < 50.0
    ctx.close()


def test_gpu_tree_errors(sph, tables256, impact):
    ic = impact(2000, 5)
    ic["pos"][17, 1] = np.nan
    ctx = sph.Context(sph.default_params(), ic, tables256)
    with pytest.raises(sph.SphError, match="non-finite"):
        ctx.build_tree()
    ctx.close()
    ic = impact(2000, 5)
    ctx = sph.Context(sph.default_params(max_tree_depth=11), ic, tables256)
    with pytest.raises(sph.SphError, match="max_tree_depth"):
        ctx.build_tree()
    ctx.close()
