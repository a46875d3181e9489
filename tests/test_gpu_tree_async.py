"""The asynchronous octree hand-off (Compute.cpp:456-516, Compute.hpp:74):
a CPU thread builds the next tree from a snapshot of the positions while the
GPU steps with the current one, and the call that swaps a finished tree in
starts the next build.  Which tree a step uses depends on thread timing (the
reference's own behaviour, Q12), so the test records the tree of every step
and replays that sequence through sph_upload_tree on a fresh context: given
the same trees, the steps must be bit-identical."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "rho", "u", "pressure", "h", "gradterm", "soundspeed", "du", "balsara", "dhdt",
         "active", "next_active", "gravnorm", "cell_array", "global_time", "dt")


def _tree_key(tree):
    h = hashlib.sha1()
    for k in ("octree", "level_nodes", "level_counts"):
        h.update(np.ascontiguousarray(tree[k], np.int32).tobytes())
    return h.hexdigest()[:12]


@pytest.mark.parametrize("n,steps,threads", [(100000, 30, 4)])
def test_async_tree_handoff_replays_bit_identically(sph, tables256, n, steps, threads):
    prm = sph.default_params()
    cfg = sph.impact_config(n, 21)
    hl = sph.Headless(prm, cfg, eos_res=256, tree_mode=sph.TREE_ASYNC, tree_threads=threads)
    ctx = hl.ctx
    ic = {k: ctx.download(k) for k in ("pos", "vel", "rho", "u", "mass", "mat", "pressure", "h")}
    used, trees = [], {}
    for i in range(steps):
        hl.step()
        t = ctx.download_tree()
        key = _tree_key(t)
        trees.setdefault(key, t)
        used.append(key)
    swapped, _ = hl.tree_stats()
    final = {k: ctx.download(k) for k in STATE}
    hl.close()
    distinct = len(set(used))
    print(f"[tree async] {swapped} trees swapped in, {distinct} distinct trees used over {steps} steps: "
          f"{' '.join(u[:4] for u in used)}", flush=True)
    assert swapped > 1 and distinct > 1
    for k in ("pos", "vel", "acc", "rho", "h", "u"):
        assert np.all(np.isfinite(final[k])), k
    # the initial tree is the one built from the initial conditions
    # replay: the same trees at the same steps, uploaded through the C ABI
    rep = sph.Context(prm, ic, tables256)
    last = None
    for i in range(steps):
        if used[i] != last:
            rep.upload_tree(trees[used[i]])
            last = used[i]
        rep.sort()
        rep.gravity()
        rep.density()
        rep.acceleration()
        rep.step()
    rep.sync()
    for k in STATE:
        assert np.array_equal(rep.download(k), final[k]), k
    rep.close()
