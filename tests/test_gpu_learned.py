"""Learned smoothing-length guess on the GPU: BF16 MFMA forward and backward
against the FP64 restatement (oracle/mlp.py), features against their FP64
restatement, the sample dump, training, and what the trained network gains
over plain Newton-Raphson at identical converged h (|dh/h| <= 1e-5).

The comparisons stop the iteration with density_stop = 1 at density_eta =
2.5e-6: each run ends within 2.5e-6 of the fixed point of the update, so
the plain and the learned run end within 1e-5 of each other for every
particle.  The reference's successive-iterate rule cannot give that: it
stops density_eta / (1 - k) from the fixed point, and weakly contracting
particles (k near 1) end far apart when started from different guesses."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

TIGHT = dict(density_eta=2.5e-6, density_stop=1, max_nr_iterations=400)


def _random_norm(rng):
    return np.r_[rng.normal(0, 1, 12), rng.uniform(0.5, 2.0, 12), 0.01, 0.05].astype(np.float32)


def _blocks(g):
    o = omlp
    return {"W1": g[o.OFF_W1:o.OFF_B1], "b1": g[o.OFF_B1:o.OFF_W2], "W2": g[o.OFF_W2:o.OFF_B2],
            "b2": g[o.OFF_B2:o.OFF_W3], "w3": g[o.OFF_W3:o.OFF_B3], "b3": g[o.OFF_B3:]}


def test_mlp_forward_matches_fp64(sph):
    rng = np.random.default_rng(1)
    net = sph.Mlp(seed=13)
    p, _ = net.get()
    p = (p * 3.0).astype(np.float32)   # larger weights: a layout slip cannot hide in small outputs
    norm = _random_norm(rng)
    net.set(p, norm)
    x = (norm[:12] + norm[12:24] * rng.normal(0, 1, (1000, 12))).astype(np.float32)
    d_gpu = net.forward(x).astype(np.float64)
    d_ref, _ = omlp.forward(p, norm, x)
    err = np.abs(d_gpu - d_ref)
    scale = np.sqrt(np.mean((d_ref - norm[24]) ** 2))
    print(f"[mlp forward] rms |d| {scale:.3e} max err {err.max():.3e} rms err {np.sqrt(np.mean(err ** 2)):.3e}",
          flush=True)
    # BF16 operands (8-bit mantissa) through two layers: ~1% of the output scale
    assert np.sqrt(np.mean(err ** 2)) < 0.02 * scale and err.max() < 0.1 * scale


@pytest.mark.parametrize("n", [50, 4096])
def test_mlp_gradient_matches_fp64(sph, n):
    rng = np.random.default_rng(2 + n)
    net = sph.Mlp(seed=5)
    p, _ = net.get()
    norm = _random_norm(rng)
    net.set(p, norm)
    x = (norm[:12] + norm[12:24] * rng.normal(0, 1, (n, 12))).astype(np.float32)
    d = (norm[24] + norm[25] * rng.normal(0, 1, n)).astype(np.float32)
    w = (rng.uniform(size=n) < 0.9).astype(np.float32)
    loss, g = net.loss_grad(x, d, w)
    loss_ref, g_ref = omlp.loss_grad(p, norm, x, d, w)
    assert loss == pytest.approx(loss_ref, rel=2e-2)
    # error of each block relative to its own norm, or to the gradient's RMS
    # block norm where a block (b3 = sum of the residuals) nearly cancels
    floor = 0.1 * np.linalg.norm(g_ref) / np.sqrt(6)
    for name, (a, b) in {k: (v, _blocks(g_ref)[k]) for k, v in _blocks(g).items()}.items():
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), floor)
        print(f"[mlp grad n={n}] {name}: rel L2 err {rel:.3e}", flush=True)
        assert rel < 5e-2, name
    cos = float(g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    assert cos > 0.999


def _run_steps(ctx, orc, prm, steps):
    for _ in range(steps):
        tree = orc.build_octree(ctx.download("pos"), prm.max_children_in_leaf, prm.max_tree_depth)
        ctx.upload_tree(tree)
        ctx.sort()
        ctx.gravity()
        ctx.density()
        ctx.acceleration()
        ctx.step()
    ctx.sync()


def test_features_match_fp64(sph, orc, tables256, impact):
    prm = sph.default_params()
    ctx = sph.Context(prm, impact(20000, 3), tables256)
    _run_steps(ctx, orc, prm, 2)
    ctx.sort()
    f = ctx.features().astype(np.float64)
    ref = omlp.features(ctx.download("pos"), ctx.download("h"), ctx.download("dhdt"),
                        int(ctx.download("dt")[0]), ctx.download("rho"), ctx.download("mat"))
    assert np.array_equal(f[:, 3:8], np.float32(ref[:, 3:8]).astype(np.float64)) or \
        np.abs(f[:, 3:8] - ref[:, 3:8]).max() < 1e-6
    err = np.abs(f - ref) / np.maximum(np.abs(ref), 1e-3)
    print(f"[features] max rel err per column {err.max(axis=0)}", flush=True)
    assert err.max() < 1e-5


def test_dump_appends_converged_targets(sph, orc, tables256, impact):
    prm = sph.default_params()
    n = 20000
    ctx = sph.Context(prm, impact(n, 4), tables256)
    _run_steps(ctx, orc, prm, 1)
    smp = sph.Samples(4 * n)
    ctx.sort()
    f_before = ctx.features()
    h_before = ctx.download("h")
    ctx.set_sample_dump(smp)
    ctx.density()
    ctx.sync()
    ctx.set_sample_dump(None)
    assert len(smp) == n
    rec = smp.read()
    order = ctx.download("cell_array")[:, 1]          # records are in sorted order
    assert np.array_equal(rec[:, :12], f_before[order])
    d = np.log(ctx.download("h")[order].astype(np.float64)) - np.log(h_before[order].astype(np.float64))
    assert np.abs(rec[:, 12] - d).max() < 1e-5
    assert np.all(rec[:, 13] == 1.0)


def _density_ab(ctx, mlp):
    """Density from the same state, plain and with the learned guess."""
    keep = {k: ctx.download(k) for k in ("h", "rho", "global_time", "dt")}
    ctx.density()
    ctx.sync()
    plain = {"h": ctx.download("h"), "it": ctx.download("nr_iters")}
    for k, v in keep.items():
        ctx.upload(k, v)
    ctx.set_learned_guess(mlp)
    ctx.density()
    ctx.sync()
    ctx.set_learned_guess(None)
    learned = {"h": ctx.download("h"), "it": ctx.download("nr_iters")}
    return plain, learned


def test_training_and_learned_guess_gain(sph, orc, tables256, impact):
    """Dump (features, converged h) from a plain run at the tight stopping
    rule, train on the GPU, then start the iteration from the network's
    guess: the same converged h (|dh/h| <= 1e-5), fewer iterations."""
    prm = sph.default_params(**TIGHT)
    n = 100000
    ctx = sph.Context(prm, impact(n, 11), tables256)
    _run_steps(ctx, orc, prm, 1)
    smp = sph.Samples(8 * n)
    ctx.set_sample_dump(smp)
    _run_steps(ctx, orc, prm, 6)
    ctx.set_sample_dump(None)
    assert len(smp) == 6 * n
    net = sph.Mlp(seed=13)
    net.fit_norm(smp)
    losses = net.train(smp, steps=400, batch=65536, lr=2e-3, seed=14)
    print(f"[train] loss first {losses[:3]} last {losses[-3:]}", flush=True)
    assert np.all(np.isfinite(losses)) and losses[-20:].mean() < 0.3 * losses[:5].mean()
    # evaluation on the next state
    tree = orc.build_octree(ctx.download("pos"), prm.max_children_in_leaf, prm.max_tree_depth)
    ctx.upload_tree(tree)
    ctx.sort()
    ctx.gravity()
    plain, learned = _density_ab(ctx, net)
    dh = np.abs(learned["h"].astype(np.float64) / plain["h"] - 1)
    print(f"[learned] NR iterations plain {plain['it'].mean():.3f} learned {learned['it'].mean():.3f}; "
          f"|dh/h| max {dh.max():.2e} q99.9 {np.quantile(dh, 0.999):.2e}", flush=True)
    assert dh.max() <= 1e-5
    assert learned["it"].mean() < 0.85 * plain["it"].mean()


def test_bad_guess_still_converges(sph, orc, tables256, impact):
    """A network that always predicts the clamp (h x e^0.5) costs at most one
    extra iteration and ends at the plain iteration's h: its first residual
    is far above the trust bound, so the iteration restarts from h."""
    prm = sph.default_params(**TIGHT)
    ctx = sph.Context(prm, impact(20000, 12), tables256)
    _run_steps(ctx, orc, prm, 1)
    net = sph.Mlp(seed=1)
    p, norm = net.get()
    p[:] = 0.0
    p[omlp.OFF_B3] = 10.0
    net.set(p, norm)
    tree = orc.build_octree(ctx.download("pos"), prm.max_children_in_leaf, prm.max_tree_depth)
    ctx.upload_tree(tree)
    ctx.sort()
    plain, learned = _density_ab(ctx, net)
    dh = np.abs(learned["h"].astype(np.float64) / plain["h"] - 1)
    print(f"[bad guess] iterations plain {plain['it'].mean():.2f} learned {learned['it'].mean():.2f} "
          f"|dh/h| max {dh.max():.2e}", flush=True)
    assert dh.max() <= 1e-5
    assert learned["it"].mean() <= plain["it"].mean() + 1.0 + 1e-9


def test_learned_guess_with_individual_timesteps(sph, orc, tables256, impact):
    """README rule (up to 8 rungs): only the active particles' features and
    guesses are computed (compacted sorted slots).  Active particles converge
    to the plain iteration's h; inactive ones keep theirs in both runs."""
    prm = sph.default_params(activity_rule=sph.SPH_ACTIVE_README, max_rungs=8, fixes=sph.SPH_FIX_INACTIVE_DRIFT,
                             **TIGHT)
    ctx = sph.Context(prm, impact(20000, 13), tables256)
    _run_steps(ctx, orc, prm, 3)
    net = sph.Mlp(seed=3)
    tree = orc.build_octree(ctx.download("pos"), prm.max_children_in_leaf, prm.max_tree_depth)
    ctx.upload_tree(tree)
    ctx.sort()
    ctx.gravity()
    act = ctx.download("active").astype(bool)
    h0 = ctx.download("h")
    plain, learned = _density_ab(ctx, net)
    dh = np.abs(learned["h"].astype(np.float64) / plain["h"] - 1)
    print(f"[learned README] active {act.mean():.4f}; |dh/h| active max {dh[act].max():.2e}", flush=True)
    assert 0 < act.sum() < len(act)
    assert np.array_equal(learned["h"][~act], h0[~act]) and np.array_equal(plain["h"][~act], h0[~act])
    assert dh[act].max() <= 1e-5


def test_learned_guess_1e6_c4_recipe(sph):
    """The learned-h form at 1e6 particles: (features, converged h) dumped by
    8 plain steps each of seeded 1e5 (seed 11) and 1e6 (seed 12) impacts,
    2000 Adam steps at batch 65536 on the GPU, then one density pass of a
    third 1e6 impact (seed 3, after 3 steps) run plain and from the network's
    guess: every particle's converged h the same to 1e-5, fewer iterations,
    a shorter pass."""
    prm = sph.default_params(**TIGHT)
    smp = sph.Samples(8 * 1_100_000)
    for n, seed in ((100_000, 11), (1_000_000, 12)):
        hl = sph.Headless(prm, sph.impact_config(n, seed), eos_res=256, tree_mode=sph.TREE_EVERY, tree_threads=8)
        hl.step()
        hl.ctx.set_sample_dump(smp)
        for _ in range(8):
            hl.step()
        hl.ctx.set_sample_dump(None)
        hl.close()
    assert len(smp) == 8 * 1_100_000
    net = sph.Mlp(seed=13)
    net.fit_norm(smp)
    losses = net.train(smp, steps=2000, batch=65536, lr=2e-3, seed=14)
    smp.close()
    print(f"[C4] loss first {losses[0]:.4f} last10 {np.mean(losses[-10:]):.5f}", flush=True)
    assert np.all(np.isfinite(losses)) and np.mean(losses[-10:]) < 0.05 * losses[0]
    hl = sph.Headless(prm, sph.impact_config(1_000_000, 3), eos_res=256, tree_mode=sph.TREE_EVERY, tree_threads=8)
    for _ in range(3):
        hl.step()
    ctx = hl.ctx
    ctx.sort()
    ctx.sync()
    keep = {k: ctx.download(k) for k in ("h", "rho", "global_time", "dt")}
    out = {}
    for mode in ("plain", "learned"):
        ctx.set_learned_guess(net if mode == "learned" else None)
        for k, v in keep.items():
            ctx.upload(k, v)
        ctx.reset_pass_times()
        ctx.set_timing(True)
        ctx.density()
        t = ctx.pass_times()["density"]
        ctx.set_timing(False)
        out[mode] = {"ms": t["ms"], "h": ctx.download("h").astype(np.float64), "it": ctx.download("nr_iters")}
    ctx.set_learned_guess(None)
    hl.close()
    net.close()
    dh = np.abs(out["learned"]["h"] / out["plain"]["h"] - 1)
    print(f"[C4] NR iterations plain {out['plain']['it'].mean():.3f} learned {out['learned']['it'].mean():.3f}; "
          f"density ms plain {out['plain']['ms']:.3f} learned {out['learned']['ms']:.3f}; "
          f"|dh/h| max {dh.max():.2e} q99.9 {np.quantile(dh, 0.999):.2e}", flush=True)
    assert out["plain"]["it"].max() < 400 and out["learned"]["it"].max() < 400
    assert dh.max() <= 1e-5
    assert out["learned"]["it"].mean() < 0.8 * out["plain"]["it"].mean()
    assert out["learned"]["ms"] < out["plain"]["ms"]
