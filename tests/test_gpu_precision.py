"""Precision form on the GPU: potential and acceleration against the FP64
direct sum, the FP64 conservation sums against numpy, the FP64 time view,
and conservation drift over a run with individual timesteps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _prec_params(sph, **kw):
    return sph.default_params(fixes=sph.SPH_FIX_ALL, precision=1, **kw)


def _gravity(sph, orc, prm, ic, tables):
    ctx = sph.Context(prm, ic, tables)
    ctx.upload_tree(orc.build_octree(ic["pos"], prm.max_children_in_leaf, prm.max_tree_depth))
    ctx.gravity()
    ctx.gravity()          # warm: minGrav from the first evaluation
    ctx.sync()
    return ctx


def _errors(ctx, orc, ic, prm, targets):
    acc_ref, phi_ref = orc.direct_sum(ic["pos"], ic["mass"], targets, prm.gravity_g, -prm.gravity_softening)
    a = ctx.download("acc")[targets, :3].astype(np.float64)
    rel = np.linalg.norm(a - acc_ref, axis=1) / np.linalg.norm(acc_ref, axis=1)
    phi = ctx.download("potential")[targets].astype(np.float64)
    prel = np.abs(phi - phi_ref) / np.abs(phi_ref)
    return rel, prel


def test_potential_and_field_against_fp64(sph, orc, tables256, impact):
    n = 100000
    ic = impact(n, 2)
    prm = _prec_params(sph)
    ctx = _gravity(sph, orc, prm, ic, tables256)
    targets = np.random.default_rng(4096).choice(n, 4096, replace=False).astype(np.int32)
    rel, prel = _errors(ctx, orc, ic, prm, targets)
    print(f"[precision eta {prm.gravity_eta}] field rms {np.sqrt(np.mean(rel ** 2)):.3e} max {rel.max():.3e}; "
          f"potential rms {np.sqrt(np.mean(prel ** 2)):.3e} max {prel.max():.3e}", flush=True)
    assert np.sqrt(np.mean(rel ** 2)) < 2e-2
    assert np.sqrt(np.mean(prel ** 2)) < 1e-2


def test_precision_gravity_rms_below_1e3(sph, orc, tables256, impact):
    """At GRAVITY_ETA 2e-5 the P = 2 expansions reach RMS < 1e-3 against FP64
    (4096 sampled particles, seed 4096).  The unchecked-node lists must then
    hold 16384 entries: with 4096 some fill up and their force-accepted
    entries leave 20-60% errors on a few particles (the oracle shows the
    same)."""
    n = 100000
    ic = impact(n, 2)
    prm = _prec_params(sph, gravity_eta=2e-5, max_unchecked_pointers=16384)
    ctx = _gravity(sph, orc, prm, ic, tables256)
    targets = np.random.default_rng(4096).choice(n, 4096, replace=False).astype(np.int32)
    rel, prel = _errors(ctx, orc, ic, prm, targets)
    rms = float(np.sqrt(np.mean(rel ** 2)))
    print(f"[precision eta 2e-5] field rms {rms:.3e} median {np.median(rel):.3e} max {rel.max():.3e}; "
          f"potential rms {np.sqrt(np.mean(prel ** 2)):.3e}", flush=True)
    assert rms < 1e-3 and rel.max() < 2e-2


def test_conservation_sums_match_numpy(sph, orc, tables256, impact):
    n = 20000
    ic = impact(n, 5)
    prm = _prec_params(sph)
    ctx = sph.Context(prm, ic, tables256)
    ctx.upload_tree(orc.build_octree(ic["pos"], 4, 7))
    ctx.sort()
    ctx.gravity()
    ctx.density()
    ctx.acceleration()
    ctx.step()
    ctx.sync()
    c = ctx.conservation()
    m = ctx.download("mass").astype(np.float64)
    x = ctx.download("pos")[:, :3].astype(np.float64)
    v = ctx.download("vel")[:, :3].astype(np.float64)
    u = ctx.download("u").astype(np.float64)
    phi = ctx.download("potential").astype(np.float64)
    assert np.any(phi != 0)
    assert c["kinetic"] == pytest.approx(0.5 * (m * (v * v).sum(1)).sum(), rel=1e-12)
    assert c["internal"] == pytest.approx((m * u).sum(), rel=1e-12)
    assert c["potential"] == pytest.approx(0.5 * (m * phi).sum(), rel=1e-12)
    assert np.allclose(c["momentum"], (m[:, None] * v).sum(0), rtol=1e-10, atol=1e-16)
    assert np.allclose(c["angular_momentum"], (m[:, None] * np.cross(x, v)).sum(0), rtol=1e-10, atol=1e-16)
    assert c["mass"] == pytest.approx(m.sum(), rel=1e-12)
    ticks = int(ctx.download("global_time")[0])
    assert ctx.time() == ticks * (prm.dt0 / 2 ** prm.r_max)


def test_conservation_drift_with_rungs(sph, orc, tables256, impact):
    """Forty steps of the 1e5 impact in the precision form with individual
    timesteps (active iff nextActiveTime <= globalTime, at most 8 rungs):
    total energy and angular momentum drift, stated.  The step update
    (step.metal:12-46) is not symplectic and inactive particles keep the
    potential of their last active step, so the energy drifts by ~2% over
    136 s of the encounter; angular momentum holds to ~4e-4."""
    n = 100000
    prm = _prec_params(sph, activity_rule=sph.SPH_ACTIVE_README, max_rungs=8)
    ctx = sph.Context(prm, impact(n, 2), tables256)
    e0 = l0 = None
    for s in range(40):
        ctx.upload_tree(orc.build_octree(ctx.download("pos"), 4, 7))
        ctx.sort()
        ctx.gravity()
        ctx.density()
        ctx.acceleration()
        ctx.step()
        if s == 0:
            ctx.sync()
            c = ctx.conservation()
            e0 = c["kinetic"] + c["internal"] + c["potential"]
            l0 = c["angular_momentum"]
    ctx.sync()
    c = ctx.conservation()
    e1 = c["kinetic"] + c["internal"] + c["potential"]
    de = abs(e1 - e0) / abs(e0)
    dl = np.linalg.norm(c["angular_momentum"] - l0) / np.linalg.norm(l0)
    print(f"[drift] t = {ctx.time():.4f} s: |dE/E| {de:.3e}, |dL|/|L| {dl:.3e} "
          f"(E0 {e0:.6e}: K {c['kinetic']:.4e} U {c['internal']:.4e} W {c['potential']:.4e})", flush=True)
    assert np.isfinite(de) and np.isfinite(dl)
    # twice the measured drifts (|dE/E| 1.6e-2, |dL|/|L| 4.0e-4)
    assert de < 3.2e-2 and dl < 8e-4
