"""Pins of the FP64 restatement of the learned smoothing-length network
(oracle/mlp.py), which checks the GPU kernels: its gradient against finite
differences, Adam against the published update, and the feature
definitions on hand-built particle sets."""
import math

import numpy as np
import pytest

from oracle import mlp


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = mlp.init_params(3)
    norm = np.r_[rng.normal(0, 1, 12), rng.uniform(0.5, 2, 12), 0.1, 0.3]
    x = norm[:12] + norm[12:24] * rng.normal(size=(40, 12))
    d = norm[24] + norm[25] * rng.normal(size=40)
    w = (rng.uniform(size=40) < 0.8).astype(float)
    _, g = mlp.loss_grad(p, norm, x, d, w)
    idx = np.r_[rng.choice(mlp.N_PARAMS - 1, 60, replace=False), mlp.OFF_B3]
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp.loss_grad(q, norm, x, d, w)
        q[i] -= 2 * h
        lm, _ = mlp.loss_grad(q, norm, x, d, w)
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9), i


def test_forward_is_the_stated_network():
    """One sample by hand: SiLU(W2 SiLU(W1 xn + b1) + b2) . w3 + b3."""
    rng = np.random.default_rng(8)
    p = rng.normal(size=mlp.N_PARAMS) * 0.2
    norm = np.r_[np.zeros(12), np.ones(12), 0.5, 2.0]
    x = rng.normal(size=(1, 12))
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p)
    silu = lambda z: z / (1 + np.exp(-z))
    y = silu(W2 @ silu(W1 @ x[0] + b1) + b2) @ w3 + b3
    d, _ = mlp.forward(p, norm, x)
    assert d[0] == pytest.approx(y * 2.0 + 0.5, rel=1e-12)
    assert mlp.N_PARAMS == 5057


def test_adam_first_step_moves_each_parameter_by_lr():
    """Bias-corrected Adam: the first update of every parameter with a
    nonzero gradient is lr * sign(g) (to eps)."""
    p = np.zeros(5)
    g = np.array([1e-3, -2.0, 5.0, -1e-6, 0.0])
    m = np.zeros(5)
    v = np.zeros(5)
    mlp.adam_step(p, g, m, v, 1, 0.01)
    assert np.allclose(p[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert p[4] == 0.0


def _one_particle_per_cell(cells, cw=0.18):
    pos = np.zeros((len(cells), 4), np.float32)
    pos[:, :3] = (np.asarray(cells, np.float64) + 0.5) * cw
    return pos


def test_features_count_statistics_by_hand():
    block = [(x, y, z) for x in (10, 11, 12) for y in (20, 21, 22) for z in (30, 31, 32)]
    pos = _one_particle_per_cell(block + [(11, 21, 31)])          # centre cell holds 2
    n = pos.shape[0]
    f = mlp.features(pos, np.full(n, 0.2), np.full(n, 0.01), 64, np.full(n, 2.0), [400] * (n - 1) + [402])
    c = block.index((11, 21, 31))
    assert f[c, 3] == pytest.approx(math.log1p(28))          # all 27 cells
    assert f[c, 4] == pytest.approx(math.log1p(2))           # own
    assert f[c, 5] == pytest.approx(math.log1p(6))           # faces
    assert f[c, 6] == pytest.approx(math.log1p(12))          # edges
    assert f[c, 7] == pytest.approx(math.log1p(8))           # corners
    assert f[c, 8] == pytest.approx(0.0)                     # symmetric: no dipole
    assert f[c, 9] == pytest.approx((2 - 1) / 29)
    corner = block.index((10, 20, 30))
    # the corner cell sees the 8 cells of the block's corner octant (one holds 2)
    assert f[corner, 3] == pytest.approx(math.log1p(9))
    # dipole by brute force for the corner
    cnt = {tuple(b): 1 for b in block}
    cnt[(11, 21, 31)] = 2
    dip = np.zeros(3)
    tot = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                k = cnt.get((10 + dx, 20 + dy, 30 + dz), 0)
                dip += k * np.array([dx, dy, dz])
                tot += k
    assert tot == 9 and f[corner, 8] == pytest.approx(np.linalg.norm(dip) / (tot + 1))
    assert f[c, 0] == pytest.approx(math.log(0.2)) and f[c, 2] == pytest.approx(math.log(2.0))
    assert f[c, 1] == pytest.approx(0.01 * 64 * (32.0 / 65536.0) / 0.2)
    assert f[0, 10] == 1 and f[0, 11] == 0 and f[-1, 10] == 0 and f[-1, 11] == 1


def test_features_wrap_the_periodic_cells():
    """Cells 255 and 0 are neighbours (sort.metal:15-44 wraps at 2^CELL_POWER)."""
    pos = _one_particle_per_cell([(255, 5, 5), (0, 5, 5), (256, 5, 5)])
    f = mlp.features(pos, np.full(3, 0.2), np.zeros(3), 1, np.ones(3), [400] * 3)
    # (0,5,5) and (256,5,5) are the same cell; (255,5,5) is its face neighbour
    assert f[1, 4] == pytest.approx(math.log1p(2)) and f[1, 5] == pytest.approx(math.log1p(1))
    assert f[0, 5] == pytest.approx(math.log1p(2))
