"""FP64 CPU restatement of the learned smoothing-length network.

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward/backward kernels
(BF16 MFMA, FP32 accumulate).  The network has no counterpart in the
reference; it is the learned Newton-Raphson starting guess for the density
pass (density.metal:73-143 is the iteration it seeds).

Network: 12 -> 64 -> 64 -> 1, SiLU activations, on normalised features
    xn = (x - mu) / sigma,  prediction  d = y * y_sigma + y_mu,
where d estimates log(h_converged) - log(h_start).  Flat parameter order
(5,057 floats): W1[64][12], b1[64], W2[64][64], b2[64], w3[64], b3.
Loss: mean over the samples of w (y - (d_true - y_mu) / y_sigma)^2.
"""
import numpy as np

N_IN, N_HID = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1
N_NORM = 2 * N_IN + 2


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_HID, N_IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(N_HID, N_HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def normalise(norm, x):
    norm = np.asarray(norm, np.float64)
    return (np.asarray(x, np.float64) - norm[:N_IN]) / norm[N_IN:2 * N_IN]


def forward(params, norm, x):
    """Returns (prediction d, cache) for raw features x [n, 12]."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    xn = normalise(norm, x)
    z1 = xn @ W1.T + b1
    s1 = _sig(z1)
    a1 = z1 * s1
    z2 = a1 @ W2.T + b2
    s2 = _sig(z2)
    a2 = z2 * s2
    y = a2 @ w3 + b3
    norm = np.asarray(norm, np.float64)
    return y * norm[2 * N_IN + 1] + norm[2 * N_IN], (xn, z1, s1, a1, z2, s2, a2, y)


def loss_grad(params, norm, x, d_true, w=None):
    """Weighted MSE in the normalised target space and its gradient (FP64)."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    norm = np.asarray(norm, np.float64)
    _, (xn, z1, s1, a1, z2, s2, a2, y) = forward(params, norm, x)
    t = (np.asarray(d_true, np.float64) - norm[2 * N_IN]) / norm[2 * N_IN + 1]
    w = np.ones_like(t) if w is None else np.asarray(w, np.float64)
    W = float(t.size)
    r = y - t
    loss = float((w * r * r).sum() / W)
    dy = 2.0 * w * r / W
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = a2.T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, w3) * (s2 * (1 + z2 * (1 - s2)))
    g[OFF_W2:OFF_B2] = (dz2.T @ a1).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2) * (s1 * (1 + z1 * (1 - s1)))
    g[OFF_W1:OFF_B1] = (dz1.T @ xn).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def init_params(seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias
    (the same distribution the library draws; its generator differs)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    for off, n, fan in ((OFF_W1, N_HID * N_IN, N_IN), (OFF_B1, N_HID, N_IN), (OFF_W2, N_HID * N_HID, N_HID),
                        (OFF_B2, N_HID, N_HID), (OFF_W3, N_HID, N_HID), (OFF_B3, 1, N_HID)):
        p[off:off + n] = rng.uniform(-1, 1, n) / np.sqrt(fan)
    return p


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1) in FP64, in place."""
    m *= b1
    m += (1 - b1) * g
    v *= b2
    v += (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p -= lr * mh / (np.sqrt(vh) + eps)


def features(pos4, h, dhdt, dt_ticks, rho, mat, cell_width=0.18, cell_power=8, dt_min1=32.0 / 65536.0):
    """The 12 input features of particle i at the start of a density pass
    (FP64), from the state the pass starts from:
      0 log h, 1 (dh/dt) dt / h with dt = DT_MIN1 x the last tick step,
      2 log rho,
      3-7 log(1 + n) of the particle counts in the 27 cells around its own
          cell (sort.metal:15-44 cells, periodic): all 27, own, 6 faces,
          12 edges, 8 corners,
      8 |sum_k n_k offset_k| / (n_27 + 1), 9 (max_k n_k - min_k n_k) / (n_27 + 1),
      10-11 material one-hot (400, other)."""
    pos = np.asarray(pos4, np.float32)[:, :3]
    side = 1 << cell_power
    mask = side - 1
    c = np.floor(pos / np.float32(cell_width)).astype(np.int64)
    c = np.fmod(c, side)
    c[c < 0] += side
    key = (c[:, 0] * side + c[:, 1]) * side + c[:, 2]
    uniq, cnt = np.unique(key, return_counts=True)
    lookup = dict(zip(uniq.tolist(), cnt.tolist()))
    n = pos.shape[0]
    offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    counts = np.zeros((n, 27))
    for k, (dx, dy, dz) in enumerate(offs):
        kk = ((((c[:, 0] + dx) & mask) * side + ((c[:, 1] + dy) & mask)) * side + ((c[:, 2] + dz) & mask))
        counts[:, k] = [lookup.get(int(v), 0) for v in kk]
    order = np.array([abs(dx) + abs(dy) + abs(dz) for dx, dy, dz in offs])
    off = np.array(offs, np.float64)
    T = counts.sum(1)
    f = np.zeros((n, N_IN))
    h = np.asarray(h, np.float64)
    f[:, 0] = np.log(h)
    f[:, 1] = np.asarray(dhdt, np.float64) * (dt_min1 * float(dt_ticks)) / h
    f[:, 2] = np.log(np.asarray(rho, np.float64))
    f[:, 3] = np.log1p(T)
    for j, o in enumerate((0, 1, 2, 3)):
        f[:, 4 + j] = np.log1p(counts[:, order == o].sum(1))
    f[:, 8] = np.linalg.norm(counts @ off, axis=1) / (T + 1)
    f[:, 9] = (counts.max(1) - counts.min(1)) / (T + 1)
    mat = np.asarray(mat)
    f[:, 10] = mat == 400
    f[:, 11] = mat != 400
    return f
