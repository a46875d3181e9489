"""ANEOS table ingestion (host/ANEOS.cpp, loadANEOSDataFromFile of
ANEOS.cpp:21-175) on synthetic tables written in the reference's text layout.
The reference's own table files are not in its checkout, so the pins are
closed forms:

  fixed resampler    reproduces a field linear in (rho, u) exactly (up to
                     FP32 rounding), inside the table and in extrapolation;
  as written (Q13)   returns the same linear field at the node-mirrored point
                     of each bracket (weights inverted), computed here in FP64
                     from brackets found by searchsorted;
  the u-search quirk when the upper density row ends below u, its search
                     runs into the next row of the contiguous array: the
                     expected brackets come from a flat-index search;
  parsing            13 header lines, lists over several lines with the rest
                     of a line dropped, too little data reported.
"""
import numpy as np
import pytest

HEADER = ["synthetic ANEOS-format table written by tests/test_aneos.py"] + [f"header line {i}" for i in range(2, 14)]


def _write(path, rho, T, U, P, C, per_line=5):
    """The reference's layout: header, "numRho numT", lists, then one line per
    (rho, T) with rho fastest: u P c s."""
    lines = list(HEADER) + [f"{len(rho)} {len(T)}"]
    for arr in (rho, T):
        for i in range(0, len(arr), per_line):
            lines.append(" ".join(f"{v:.9e}" for v in arr[i:i + per_line]))
    for t in range(len(T)):
        for r in range(len(rho)):
            lines.append(f"{U[r, t]:.9e} {P[r, t]:.9e} {C[r, t]:.9e} 0.0")
    path.write_text("\n".join(lines) + "\n")


def _axes(sph, prm, res):
    """The reference's FP32 axes (ANEOS.cpp:123-136)."""
    f32 = np.float32
    mr, Mr = f32(np.float64(prm.eos_min_rho) * 1e6), f32(np.float64(prm.eos_max_rho) * 1e6)
    mu, Mu = f32(np.float64(prm.eos_min_u) * 1e12), f32(np.float64(prm.eos_max_u) * 1e12)
    i = np.arange(res, dtype=np.float32) / f32(res)
    return (mr * np.power(Mr / mr, i, dtype=np.float32)).astype(np.float64), \
           (mu * np.power(Mu / mu, i, dtype=np.float32)).astype(np.float64)


def _grid(n_rho=24, n_t=20):
    rho = np.logspace(-0.5, 6.5, n_rho)                # kg/m^3, beyond the axes on both sides
    T = np.logspace(0.5, 7.0, n_t)                     # K
    # u(rho, T) increasing in T, differently for each row
    U = 50.0 * T[None, :] ** 1.6 * (1.0 + 0.05 * np.log10(rho)[:, None] ** 2)
    return rho, T, U


def _bracket(grid, x):
    """The reference's monotone search: largest i <= n-2 with x > grid[i]."""
    return np.minimum(np.searchsorted(grid[1:], x, side="left"), len(grid) - 2)


@pytest.fixture
def linear_table(tmp_path):
    rho, T, U = _grid()
    a, b, c0, c1 = 3.0e9, 2.5e3, 4.0e3, 0.7     # P = a rho + b u (Pa); c linear in rho and u (m/s)
    P = a * rho[:, None] + b * U
    C = c0 + 1e-3 * rho[:, None] + c1 * 1e-6 * U
    path = tmp_path / "linear.txt"
    _write(path, rho, T, U, P, C)
    return path, rho, T, U, (lambda r, u: (a * r + b * u) * 1e-18, lambda r, u: (c0 + 1e-3 * r + c1 * 1e-6 * u) * 1e-6)


def test_fixed_resampler_reproduces_a_linear_field(sph, linear_table):
    path, rho, T, U, (fp, fc) = linear_table
    prm = sph.default_params()
    res = 64
    tab = sph.load_aneos(str(path), prm, res, fixed=True).astype(np.float64)
    ra, ua = _axes(sph, prm, res)
    R, Uu = np.meshgrid(ra, ua, indexing="ij")
    for k, f in ((0, fp), (1, fc)):
        want = f(R, Uu)
        err = np.abs(tab[:, :, k] - want) / np.abs(want).max(axis=1, keepdims=True)
        print(f"component {k}: max error / row max {err.max():.2e}")
        assert err.max() < 2e-5


def test_as_written_resampler_mirrors_each_bracket(sph, linear_table):
    """Weights inverted (Q13): the lower node gets (x - x_lo)/(x_hi - x_lo).
    For a field linear along each row and across rows, that is the field at
    x_lo + x_hi - x within the bracket, along u in each of the two rows, then
    mixed across the rows with the inverted density weight."""
    path, rho, T, U, (fp, fc) = linear_table
    prm = sph.default_params()
    res = 48
    tab = sph.load_aneos(str(path), prm, res, fixed=False).astype(np.float64)
    ra, ua = _axes(sph, prm, res)
    for i in range(res):
        r = _bracket(rho, ra[i])
        w = (ra[i] - rho[r]) / (rho[r + 1] - rho[r])               # belongs to rho[r+1], multiplies row r
        jl = _bracket(U[r], ua)
        jh = _bracket(U[r + 1], ua)
        assert np.all(jh <= len(T) - 2)                              # the u-search quirk is inert here
        ul = U[r, jl] + U[r, jl + 1] - ua                            # mirrored u in the lower row
        uh = U[r + 1, jh] + U[r + 1, jh + 1] - ua
        for k, f in ((0, fp), (1, fc)):
            want = w * f(rho[r], ul) + (1 - w) * f(rho[r + 1], uh)
            got = tab[i, :, k]
            scale = np.maximum(np.abs(f(rho[r], U[r, jl + 1])), np.abs(f(rho[r + 1], U[r + 1, jh + 1])))
            assert np.all(np.abs(got - want) <= 2e-5 * scale + 1e-30), (i, k, np.abs(got - want).max())


def test_as_written_upper_row_search_runs_into_the_next_row(sph, tmp_path):
    """Upper density row ending below the lower row's energies: the reference
    bounds its search by the lower row's index, so the upper row's index runs
    past the row end and reads the next row of the contiguous array; the
    fixed form stops at the last interval."""
    rho = np.array([0.5, 2.0, 1e3, 1e5, 1e7])
    T = np.logspace(1, 4, 6)
    U = np.empty((5, 6))
    U[0] = np.logspace(2, 13, 6)        # row 0 reaches 1e13 J/kg
    U[1] = np.logspace(1, 6, 6)         # row 1 ends at 1e6 J/kg
    U[2:] = np.logspace(1, 13, 6)
    P = 1e18 * (1.0 + np.arange(30).reshape(5, 6))          # distinct per node: P in code units = 1 + flat index
    C = 1e6 * np.ones((5, 6))
    path = tmp_path / "quirk.txt"
    _write(path, rho, T, U, P, C)
    prm = sph.default_params()
    res = 32
    ra, ua = _axes(sph, prm, res)
    assert ra[0] > rho[0] and ra[1] < rho[1]               # the first density samples bracket rows 0 and 1
    flatU = U.reshape(-1)
    got = sph.load_aneos(str(path), prm, res, fixed=False)[0, :, 0].astype(np.float64)
    fixed = sph.load_aneos(str(path), prm, res, fixed=True)[0, :, 0].astype(np.float64)
    w = (ra[0] - rho[0]) / (rho[1] - rho[0])
    jl, jh, want = 0, 0, []
    for u in ua:
        while u > flatU[0 * 6 + jl + 1] and jl < 6 - 2:
            jl += 1
        while u > flatU[1 * 6 + jh + 1] and jl < 6 - 2:     # the lower row's index bounds the upper search
            jh += 1
        fl = (u - flatU[jl]) / (flatU[jl + 1] - flatU[jl])
        fh = (u - flatU[6 + jh]) / (flatU[6 + jh + 1] - flatU[6 + jh])
        Pf = 1.0 + np.arange(30)
        want.append(w * (Pf[jl] * fl + Pf[jl + 1] * (1 - fl)) + (1 - w) * (Pf[6 + jh] * fh + Pf[6 + jh + 1] * (1 - fh)))
    want = np.array(want)
    assert np.max(np.abs(got - want) / np.abs(want)) < 1e-5
    assert np.max(np.abs(got - fixed)) > 0.1                  # the quirk changes the table


def test_parse_layout_and_errors(sph, tmp_path):
    prm = sph.default_params()
    rho, T, U = _grid(6, 5)
    P = 1e18 * np.ones_like(U)
    C = 1e6 * np.ones_like(U)
    # lists over several lines; a value after the last density on its line is dropped
    lines = list(HEADER) + ["6 5", " ".join(f"{v:.9e}" for v in rho[:4]),
                            " ".join(f"{v:.9e}" for v in rho[4:]) + " 123.0",
                            " ".join(f"{v:.9e}" for v in T)]
    for t in range(5):
        for r in range(6):
            lines.append(f"{U[r, t]:.9e} {P[r, t]:.9e} {C[r, t]:.9e}")
    p = tmp_path / "ok.txt"
    p.write_text("\n".join(lines) + "\n")
    tab = sph.load_aneos(str(p), prm, 8, fixed=True)
    assert np.allclose(tab[..., 0], 1.0, rtol=1e-5) and np.allclose(tab[..., 1], 1.0, rtol=1e-5)
    short = tmp_path / "short.txt"
    short.write_text("\n".join(lines[:-3]) + "\n")
    with pytest.raises(sph.SphError, match="the table ends after"):
        sph.load_aneos(str(short), prm, 8)
    with pytest.raises(sph.SphError, match="cannot open the table file"):
        sph.load_aneos(str(tmp_path / "absent.txt"), prm, 8)
    bad = tmp_path / "bad.txt"
    bad.write_text("\n".join(HEADER + ["1 5"]) + "\n")
    with pytest.raises(sph.SphError, match="bad table size"):
        sph.load_aneos(str(bad), prm, 8)


def test_tillotson_table_through_the_ingestion_path(sph, tmp_path):
    """A table sampled from the granite Tillotson form on a (rho, T) grid and
    resampled by the fixed resampler agrees with the directly evaluated
    table where both are smooth (compressed states, below incipient
    vaporisation): the file path carries a realistic table end to end."""
    prm = sph.default_params()
    res = 128
    direct = sph.make_eos_tables(prm, res)[400].astype(np.float64)
    ra, ua = _axes(sph, prm, res)
    rho = np.logspace(3.0, 4.2, 60)
    T = np.logspace(0, 4.5, 60)
    U = np.outer(np.ones_like(rho), 1.0e3 * T)                       # u = cv T, cv = 1000 J/(kg K)
    # pressure and sound speed from the direct table at the nodes (bilinear in log axes)
    from scipy.interpolate import RegularGridInterpolator
    lr, lu = np.log(ra), np.log(ua)
    ip = RegularGridInterpolator((lr, lu), direct[:, :, 0] * 1e18, bounds_error=False, fill_value=None)
    ic = RegularGridInterpolator((lr, lu), direct[:, :, 1] * 1e6, bounds_error=False, fill_value=None)
    pts = np.stack([np.log(np.repeat(rho[:, None], len(T), 1)), np.log(U)], -1)
    P, C = ip(pts), ic(pts)
    path = tmp_path / "granite.txt"
    _write(path, rho, T, U, P, C)
    tab = sph.load_aneos(str(path), prm, res, fixed=True).astype(np.float64)
    sel_r = (ra > 2680.0) & (ra < 1.2e4)
    sel_u = (ua > 1e3 * 3) & (ua < 2.0e6)
    a, b = tab[np.ix_(sel_r, sel_u)][..., 0], direct[np.ix_(sel_r, sel_u)][..., 0]
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
    print(f"Tillotson through the file: median rel {np.median(rel):.2e}, q95 {np.quantile(rel, 0.95):.2e}")
    assert np.median(rel) < 1e-2
