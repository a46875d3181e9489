"""Host layer and C ABI on the CPU: synthetic inputs, the product octree
builder against the oracle's restatement of octree.cpp, and the shared
libraries' exported symbols.  No GPU calls."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# -------------------------------------------------------------- C ABI
def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header,lib", [("sph_capi.h", "libsphhip.so"), ("sph_host.h", "libsphhost.so")])
def test_libraries_export_every_declared_entry_point(sph, header, lib):
    path = os.path.join(sph.LIB_DIR, lib)
    declared = _declared(header)
    assert declared, header
    missing = [n for n in declared if n not in _exported(path)]
    assert not missing, missing
    handle = ctypes.CDLL(path)
    for n in declared:
        getattr(handle, n)


def test_python_binding_lists_match_headers(sph):
    assert sorted(sph.CAPI_SYMBOLS) == _declared("sph_capi.h")
    assert sorted(sph.HOST_SYMBOLS) == _declared("sph_host.h")


def test_default_params_are_the_reference_values(sph):
    p = sph.default_params()
    # Parameters.h:11-53
    assert p.viscosity_alpha == pytest.approx(1.5) and p.density_eta == pytest.approx(0.005)
    assert p.max_smoothing_length == 1.0 and p.max_nr_iterations == 50 and p.n_neighbours_estim == 128
    assert p.resolution_eta == pytest.approx(1.2348) and p.cfl == pytest.approx(0.2) and p.dt0 == 32.0
    assert p.r_max == 16 and p.gamma == pytest.approx(1.825742)
    assert p.gravity_eta == pytest.approx(0.005) and p.gravity_softening == pytest.approx(0.04)
    assert p.gravity_max_recursion == 1 and p.max_unchecked_pointers == 1024 and p.max_children_in_leaf == 4
    assert p.max_tree_depth == 7 and p.cell_width == pytest.approx(0.18) and p.cell_power == 8
    assert p.gravity_g == pytest.approx(6.67e-5) and p.fixes == 0
    assert p.activity_rule == sph.SPH_ACTIVE_AS_WRITTEN and p.max_rungs == 0
    assert p.precision == 0 and p.density_lanes == 8 and p.accel_lanes == 0 and p.lane_split == 0
    assert p.small_n_lanes == 262144
    assert p.expansion_order == 2 and p.mac_order == 0


def test_params_struct_layout_matches_c(sph):
    # 26 four-byte fields, precision, density_lanes, accel_lanes, lane_split, expansion_order, mac_order,
    # density_flat, density_stop, density_order, list_density_lanes, list_accel_lanes, gather_reach,
    # near_list_cap, small_n_lanes
    assert ctypes.sizeof(sph.SphParams) == 4 * (26 + 14)


# ---------------------------------------------------------------- EOS
def test_eos_tables_are_physical(sph, tables256):
    res = 256

    def row(rho_si):  # density sample i sits at 1 kg/m^3 * 1e6^(i/res)
        return math.log(rho_si) / math.log(1e6) * res

    for mat, tab in tables256.items():
        P, c = tab[..., 0], tab[..., 1]
        assert np.all(np.isfinite(tab)) and np.all(P >= 0) and np.all(c > 0)
        # pressure grows with internal energy at fixed compressed density
        assert np.all(np.diff(P[200, :]) >= -1e-12)
    j = 10  # u ~ 250 J/kg: cold, so w = 1 + u / (E0 eta^2) = 1 to 1e-4
    for mat, rho0, A, B, ab in ((400, 2680.0, 1.8e10, 1.8e10, 1.8), (402, 7800.0, 1.28e11, 1.05e11, 2.0)):
        below, above = math.floor(row(rho0)), math.ceil(row(rho0))
        # expanded cold material: tension floored at zero; compressed: P = (a + b) rho u + A mu + B mu^2
        assert tables256[mat][below, j, 0] == 0.0
        for i in (below, above):
            rho = 1e6 ** (i / res)
            mu = rho / rho0 - 1
            u = 100.0 * 1e10 ** (j / res)
            P = ab * rho * u + A * mu + B * mu * mu
            if i == above:
                assert tables256[mat][i, j, 0] == pytest.approx(P * 1e-18, rel=1e-3)
            # c^2 = dP/drho + (P / rho^2) dP/du, with the unfloored P (code velocity unit 1000 km/s)
            c2 = (A + 2 * B * mu) / rho0 + P * ab / rho
            assert tables256[mat][i, j, 1] == pytest.approx(math.sqrt(c2) * 1e-6, rel=2e-3)


def test_eos_tables_are_deterministic(sph):
    prm = sph.default_params()
    a = sph.make_eos_tables(prm, 64)
    b = sph.make_eos_tables(prm, 64)
    assert all(np.array_equal(a[k], b[k]) for k in a)


# ----------------------------------------------------- initial conditions
@pytest.mark.parametrize("headon", [False, True])
def test_impact_initial_conditions(sph, tables256, headon):
    prm = sph.default_params()
    n = 20000
    cfg = sph.impact_config(n, 9, headon=headon)
    ic = sph.make_impact(prm, cfg, tables256)
    m, pos, vel = ic["mass"].astype(np.float64), ic["pos"][:, :3].astype(np.float64), ic["vel"][:, :3].astype(np.float64)
    n_t = int(round(0.89 * n))
    assert m[:n_t].sum() == pytest.approx(5.3, rel=1e-5) and m[n_t:].sum() == pytest.approx(0.8, rel=1e-5)
    # centre-of-mass frame (sampling noise only)
    assert np.linalg.norm((m[:, None] * pos).sum(0) / m.sum()) < 0.05
    assert np.linalg.norm((m[:, None] * vel).sum(0) / m.sum()) < 1e-6
    # bodies: radii, cores, materials
    ct = (m[:n_t, None] * pos[:n_t]).sum(0) / m[:n_t].sum()
    ci = (m[n_t:, None] * pos[n_t:]).sum(0) / m[n_t:].sum()
    rt = np.linalg.norm(pos[:n_t] - ct, axis=1)
    ri = np.linalg.norm(pos[n_t:] - ci, axis=1)
    assert rt.max() <= 6.4 + 0.1 and ri.max() <= 3.6 + 0.1
    core_t = ic["mat"][:n_t] == 402
    assert core_t.sum() == round(0.3 * n_t)
    # (the mass-weighted centre carries sampling noise of ~R / sqrt(n))
    assert rt[core_t].max() <= 0.55 * 6.4 + 0.1 and rt[~core_t].min() >= 0.55 * 6.4 - 0.1
    # geometry: separation 1.5 (Rt + Ri), impact parameter b (Rt + Ri), closing speed
    d = ci - ct
    b = 0.3 if headon else math.sin(math.pi / 4)
    assert np.linalg.norm(d) == pytest.approx(1.5 * 10.0, rel=0.02)
    assert abs(d[1]) == pytest.approx(b * 10.0, rel=0.03)
    dv = vel[n_t:].mean(0) - vel[:n_t].mean(0)
    assert dv[0] == pytest.approx(0.012 if headon else 0.009, rel=1e-4)
    assert np.all(ic["u"] == np.float32(5e-6 if headon else 2e-6))
    # h0 = 1.2348 (m / rho0)^(1/3)
    assert np.allclose(ic["h"], 1.2348 * (ic["mass"] / ic["rho"]) ** (1 / 3), rtol=1e-5)


def test_impact_is_deterministic_and_seeded(sph, tables256):
    prm = sph.default_params()
    a = sph.make_impact(prm, sph.impact_config(1000, 5), tables256)
    b = sph.make_impact(prm, sph.impact_config(1000, 5), tables256)
    c = sph.make_impact(prm, sph.impact_config(1000, 6), tables256)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert not np.array_equal(a["pos"], c["pos"])


# ------------------------------------------------------------- octree
@pytest.mark.parametrize("n,seed", [(1, 1), (3, 2), (4, 3), (5, 4), (777, 5), (20000, 6)])
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_product_octree_equals_restatement(sph, orc, tables256, n, seed, threads):
    prm = sph.default_params()
    pos = sph.make_impact(prm, sph.impact_config(max(n, 10), seed), tables256)["pos"][:n]
    a = sph.octree_build(pos, 4, 7, threads)
    b = orc.build_octree(pos, 4, 7)
    for k in ("octree", "level_nodes", "level_counts"):
        assert np.array_equal(a[k], b[k]), k
    assert a["node_count"] == b["node_count"]


def test_octree_with_coincident_particles_stops_at_max_depth(sph, orc):
    pos = np.zeros((50, 4), np.float32)
    pos[25:, 0] = 1.0
    a = sph.octree_build(pos, 4, 7, 4)
    b = orc.build_octree(pos, 4, 7)
    assert np.array_equal(a["octree"], b["octree"])
    assert len(a["level_counts"]) == 8


def test_create_rejects_empty_and_invalid_inputs_before_touching_a_device(sph):
    """sph_create checks its arguments before it looks for a GPU: an empty
    particle set and out-of-range parameters are SPH_E_ARG on any machine."""
    h, _ = sph.libs()
    prm = sph.default_params()
    tabs = sph.make_eos_tables(prm, 8)
    t400 = np.ascontiguousarray(tabs[400])
    arr = (sph.SphEosTable * 2)(sph.SphEosTable(400, 8, sph._fp(t400)), sph.SphEosTable(402, 8, sph._fp(t400)))
    one = np.zeros((1, 4), np.float32)
    s1 = np.ones(1, np.float32)
    mat = np.full(1, 400, np.int32)
    f = sph._fp
    out = ctypes.c_void_p()
    for n in (0, -3):
        view = sph.SphIcView(n, f(one), f(one), f(s1), f(s1), f(s1), f(mat, ctypes.c_int32), f(s1), f(s1))
        assert h.sph_create(ctypes.byref(prm), ctypes.byref(view), arr, 2, ctypes.byref(out)) == sph.SPH_E_ARG
        assert not out.value
    view = sph.SphIcView(1, f(one), f(one), f(s1), f(s1), f(s1), f(mat, ctypes.c_int32), f(s1), f(s1))
    for field, bad in (("cell_power", 0), ("max_unchecked_pointers", 1), ("expansion_order", 4), ("density_lanes", 3),
                       ("density_order", 8), ("density_order", 48), ("density_order", 1024), ("small_n_lanes", -1)):
        p = sph.default_params(**{field: bad})
        assert h.sph_create(ctypes.byref(p), ctypes.byref(view), arr, 2, ctypes.byref(out)) == sph.SPH_E_ARG, field
    assert h.sph_create(ctypes.byref(prm), ctypes.byref(view), arr, 0, ctypes.byref(out)) == sph.SPH_E_ARG


def test_octree_builder_threads_under_thread_sanitizer(tmp_path):
    """ThreadSanitizer run of the host octree builder (the reference builds
    its tree on a separate CPU thread, Compute.cpp:456-516; this builder also
    splits the subtrees over threads): 16 threads on 1e5 particles, three
    times, with no race reported and the 1-thread tree reproduced."""
    exe = tmp_path / "octree_threads"
    pkg = os.path.join(ROOT, "real-time-planetary-impacts-mps_amd")
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(pkg, "host"), os.path.join(ROOT, "tests", "tsan", "octree_threads.cpp"),
         os.path.join(pkg, "host", "Octree.cpp"), "-o", str(exe)], capture_output=True, text=True)
    if build.returncode != 0 and "tsan" in build.stderr.lower():
        pytest.skip("ThreadSanitizer runtime not available: " + build.stderr[-200:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    run = subprocess.run([str(exe), "100000"], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr[-2000:]
    assert "ThreadSanitizer" not in run.stderr
    assert run.stdout.startswith("ok")
