"""HDF5 initial conditions and snapshots without libhdf5 (host/Hdf5.cpp,
host/Snapshot.cpp), on the CPU.

The reader is checked against the reference's demo file
(SPH/demo_impact_n40.hdf5, written by libhdf5), whose values the golden
fixture holds as converted by h5dump; the writer is checked by h5dump, which
must list and dump what it wrote.  Both checks need files or tools present
only where the reference checkout and the HDF5 tools are; they are skipped
elsewhere.  A round trip through the writer and reader, and malformed files,
run everywhere."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = "/root/reference/SPH/demo_impact_n40.hdf5"
H5DUMP = "/opt/conda/bin/h5dump"
FIELDS = ("pos", "vel", "rho", "u", "mass", "mat", "pressure", "h")


def _demo_golden():
    d = np.load(os.path.join(ROOT, "tests", "golden", "demo_impact_n40.npz"))
    return {k: d[k] for k in d.files}


def _ic(n, seed=0):
    rng = np.random.default_rng(seed)
    ic = {"pos": np.zeros((n, 4), np.float32), "vel": np.zeros((n, 4), np.float32)}
    ic["pos"][:, :3] = rng.normal(size=(n, 3))
    ic["vel"][:, :3] = rng.normal(size=(n, 3)) * 1e-2
    for k in ("rho", "u", "mass", "pressure", "h"):
        ic[k] = rng.random(n).astype(np.float32)
    ic["mat"] = rng.choice([400, 402], n).astype(np.int32)
    return ic


@pytest.mark.skipif(not os.path.exists(DEMO), reason="the reference's demo file is not present")
def test_reader_reproduces_the_reference_demo_file(sph):
    """readHDFFile on the reference's own file: FP64 coordinates rounded to
    FP32, every other dataset as stored; equal to the golden fixture."""
    d = sph.read_hdf(DEMO)
    g = _demo_golden()
    for k in FIELDS:
        a = d[k][:, :3] if k in ("pos", "vel") else d[k]
        assert a.dtype == g[k].dtype and np.array_equal(a, g[k]), k
    assert np.all(d["pos"][:, 3] == 0) and np.all(d["vel"][:, 3] == 0)


def test_write_read_round_trip(sph, tmp_path):
    ic = _ic(1000, 1)
    path = str(tmp_path / "snap.hdf5")
    sph.write_hdf(path, ic, potentials=np.linspace(-1, 0, 1000, dtype=np.float32))
    back = sph.read_hdf(path)
    for k in FIELDS:
        want = ic[k].copy()
        if k in ("pos", "vel"):
            want[:, 3] = 0          # the file holds 3 columns
        assert np.array_equal(back[k], want), k


def test_round_trip_of_the_demo_values(sph, tmp_path):
    g = _demo_golden()
    n = g["mass"].shape[0]
    ic = {k: g[k] for k in FIELDS if k not in ("pos", "vel")}
    for k in ("pos", "vel"):
        ic[k] = np.zeros((n, 4), np.float32)
        ic[k][:, :3] = g[k]
    path = str(tmp_path / "demo.hdf5")
    sph.write_hdf(path, ic)
    back = sph.read_hdf(path)
    for k in FIELDS:
        assert np.array_equal(back[k], ic[k]), k


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump is not installed")
def test_h5dump_reads_what_the_writer_wrote(sph, tmp_path):
    """writeHDFFile's layout (hdfHandler.cpp:232-248): h5dump lists every
    dataset with its type and shape and dumps the same bytes."""
    n = 777
    ic = _ic(n, 2)
    pot = np.arange(n, dtype=np.float32)
    path = str(tmp_path / "w.hdf5")
    sph.write_hdf(path, ic, potentials=pot)
    head = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True, cwd=str(tmp_path)).stdout
    for name, typ, shape in [("Coordinates", "H5T_IEEE_F32LE", f"( {n}, 3 )"), ("Velocities", "H5T_IEEE_F32LE", f"( {n}, 3 )"),
                             ("Densities", "H5T_IEEE_F32LE", f"( {n} )"), ("Entropies", "H5T_IEEE_F32LE", "( 0 )"),
                             ("MaterialIDs", "H5T_STD_I32LE", f"( {n} )"), ("ParticleIDs", "H5T_STD_I32LE", f"( {n} )"),
                             ("Potentials", "H5T_IEEE_F32LE", f"( {n} )")]:
        block = head.split(f'DATASET "{name}"')[1].split("}")[0]
        assert typ in block and shape in block, (name, block)
    want = {"Coordinates": (ic["pos"][:, :3], "<f4"), "Velocities": (ic["vel"][:, :3], "<f4"),
            "Densities": (ic["rho"], "<f4"), "InternalEnergies": (ic["u"], "<f4"), "Masses": (ic["mass"], "<f4"),
            "MaterialIDs": (ic["mat"], "<i4"), "Pressures": (ic["pressure"], "<f4"),
            "SmoothingLengths": (ic["h"], "<f4"), "Potentials": (pot, "<f4"),
            "ParticleIDs": (np.arange(n, dtype=np.int32), "<i4")}
    for name, (arr, dt) in want.items():
        out = str(tmp_path / (name + ".bin"))
        subprocess.run([H5DUMP, "-d", f"/PartType0/{name}", "-b", "LE", "-o", out, path], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=str(tmp_path))
        got = np.fromfile(out, dtype=dt).reshape(arr.shape)
        assert np.array_equal(got, arr), name


def test_reader_reports_missing_and_foreign_files(sph, tmp_path):
    with pytest.raises(sph.SphError, match="cannot open"):
        sph.read_hdf(str(tmp_path / "absent.hdf5"))
    p = tmp_path / "text.hdf5"
    p.write_bytes(b"not an hdf5 file" * 100)
    with pytest.raises(sph.SphError, match="not an HDF5 file"):
        sph.read_hdf(str(p))


def test_reader_reports_unsupported_versions(sph, tmp_path):
    ic = _ic(50, 3)
    path = tmp_path / "v.hdf5"
    sph.write_hdf(str(path), ic)
    b = bytearray(path.read_bytes())
    b[8] = 2                                   # superblock version 2
    path.write_bytes(bytes(b))
    with pytest.raises(sph.SphError, match="superblock version 2"):
        sph.read_hdf(str(path))


def test_reader_survives_truncated_and_corrupted_files(sph, tmp_path):
    """Every truncation and 300 random byte corruptions of a small file either
    read or raise SphError; none crashes or hangs."""
    ic = _ic(40, 4)
    src = tmp_path / "s.hdf5"
    sph.write_hdf(str(src), ic)
    data = src.read_bytes()
    bad = tmp_path / "bad.hdf5"
    outcomes = {"read": 0, "error": 0}
    for cut in list(range(0, len(data), max(1, len(data) // 97))) + [len(data) - 1]:
        bad.write_bytes(data[:cut])
        try:
            sph.read_hdf(str(bad))
            outcomes["read"] += 1
        except sph.SphError:
            outcomes["error"] += 1
    rng = np.random.default_rng(5)
    for _ in range(300):
        b = bytearray(data)
        for at in rng.integers(0, len(b), rng.integers(1, 6)):
            b[at] = int(rng.integers(0, 256))
        bad.write_bytes(bytes(b))
        try:
            sph.read_hdf(str(bad))
            outcomes["read"] += 1
        except sph.SphError:
            outcomes["error"] += 1
    print(outcomes)
    assert outcomes["error"] > 0


def test_headless_from_a_missing_file_reports_the_reader_error(sph, tmp_path):
    with pytest.raises(sph.SphError, match="cannot open"):
        sph.Headless.from_file(sph.default_params(), str(tmp_path / "absent.hdf5"))


@pytest.mark.skipif(not shutil.which("nm"), reason="nm is not installed")
def test_hdf_entry_points_exported(sph):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(sph.LIB_DIR, "libsphhost.so")],
                         capture_output=True, text=True, check=True).stdout
    for name in ("sph_read_hdf", "sph_write_hdf", "sph_particles_free", "sph_checkpoint_write",
                 "sph_checkpoint_read", "sph_hdf_last_error", "sph_headless_create_from_file"):
        assert f" T {name}" in out, name
