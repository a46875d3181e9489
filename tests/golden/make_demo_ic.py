#!/usr/bin/env python3
"""Writes tests/golden/demo_impact_n40.npz from the reference's demo initial
conditions (SPH/demo_impact_n40.hdf5), the one data fixture the reference
ships.

The datasets the reference reads (hdfHandler.cpp:120-129) are dumped as
little-endian binary with h5dump (the HDF5 tools of /opt/conda, present only
in the build container; there is no h5py) and converted the way the reader
converts them: H5Dread into H5T_NATIVE_FLOAT (FP64 coordinates rounded to
nearest FP32) and H5T_NATIVE_INT32 (hdfHandler.cpp:53,85,110).

    python tests/golden/make_demo_ic.py [path/to/demo_impact_n40.hdf5]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/SPH/demo_impact_n40.hdf5"
H5DUMP = "/opt/conda/bin/h5dump"

# dataset -> (on-disk dtype, components, output key, output dtype)
DATASETS = {
    "Coordinates": ("<f8", 3, "pos", np.float32),
    "Velocities": ("<f4", 3, "vel", np.float32),
    "Densities": ("<f4", 1, "rho", np.float32),
    "InternalEnergies": ("<f4", 1, "u", np.float32),
    "Masses": ("<f4", 1, "mass", np.float32),
    "MaterialIDs": ("<i4", 1, "mat", np.int32),
    "Pressures": ("<f4", 1, "pressure", np.float32),
    "SmoothingLengths": ("<f4", 1, "h", np.float32),
}


def dump(src, name, dtype, comps, tmp):
    out = os.path.join(tmp, name + ".bin")
    subprocess.run([H5DUMP, "-d", f"/PartType0/{name}", "-b", "LE", "-o", out, src], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tmp)
    a = np.fromfile(out, dtype=dtype)
    return a.reshape(-1, comps) if comps > 1 else a


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    # the on-disk types are checked against the header before conversion
    header = subprocess.run([H5DUMP, "-H", "-d", "/PartType0/Coordinates", src], capture_output=True, text=True,
                            check=True, cwd="/tmp").stdout
    assert "H5T_IEEE_F64LE" in header, header
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (dtype, comps, key, odt) in DATASETS.items():
            out[key] = dump(src, name, dtype, comps, tmp).astype(odt)
    n = out["mass"].shape[0]
    assert all(v.shape[0] == n for v in out.values())
    np.savez_compressed(os.path.join(HERE, "demo_impact_n40.npz"), **out)
    with open(src, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    meta = {"source": "SPH/demo_impact_n40.hdf5", "sha256": digest, "n": int(n),
            "datasets": {k: v[2] for k, v in DATASETS.items()},
            "conversion": "FP64 -> FP32 round to nearest, as H5Dread into H5T_NATIVE_FLOAT"}
    with open(os.path.join(HERE, "demo_impact_n40.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {n} particles")


if __name__ == "__main__":
    main()
