"""CPU restatement of the reference's SPH + FMM step core (ctypes bindings).

TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  Only the
tests, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package; the product path never does.  See ``oracle/README.md``
for what pins each function.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "liboracle.so")
_lib: Optional[ctypes.CDLL] = None

_P = ctypes.POINTER


class OrcState(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("pos", _P(ctypes.c_float)), ("vel", _P(ctypes.c_float)), ("acc", _P(ctypes.c_float)),
        ("rho", _P(ctypes.c_float)), ("u", _P(ctypes.c_float)), ("m", _P(ctypes.c_float)),
        ("pressure", _P(ctypes.c_float)), ("h", _P(ctypes.c_float)), ("mat", _P(ctypes.c_int32)),
        ("gradterm", _P(ctypes.c_float)), ("soundspeed", _P(ctypes.c_float)), ("du", _P(ctypes.c_float)),
        ("balsara", _P(ctypes.c_float)), ("dhdt", _P(ctypes.c_float)), ("active", _P(ctypes.c_uint8)),
        ("next_active", _P(ctypes.c_int32)), ("gravnorm", _P(ctypes.c_float)),
        ("cell_array", _P(ctypes.c_uint32)), ("cell_start", _P(ctypes.c_uint32)), ("cell_end", _P(ctypes.c_uint32)),
        ("global_time", _P(ctypes.c_int32)), ("dt", _P(ctypes.c_int32)), ("nr_iters", _P(ctypes.c_int32)),
        ("eos_forsterite", _P(ctypes.c_float)), ("eos_fe", _P(ctypes.c_float)), ("eos_res", ctypes.c_int32),
        ("dlnh", _P(ctypes.c_float)),
    ]


class OrcTree(ctypes.Structure):
    _fields_ = [
        ("octree", _P(ctypes.c_int32)), ("n_ints", ctypes.c_int64), ("level_nodes", _P(ctypes.c_int32)),
        ("level_counts", _P(ctypes.c_int32)), ("n_levels", ctypes.c_int32), ("node_count", ctypes.c_int32),
    ]


class OrcGravityBuffers(ctypes.Structure):
    _fields_ = [
        ("multipoles", _P(ctypes.c_float)), ("locals", _P(ctypes.c_float)), ("parent_index", _P(ctypes.c_uint32)),
        ("list_a", _P(ctypes.c_int32)), ("list_b", _P(ctypes.c_int32)), ("list_len", ctypes.c_int64),
    ]


def build():
    """Compiles liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB)
        L = _lib
        L.orc_morton3d.argtypes = [ctypes.c_uint32] * 3
        L.orc_morton3d.restype = ctypes.c_uint32
        for name in ("orc_hash", "orc_radix_sort", "orc_find_cell_positions", "orc_sort", "orc_density",
                     "orc_acceleration", "orc_step"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _P(OrcState)]
            getattr(L, name).restype = None
        L.orc_W.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_W.restype = ctypes.c_float
        L.orc_dW_dh.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_dW_dh.restype = ctypes.c_float
        L.orc_gradW.argtypes = [_P(ctypes.c_float), ctypes.c_float, ctypes.c_float, ctypes.c_float, _P(ctypes.c_float)]
        L.orc_gradW.restype = None
        L.orc_dphi_dr.argtypes = [_P(ctypes.c_float), ctypes.c_float, ctypes.c_int]
        L.orc_dphi_dr.restype = ctypes.c_float
        L.orc_cell_range.argtypes = [ctypes.c_void_p, _P(ctypes.c_float), ctypes.c_float, _P(ctypes.c_int32)]
        L.orc_cell_range.restype = None
        L.orc_eos_sample.argtypes = [ctypes.c_void_p, _P(ctypes.c_float), ctypes.c_int32, ctypes.c_float,
                                     ctypes.c_float, _P(ctypes.c_float)]
        L.orc_eos_sample.restype = None
        L.orc_build_octree.argtypes = [_P(ctypes.c_float), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P(OrcTree)]
        L.orc_build_octree_morton.argtypes = [_P(ctypes.c_float), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              _P(OrcTree)]
        L.orc_free_tree.argtypes = [_P(OrcTree)]
        L.orc_free_tree.restype = None
        for name in ("orc_gravity_up", "orc_gravity_down"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _P(OrcState), _P(OrcTree), _P(OrcGravityBuffers)]
            getattr(L, name).restype = None
        L.orc_gravity_account.argtypes = [ctypes.c_void_p, _P(OrcState), _P(OrcTree), _P(OrcGravityBuffers),
                                          ctypes.c_int, _P(ctypes.c_int64), _P(ctypes.c_double),
                                          _P(ctypes.c_int64), _P(ctypes.c_double)]
        L.orc_gravity_account.restype = ctypes.c_int
        L.orc_direct_sum_fp64.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int32, _P(ctypes.c_int32),
                                          ctypes.c_int32, ctypes.c_double, ctypes.c_double, _P(ctypes.c_double),
                                          _P(ctypes.c_double)]
        L.orc_direct_sum_fp64.restype = None
        L.orc_pole_m2l.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_float]
        L.orc_pole_m2l.restype = None
        L.orc_pole_l2l.argtypes = [_P(ctypes.c_float)] * 3
        L.orc_pole_l2l.restype = None
        L.orc_pole_l2p.argtypes = [_P(ctypes.c_float)] * 3
        L.orc_pole_l2p.restype = None
        L.orc_pole_mac.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_float]
        L.orc_pole_mac.restype = ctypes.c_int
        L.orc_pole_shift_multipole.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int,
                                               _P(ctypes.c_float)]
        L.orc_pole_shift_multipole.restype = None
        L.orc_pole_derivatives.argtypes = [_P(ctypes.c_float), ctypes.c_float, _P(ctypes.c_float)]
        L.orc_pole_derivatives.restype = None
        L.orc_pole3_m2l.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_float]
        L.orc_pole3_m2l.restype = None
        L.orc_pole3_l2l.argtypes = [_P(ctypes.c_float)] * 3
        L.orc_pole3_l2l.restype = None
        L.orc_pole3_l2p.argtypes = [_P(ctypes.c_float)] * 3
        L.orc_pole3_l2p.restype = None
        L.orc_pole3_mac.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_float, ctypes.c_int]
        L.orc_pole3_mac.restype = ctypes.c_int
        L.orc_pole3_shift_multipole.argtypes = [_P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                                _P(ctypes.c_float)]
        L.orc_pole3_shift_multipole.restype = None
        L.orc_pole3_derivatives.argtypes = [_P(ctypes.c_float), ctypes.c_float, _P(ctypes.c_float)]
        L.orc_pole3_derivatives.restype = None
    return _lib


def _p(a: np.ndarray, t=ctypes.c_float):
    return a.ctypes.data_as(_P(t))


def _pp(params) -> ctypes.c_void_p:
    """Address of an sph_params ctypes structure (the product package's SphParams)."""
    return ctypes.cast(ctypes.byref(params), ctypes.c_void_p)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


def pole_m2l(local16, mp24, eps):
    out = f32(local16).copy()
    lib().orc_pole_m2l(_p(out), _p(f32(mp24)), eps)
    return out


def pole_l2l(local16, r3):
    out = np.zeros(16, np.float32)
    lib().orc_pole_l2l(_p(f32(local16)), _p(f32(r3)), _p(out))
    return out


def pole_l2p(local16, x3):
    out = np.zeros(3, np.float32)
    lib().orc_pole_l2p(_p(f32(local16)), _p(f32(x3)), _p(out))
    return out


def pole_mac(A24, B24, eta):
    return bool(lib().orc_pole_mac(_p(f32(A24)), _p(f32(B24)), eta))


def pole_shift_multipole(mp24, r3, fix_zz):
    out = np.zeros(24, np.float32)
    lib().orc_pole_shift_multipole(_p(f32(mp24)), _p(f32(r3)), 1 if fix_zz else 0, _p(out))
    return out


def pole_derivatives(r3, eps):
    out = np.zeros(10, np.float32)
    lib().orc_pole_derivatives(_p(f32(r3)), eps, _p(out))
    return out


def pole3_m2l(local32, mp40, eps):
    out = f32(local32).copy()
    lib().orc_pole3_m2l(_p(out), _p(f32(mp40)), eps)
    return out


def pole3_l2l(local32, r3):
    out = np.zeros(32, np.float32)
    lib().orc_pole3_l2l(_p(f32(local32)), _p(f32(r3)), _p(out))
    return out


def pole3_l2p(local32, x3):
    out = np.zeros(3, np.float32)
    lib().orc_pole3_l2p(_p(f32(local32)), _p(f32(x3)), _p(out))
    return out


def pole3_mac(A40, B40, eta, fix_p3=False):
    return bool(lib().orc_pole3_mac(_p(f32(A40)), _p(f32(B40)), eta, 1 if fix_p3 else 0))


def pole3_shift_multipole(mp40, r3, fix_zz, fix_p3=False):
    out = np.zeros(40, np.float32)
    lib().orc_pole3_shift_multipole(_p(f32(mp40)), _p(f32(r3)), 1 if fix_zz else 0, 1 if fix_p3 else 0, _p(out))
    return out


def pole3_derivatives(r3, eps):
    out = np.zeros(20, np.float32)
    lib().orc_pole3_derivatives(_p(f32(r3)), eps, _p(out))
    return out


def W(r, h1):
    return float(lib().orc_W(r, h1))


def dW_dh(r, h1):
    return float(lib().orc_dW_dh(r, h1))


def gradW(x3, r, r1, h1):
    out = np.zeros(3, np.float32)
    lib().orc_gradW(_p(f32(x3)), r, r1, h1, _p(out))
    return out


def dphi_dr(x3, h, fixed):
    return float(lib().orc_dphi_dr(_p(f32(x3)), h, 1 if fixed else 0))


def cell_range(params, x3, h):
    out = np.zeros(6, np.int32)
    lib().orc_cell_range(_pp(params), _p(f32(x3)), h, _p(out, ctypes.c_int32))
    return out[:3], out[3:]


def eos_sample(params, table, u, rho):
    out = np.zeros(2, np.float32)
    t = f32(table)
    lib().orc_eos_sample(_pp(params), _p(t), t.shape[0], u, rho, _p(out))
    return out


def morton3d(x, y, z):
    return int(lib().orc_morton3d(x, y, z))


def build_octree(pos4: np.ndarray, max_leaf: int = 4, max_depth: int = 7) -> Dict[str, object]:
    """octree.cpp:14-141 restated: octreeData, level lists, nodeValues."""
    L = lib()
    pos4 = np.ascontiguousarray(pos4, np.float32)
    t = OrcTree()
    if L.orc_build_octree(_p(pos4), pos4.shape[0], max_leaf, max_depth, ctypes.byref(t)) != 0:
        raise ValueError("build_octree: non-finite particle position")
    try:
        counts = np.ctypeslib.as_array(t.level_counts, shape=(t.n_levels,)).copy()
        return {
            "octree": np.ctypeslib.as_array(t.octree, shape=(t.n_ints,)).copy(),
            "level_nodes": np.ctypeslib.as_array(t.level_nodes, shape=(int(counts.sum()),)).copy(),
            "level_counts": counts,
            "node_count": int(t.node_count),
        }
    finally:
        L.orc_free_tree(ctypes.byref(t))


def build_octree_morton(pos4: np.ndarray, max_leaf: int = 4, max_depth: int = 7) -> Dict[str, object]:
    """The build's GPU octree restated (Morton keys of the bounding cube), in
    octree.cpp's flat layout."""
    L = lib()
    pos4 = np.ascontiguousarray(pos4, np.float32)
    t = OrcTree()
    rc = L.orc_build_octree_morton(_p(pos4), pos4.shape[0], max_leaf, max_depth, ctypes.byref(t))
    if rc != 0:
        raise ValueError(f"build_octree_morton failed ({rc})")
    try:
        counts = np.ctypeslib.as_array(t.level_counts, shape=(t.n_levels,)).copy()
        return {
            "octree": np.ctypeslib.as_array(t.octree, shape=(t.n_ints,)).copy(),
            "level_nodes": np.ctypeslib.as_array(t.level_nodes, shape=(int(counts.sum()),)).copy(),
            "level_counts": counts,
            "node_count": int(t.node_count),
        }
    finally:
        L.orc_free_tree(ctypes.byref(t))


def direct_sum(pos4: np.ndarray, mass: np.ndarray, targets: np.ndarray, G: float, softening: float = 0.0):
    """FP64 direct-sum acceleration (G included) and potential at `targets`."""
    L = lib()
    pos4 = np.ascontiguousarray(pos4, np.float32)
    mass = np.ascontiguousarray(mass, np.float32)
    targets = np.ascontiguousarray(targets, np.int32)
    acc = np.zeros((targets.size, 3), np.float64)
    phi = np.zeros(targets.size, np.float64)
    L.orc_direct_sum_fp64(_p(pos4), _p(mass), pos4.shape[0], _p(targets, ctypes.c_int32), targets.size,
                          G, softening, _p(acc, ctypes.c_double), _p(phi, ctypes.c_double))
    return acc, phi


class State:
    """The reference's device buffers as numpy arrays, advanced by the oracle."""

    def __init__(self, params, ic: Dict[str, np.ndarray], tables: Dict[int, np.ndarray]):
        self.params = params
        n = int(ic["mass"].shape[0])
        self.n = n
        ncells = 1 << (3 * params.cell_power)
        f32 = lambda a: np.ascontiguousarray(a, np.float32).copy()
        self.a = {
            "pos": f32(ic["pos"]).reshape(n, 4), "vel": f32(ic["vel"]).reshape(n, 4),
            "acc": np.zeros((n, 4), np.float32), "rho": f32(ic["rho"]), "u": f32(ic["u"]),
            "mass": f32(ic["mass"]), "pressure": f32(ic["pressure"]), "h": f32(ic["h"]),
            "mat": np.ascontiguousarray(ic["mat"], np.int32).copy(),
            "gradterm": np.zeros(n, np.float32), "soundspeed": np.zeros(n, np.float32),
            "du": np.zeros(n, np.float32), "balsara": np.zeros(n, np.float32), "dhdt": np.zeros(n, np.float32),
            "active": np.ones(n, np.uint8), "next_active": np.zeros(n, np.int32),
            "gravnorm": np.zeros(n, np.float32), "cell_array": np.zeros((n, 2), np.uint32),
            "cell_start": np.full(ncells, 0xFFFFFFFF, np.uint32), "cell_end": np.zeros(ncells, np.uint32),
            "global_time": np.zeros(1, np.int32), "dt": np.zeros(1, np.int32), "nr_iters": np.zeros(n, np.int32),
            "dlnh": np.zeros(n, np.float32),
        }
        self.tab400 = np.ascontiguousarray(tables[400], np.float32)
        self.tab402 = np.ascontiguousarray(tables[402], np.float32)
        self.tree = None
        self.grav = None

    def __getitem__(self, k):
        return self.a[k]

    def __setitem__(self, k, v):
        self.a[k][...] = v
        if k == "gravnorm":
            self._grav_known = True

    def _s(self) -> OrcState:
        a = self.a
        return OrcState(
            self.n, _p(a["pos"]), _p(a["vel"]), _p(a["acc"]), _p(a["rho"]), _p(a["u"]), _p(a["mass"]),
            _p(a["pressure"]), _p(a["h"]), _p(a["mat"], ctypes.c_int32), _p(a["gradterm"]), _p(a["soundspeed"]),
            _p(a["du"]), _p(a["balsara"]), _p(a["dhdt"]), _p(a["active"], ctypes.c_uint8),
            _p(a["next_active"], ctypes.c_int32), _p(a["gravnorm"]), _p(a["cell_array"], ctypes.c_uint32),
            _p(a["cell_start"], ctypes.c_uint32), _p(a["cell_end"], ctypes.c_uint32),
            _p(a["global_time"], ctypes.c_int32), _p(a["dt"], ctypes.c_int32), _p(a["nr_iters"], ctypes.c_int32),
            _p(self.tab400), _p(self.tab402), self.tab400.shape[0], _p(a["dlnh"]))

    def _run(self, name):
        s = self._s()
        getattr(lib(), name)(_pp(self.params), ctypes.byref(s))

    def hash(self): self._run("orc_hash")
    def radix_sort(self): self._run("orc_radix_sort")
    def find_cell_positions(self): self._run("orc_find_cell_positions")
    def sort(self): self._run("orc_sort")
    def density(self): self._run("orc_density")
    def acceleration(self): self._run("orc_acceleration")
    def step(self): self._run("orc_step")

    def set_tree(self, tree: Dict[str, object]):
        self.tree = {k: (np.ascontiguousarray(v, np.int32) if isinstance(v, np.ndarray) else v)
                     for k, v in tree.items()}
        K = self.tree["node_count"]
        counts = self.tree["level_counts"]
        max_nonlast = int(counts[:-1].max()) if counts.size > 1 else 1
        ll = max(max_nonlast, 1) * self.params.max_unchecked_pointers
        self.grav = {
            "multipoles": np.zeros((K, 40), np.float32), "locals": np.zeros((K, 32), np.float32),
            "parent_index": np.zeros(K, np.uint32), "list_a": np.zeros(ll, np.int32),
            "list_b": np.zeros(ll, np.int32), "list_len": ll,
        }

    def _tree_structs(self):
        t = self.tree
        ot = OrcTree(_p(t["octree"], ctypes.c_int32), t["octree"].size, _p(t["level_nodes"], ctypes.c_int32),
                     _p(t["level_counts"], ctypes.c_int32), t["level_counts"].size, t["node_count"])
        g = self.grav
        gb = OrcGravityBuffers(_p(g["multipoles"]), _p(g["locals"]), _p(g["parent_index"], ctypes.c_uint32),
                               _p(g["list_a"], ctypes.c_int32), _p(g["list_b"], ctypes.c_int32), g["list_len"])
        return ot, gb

    def gravity_up(self):
        ot, gb = self._tree_structs()
        s = self._s()
        lib().orc_gravity_up(_pp(self.params), ctypes.byref(s), ctypes.byref(ot), ctypes.byref(gb))

    def gravity_down(self):
        ot, gb = self._tree_structs()
        s = self._s()
        lib().orc_gravity_down(_pp(self.params), ctypes.byref(s), ctypes.byref(ot), ctypes.byref(gb))

    def gravity_account(self, plant: int = 0) -> Dict[str, object]:
        """The down pass with the interaction account (orc_gravity_account),
        after gravity_up: per particle the particles and FP64 mass it received
        through M2L, its leaf and its P2P source leaves, and the part counted
        twice through accepted ancestors.  `plant` 1/2 plants a dropped /
        doubled list entry (test use)."""
        ot, gb = self._tree_structs()
        s = self._s()
        n = self.n
        out = {"count": np.zeros(n, np.int64), "mass": np.zeros(n, np.float64),
               "excess_count": np.zeros(n, np.int64), "excess_mass": np.zeros(n, np.float64)}
        unexpected = lib().orc_gravity_account(
            _pp(self.params), ctypes.byref(s), ctypes.byref(ot), ctypes.byref(gb), int(plant),
            _p(out["count"], ctypes.c_int64), _p(out["mass"], ctypes.c_double),
            _p(out["excess_count"], ctypes.c_int64), _p(out["excess_mass"], ctypes.c_double))
        out["unexpected"] = int(unexpected)
        return out

    def gravity_down_on(self, multipoles24: np.ndarray):
        """The down pass alone on given multipole records (node_count x 24,
        the SPH_F_MULTIPOLES layout, order 2), after a gravity_up of the same
        tree (which supplies the parent indexes and activity flags)."""
        assert self.params.expansion_order != 3
        self.grav["multipoles"][:, :24] = multipoles24
        self.gravity_down()

    def gravity(self):
        """gravity.metal upPass + downPass.  The near-field form first
        estimates |a| with one as-written pass (P = 2, leaves force-accept)
        while no field is known, as the GPU path does: with gravNorm = 0 every
        MAC test rejects and the near field would sum every pair."""
        if (self.params.fixes & 16) and not getattr(self, "_grav_known", False):
            real = self.params
            boot = type(real).from_buffer_copy(real)
            boot.fixes = real.fixes & ~16
            boot.expansion_order = 2
            boot.mac_order = 0
            boot.precision = 0
            self.params = boot
            try:
                self.gravity_up()
                self.gravity_down()
            finally:
                self.params = real
        self.gravity_up()
        self.gravity_down()
        self._grav_known = True

    def full_step(self, tree: Optional[Dict[str, object]] = None):
        """Headless::step order: sort, gravity, density, acceleration, step."""
        if tree is not None:
            self.set_tree(tree)
        self.sort()
        self.gravity()
        self.density()
        self.acceleration()
        self.step()
