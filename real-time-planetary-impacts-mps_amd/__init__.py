"""MI355X (gfx950) SPH + FMM step core -- ctypes bindings.

The compute path is native: ``lib/libsphhip.so`` (hand-written HIP kernels
behind the C ABI of ``include/sph_capi.h``) and ``lib/libsphhost.so`` (the
C++ host layer of ``include/sph_host.h``: synthetic inputs, the CPU octree
builder and the ``Compute``/``Headless`` driver).  This module only moves
numpy arrays across that boundary.  Loading fails loudly when the libraries
have not been built; there is no Python fallback for any pass.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libsphhip.so")
HOST_LIB = os.path.join(LIB_DIR, "libsphhost.so")

# ---------------------------------------------------------------- constants
SPH_OK = 0
SPH_E_ARG, SPH_E_STATE, SPH_E_NOTREE, SPH_E_CAPACITY, SPH_E_NODEVICE = -1, -2, -3, -4, -5
SPH_FIX_M2M_ZZ = 1 << 0
SPH_FIX_SYNC_RANGE = 1 << 1
SPH_FIX_SOFTENING = 1 << 2
SPH_FIX_DENSITY_CACHE = 1 << 3
SPH_FIX_NEAR_FIELD = 1 << 4
SPH_FIX_ORDER3 = 32
SPH_FIX_INACTIVE_DRIFT = 64
SPH_FIX_ALL = 0x7F
SPH_ACTIVE_AS_WRITTEN = 0
SPH_ACTIVE_README = 1
TREE_ASYNC, TREE_EVERY, TREE_FIXED, TREE_GPU = 0, 1, 2, 3

FIELDS = [
    "pos", "vel", "acc", "rho", "u", "mass", "pressure", "h", "mat", "gradterm",
    "soundspeed", "du", "balsara", "dhdt", "active", "next_active", "gravnorm",
    "cell_array", "cell_start", "cell_end", "global_time", "dt", "nr_iters",
    "multipoles", "locals", "density_work", "potential", "dlnh",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
_F4 = {"pos", "vel", "acc"}
_I32 = {"mat", "next_active", "global_time", "dt", "nr_iters"}
_U32 = {"cell_array", "cell_start", "cell_end"}
PASSES = ["sort", "gravity_up", "gravity_down", "density", "acceleration", "step", "tree_build"]


class SphParams(ctypes.Structure):
    _fields_ = [
        ("viscosity_alpha", ctypes.c_float), ("density_eta", ctypes.c_float),
        ("max_smoothing_length", ctypes.c_float), ("max_nr_iterations", ctypes.c_int32),
        ("n_neighbours_estim", ctypes.c_int32), ("resolution_eta", ctypes.c_float),
        ("eos_min_rho", ctypes.c_float), ("eos_max_rho", ctypes.c_float),
        ("eos_min_u", ctypes.c_float), ("eos_max_u", ctypes.c_float),
        ("cfl", ctypes.c_float), ("dt0", ctypes.c_float), ("gamma", ctypes.c_float),
        ("r_max", ctypes.c_int32), ("activity_rule", ctypes.c_int32), ("max_rungs", ctypes.c_int32),
        ("gravity_eta", ctypes.c_float), ("gravity_g", ctypes.c_float), ("gravity_softening", ctypes.c_float),
        ("gravity_max_recursion", ctypes.c_int32), ("max_unchecked_pointers", ctypes.c_int32),
        ("max_children_in_leaf", ctypes.c_int32), ("max_tree_depth", ctypes.c_int32),
        ("cell_width", ctypes.c_float), ("cell_power", ctypes.c_int32),
        ("fixes", ctypes.c_uint32), ("precision", ctypes.c_int32), ("density_lanes", ctypes.c_int32),
        ("accel_lanes", ctypes.c_int32), ("lane_split", ctypes.c_int32),
        ("expansion_order", ctypes.c_int32), ("mac_order", ctypes.c_int32), ("density_flat", ctypes.c_int32),
        ("density_stop", ctypes.c_int32),
        ("density_order", ctypes.c_int32), ("list_density_lanes", ctypes.c_int32),
        ("list_accel_lanes", ctypes.c_int32), ("gather_reach", ctypes.c_int32),
        ("near_list_cap", ctypes.c_int32), ("small_n_lanes", ctypes.c_int32),
    ]


class SphEosTable(ctypes.Structure):
    _fields_ = [("material_id", ctypes.c_int32), ("res", ctypes.c_int32), ("pc", ctypes.POINTER(ctypes.c_float))]


class SphIcView(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("pos4", ctypes.POINTER(ctypes.c_float)), ("vel4", ctypes.POINTER(ctypes.c_float)),
        ("rho", ctypes.POINTER(ctypes.c_float)), ("u", ctypes.POINTER(ctypes.c_float)),
        ("m", ctypes.POINTER(ctypes.c_float)), ("mat", ctypes.POINTER(ctypes.c_int32)),
        ("p", ctypes.POINTER(ctypes.c_float)), ("h", ctypes.POINTER(ctypes.c_float)),
    ]


class SphTillotson(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("rho0", "A", "B", "E0", "Eiv", "Ecv", "a", "b", "alpha", "beta")]


class SphImpactConfig(ctypes.Structure):
    _fields_ = [
        ("n_total", ctypes.c_int32), ("target_fraction", ctypes.c_float), ("target_radius", ctypes.c_float),
        ("target_mass", ctypes.c_float), ("impactor_radius", ctypes.c_float), ("impactor_mass", ctypes.c_float),
        ("core_mass_fraction", ctypes.c_float), ("core_radius_fraction", ctypes.c_float), ("u0", ctypes.c_float),
        ("separation_factor", ctypes.c_float), ("impact_parameter", ctypes.c_float),
        ("closing_speed", ctypes.c_float), ("seed", ctypes.c_uint64),
    ]


class SphOctree(ctypes.Structure):
    _fields_ = [
        ("octree", ctypes.POINTER(ctypes.c_int32)), ("n_ints", ctypes.c_int64),
        ("level_nodes", ctypes.POINTER(ctypes.c_int32)), ("level_counts", ctypes.POINTER(ctypes.c_int32)),
        ("n_levels", ctypes.c_int32), ("node_count", ctypes.c_int32),
    ]


class SphParticles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32)] + [(k, ctypes.POINTER(ctypes.c_float)) for k in
                                          ("pos4", "vel4", "rho", "u", "m")] + [
        ("mat", ctypes.POINTER(ctypes.c_int32)), ("p", ctypes.POINTER(ctypes.c_float)),
        ("h", ctypes.POINTER(ctypes.c_float))]


# C-ABI symbols of include/sph_capi.h and include/sph_host.h
CAPI_SYMBOLS = [
    "sph_params_default", "sph_create", "sph_destroy", "sph_last_error", "sph_num_particles",
    "sph_host_alloc", "sph_host_free", "sph_snapshot_positions", "sph_wait_snapshot", "sph_upload_tree",
    "sph_tree_upload_done", "sph_sort", "sph_gravity", "sph_density", "sph_acceleration", "sph_step",
    "sph_sync", "sph_field_bytes", "sph_download", "sph_upload", "sph_set_timing", "sph_pass_times",
    "sph_reset_pass_times",
    "sph_mlp_create", "sph_mlp_destroy", "sph_mlp_last_error", "sph_mlp_get", "sph_mlp_set", "sph_mlp_forward",
    "sph_mlp_loss_grad", "sph_mlp_fit_norm", "sph_mlp_train", "sph_samples_create", "sph_samples_destroy",
    "sph_samples_count", "sph_samples_read", "sph_samples_append", "sph_samples_clear", "sph_set_sample_dump",
    "sph_set_learned_guess", "sph_features", "sph_conservation", "sph_time",
    "sph_build_tree", "sph_tree_info", "sph_download_tree", "sph_draw_snapshot", "sph_download_snapshot",
    "sph_snapshot_resolution", "sph_device_field", "sph_sort_noalloc", "sph_set_work_counters",
]
HOST_SYMBOLS = [
    "sph_tillotson_granite", "sph_tillotson_iron", "sph_make_eos_table", "sph_impact_config_default",
    "sph_impact_config_headon", "sph_make_impact", "sph_octree_build", "sph_octree_free",
    "sph_headless_create", "sph_headless_destroy", "sph_headless_step", "sph_headless_enqueue_step",
    "sph_headless_ctx", "sph_headless_tree_stats", "sph_headless_last_error",
    "sph_read_hdf", "sph_particles_free", "sph_write_hdf", "sph_checkpoint_write", "sph_checkpoint_read",
    "sph_hdf_last_error", "sph_headless_create_from_file", "sph_load_aneos", "sph_aneos_last_error",
    "sph_headless_create_with_aneos", "sph_headless_set_snapshots", "sph_write_snapshot_png",
    "sph_headless_render_buffers",
]

_hip: Optional[ctypes.CDLL] = None
_host: Optional[ctypes.CDLL] = None


class SphError(RuntimeError):
    pass


def _fp(a: np.ndarray, t=ctypes.c_float):
    return a.ctypes.data_as(ctypes.POINTER(t))


def libs():
    """Loads libsphhip.so and libsphhost.so (built by __graft_entry__.build())."""
    global _hip, _host
    if _hip is None:
        for p in (HIP_LIB, HOST_LIB):
            if not os.path.exists(p):
                raise SphError(f"{p} is missing: build the HIP extension first (__graft_entry__.build())")
        _hip = ctypes.CDLL(HIP_LIB, mode=ctypes.RTLD_GLOBAL)
        _host = ctypes.CDLL(HOST_LIB, mode=ctypes.RTLD_GLOBAL)
        h, o = _hip, _host
        h.sph_params_default.argtypes = [ctypes.POINTER(SphParams)]
        h.sph_params_default.restype = None
        h.sph_create.argtypes = [ctypes.POINTER(SphParams), ctypes.POINTER(SphIcView), ctypes.POINTER(SphEosTable),
                                 ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        h.sph_destroy.argtypes = [ctypes.c_void_p]
        h.sph_destroy.restype = None
        h.sph_last_error.argtypes = [ctypes.c_void_p]
        h.sph_last_error.restype = ctypes.c_char_p
        h.sph_num_particles.argtypes = [ctypes.c_void_p]
        for name in ("sph_sort", "sph_gravity", "sph_density", "sph_acceleration", "sph_step", "sph_sync",
                     "sph_wait_snapshot", "sph_tree_upload_done", "sph_reset_pass_times", "sph_build_tree"):
            getattr(h, name).argtypes = [ctypes.c_void_p]
        h.sph_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.sph_set_work_counters.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.sph_pass_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
        h.sph_field_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.sph_field_bytes.restype = ctypes.c_int64
        h.sph_download.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        h.sph_upload.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        h.sph_device_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int64)]
        h.sph_upload_tree.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
                                      ctypes.c_int32]
        I32 = ctypes.POINTER(ctypes.c_int32)
        h.sph_tree_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), I32, I32, I32, ctypes.c_int32]
        h.sph_download_tree.argtypes = [ctypes.c_void_p, I32, I32]
        h.sph_draw_snapshot.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.sph_download_snapshot.argtypes = [ctypes.c_void_p, I32]
        h.sph_snapshot_resolution.argtypes = [ctypes.c_void_p]
        h.sph_snapshot_positions.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        h.sph_sort_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        h.sph_host_alloc.argtypes = [ctypes.c_size_t]
        h.sph_host_alloc.restype = ctypes.c_void_p
        h.sph_host_free.argtypes = [ctypes.c_void_p]
        h.sph_host_free.restype = None
        F = ctypes.POINTER(ctypes.c_float)
        h.sph_mlp_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        h.sph_mlp_destroy.argtypes = [ctypes.c_void_p]
        h.sph_mlp_destroy.restype = None
        h.sph_mlp_last_error.argtypes = [ctypes.c_void_p]
        h.sph_mlp_last_error.restype = ctypes.c_char_p
        h.sph_mlp_get.argtypes = [ctypes.c_void_p, F, F]
        h.sph_mlp_set.argtypes = [ctypes.c_void_p, F, F]
        h.sph_mlp_forward.argtypes = [ctypes.c_void_p, F, ctypes.c_int64, F]
        h.sph_mlp_loss_grad.argtypes = [ctypes.c_void_p, F, F, F, ctypes.c_int64, F, F]
        h.sph_mlp_fit_norm.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.sph_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_uint64, F]
        h.sph_samples_create.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]
        h.sph_samples_destroy.argtypes = [ctypes.c_void_p]
        h.sph_samples_destroy.restype = None
        h.sph_samples_count.argtypes = [ctypes.c_void_p]
        h.sph_samples_count.restype = ctypes.c_int64
        h.sph_samples_read.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, F]
        h.sph_samples_append.argtypes = [ctypes.c_void_p, F, ctypes.c_int64]
        h.sph_samples_clear.argtypes = [ctypes.c_void_p]
        h.sph_set_sample_dump.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.sph_set_learned_guess.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.sph_features.argtypes = [ctypes.c_void_p, F]
        h.sph_conservation.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        h.sph_time.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        o.sph_tillotson_granite.argtypes = [ctypes.POINTER(SphTillotson)]
        o.sph_tillotson_iron.argtypes = [ctypes.POINTER(SphTillotson)]
        o.sph_make_eos_table.argtypes = [ctypes.POINTER(SphParams), ctypes.POINTER(SphTillotson), ctypes.c_int,
                                         ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]
        o.sph_impact_config_default.argtypes = [ctypes.POINTER(SphImpactConfig), ctypes.c_int32, ctypes.c_uint64]
        o.sph_impact_config_headon.argtypes = [ctypes.POINTER(SphImpactConfig), ctypes.c_int32, ctypes.c_uint64]
        o.sph_make_impact.argtypes = [ctypes.POINTER(SphParams), ctypes.POINTER(SphImpactConfig),
                                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int] + \
            [ctypes.POINTER(ctypes.c_float)] * 5 + [ctypes.POINTER(ctypes.c_int32)] + \
            [ctypes.POINTER(ctypes.c_float)] * 2
        o.sph_octree_build.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.POINTER(SphOctree)]
        o.sph_octree_free.argtypes = [ctypes.POINTER(SphOctree)]
        o.sph_octree_free.restype = None
        o.sph_headless_create.argtypes = [ctypes.POINTER(SphParams), ctypes.POINTER(SphImpactConfig), ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        o.sph_headless_destroy.argtypes = [ctypes.c_void_p]
        o.sph_headless_destroy.restype = None
        o.sph_headless_step.argtypes = [ctypes.c_void_p]
        o.sph_headless_enqueue_step.argtypes = [ctypes.c_void_p]
        o.sph_headless_ctx.argtypes = [ctypes.c_void_p]
        o.sph_headless_ctx.restype = ctypes.c_void_p
        o.sph_headless_tree_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(ctypes.c_double)]
        o.sph_headless_render_buffers.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_void_p)] * 3
        o.sph_headless_last_error.argtypes = [ctypes.c_void_p]
        o.sph_headless_last_error.restype = ctypes.c_char_p
        o.sph_read_hdf.argtypes = [ctypes.c_char_p, ctypes.POINTER(SphParticles)]
        o.sph_particles_free.argtypes = [ctypes.POINTER(SphParticles)]
        o.sph_particles_free.restype = None
        o.sph_write_hdf.argtypes = [ctypes.c_char_p, ctypes.POINTER(SphIcView), ctypes.POINTER(ctypes.c_float)]
        o.sph_checkpoint_write.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        o.sph_checkpoint_read.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        o.sph_headless_create_with_aneos.argtypes = [ctypes.POINTER(SphParams), ctypes.c_char_p, ctypes.c_char_p,
                                                     ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        o.sph_headless_set_snapshots.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        o.sph_write_snapshot_png.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.c_char_p]
        o.sph_load_aneos.argtypes = [ctypes.c_char_p, ctypes.POINTER(SphParams), ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]
        o.sph_aneos_last_error.argtypes = []
        o.sph_aneos_last_error.restype = ctypes.c_char_p
        o.sph_hdf_last_error.argtypes = []
        o.sph_hdf_last_error.restype = ctypes.c_char_p
        o.sph_headless_create_from_file.argtypes = [ctypes.POINTER(SphParams), ctypes.c_char_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    return _hip, _host


# ------------------------------------------------------------------ inputs
def default_params(**overrides) -> SphParams:
    h, _ = libs()
    p = SphParams()
    h.sph_params_default(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def params_dict(p: SphParams) -> Dict[str, float]:
    return {name: getattr(p, name) for name, _ in SphParams._fields_ if name != "reserved"}


def make_eos_tables(params: SphParams, res: int = 256) -> Dict[int, np.ndarray]:
    """Synthetic Tillotson (P, c) tables: {400: granite, 402: iron}, each (res, res, 2) float32."""
    _, o = libs()
    out = {}
    for mat, fn in ((400, o.sph_tillotson_granite), (402, o.sph_tillotson_iron)):
        t = SphTillotson()
        fn(ctypes.byref(t))
        tab = np.zeros((res, res, 2), np.float32)
        rc = o.sph_make_eos_table(ctypes.byref(params), ctypes.byref(t), res, mat, _fp(tab))
        if rc != SPH_OK:
            raise SphError(f"sph_make_eos_table failed: {rc}")
        out[mat] = tab
    return out


def impact_config(n_total: int, seed: int, headon: bool = False, **overrides) -> SphImpactConfig:
    _, o = libs()
    c = SphImpactConfig()
    (o.sph_impact_config_headon if headon else o.sph_impact_config_default)(ctypes.byref(c), n_total, seed)
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def make_impact(params: SphParams, cfg: SphImpactConfig, tables: Dict[int, np.ndarray]) -> Dict[str, np.ndarray]:
    """Two-body initial conditions in DataStruct layout (hdfHandler.hpp:15-26)."""
    _, o = libs()
    n = cfg.n_total
    ic = {
        "pos": np.zeros((n, 4), np.float32), "vel": np.zeros((n, 4), np.float32),
        "rho": np.zeros(n, np.float32), "u": np.zeros(n, np.float32), "mass": np.zeros(n, np.float32),
        "mat": np.zeros(n, np.int32), "pressure": np.zeros(n, np.float32), "h": np.zeros(n, np.float32),
    }
    t400 = np.ascontiguousarray(tables[400])
    t402 = np.ascontiguousarray(tables[402])
    rc = o.sph_make_impact(ctypes.byref(params), ctypes.byref(cfg), _fp(t400), _fp(t402), t400.shape[0],
                           _fp(ic["pos"]), _fp(ic["vel"]), _fp(ic["rho"]), _fp(ic["u"]), _fp(ic["mass"]),
                           _fp(ic["mat"], ctypes.c_int32), _fp(ic["pressure"]), _fp(ic["h"]))
    if rc != SPH_OK:
        raise SphError(f"sph_make_impact failed: {rc}")
    return ic


def load_aneos(path: str, params: SphParams, res: int, fixed: bool = False) -> np.ndarray:
    """loadANEOSDataFromFile (ANEOS.cpp:21-175) through sph_load_aneos: the
    (P, c) table, shape (res, res, 2), row = density sample."""
    _, o = libs()
    out = np.zeros((res, res, 2), np.float32)
    if o.sph_load_aneos(os.fsencode(path), ctypes.byref(params), res, int(fixed), _fp(out)) != SPH_OK:
        raise SphError(f"sph_load_aneos: {o.sph_aneos_last_error().decode()}")
    return out


def write_snapshot_png(materials: np.ndarray, path: str):
    """sph_write_snapshot_png: the reference's snapshot colouring, as PNG."""
    _, o = libs()
    m = np.ascontiguousarray(materials, np.int32)
    if m.ndim != 2 or m.shape[0] != m.shape[1]:
        raise SphError("materials must be a square image")
    if o.sph_write_snapshot_png(_fp(m, ctypes.c_int32), m.shape[0], os.fsencode(path)) != SPH_OK:
        raise SphError(f"sph_write_snapshot_png: {o.sph_hdf_last_error().decode()}")


def read_hdf(path: str) -> Dict[str, np.ndarray]:
    """readHDFFile (hdfHandler.cpp:12-136) through sph_read_hdf: the
    DataStruct datasets of PartType0, as float32 / int32 arrays (pos and vel
    with 4 floats per particle)."""
    _, o = libs()
    t = SphParticles()
    if o.sph_read_hdf(os.fsencode(path), ctypes.byref(t)) != SPH_OK:
        raise SphError(f"sph_read_hdf: {o.sph_hdf_last_error().decode()}")
    try:
        n = t.n
        f = lambda p, shape: np.ctypeslib.as_array(p, shape=shape).copy()
        return {"pos": f(t.pos4, (n, 4)), "vel": f(t.vel4, (n, 4)), "rho": f(t.rho, (n,)), "u": f(t.u, (n,)),
                "mass": f(t.m, (n,)), "mat": f(t.mat, (n,)), "pressure": f(t.p, (n,)), "h": f(t.h, (n,))}
    finally:
        o.sph_particles_free(ctypes.byref(t))


def write_hdf(path: str, ic: Dict[str, np.ndarray], potentials: Optional[np.ndarray] = None):
    """writeHDFFile (hdfHandler.cpp:139-254) through sph_write_hdf."""
    _, o = libs()
    keep = {k: np.ascontiguousarray(ic[k], np.int32 if k == "mat" else np.float32)
            for k in ("pos", "vel", "rho", "u", "mass", "mat", "pressure", "h")}
    n = int(keep["mass"].shape[0])
    view = SphIcView(n, _fp(keep["pos"]), _fp(keep["vel"]), _fp(keep["rho"]), _fp(keep["u"]), _fp(keep["mass"]),
                     _fp(keep["mat"], ctypes.c_int32), _fp(keep["pressure"]), _fp(keep["h"]))
    pot = None if potentials is None else np.ascontiguousarray(potentials, np.float32)
    rc = o.sph_write_hdf(os.fsencode(path), ctypes.byref(view), None if pot is None else _fp(pot))
    if rc != SPH_OK:
        raise SphError(f"sph_write_hdf: {o.sph_hdf_last_error().decode()}")


def octree_build(pos4: np.ndarray, max_leaf: int = 4, max_depth: int = 7, nthreads: int = 1) -> Dict[str, object]:
    """CPU octree (octree.cpp layout): octree ints, per-level node offsets, node count."""
    _, o = libs()
    pos4 = np.ascontiguousarray(pos4, np.float32)
    t = SphOctree()
    rc = o.sph_octree_build(_fp(pos4), pos4.shape[0], max_leaf, max_depth, nthreads, ctypes.byref(t))
    if rc != SPH_OK:
        raise SphError(f"sph_octree_build failed: {rc}")
    try:
        octree = np.ctypeslib.as_array(t.octree, shape=(t.n_ints,)).copy()
        counts = np.ctypeslib.as_array(t.level_counts, shape=(t.n_levels,)).copy()
        nodes = np.ctypeslib.as_array(t.level_nodes, shape=(int(counts.sum()),)).copy()
        return {"octree": octree, "level_nodes": nodes, "level_counts": counts, "node_count": int(t.node_count)}
    finally:
        o.sph_octree_free(ctypes.byref(t))


# ----------------------------------------------------------------- context
SORT_NOALLOC_MAX_N = 4194304  # SPH_SORT_NOALLOC_MAX_N


def sort_noalloc(pos4_ptr: int, n: int, cell_width: float, cell_power: int, pairs_ptr: int, cell_se_ptr: int,
                 stream: int = 0) -> None:
    """sph_sort_noalloc on device buffers the caller owns: positions (16-byte
    stride), the sorted (key, index) pairs (2n uint32) and the interleaved
    (first, last) cell table (2 x 2^(3 cell_power) uint32).  Enqueues on
    `stream` (0 = the null stream) and does not wait."""
    rc = libs()[0].sph_sort_noalloc(ctypes.c_void_p(pos4_ptr), int(n), float(cell_width), int(cell_power),
                                    ctypes.c_void_p(pairs_ptr), ctypes.c_void_p(cell_se_ptr),
                                    ctypes.c_void_p(stream or None))
    if rc != 0:
        raise SphError(f"sph_sort_noalloc failed ({rc}): {libs()[0].sph_last_error(None).decode()}")


class Context:
    """One sph_ctx: device state of a simulation, driven pass by pass."""

    def __init__(self, params: SphParams, ic: Dict[str, np.ndarray], tables: Dict[int, np.ndarray]):
        h, _ = libs()
        self._h = h
        self.params = params
        self.n = int(ic["mass"].shape[0])
        self._keep = {k: np.ascontiguousarray(v) for k, v in ic.items()}
        view = SphIcView(self.n, _fp(self._keep["pos"]), _fp(self._keep["vel"]), _fp(self._keep["rho"]),
                         _fp(self._keep["u"]), _fp(self._keep["mass"]), _fp(self._keep["mat"], ctypes.c_int32),
                         _fp(self._keep["pressure"]), _fp(self._keep["h"]))
        self._tabs = [np.ascontiguousarray(tables[400]), np.ascontiguousarray(tables[402])]
        arr = (SphEosTable * 2)(SphEosTable(400, self._tabs[0].shape[0], _fp(self._tabs[0])),
                                SphEosTable(402, self._tabs[1].shape[0], _fp(self._tabs[1])))
        ptr = ctypes.c_void_p()
        rc = h.sph_create(ctypes.byref(params), ctypes.byref(view), arr, 2, ctypes.byref(ptr))
        if rc != SPH_OK:
            raise SphError(f"sph_create failed: {rc}")
        self.ptr = ptr

    @classmethod
    def from_handle(cls, ptr: int, params: SphParams):
        obj = cls.__new__(cls)
        obj._h = libs()[0]
        obj.params = params
        obj.ptr = ctypes.c_void_p(ptr)
        obj.n = obj._h.sph_num_particles(obj.ptr)
        obj._borrowed = True
        return obj

    def close(self):
        if getattr(self, "ptr", None) and not getattr(self, "_borrowed", False):
            self._h.sph_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != SPH_OK:
            raise SphError(f"{what} failed ({rc}): {self._h.sph_last_error(self.ptr).decode()}")

    def sort(self): self._check(self._h.sph_sort(self.ptr), "sph_sort")
    def gravity(self): self._check(self._h.sph_gravity(self.ptr), "sph_gravity")
    def density(self): self._check(self._h.sph_density(self.ptr), "sph_density")
    def acceleration(self): self._check(self._h.sph_acceleration(self.ptr), "sph_acceleration")
    def step(self): self._check(self._h.sph_step(self.ptr), "sph_step")
    def sync(self): self._check(self._h.sph_sync(self.ptr), "sph_sync")

    def upload_tree(self, tree: Dict[str, object]):
        oc = np.ascontiguousarray(tree["octree"], np.int32)
        ln = np.ascontiguousarray(tree["level_nodes"], np.int32)
        lc = np.ascontiguousarray(tree["level_counts"], np.int32)
        rc = self._h.sph_upload_tree(self.ptr, _fp(oc, ctypes.c_int32), oc.size, _fp(ln, ctypes.c_int32),
                                     _fp(lc, ctypes.c_int32), lc.size, int(tree["node_count"]))
        self._check(rc, "sph_upload_tree")
        self.sync()  # pageable source buffers: wait until the copy has landed

    def checkpoint_write(self, path: str):
        """sph_checkpoint_write: the context's state as an HDF5 checkpoint."""
        o = libs()[1]
        if o.sph_checkpoint_write(self.ptr, os.fsencode(path)) != SPH_OK:
            raise SphError(f"sph_checkpoint_write: {o.sph_hdf_last_error().decode()}")

    def checkpoint_read(self, path: str):
        """sph_checkpoint_read: restores a checkpoint into this context."""
        o = libs()[1]
        if o.sph_checkpoint_read(self.ptr, os.fsencode(path)) != SPH_OK:
            raise SphError(f"sph_checkpoint_read: {o.sph_hdf_last_error().decode()}")

    def draw_snapshot(self, res: int = 1024) -> np.ndarray:
        """sph_draw_snapshot + sph_download_snapshot: (res, res) int32 material ids."""
        self._check(self._h.sph_draw_snapshot(self.ptr, res), "sph_draw_snapshot")
        img = np.zeros((res, res), np.int32)
        self._check(self._h.sph_download_snapshot(self.ptr, _fp(img, ctypes.c_int32)), "sph_download_snapshot")
        return img

    def build_tree(self):
        """Octree built on the GPU from the current positions (sph_build_tree)."""
        self._check(self._h.sph_build_tree(self.ptr), "sph_build_tree")

    def download_tree(self) -> Dict[str, object]:
        """The current tree in the layout upload_tree takes."""
        n_ints, node_count, n_levels = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        lc = np.zeros(64, np.int32)
        rc = self._h.sph_tree_info(self.ptr, ctypes.byref(n_ints), ctypes.byref(node_count), ctypes.byref(n_levels),
                                   _fp(lc, ctypes.c_int32), lc.size)
        self._check(rc, "sph_tree_info")
        oc = np.zeros(n_ints.value, np.int32)
        ln = np.zeros(node_count.value, np.int32)
        self._check(self._h.sph_download_tree(self.ptr, _fp(oc, ctypes.c_int32), _fp(ln, ctypes.c_int32)),
                    "sph_download_tree")
        return {"octree": oc, "level_nodes": ln, "level_counts": lc[:min(n_levels.value, lc.size)].copy(),
                "node_count": int(node_count.value)}

    def _shape_dtype(self, field: str, nbytes: int):
        if field in _F4:
            return (self.n, 4), np.float32
        if field in ("multipoles", "locals"):
            # records of 24 / 16 floats, 40 / 32 with third-order expansions
            w = {"multipoles": 24, "locals": 16}[field] + (16 if self.params.expansion_order == 3 else 0)
            return (nbytes // (4 * w), w), np.float32
        if field == "cell_array":
            return (self.n, 2), np.uint32
        if field == "density_work":
            return (self.n, 4), np.int32
        if field in _U32:
            return (nbytes // 4,), np.uint32
        if field in _I32:
            return (nbytes // 4,), np.int32
        if field == "active":
            return (nbytes,), np.uint8
        return (nbytes // 4,), np.float32

    def download(self, field: str) -> np.ndarray:
        fid = FIELD_ID[field]
        nbytes = self._h.sph_field_bytes(self.ptr, fid)
        if nbytes < 0:
            raise SphError(f"unknown field {field}")
        shape, dt = self._shape_dtype(field, nbytes)
        out = np.zeros(shape, dt)
        self._check(self._h.sph_download(self.ptr, fid, out.ctypes.data_as(ctypes.c_void_p)), f"download {field}")
        return out

    def device_field(self, field: str):
        """(device address, bytes) of a per-particle field (sph_device_field)."""
        ptr, nb = ctypes.c_void_p(), ctypes.c_int64()
        self._check(self._h.sph_device_field(self.ptr, FIELD_ID[field], ctypes.byref(ptr), ctypes.byref(nb)),
                    "sph_device_field")
        return int(ptr.value or 0), int(nb.value)

    def upload(self, field: str, arr: np.ndarray):
        fid = FIELD_ID[field]
        nbytes = self._h.sph_field_bytes(self.ptr, fid)
        shape, dt = self._shape_dtype(field, nbytes)
        a = np.ascontiguousarray(arr, dt).reshape(shape)
        self._check(self._h.sph_upload(self.ptr, fid, a.ctypes.data_as(ctypes.c_void_p)), f"upload {field}")

    def set_timing(self, on: bool = True):
        self._check(self._h.sph_set_timing(self.ptr, 1 if on else 0), "sph_set_timing")

    def pass_times(self) -> Dict[str, Dict[str, float]]:
        ms = (ctypes.c_double * len(PASSES))()
        calls = (ctypes.c_int64 * len(PASSES))()
        self._check(self._h.sph_pass_times(self.ptr, ms, calls), "sph_pass_times")
        return {p: {"ms": ms[i], "calls": calls[i]} for i, p in enumerate(PASSES)}

    def set_work_counters(self, on: bool = True):
        """Records the reference algorithm's density work per particle
        (field "density_work") in the following density passes."""
        self._check(self._h.sph_set_work_counters(self.ptr, 1 if on else 0), "sph_set_work_counters")

    def reset_pass_times(self):
        self._check(self._h.sph_reset_pass_times(self.ptr), "sph_reset_pass_times")

    # learned smoothing-length guess
    def set_sample_dump(self, samples: Optional["Samples"]):
        self._check(self._h.sph_set_sample_dump(self.ptr, samples.ptr if samples is not None else None), "sph_set_sample_dump")

    def set_learned_guess(self, mlp: Optional["Mlp"]):
        self._check(self._h.sph_set_learned_guess(self.ptr, mlp.ptr if mlp is not None else None), "sph_set_learned_guess")

    def conservation(self) -> Dict[str, object]:
        """FP64 device sums: energies, momentum, angular momentum, mass."""
        out = (ctypes.c_double * 10)()
        self._check(self._h.sph_conservation(self.ptr, out), "sph_conservation")
        v = list(out)
        return {"kinetic": v[0], "internal": v[1], "potential": v[2], "momentum": np.array(v[3:6]),
                "angular_momentum": np.array(v[6:9]), "mass": v[9]}

    def time(self) -> float:
        t = ctypes.c_double()
        self._check(self._h.sph_time(self.ptr, ctypes.byref(t)), "sph_time")
        return float(t.value)

    def features(self) -> np.ndarray:
        out = np.zeros((self.n, MLP_INPUTS), np.float32)
        self._check(self._h.sph_features(self.ptr, _fp(out)), "sph_features")
        return out


MLP_INPUTS, MLP_HIDDEN, MLP_PARAMS, MLP_NORM, SAMPLE_FLOATS = 12, 64, 5057, 26, 16


class Samples:
    """Device buffer of (features, log(h_converged / h_start), weight) records."""

    def __init__(self, capacity: int):
        h, _ = libs()
        self._h = h
        ptr = ctypes.c_void_p()
        rc = h.sph_samples_create(int(capacity), ctypes.byref(ptr))
        if rc != SPH_OK:
            raise SphError(f"sph_samples_create failed: {rc}")
        self.ptr = ptr

    def __len__(self):
        return int(self._h.sph_samples_count(self.ptr))

    def read(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = len(self) - first if count is None else count
        out = np.zeros((count, SAMPLE_FLOATS), np.float32)
        rc = self._h.sph_samples_read(self.ptr, first, count, _fp(out))
        if rc != SPH_OK:
            raise SphError(f"sph_samples_read failed: {rc}")
        return out

    def append(self, rec: np.ndarray):
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, SAMPLE_FLOATS)
        rc = self._h.sph_samples_append(self.ptr, _fp(rec), rec.shape[0])
        if rc != SPH_OK:
            raise SphError(f"sph_samples_append failed: {rc}")

    def clear(self):
        self._h.sph_samples_clear(self.ptr)

    def close(self):
        if getattr(self, "ptr", None):
            self._h.sph_samples_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The learned smoothing-length network (BF16 MFMA forward/backward, Adam)."""

    def __init__(self, seed: int = 13):
        h, _ = libs()
        self._h = h
        ptr = ctypes.c_void_p()
        rc = h.sph_mlp_create(int(seed), ctypes.byref(ptr))
        if rc != SPH_OK:
            raise SphError(f"sph_mlp_create failed: {rc}")
        self.ptr = ptr

    def _check(self, rc, what):
        if rc != SPH_OK:
            raise SphError(f"{what} failed ({rc}): {self._h.sph_mlp_last_error(self.ptr).decode()}")

    def get(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        nm = np.zeros(MLP_NORM, np.float32)
        self._check(self._h.sph_mlp_get(self.ptr, _fp(p), _fp(nm)), "sph_mlp_get")
        return p, nm

    def set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._check(self._h.sph_mlp_set(self.ptr, None if p is None else _fp(p), None if nm is None else _fp(nm)),
                    "sph_mlp_set")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        out = np.zeros(x.shape[0], np.float32)
        self._check(self._h.sph_mlp_forward(self.ptr, _fp(x), x.shape[0], _fp(out)), "sph_mlp_forward")
        return out

    def loss_grad(self, x, d, w=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        d = np.ascontiguousarray(d, np.float32)
        wa = None if w is None else np.ascontiguousarray(w, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        self._check(self._h.sph_mlp_loss_grad(self.ptr, _fp(x), _fp(d), None if wa is None else _fp(wa), x.shape[0],
                                              _fp(loss), _fp(g)), "sph_mlp_loss_grad")
        return float(loss[0]), g

    def fit_norm(self, samples: Samples):
        self._check(self._h.sph_mlp_fit_norm(self.ptr, samples.ptr), "sph_mlp_fit_norm")

    def train(self, samples: Samples, steps: int, batch: int = 65536, lr: float = 1e-3, seed: int = 14):
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(self._h.sph_mlp_train(self.ptr, samples.ptr, steps, batch, lr, seed, _fp(losses)),
                    "sph_mlp_train")
        return losses[:steps]

    def close(self):
        if getattr(self, "ptr", None):
            self._h.sph_mlp_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Headless:
    """Headless::step driver (Headless.cpp:37-55) on synthetic inputs."""

    def __init__(self, params: SphParams, cfg: SphImpactConfig, eos_res: int = 256, tree_mode: int = TREE_ASYNC,
                 tree_threads: int = 8):
        _, o = libs()
        self._o = o
        self.params = params
        ptr = ctypes.c_void_p()
        rc = o.sph_headless_create(ctypes.byref(params), ctypes.byref(cfg), eos_res, tree_mode, tree_threads,
                                   ctypes.byref(ptr))
        if rc != SPH_OK:
            raise SphError(f"sph_headless_create failed: {rc}")
        self.ptr = ptr
        self.ctx = Context.from_handle(o.sph_headless_ctx(ptr), params)

    @classmethod
    def from_file(cls, params: SphParams, path: str, eos_res: int = 256, tree_mode: int = TREE_ASYNC,
                  tree_threads: int = 8, aneos: Optional[Dict[int, str]] = None, fixed_resampler: bool = False):
        """Headless::Headless(filepath) (Headless.cpp:19-23): initial
        conditions from an HDF5 file; EOS tables from ANEOS files
        ({400: path, 402: path}) or, without them, the Tillotson tables."""
        _, o = libs()
        obj = cls.__new__(cls)
        obj._o = o
        obj.params = params
        ptr = ctypes.c_void_p()
        if aneos:
            rc = o.sph_headless_create_with_aneos(ctypes.byref(params), os.fsencode(path), os.fsencode(aneos[400]),
                                                  os.fsencode(aneos[402]), eos_res, int(fixed_resampler),
                                                  tree_mode, tree_threads, ctypes.byref(ptr))
        else:
            rc = o.sph_headless_create_from_file(ctypes.byref(params), os.fsencode(path), eos_res, tree_mode,
                                                 tree_threads, ctypes.byref(ptr))
        if rc != SPH_OK:
            obj.ptr = None
            raise SphError(f"sph_headless_create_from_file failed ({rc}): {o.sph_hdf_last_error().decode()}")
        obj.ptr = ptr
        obj.ctx = Context.from_handle(o.sph_headless_ctx(ptr), params)
        return obj

    def set_snapshots(self, directory: Optional[str], res: int = 1024, interval: int = 1000):
        """sph_headless_set_snapshots: PNG snapshots into `directory` (None: off)."""
        d = None if directory is None else os.fsencode(directory)
        if self._o.sph_headless_set_snapshots(self.ptr, d, res, interval) != SPH_OK:
            raise SphError("sph_headless_set_snapshots failed")

    def step(self):
        rc = self._o.sph_headless_step(self.ptr)
        if rc != SPH_OK:
            raise SphError(f"Headless::step failed ({rc}): {self._o.sph_headless_last_error(self.ptr).decode()}")

    def enqueue_step(self):
        rc = self._o.sph_headless_enqueue_step(self.ptr)
        if rc != SPH_OK:
            raise SphError(f"Headless::step failed ({rc}): {self._o.sph_headless_last_error(self.ptr).decode()}")

    def render_buffers(self):
        """Compute's positionBuffer, materialIdBuffer, densityBuffer (device addresses)."""
        p = [ctypes.c_void_p() for _ in range(3)]
        if self._o.sph_headless_render_buffers(self.ptr, *[ctypes.byref(x) for x in p]) != SPH_OK:
            raise SphError("sph_headless_render_buffers failed")
        return tuple(int(x.value or 0) for x in p)

    def tree_stats(self):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self._o.sph_headless_tree_stats(self.ptr, ctypes.byref(n), ctypes.byref(ms))
        return int(n.value), float(ms.value)

    def close(self):
        if getattr(self, "ptr", None):
            self._o.sph_headless_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
