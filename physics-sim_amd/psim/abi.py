"""ctypes mirror of include/dem.h (struct layouts and entry points of libpsim.so)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

DEM_OK = 0
DEM_EINVAL = -1
DEM_ENOMEM = -2
DEM_EDEVICE = -3
DEM_ECAPACITY = -4
DEM_ESTATE = -5
DEM_EBUSY = -6
DEM_KE_NOALLOC_MAX = 4194304   # include/dem.h

UNIFORM_FIELDS = [
    ("walls", C.c_int32), ("hor_bound", C.c_float), ("vert_bound", C.c_float), ("round_bounds", C.c_int32),
    ("bound_radius", C.c_float), ("wall_friction", C.c_float), ("gravity", C.c_int32),
    ("planet_mode", C.c_int32), ("bonds", C.c_int32), ("collisions", C.c_int32),
    ("friction_coefficient", C.c_float), ("gravity_acc", C.c_float), ("bond_normal_stiffness", C.c_float),
    ("bonds_tear", C.c_int32), ("bond_normal_strength", C.c_float), ("contact_damping", C.c_float),
    ("bond_damping", C.c_float), ("drag", C.c_float), ("bond_shear_strength", C.c_float), ("dT", C.c_float),
    ("bond_shear_stiffness", C.c_float), ("gravity_x", C.c_float), ("gravity_y", C.c_float),
    ("mouse_gravity", C.c_int32), ("moment_contribution_factor", C.c_float),
    ("collision_interval", C.c_int32), ("local_damping", C.c_int32), ("local_damping_alpha", C.c_float),
    ("particles", C.c_int32),
]


class DemUniform(C.Structure):
    _fields_ = UNIFORM_FIELDS


assert C.sizeof(DemUniform) == 29 * 4


class DemMaterial(C.Structure):
    _fields_ = [("red", C.c_float), ("green", C.c_float), ("blue", C.c_float), ("density", C.c_float),
                ("normal_stiffness", C.c_float), ("shear_stiffness", C.c_float)]


class DemGrid(C.Structure):
    _fields_ = [("cell_size", C.c_float), ("cell_cap", C.c_int32), ("w", C.c_int32), ("h", C.c_int32)]


class DemParams(C.Structure):
    _fields_ = [("u", DemUniform), ("materials", C.POINTER(DemMaterial)), ("n_materials", C.c_int32),
                ("grid", DemGrid), ("use_f64", C.c_int32), ("record_data", C.c_int32),
                ("reorder_interval", C.c_int32), ("block_size", C.c_int32), ("graph", C.c_int32),
                ("restitution", C.c_float)]


FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)


class DemStateView(C.Structure):
    _fields_ = [("n", C.c_int32), ("pos", FP), ("vel", FP), ("acc", FP), ("rot", FP), ("rot_vel", FP),
                ("rot_acc", FP), ("forces", FP), ("radii", FP), ("fixity", IP), ("material_pointers", IP),
                ("bonds", IP), ("n_bonds", C.c_int32), ("bond_strengths", FP), ("contacts", FP),
                ("contact_slots", C.c_int32), ("data", FP)]


class DemServerParams(C.Structure):
    _fields_ = [("slot_capacity", C.c_int64), ("cell_capacity", C.c_int64), ("bond_capacity", C.c_int64),
                ("record_capacity", C.c_int64), ("scene_capacity", C.c_int32), ("material_capacity", C.c_int32),
                ("block_size", C.c_int32), ("record_data", C.c_int32), ("substeps", C.c_int32),
                ("max_resident", C.c_int32), ("bypass", C.c_int32), ("contact_slots", C.c_int32),
                ("graph", C.c_int32)]


DEM_REQ_QUEUED, DEM_REQ_RESIDENT, DEM_REQ_DONE, DEM_REQ_FAILED = 0, 1, 2, 3


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous, "need contiguous float32"
    return a.ctypes.data_as(FP)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous, "need contiguous int32"
    return a.ctypes.data_as(IP)


def uniform_from_dict(d) -> DemUniform:
    u = DemUniform()
    for name, _ in UNIFORM_FIELDS:
        if name in d:
            setattr(u, name, d[name])
    return u


def uniform_to_dict(u: DemUniform):
    return {name: getattr(u, name) for name, _ in UNIFORM_FIELDS}


def materials_array(mats) -> "C.Array":
    """mats: flat list/array of 6*M floats (settings.materials layout)."""
    m = np.asarray(mats, dtype=np.float32).reshape(-1, 6)
    arr = (DemMaterial * len(m))()
    for i, row in enumerate(m):
        arr[i] = DemMaterial(*[float(x) for x in row])
    return arr


# ---------------------------------------------------------------------------
# Product library (libpsim.so, built from physics-sim_amd/csrc)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG_DIR), "csrc", "libpsim.so")

EXPORTS = [
    "dem_last_error", "dem_version", "dem_device_count", "dem_create", "dem_destroy", "dem_set_params",
    "dem_upload_state", "dem_step", "dem_download_state", "dem_sync", "dem_kinetic_energy",
    "dem_kinetic_energy_noalloc",
    "dem_broken_bond_count", "dem_counters", "dem_get_datum", "dem_debug_broad_phase",
    "dem_debug_contacts", "dem_build_bond_pairs", "dem_group_sums", "dem_last_step_timing",
    "dem_server_create", "dem_server_destroy", "dem_server_submit", "dem_server_tick", "dem_server_status",
    "dem_server_take", "dem_server_stats",
]

_lib = None


class DemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dem error %d: %s" % (code, msg))
        self.code = code


def load_library(path: str = LIB_PATH):
    """Loads libpsim.so.  Raises if it is missing: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libpsim.so not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    lib.dem_last_error.restype = C.c_char_p
    lib.dem_version.restype = C.c_char_p
    lib.dem_device_count.restype = C.c_int
    lib.dem_create.argtypes = [C.POINTER(DemParams), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
    lib.dem_destroy.argtypes = [C.c_void_p]
    lib.dem_set_params.argtypes = [C.c_void_p, C.POINTER(DemParams)]
    lib.dem_upload_state.argtypes = [C.c_void_p, C.POINTER(DemStateView)]
    lib.dem_step.argtypes = [C.c_void_p, C.c_int32]
    lib.dem_download_state.argtypes = [C.c_void_p, C.POINTER(DemStateView)]
    lib.dem_sync.argtypes = [C.c_void_p]
    lib.dem_kinetic_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    lib.dem_kinetic_energy_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    lib.dem_broken_bond_count.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.dem_counters.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32]
    lib.dem_get_datum.argtypes = [C.c_void_p, IP, C.POINTER(C.c_double)]
    lib.dem_debug_broad_phase.argtypes = [C.c_void_p, IP, IP, IP]
    lib.dem_debug_contacts.argtypes = [C.c_void_p, C.c_int32, IP, IP]
    lib.dem_build_bond_pairs.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_float, IP, C.c_int32, IP]
    lib.dem_group_sums.argtypes = [C.c_void_p, IP, C.c_int32, C.POINTER(C.c_double)]
    lib.dem_last_step_timing.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.dem_server_create.argtypes = [C.POINTER(DemServerParams), C.POINTER(C.c_void_p)]
    lib.dem_server_destroy.argtypes = [C.c_void_p]
    lib.dem_server_submit.argtypes = [C.c_void_p, C.POINTER(DemParams), C.POINTER(DemStateView), C.c_int32,
                                      C.POINTER(C.c_int64)]
    lib.dem_server_tick.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.dem_server_status.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]
    lib.dem_server_take.argtypes = [C.c_void_p, C.c_int64, C.POINTER(DemStateView), C.POINTER(C.c_double)]
    lib.dem_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int32]
    _lib = lib
    return lib


def check(rc):
    if rc != DEM_OK:
        msg = _lib.dem_last_error().decode() if _lib is not None else ""
        raise DemError(rc, msg)
    return rc
