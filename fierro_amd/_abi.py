"""ctypes binding of include/fsgh.h (libfsgh.so, built in-tree by build.py).

There is no fallback: if the HIP library is missing or was built without
device code the import of the solver fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfsgh.so")

FSGH_OK = 0
RUN_GRAPH, RUN_STOP_ON_OUTPUT, RUN_PERSISTENT = 1, 2, 4
ERRORS = {-1: "FSGH_E_INVALID", -2: "FSGH_E_HIP", -3: "FSGH_E_NOMEM", -4: "FSGH_E_NONFINITE",
          -5: "FSGH_E_STATE", -6: "FSGH_E_UNSUPPORTED", -7: "FSGH_E_NOTBUILT"}

# enum fsgh_field
NODE_COORDS, NODE_VEL, NODE_MASS = 0, 1, 2
ELEM_DEN, ELEM_PRES, ELEM_SSPD, ELEM_SIE, ELEM_VOL, ELEM_DIV, ELEM_MASS = 3, 4, 5, 6, 7, 8, 9
ELEM_MAT_ID, ELEM_GAMMA, ELEM_CSMIN, CORNER_FORCE, ELEM_STRESS = 10, 11, 12, 13, 14

EOS_IDEAL_GAS, EOS_USER = 0, 1
STRENGTH_NONE, STRENGTH_HYPO, STRENGTH_HYPER = 0, 1, 2
STRENGTH_PARAMS = 8

# every symbol include/fsgh.h declares
EXPORTS = [
    "fsgh_create", "fsgh_destroy", "fsgh_last_error", "fsgh_device_count",
    "fsgh_get_corners_in_node", "fsgh_get_renumbering", "fsgh_get_device_csr", "fsgh_set_materials", "fsgh_set_bcs",
    "fsgh_bc_set_size", "fsgh_get_bc_set", "fsgh_setup_fills", "fsgh_set_material_models", "fsgh_upload", "fsgh_download", "fsgh_rk_init",
    "fsgh_get_timestep", "fsgh_get_divergence", "fsgh_get_force_sgh",
    "fsgh_update_velocity_sgh", "fsgh_boundary_velocity", "fsgh_update_energy_sgh",
    "fsgh_update_position_sgh", "fsgh_get_vol", "fsgh_update_state", "fsgh_energy",
    "fsgh_run_begin", "fsgh_run", "fsgh_run_status", "fsgh_profile_enable",
    "fsgh_profile_read", "fsgh_check", "fsgh_stream",
    "fsgh_forward", "fsgh_dt_history", "fsgh_backward", "fsgh_tangent",
    "fsgh_autotune", "fsgh_set_tuning", "fsgh_get_tuning",
    "fsgh_rz_extensive_mass", "fsgh_rz_update_node_mass",
    "fsgh_set_strength_params", "fsgh_get_velgrad", "fsgh_get_stage_info",
    "fsgh_energy_noalloc",
]
NOALLOC_MAX_ENTRIES = 2147221503

KERNEL_DIV_EXTREMES, KERNEL_FORCE, KERNEL_VEL_BC_POS, KERNEL_ENERGY_STATE_DIV = 0, 1, 2, 3
NUM_TUNED_KERNELS = 4
TUNED_KERNEL_NAMES = ["div_extremes", "force", "vel_bc_pos", "energy_state_div"]


class MeshDesc(C.Structure):
    _fields_ = [("num_nodes", C.c_int64), ("num_elems", C.c_int64),
                ("num_instances", C.c_int32), ("renumber", C.c_int32),
                ("instance_node_off", C.c_void_p), ("instance_elem_off", C.c_void_p),
                ("num_dims", C.c_int32), ("reserved_", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("q1", C.c_double), ("q1ex", C.c_double), ("q2", C.c_double),
                ("q2ex", C.c_double)]


class Boundary(C.Structure):
    _fields_ = [("surface", C.c_int32), ("hydro_bc", C.c_int32), ("value", C.c_double)]


class Fill(C.Structure):
    _fields_ = [("volume", C.c_int32), ("mat_id", C.c_int32),
                ("x1", C.c_double), ("x2", C.c_double), ("y1", C.c_double),
                ("y2", C.c_double), ("z1", C.c_double), ("z2", C.c_double),
                ("radius1", C.c_double), ("radius2", C.c_double),
                ("velocity", C.c_int32), ("pad_", C.c_int32),
                ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("speed", C.c_double), ("sie", C.c_double), ("den", C.c_double)]


class RunParams(C.Structure):
    _fields_ = [("time_final", C.c_double), ("dt_start", C.c_double),
                ("dt_max", C.c_double), ("dt_min", C.c_double), ("dt_cfl", C.c_double),
                ("graphics_dt_ival", C.c_double), ("graphics_cyc_ival", C.c_int64),
                ("cycle_stop", C.c_int64), ("fuzz", C.c_double), ("small", C.c_double),
                ("tiny", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("block", C.c_int32), ("min_waves", C.c_int32),
                ("xcd_order", C.c_int32), ("ms", C.c_double), ("ms_default", C.c_double)]


class InstanceStatus(C.Structure):
    _fields_ = [("cycles", C.c_int64), ("time_value", C.c_double), ("dt", C.c_double),
                ("graphics_time", C.c_double), ("stopped", C.c_int32),
                ("write_pending", C.c_int32)]


_lib = None


# test-only variant of the library: the user strength hook runs the
# hypoelastic test model (built with FSGH_TEST_STRENGTH_MODEL by build.py)
STRENGTH_TEST_LIB_PATH = os.path.join(_HERE, "libfsgh_strength_test.so")
_libs: dict = {}


def load(path: str | None = None) -> C.CDLL:
    """Load libfsgh.so (or the library at ``path``); raise if it is not built."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    key = os.path.abspath(path or LIB_PATH)
    if key in _libs:
        return _libs[key]
    if not os.path.exists(key):
        raise RuntimeError(f"{key} is not built: run fierro_amd/build.py "
                           "(or __graft_entry__.build()) first")
    L = C.CDLL(key)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "fsgh_create": [C.POINTER(MeshDesc), vp, vp, C.c_int, C.POINTER(vp)],
        "fsgh_destroy": [vp],
        "fsgh_get_corners_in_node": [vp, vp, vp],
        "fsgh_get_renumbering": [vp, vp, vp],
        "fsgh_get_device_csr": [vp, vp, vp],
        "fsgh_set_materials": [vp, i32, vp],
        "fsgh_set_material_models": [vp, i32, i32, i32],
        "fsgh_set_bcs": [vp, i32, i32, vp],
        "fsgh_bc_set_size": [vp, i32, i32, C.POINTER(i64)],
        "fsgh_get_bc_set": [vp, i32, i32, vp],
        "fsgh_setup_fills": [vp, i32, i32, vp, i32, i32, vp, vp],
        "fsgh_upload": [vp, i32, vp, C.c_size_t],
        "fsgh_download": [vp, i32, vp, C.c_size_t],
        "fsgh_rk_init": [vp],
        "fsgh_get_timestep": [vp, dbl, dbl, dbl, dbl, dbl, dbl, dbl, C.POINTER(dbl)],
        "fsgh_get_divergence": [vp],
        "fsgh_get_force_sgh": [vp, dbl, dbl, dbl, dbl],
        "fsgh_update_velocity_sgh": [vp, dbl, dbl],
        "fsgh_boundary_velocity": [vp],
        "fsgh_update_energy_sgh": [vp, dbl, dbl],
        "fsgh_update_position_sgh": [vp, dbl, dbl],
        "fsgh_get_vol": [vp],
        "fsgh_update_state": [vp, dbl, dbl],
        "fsgh_energy": [vp, i32, C.POINTER(dbl), C.POINTER(dbl)],
        "fsgh_run_begin": [vp, vp],
        "fsgh_run": [vp, i64, i32],
        "fsgh_run_status": [vp, i32, C.POINTER(InstanceStatus)],
        "fsgh_profile_enable": [vp, i32],
        "fsgh_profile_read": [vp, i32, C.c_char_p, C.c_size_t, vp, vp, C.POINTER(i32)],
        "fsgh_check": [vp],
        "fsgh_forward": [vp, vp, i64, i64, vp],
        "fsgh_dt_history": [vp, vp, i64],
        "fsgh_backward": [vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "fsgh_tangent": [vp, vp, vp, vp, vp, vp, vp, vp],
        "fsgh_autotune": [vp, i32, C.POINTER(Tuning)],
        "fsgh_set_tuning": [vp, C.POINTER(Tuning), i32],
        "fsgh_get_tuning": [vp, C.POINTER(Tuning), i32],
        "fsgh_rz_extensive_mass": [vp],
        "fsgh_set_strength_params": [vp, i32, i32, vp],
        "fsgh_get_stage_info": [vp, i32, vp, vp],
        "fsgh_get_velgrad": [vp, vp],
        "fsgh_rz_update_node_mass": [vp, dbl],
        "fsgh_energy_noalloc": [i32, i32, i64, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.fsgh_last_error.argtypes = [vp]
    L.fsgh_last_error.restype = C.c_char_p
    L.fsgh_device_count.argtypes = []
    L.fsgh_device_count.restype = C.c_int
    L.fsgh_stream.argtypes = [vp]
    L.fsgh_stream.restype = vp
    _libs[key] = L
    if path is None:
        _lib = L
    return L


class FsghError(RuntimeError):
    pass


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != FSGH_OK:
        msg = ""
        if ctx:
            msg = (load().fsgh_last_error(ctx) or b"").decode(errors="replace")
        raise FsghError(f"{what}: {ERRORS.get(rc, rc)} {msg}")
