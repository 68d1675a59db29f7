"""ctypes bindings of the in-tree native libraries.

``libludwig_hip.so`` is the product path (gfx950 kernels behind the C ABI of
include/ludwig_hip.h); ``libludwig_setup.so`` holds the host table producers
(include/ludwig_setup.h).  Loading fails loudly when a library is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_BUILD = Path(__file__).resolve().parent / "_build"

LB_OK, LB_EINVAL, LB_EHIP, LB_ENOMEM, LB_ESTATE, LB_ENONFINITE = 0, -1, -2, -3, -4, -5
LB_STORAGE_FP32, LB_STORAGE_FP16, LB_STORAGE_BF16 = 0, 1, 2
LB_FIELD_F, LB_FIELD_RHO, LB_FIELD_VEL, LB_FIELD_G, LB_FIELD_T = 0, 1, 2, 3, 4
LB_F_N = 18
FORCE_KEYS = ("Fx_pressure", "Fy_pressure", "Fz_pressure", "Fx_viscous", "Fy_viscous", "Fz_viscous",
              "Mx", "My", "Mz", "Fx", "Fy", "Fz", "Cd", "Cl", "Cs", "Cmx", "Cmy", "Cmz")

i32p = C.POINTER(C.c_int32)
i8p = C.POINTER(C.c_int8)
u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)


class LbConfig(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("domain_nx", C.c_int32), ("domain_ny", C.c_int32),
                ("domain_nz", C.c_int32), ("temporal_interpolation", C.c_int32), ("ddf_enabled", C.c_int32),
                ("q_min_threshold", C.c_float), ("storage", C.c_int32), ("strict_fp", C.c_int32),
                ("wide_offsets", C.c_int32), ("reserved", C.c_int32 * 6)]


class LbLevelDesc(C.Structure):
    _fields_ = [("level_id", C.c_int32), ("n_blocks", C.c_int32),
                ("map_x", i32p), ("map_y", i32p), ("map_z", i32p),
                ("neighbor_table", i32p), ("block_pointer", i32p),
                ("ptr_dim_x", C.c_int32), ("ptr_dim_y", C.c_int32), ("ptr_dim_z", C.c_int32),
                ("obstacle", u8p), ("sponge", f32p), ("wall_dist", f32p),
                ("tau", C.c_float), ("tau_g", C.c_float), ("dx", C.c_double),
                ("n_boundary_cells", C.c_int64),
                ("bouzidi_cell_block", i32p), ("bouzidi_cell_x", i8p), ("bouzidi_cell_y", i8p),
                ("bouzidi_cell_z", i8p), ("bouzidi_q_dense", u16p), ("bouzidi_q_cells", u16p)]


class LbStepParams(C.Structure):
    _fields_ = [("c_wale", C.c_float), ("nu_sgs_background", C.c_float), ("inlet_turbulence", C.c_float),
                ("symmetric", C.c_int32), ("wall_model", C.c_int32), ("sponge_blend_distributions", C.c_int32),
                ("omega_bulk", C.c_float), ("omega_3", C.c_float), ("omega_4", C.c_float),
                ("adaptive_omega4", C.c_int32), ("lambda_param", C.c_float), ("limiter", C.c_int32),
                ("compressible", C.c_int32), ("T_inlet", C.c_float), ("T_wall", C.c_float),
                ("wall_bc_adiabatic", C.c_int32), ("reserved", C.c_int32 * 8)]


class LbForceParams(C.Structure):
    _fields_ = [("mesh_offset", C.c_double * 3), ("velocity_scale", C.c_double), ("rho_physical", C.c_double),
                ("rho_ref", C.c_double), ("u_ref", C.c_double), ("area_ref", C.c_double),
                ("chord_ref", C.c_double), ("moment_center", C.c_double * 3), ("c_wale", C.c_float),
                ("nu_sgs_background", C.c_float), ("search_radius", C.c_int32), ("symmetric", C.c_int32),
                ("reserved", C.c_int32 * 8)]


HIP_EXPORTS = {
    "lb_last_error": (C.c_char_p, [C.c_void_p]),
    "lb_version": (C.c_char_p, []),
    "lb_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(LbConfig)]),
    "lb_destroy": (None, [C.c_void_p]),
    "lb_level_upload": (C.c_int, [C.c_void_p, C.POINTER(LbLevelDesc)]),
    "lb_init_equilibrium": (C.c_int, [C.c_void_p, C.c_float]),
    "lb_step_batch": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_float, C.POINTER(LbStepParams)]),
    "lb_synchronize": (C.c_int, [C.c_void_p]),
    "lb_mesh_upload": (C.c_int, [C.c_void_p, C.c_int64, f32p, f32p, f32p, f32p, f32p, f32p, f32p]),
    "lb_forces": (C.c_int, [C.c_void_p, C.POINTER(LbForceParams), f64p, f32p, f32p, f32p, f32p]),
    "lb_surface_maps": (C.c_int, [C.c_void_p, C.POINTER(LbForceParams), C.c_int32, f32p, f32p, f32p, f32p]),
    "lb_download_level": (C.c_int, [C.c_void_p, C.c_int, C.c_int, f32p]),
    "lb_upload_level": (C.c_int, [C.c_void_p, C.c_int, C.c_int, f32p]),
    "lb_stats": (C.c_int, [C.c_void_p, C.c_int, f64p]),
    "lb_check_finite": (C.c_int, [C.c_void_p, i64p]),
    "lb_stats_noalloc": (C.c_int, [C.c_void_p, C.c_int, f64p]),
    "lb_device_bytes": (C.c_int64, [C.c_void_p]),
    "lb_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "lb_profile_read": (C.c_int, [C.c_void_p, C.c_int, f64p, i64p]),
    "lb_profile_read_kernel": (C.c_int, [C.c_void_p, C.c_int, C.c_int, f64p, i64p]),
    "lb_bandwidth_probe": (C.c_int, [C.c_int, C.c_int64, C.c_int, f64p]),
    "lb_engine_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int32, C.c_int32]),
    "lb_engine_destroy": (None, [C.c_void_p]),
    "lb_engine_last_error": (C.c_char_p, [C.c_void_p]),
    "lb_engine_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, f32p, C.c_int32, C.c_int32,
                                C.POINTER(LbStepParams)]),
    "lb_engine_remove": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lb_engine_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "lb_engine_synchronize": (C.c_int, [C.c_void_p]),
    "lb_engine_case_progress": (C.c_int, [C.c_void_p, C.c_void_p, i64p]),
    "lb_engine_sample_forces": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(LbForceParams), C.c_int64, C.c_int64]),
    "lb_engine_case_forces": (C.c_int, [C.c_void_p, C.c_void_p, f64p, i64p]),
    "lb_engine_check_finite": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_int32)]),
    "lb_engine_set_options": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "lb_footprint": (C.c_int, [C.POINTER(LbConfig), C.POINTER(LbLevelDesc), C.c_int32, C.c_int64, C.c_int32,
                               i64p]),
    "lb_device_memory": (C.c_int, [C.c_int, i64p, i64p]),
    "lb_bouzidi_qmap_gpu": (C.c_int, [C.c_int, C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, f64p, C.c_int64,
                                      C.c_double, f64p, C.POINTER(i32p), C.POINTER(u16p), C.POINTER(i32p), i64p]),
    "lb_voxelize_gpu": (C.c_int, [C.c_int, C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, f64p, C.c_int64,
                                  C.c_double, f64p, u8p]),
    "lb_wall_distances_gpu": (C.c_int, [C.c_int, C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, u8p, C.c_double,
                                        f32p, i64p]),
    "lb_free": (None, [C.c_void_p]),
    "lb_setup_gpu_last_error": (C.c_char_p, []),
}

LB_SCHED_XCD = 0
LB_SCHED_LINEAR = 1
LB_SCHED_PERSISTENT = 2

SETUP_EXPORTS = {
    "lbs_last_error": (C.c_char_p, []),
    "lbs_free": (None, [C.c_void_p]),
    "lbs_stl_read": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(f64p), i64p]),
    "lbs_geometry_properties": (C.c_int, [f64p, C.c_int64, f64p, f64p, f64p]),
    "lbs_subdivide": (C.c_int, [f64p, C.c_int64, C.c_double, C.POINTER(f64p), C.POINTER(i64p), i64p]),
    "lbs_voxelize": (C.c_int, [C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, f64p, C.c_int64, C.c_double,
                               f64p, u8p]),
    "lbs_flood_fill": (C.c_int64, [C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, u8p]),
    "lbs_sponge": (C.c_int, [C.c_int64, i32p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_int, f32p]),
    "lbs_wall_distances": (C.c_int64, [C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, u8p, C.c_double, f32p]),
    "lbs_bouzidi_qmap": (C.c_int, [C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, f64p, C.c_int64,
                                   C.c_double, f64p, C.POINTER(i32p), C.POINTER(u16p), C.POINTER(i32p), i64p]),
    "lbs_double_to_half": (C.c_uint16, [C.c_double]),
}

_hip = None
_setup = None


def _load(name: str, exports: dict) -> C.CDLL:
    path = _BUILD / name
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build the native libraries first "
            f"(python -c 'import __graft_entry__ as g; g.build()' or python openludwig_amd/build.py)")
    lib = C.CDLL(str(path))
    for sym, (res, args) in exports.items():
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    return lib


def hip_lib() -> C.CDLL:
    global _hip
    if _hip is None:
        _hip = _load("libludwig_hip.so", HIP_EXPORTS)
    return _hip


def setup_lib() -> C.CDLL:
    global _setup
    if _setup is None:
        _setup = _load("libludwig_setup.so", SETUP_EXPORTS)
    return _setup


def ptr(a: np.ndarray | None, ctype):
    """Pointer to a C-contiguous numpy array (or NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


class NativeError(RuntimeError):
    pass


def check_setup(rc: int, what: str) -> None:
    if rc < 0:
        msg = setup_lib().lbs_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({rc}): {msg}")
