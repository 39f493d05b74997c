"""ctypes binding of liboal.so (include/oal.h).

The library must have been built (python -m openaerolab_alpha_01_amd.build); loading fails loudly
otherwise: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liboal.so"

OAL_ABI_VERSION = 3
OAL_OK, OAL_ERR_ARG, OAL_ERR_HIP, OAL_ERR_NOMEM, OAL_ERR_UNSUPPORTED, OAL_ERR_STATE, OAL_ERR_NONFINITE = 0, -1, -2, -3, -4, -5, -6
OAL_F32, OAL_F64, OAL_F16S, OAL_BF16S = 0, 1, 2, 3
OAL_STREAM_AB, OAL_STREAM_AA = 0, 1
OAL_BB_REFERENCE, OAL_BB_HALFWAY = 0, 1
OAL_LIMITER_NONE, OAL_LIMITER_FACTORED, OAL_LIMITER_POSITIVITY = 0, 1, 2
OAL_ROUND_STOCHASTIC, OAL_ROUND_NEAREST = 0, 1
OAL_STEP_FORCE = 1
OAL_BFL_FALLBACK_HALFWAY, OAL_BFL_FALLBACK_WALL_EQ, OAL_BFL_FALLBACK_BOUNCE_BACK = 0, 1, 2

STORAGE_CODES = {"f32": OAL_F32, "f64": OAL_F64, "f16s": OAL_F16S, "bf16s": OAL_BF16S}

# every entry point declared in include/oal.h
EXPORTS = [
    "oal_abi_version", "oal_strerror", "oal_last_error", "oal_plan_create", "oal_plan_destroy",
    "oal_level_cells", "oal_level_info", "oal_upload_f", "oal_download_f", "oal_init_equilibrium", "oal_level_step",
    "oal_step", "oal_force", "oal_mass", "oal_download_macro", "oal_guard", "oal_sync", "oal_time_steps",
    "oal_get_tuning", "oal_set_tuning", "oal_autotune", "oal_copy_bandwidth", "oal_set_bouzidi",
    "oal_write_raw_state", "oal_read_raw_state", "oal_mass_noalloc", "oal_set_sponge",
    "oal_set_surface", "oal_surface_forces", "oal_download_surface",
]


class OalCollision(C.Structure):
    _fields_ = [
        ("omega_bulk", C.c_float), ("omega_3", C.c_float), ("omega_4", C.c_float), ("lambda_", C.c_float),
        ("adaptive_omega4", C.c_int32), ("limiter", C.c_int32), ("compressible", C.c_int32),
        ("c_wale", C.c_float), ("nu_sgs_background", C.c_float), ("inlet_turbulence", C.c_float),
        ("transverse_freestream", C.c_int32), ("transverse_freestream_blend", C.c_float),
    ]


class OalDomainDesc(C.Structure):
    _fields_ = [
        ("extent", C.c_int32 * 3), ("periodic", C.c_int32 * 3), ("tau", C.c_float), ("u_inlet", C.c_float),
        ("body_force", C.c_float * 3), ("amr_sgs_scale", C.c_float), ("ymin_wall_bc", C.c_int32),
        ("zwall_bc", C.c_int32), ("force_scale", C.c_float), ("reserved", C.c_int32),
    ]


class OalLevelDesc(C.Structure):
    _fields_ = [
        ("n_blocks", C.c_int32),
        ("map_x", C.POINTER(C.c_int32)), ("map_y", C.POINTER(C.c_int32)), ("map_z", C.POINTER(C.c_int32)),
        ("neighbor_table", C.POINTER(C.c_int32)), ("block_domain", C.POINTER(C.c_int32)),
        ("n_domains", C.c_int32), ("domains", C.POINTER(OalDomainDesc)), ("obstacle", C.POINTER(C.c_uint8)),
    ]


class OalPlanDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("storage", C.c_int32), ("block_size", C.c_int32), ("stream_mode", C.c_int32),
        ("bb_mode", C.c_int32), ("n_levels", C.c_int32), ("levels", C.POINTER(OalLevelDesc)),
        ("collision", OalCollision), ("device", C.c_int32), ("rounding", C.c_int32),
    ]


class OalLevelInfo(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64), ("n_fluid_cells", C.c_int64), ("n_blocks", C.c_int32), ("n_regular", C.c_int32),
        ("n_interior", C.c_int32), ("n_boundary_regular", C.c_int32), ("n_boundary", C.c_int32),
        ("n_face_cells", C.c_int32), ("n_ghost_blocks", C.c_int32), ("n_covered_blocks", C.c_int32),
        ("n_partial", C.c_int32), ("n_sponge", C.c_int32), ("n_solid_regular", C.c_int32),
    ]


class OalTuning(C.Structure):
    _fields_ = [("pipe_planes", C.c_int32), ("overlap", C.c_int32), ("face_kernel", C.c_int32),
                ("graph", C.c_int32), ("pair_kernel", C.c_int32), ("reserved", C.c_int32 * 3)]


class OalSurfaceParams(C.Structure):
    _fields_ = [
        ("dx", C.c_double), ("mesh_offset", C.c_double * 3), ("pressure_scale", C.c_double),
        ("stress_scale", C.c_double), ("tau", C.c_double), ("c_wale", C.c_double), ("nu_sgs_bg", C.c_double),
        ("cs2", C.c_double), ("moment_center", C.c_double * 3), ("rho_ref", C.c_double), ("u_ref", C.c_double),
        ("area_ref", C.c_double), ("chord_ref", C.c_double), ("search_radius", C.c_int32),
        ("recover_normal", C.c_int32), ("symmetric", C.c_int32), ("reserved", C.c_int32),
    ]


class OalSurfaceForces(C.Structure):
    _fields_ = [
        ("F_pressure", C.c_double * 3), ("F_viscous", C.c_double * 3), ("F", C.c_double * 3), ("M", C.c_double * 3),
        ("Cd", C.c_double), ("Cl", C.c_double), ("Cs", C.c_double), ("Cm", C.c_double * 3), ("n_mapped", C.c_int64),
    ]


OAL_ALL_LEVELS = -1


OAL_TUNE_SHAPES = 20
TUNE_KEYS = [((1, 2, 4, 8, 16)[i % 5], int((i // 5) % 2 == 0), int(i >= 10)) for i in range(OAL_TUNE_SHAPES)]


class OalError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {msg} (status {code})")
        self.code = code


_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load liboal.so once (OAL_LIBRARY overrides its path); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else Path(os.environ.get("OAL_LIBRARY", LIB_PATH))
    if not p.exists():
        raise RuntimeError(f"liboal.so not built ({p}); run `python -m openaerolab_alpha_01_amd.build`")
    lib = C.CDLL(str(p))
    vp, i32, i64p, dp = C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double)
    sig = {
        "oal_abi_version": ([], C.c_int),
        "oal_strerror": ([C.c_int], C.c_char_p),
        "oal_last_error": ([], C.c_char_p),
        "oal_plan_create": ([C.POINTER(OalPlanDesc), C.POINTER(vp)], C.c_int),
        "oal_plan_destroy": ([vp], C.c_int),
        "oal_level_cells": ([vp, C.c_int, i64p], C.c_int),
        "oal_level_info": ([vp, C.c_int, C.POINTER(OalLevelInfo)], C.c_int),
        "oal_upload_f": ([vp, C.c_int, vp, C.c_int], C.c_int),
        "oal_download_f": ([vp, C.c_int, vp, C.c_int], C.c_int),
        "oal_init_equilibrium": ([vp, C.c_int, dp, dp, C.c_double, dp], C.c_int),
        "oal_level_step": ([vp, C.c_int, i32, C.c_int, vp], C.c_int),
        "oal_step": ([vp, C.c_int, i32, C.c_int, vp], C.c_int),
        "oal_force": ([vp, C.c_int, dp], C.c_int),
        "oal_mass": ([vp, C.c_int, dp, vp], C.c_int),
        "oal_mass_noalloc": ([vp, C.c_int, vp, vp], C.c_int),
        "oal_set_sponge": ([vp, C.c_int, vp, i32], C.c_int),
        "oal_download_macro": ([vp, C.c_int, dp, dp], C.c_int),
        "oal_guard": ([vp, C.c_int, C.c_float, C.c_float, i64p], C.c_int),
        "oal_sync": ([vp], C.c_int),
        "oal_time_steps": ([vp, C.c_int, C.c_int, C.c_int, dp, dp], C.c_int),
        "oal_get_tuning": ([vp, C.POINTER(OalTuning)], C.c_int),
        "oal_set_tuning": ([vp, C.POINTER(OalTuning)], C.c_int),
        "oal_autotune": ([vp, C.c_int, C.c_int, C.POINTER(OalTuning), dp], C.c_int),
        "oal_copy_bandwidth": ([C.c_int, C.c_int64, C.c_int, dp], C.c_int),
        "oal_set_bouzidi": ([vp, C.c_int, vp, i32, C.c_float], C.c_int),
        "oal_write_raw_state": ([vp, C.c_int, i32, C.c_double, C.c_char_p], C.c_int),
        "oal_read_raw_state": ([vp, C.c_int, C.c_char_p, C.POINTER(C.c_int32), dp], C.c_int),
        "oal_set_surface": ([vp, C.c_int, C.c_int64, vp, vp, vp], C.c_int),
        "oal_surface_forces": ([vp, C.c_int, C.POINTER(OalSurfaceParams), C.POINTER(OalSurfaceForces), vp], C.c_int),
        "oal_download_surface": ([vp, C.c_int, vp, vp, vp, vp], C.c_int),
    }
    for name, (argt, rest) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = rest
    if lib.oal_abi_version() != OAL_ABI_VERSION:
        raise RuntimeError("liboal.so ABI version mismatch")
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != OAL_OK:
        lib = load()
        raise OalError(code, where, lib.oal_last_error().decode(errors="replace"))
