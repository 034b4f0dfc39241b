"""ctypes binding of libtinysph_hip.so (the C ABI in include/tinysph.h).

The library is loaded from this package directory only; there is no fallback:
if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtinysph_hip.so")

# constants mirrored from include/tinysph.h
OK, EINVAL, EHIP, ESTATE, ECAPACITY, EGRID = 0, -1, -2, -3, -4, -5
FLUID, WALL, INLET, OUTLET = 0, 1, 6, 11
PREC_FP64, PREC_MIXED = 0, 1
FORCES_WCSPH, FORCES_DBC, FORCES_BT = 0, 1, 2
WALL_DBC, WALL_MDBC, WALL_BT = 0, 1, 2
CASE_STILL_WATER_DBC, CASE_DAM_BREAK_MDBC, CASE_DAM_BREAK_BT, CASE_ALE_DAM_BREAK = 1, 2, 3, 4
CASE_INFLOW_SLAB, CASE_INFLOW_CHANNEL, CASE_DAM_BREAK_MDBC_REF = 5, 6, 7
GRID_SCRATCH_POINTS = 1 << 19  # TSPH_GRID_SCRATCH_POINTS
INLET1, OUTLET1 = 7, 12
SCHEME_SYMPLECTIC, SCHEME_RALE = 0, 1

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("h", "kap", "cs", "rho0", "avisc", "delta", "dp",
                                           "x_0", "x_m", "y_0", "y_m")] + [
        ("precision", C.c_int), ("capacity", C.c_int)]


class Particles(C.Structure):
    _fields_ = [("np", C.c_int), ("type", _i32p), ("index", _i32p),
                ("p", _dp), ("rho", _dp), ("drho", _dp), ("r", _dp), ("v", _dp), ("a", _dp),
                ("r_o", _dp), ("v_o", _dp), ("rho_o", _dp), ("gnr", _dp), ("n", _dp)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("pairs", "candidates", "active", "pair_passes",
                                           "density_out_of_range", "outside_grid", "removed",
                                           "appended", "zone_rejected", "zone_chained",
                                           "reinit_out_of_range", "scan_stalls")]


class History(C.Structure):
    _fields_ = [("np", C.c_int), ("v_oo", C.POINTER(C.c_double)), ("rho_oo", C.POINTER(C.c_double)),
                ("rho_temp", C.POINTER(C.c_double))]


class Zone(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x0", "xm", "y0", "ym", "vx", "vy")] + [
        ("ox", C.c_int), ("oy", C.c_int), ("btype", C.c_int)]


class CaseC(C.Structure):
    _fields_ = [("prm", Params), ("dt", C.c_double), ("wall", C.c_int), ("steps", C.c_int),
                ("scheme", C.c_int), ("n_fluid", C.c_int), ("n_wall", C.c_int),
                ("particles", Particles), ("n_zones", C.c_int), ("zones", Zone * 4)]


ALE_NAMES = ("omega", "omegarho", "omegav", "omega_o", "omegarho_o", "omegav_o", "domega",
             "domegarho", "domegav", "gradrho", "gradvx", "gradvy")
ALE_VEC = ("omegav", "omegav_o", "domegav", "gradrho", "gradvx", "gradvy")
MUSCL_NONE, MUSCL_LINEAR, MUSCL_MINMOD = 0, 1, 2


class AleFields(C.Structure):
    _fields_ = [("np", C.c_int)] + [(n, _dp) for n in ALE_NAMES]


EXPORTS = {
    # name: (restype, argtypes)
    "tsph_last_error": (C.c_char_p, []),
    "tsph_abi_version": (C.c_int, []),
    "tsph_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "tsph_destroy": (C.c_int, [C.c_void_p]),
    "tsph_grid": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tsph_mass": (C.c_int, [C.c_void_p, _dp]),
    "tsph_upload": (C.c_int, [C.c_void_p, C.POINTER(Particles)]),
    "tsph_download": (C.c_int, [C.c_void_p, C.POINTER(Particles)]),
    "tsph_get_np": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tsph_clear_cells": (C.c_int, [C.c_void_p]),
    "tsph_particle_to_cell": (C.c_int, [C.c_void_p]),
    "tsph_get_cell_lists": (C.c_int, [C.c_void_p, _i32p, _i32p]),
    "tsph_density_to_pressure": (C.c_int, [C.c_void_p]),
    "tsph_compute_forces": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_mdbc_density": (C.c_int, [C.c_void_p]),
    "tsph_bt_density": (C.c_int, [C.c_void_p]),
    "tsph_symplectic_predictor": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_symplectic_corrector": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_remove_out_of_domain": (C.c_int, [C.c_void_p]),
    "tsph_step_symplectic": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int]),
    "tsph_water_elevation": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp]),
    "tsph_interpolate_point": (C.c_int, [C.c_void_p, C.c_double, C.c_double, _dp]),
    "tsph_get_counters": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "tsph_reset_counters": (C.c_int, [C.c_void_p]),
    "tsph_set_tile_rows": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_get_tile_rows": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tsph_set_waves_per_cu": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_set_scan_wait": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_synchronize": (C.c_int, [C.c_void_p]),
    "tsph_pair_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_pair_timing_read": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int)]),
    "tsph_ale_enable": (C.c_int, [C.c_void_p]),
    "tsph_ale_upload": (C.c_int, [C.c_void_p, C.POINTER(AleFields)]),
    "tsph_ale_download": (C.c_int, [C.c_void_p, C.POINTER(AleFields)]),
    "tsph_rale_forces": (C.c_int, [C.c_void_p]),
    "tsph_rale_leapfrog_one": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_rale_leapfrog_two": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "tsph_step_rale": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int]),
    "tsph_ale_compute_density": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_ale_kernel_gradient": (C.c_int, [C.c_void_p]),
    "tsph_ale_density_pressure": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_ale_compute_acceleration": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_ale_leapfrog_one": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_ale_leapfrog_two": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "tsph_step_ale": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]),
    "tsph_bz_add_zone": (C.c_int, [C.c_void_p, C.POINTER(Zone), C.POINTER(C.c_int)]),
    "tsph_bz_integrate_inlet": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "tsph_bz_integrate_outlet": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "tsph_renumber": (C.c_int, [C.c_void_p]),
    "tsph_set_compensated": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_step_inflow": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]),
    "tsph_case_build": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_uint64,
                                  C.POINTER(C.POINTER(CaseC))]),
    "tsph_case_free": (C.c_int, [C.POINTER(CaseC)]),
    "tsph_ensemble_build": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.c_int,
                                      C.POINTER(C.POINTER(CaseC))]),
    "tsph_batch_create": (C.c_int, [C.POINTER(Params), _dp, C.c_int, C.POINTER(C.c_void_p)]),
    "tsph_batch_cases": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tsph_batch_grid": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  _dp]),
    "tsph_batch_upload": (C.c_int, [C.c_void_p, C.POINTER(Particles)]),
    "tsph_batch_get_np": (C.c_int, [C.c_void_p, _i32p]),
    "tsph_batch_download": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Particles)]),
    "tsph_batch_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "tsph_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), _dp, _dp]),
    "tsph_set_balanced": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_set_cell_order": (C.c_int, [C.c_void_p, C.c_int]),
    "tsph_write_vtk": (C.c_int, [C.c_void_p, C.c_char_p]),
    "tsph_io_wait": (C.c_int, [C.c_void_p]),
    "tsph_interpolate_grid": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp, _dp]),
    "tsph_interpolate_grid_noalloc": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                                C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                _dp, _dp, _dp]),
    "tsph_write_interpolated_vtk": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_char_p]),
    "tsph_format_vtk": (C.c_int, [C.POINTER(Particles), C.c_char_p]),
    "tsph_write_grid_vtk": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_char_p]),
    "tsph_format_grid_vtk": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.POINTER(C.c_int32), C.c_char_p]),
    "tsph_write_hydrostatic_pressure": (C.c_int, [C.c_void_p, C.c_char_p]),
    "tsph_format_hydrostatic_pressure": (C.c_int, [C.POINTER(Particles), C.c_char_p]),
    "tsph_format_interpolated_vtk": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double,
                                               C.c_double, _dp, _dp, _dp, C.c_char_p]),
    "tsph_save_state": (C.c_int, [C.c_void_p, C.c_char_p]),
    "tsph_load_state": (C.c_int, [C.c_void_p, C.c_char_p]),
    "tsph_history_enable": (C.c_int, [C.c_void_p]),
    "tsph_history_upload": (C.c_int, [C.c_void_p, C.POINTER(History)]),
    "tsph_history_download": (C.c_int, [C.c_void_p, C.POINTER(History)]),
    "tsph_verlet": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_euler": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_leapfrog_one": (C.c_int, [C.c_void_p, C.c_double]),
    "tsph_leapfrog_two": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "tsph_density_reinit_sum": (C.c_int, [C.c_void_p]),
    "tsph_density_reinit_overwrite": (C.c_int, [C.c_void_p]),
    "tsph_step_verlet": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int]),
    "tsph_step_leapfrog": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int]),
    "tsph_step_symplectic_reinit": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int,
                                              C.c_int]),
}

_LIB = None


class TsphError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load the in-tree library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise TsphError(f"{LIB_PATH} is missing: run `python -m tinysph_old_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int, what: str) -> None:
    if rc != OK:
        msg = lib().tsph_last_error().decode(errors="replace")
        raise TsphError(f"{what} failed ({rc}): {msg}")


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_i32p)
