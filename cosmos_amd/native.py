"""ctypes binding of the C-ABI in include/cosmos_fv.h.

The HIP library is the only implementation: if libcosmos_fv.so is missing
or fails to load, every call raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("COSMOS_FV_LIB") or os.path.join(HERE, "libcosmos_fv.so")

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_u8_p = ctypes.POINTER(ctypes.c_ubyte)
_handle = ctypes.c_void_p



class BatchProblem(ctypes.Structure):
    """cosmos_batch_problem"""
    _fields_ = [("physics", ctypes.c_int), ("dim", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("cell_spacing", ctypes.c_double), ("cfl", ctypes.c_double), ("final_time", ctypes.c_double),
                ("max_steps", ctypes.c_int), ("adiabatic_index", ctypes.c_double)]


class Status(ctypes.Structure):
    """cosmos_status_t"""
    _fields_ = [(n, ctypes.c_longlong) for n in ("c2p_cap", "c2p_q_floor", "c2p_density_floor",
                                                 "c2p_pressure_floor", "nan_cells", "nan_time_step",
                                                 "excised_grid", "reserved")]


class BatchSchedule(ctypes.Structure):
    """cosmos_batch_schedule"""
    _fields_ = [("whole_max_cells", ctypes.c_int), ("x_segments_per_task", ctypes.c_int), ("y_rows", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("y_tile_cols", ctypes.c_int)]


# name -> (restype, argtypes); mirrors include/cosmos_fv.h.
SIGNATURES = {
    "cosmos_create": (ctypes.c_int, [ctypes.POINTER(_handle), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double]),
    "cosmos_destroy": (ctypes.c_int, [_handle]),
    "cosmos_last_error": (ctypes.c_char_p, [_handle]),
    "cosmos_set_mode": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_set_con2prim": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_int]),
    "cosmos_set_launch": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int]),
    "cosmos_set_y_tile": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_set_step_fusion": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int]),
    "cosmos_set_spacetime": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double]),
    "cosmos_upload": (ctypes.c_int, [_handle, _c_double_p]),
    "cosmos_download": (ctypes.c_int, [_handle, _c_double_p]),
    "cosmos_upload_device": (ctypes.c_int, [_handle, ctypes.c_void_p]),
    "cosmos_download_device": (ctypes.c_int, [_handle, ctypes.c_void_p]),
    "cosmos_sweep": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int]),
    "cosmos_max_wave_speed": (ctypes.c_int, [_handle, _c_double_p]),
    "cosmos_stable_time_step": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_int, _c_double_p]),
    "cosmos_begin": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "cosmos_begin_at": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "cosmos_advance": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_progress": (ctypes.c_int, [_handle, _c_double_p, _c_int_p, _c_int_p]),
    "cosmos_solve": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_double_p, _c_int_p]),
    "cosmos_set_excision": (ctypes.c_int, [_handle, ctypes.c_double]),
    "cosmos_forcing": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "cosmos_solve_forced": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _c_double_p,
                                            _c_int_p]),
    "cosmos_solve_forced_at": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_double_p, _c_int_p]),
    "cosmos_checkpoint_save": (ctypes.c_int, [_handle, ctypes.c_char_p, ctypes.c_double, ctypes.c_int]),
    "cosmos_checkpoint_load": (ctypes.c_int, [_handle, ctypes.c_char_p, _c_double_p, _c_int_p]),
    "cosmos_write_dat": (ctypes.c_int, [_handle, ctypes.c_char_p]),
    "cosmos_batch_default_schedule": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(BatchSchedule)]),
    "cosmos_batch_create": (ctypes.c_int, [ctypes.POINTER(_handle), ctypes.POINTER(BatchProblem), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "cosmos_batch_destroy": (ctypes.c_int, [_handle]),
    "cosmos_batch_last_error": (ctypes.c_char_p, [_handle]),
    "cosmos_batch_set_con2prim": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_int]),
    "cosmos_batch_set_schedule": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.POINTER(BatchSchedule)]),
    "cosmos_batch_get_schedule": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.POINTER(BatchSchedule)]),
    "cosmos_batch_upload": (ctypes.c_int, [_handle, ctypes.c_int, _c_double_p]),
    "cosmos_batch_download": (ctypes.c_int, [_handle, ctypes.c_int, _c_double_p]),
    "cosmos_batch_run": (ctypes.c_int, [_handle]),
    "cosmos_batch_run_timed": (ctypes.c_int, [_handle, _c_double_p]),
    "cosmos_batch_stats": (ctypes.c_int, [_handle, _c_double_p, _c_int_p, _c_double_p]),
    "cosmos_batch_autotune": (ctypes.c_int, [_handle, ctypes.c_int, _c_double_p]),
    "cosmos_amr_create": (ctypes.c_int, [ctypes.POINTER(_handle), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double]),
    "cosmos_amr_destroy": (ctypes.c_int, [_handle]),
    "cosmos_amr_last_error": (ctypes.c_char_p, [_handle]),
    "cosmos_amr_set_con2prim": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_int]),
    "cosmos_amr_upload": (ctypes.c_int, [_handle, _c_double_p]),
    "cosmos_amr_download": (ctypes.c_int, [_handle, _c_double_p, _c_double_p, _c_u8_p, _c_u8_p]),
    "cosmos_amr_refine_flags": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_double]),
    "cosmos_amr_sweep": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "cosmos_amr_restrict": (ctypes.c_int, [_handle]),
    "cosmos_amr_set_spacetime": (ctypes.c_int, [_handle] + [ctypes.c_double] * 5),
    "cosmos_amr_set_excision": (ctypes.c_int, [_handle, ctypes.c_double]),
    "cosmos_amr_profile_enable": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_amr_profile_read": (ctypes.c_int, [_handle, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "cosmos_amr_level_solver": (ctypes.c_void_p, [_handle, ctypes.c_int]),
    "cosmos_amr_derefine": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_double]),
    "cosmos_amr_solve_level1": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_double_p,
                                                _c_int_p]),
    "cosmos_amr_set_flags": (ctypes.c_int, [_handle, _c_u8_p, _c_u8_p]),
    "cosmos_amr_solve_level2": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_double_p,
                                                _c_int_p]),
    "cosmos_amr_download_level2": (ctypes.c_int, [_handle, _c_double_p, _c_double_p, _c_double_p, _c_u8_p, _c_u8_p,
                                                   _c_u8_p]),
    "cosmos_adjoint_create": (ctypes.c_int, [ctypes.POINTER(_handle), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_double]),
    "cosmos_adjoint_destroy": (ctypes.c_int, [_handle]),
    "cosmos_adjoint_forcing": (ctypes.c_int, [_handle, ctypes.c_double, ctypes.c_double, _c_double_p, _c_double_p]),
    "cosmos_adjoint_set_subcycles": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_adjoint_set_tape": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_longlong]),
    "cosmos_adjoint_release": (ctypes.c_int, [_handle]),
    "cosmos_adjoint_last_error": (ctypes.c_char_p, [_handle]),
    "cosmos_adjoint_solver": (_handle, [_handle]),
    "cosmos_adjoint_sweep": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, _c_double_p, _c_double_p]),
    "cosmos_adjoint_forward": (ctypes.c_int, [_handle, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "cosmos_adjoint_backward": (ctypes.c_int, [_handle, _c_double_p, _c_double_p]),
    "cosmos_adjoint_replay": (ctypes.c_int, [_handle, _c_double_p, _c_double_p]),
    "cosmos_set_geometry_cache": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_status": (ctypes.c_int, [_handle, ctypes.c_void_p]),
    "cosmos_status_async": (ctypes.c_int, [_handle, ctypes.c_void_p]),
    "cosmos_status_reset": (ctypes.c_int, [_handle]),
    "cosmos_step_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "cosmos_noalloc_grid_limit": (ctypes.c_size_t, []),
    "cosmos_noalloc_last_error": (ctypes.c_char_p, []),
    "cosmos_adjoint_time_steps": (ctypes.c_int, [_handle, _c_double_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "cosmos_slope_limiter": (ctypes.c_int, [_c_double_p, ctypes.c_long, ctypes.c_double, ctypes.c_int, _c_double_p]),
    "cosmos_slope_vector": (ctypes.c_int, [_c_double_p, _c_double_p, _c_double_p, ctypes.c_long, ctypes.c_double,
                                            ctypes.c_int, _c_double_p]),
    "cosmos_host_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "cosmos_host_free": (ctypes.c_int, [ctypes.c_void_p]),
    "cosmos_point_last_error": (ctypes.c_char_p, []),
    "cosmos_vector_lax_friedrichs": (ctypes.c_int, [_c_double_p] * 4 + [ctypes.c_long, ctypes.c_double,
                                                                         ctypes.c_double, _c_double_p]),
    "cosmos_vector_richtmyer": (ctypes.c_int, [_c_double_p] * 4 + [ctypes.c_long, ctypes.c_double, ctypes.c_double,
                                                                    _c_double_p]),
    "cosmos_vector_force": (ctypes.c_int, [_c_double_p, _c_double_p, ctypes.c_long, _c_double_p]),
    "cosmos_vector_force_update": (ctypes.c_int, [_c_double_p] * 3 + [ctypes.c_long, ctypes.c_double,
                                                                       ctypes.c_double, _c_double_p]),
    "cosmos_vector_evolution": (ctypes.c_int, [ctypes.c_double, _c_double_p, _c_double_p, ctypes.c_long,
                                                ctypes.c_double, ctypes.c_double, _c_double_p]),
    "cosmos_vector_slope": (ctypes.c_int, [_c_double_p] * 3 + [ctypes.c_long, ctypes.c_double, ctypes.c_int,
                                                                _c_double_p]),
    "cosmos_point_slope_limiter": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_double_p]),
    "cosmos_point_conserved": (ctypes.c_int, [ctypes.c_int, _c_double_p, ctypes.c_double, _c_double_p, _c_double_p,
                                               _c_double_p]),
    "cosmos_point_primitive": (ctypes.c_int, [ctypes.c_int, _c_double_p, ctypes.c_double, _c_double_p, _c_double_p,
                                               _c_double_p]),
    "cosmos_point_flux": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_double_p, ctypes.c_double, _c_double_p,
                                          _c_double_p, _c_double_p]),
    "cosmos_point_interface_flux": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, _c_double_p,
                                                    _c_double_p, _c_double_p]),
    "cosmos_point_slic_flux": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_c_double_p] * 4 +
                               [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                _c_double_p, _c_double_p, _c_double_p]),
    "cosmos_point_half_step": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_c_double_p] * 3 +
                               [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, _c_double_p, _c_double_p, _c_double_p]),
    "cosmos_point_fractional_step": (ctypes.c_int, [ctypes.c_int, ctypes.c_double] + [_c_double_p] * 3 +
                                     [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double, _c_double_p, _c_double_p, _c_double_p]),
    "cosmos_point_slope_vector": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_c_double_p] * 3 +
                                  [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, _c_double_p,
                                   _c_double_p, _c_double_p]),
    "cosmos_point_forcing": (ctypes.c_int, [ctypes.c_int] + [_c_double_p] * 5 +
                             [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                              _c_double_p, _c_double_p, _c_double_p]),
    "cosmos_synchronize": (ctypes.c_int, [_handle]),
    "cosmos_stream": (ctypes.c_void_p, [_handle]),
    "cosmos_profile_enable": (ctypes.c_int, [_handle, ctypes.c_int]),
    "cosmos_profile_read": (ctypes.c_int, [_handle, _c_double_p, _c_int_p]),
    "cosmos_device_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]),
    "cosmos_state_planes": (ctypes.c_int, [_handle, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_long)]),
}

_lib = None


class CosmosError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CosmosError(f"{path} not found: build it with `python -m cosmos_amd.build` "
                          "(the HIP library is the only implementation of this path)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, handle=None, amr: bool = False, adjoint: bool = False, batch: bool = False) -> None:
    if rc != 0:
        lib = load()
        if batch:
            msg = lib.cosmos_batch_last_error(handle)
        elif adjoint:
            msg = lib.cosmos_adjoint_last_error(handle)
        else:
            msg = lib.cosmos_amr_last_error(handle) if amr else lib.cosmos_last_error(handle)
        raise CosmosError(f"cosmos error {rc}: {msg.decode() if msg else ''}")
