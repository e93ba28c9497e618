"""ctypes binding of libipc_warp_amd.so (the C ABI in include/ipc_warp_amd.h).

The library must have been built in-tree (``python -m ipc_warp_amd.build``);
if it is missing this module raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes
import os

from .build import LIB_PATH

_lib = None

_c_int, _c_float, _c_double, _vp = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p

_SIGNATURES = {
    "ipc_abi_version": ([], _c_int),
    "ipc_last_error": ([], ctypes.c_char_p),
    "ipc_contact_hessian_ee": ([_vp, _vp, _c_int, _vp, _c_int, _c_float, _c_double, _c_int, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ipc_contact_hessian_pt": ([_vp, _c_int, _vp, _c_int, _c_float, _c_double, _c_int, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ipc_contact_energy_ee": ([_vp, _vp, _c_int, _vp, _c_int, _c_float, _c_double, _vp, _vp], _c_int),
    "ipc_contact_energy_pt": ([_vp, _c_int, _vp, _c_int, _c_float, _c_double, _vp, _vp], _c_int),
    "ipc_collision_hessian": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _c_int, _c_float, _c_double, _vp, _vp, _vp, _vp, _vp],
                              _c_int),
    "ipc_collision_energy": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _c_int, _c_float, _c_double, _vp, _vp], _c_int),
    "ipc_contact_hessian_ee_f64": ([_vp, _vp, _c_int, _vp, _c_int, _c_double, _c_double, _c_int, _vp, _vp, _vp, _vp,
                                    _vp], _c_int),
    "ipc_contact_hessian_pt_f64": ([_vp, _c_int, _vp, _c_int, _c_double, _c_double, _c_int, _vp, _vp, _vp, _vp, _vp],
                                   _c_int),
    "ipc_contact_energy_ee_f64": ([_vp, _vp, _c_int, _vp, _c_int, _c_double, _c_double, _vp, _vp], _c_int),
    "ipc_contact_energy_pt_f64": ([_vp, _c_int, _vp, _c_int, _c_double, _c_double, _vp, _vp], _c_int),
    "ipc_collision_time_ee_f64": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_collision_time_pt_f64": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_contact_points_ee": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ipc_contact_points_pt": ([_vp, _c_int, _vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ipc_collision_time_ee": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_collision_time_pt": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_conservative_time_ee": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_conservative_time_pt": ([_vp, _vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_ctx_create": ([ctypes.POINTER(_vp), _c_int, _vp, _c_int, _vp, _c_int, _vp], _c_int),
    "ipc_ctx_destroy": ([_vp], _c_int),
    "ipc_detect_collision_ee": ([_vp, _vp, _c_float, _c_float, _c_int, _vp, _vp, _c_int, _vp, _vp], _c_int),
    "ipc_detect_collision_pt": ([_vp, _vp, _c_float, _c_float, _c_int, _vp, _vp, _c_int, _vp, _vp], _c_int),
    "ipc_ccd_candidates_ee": ([_vp, _vp, _vp, _c_float, _c_int, _vp, _vp, _c_int, _vp, _vp], _c_int),
    "ipc_ccd_candidates_pt": ([_vp, _vp, _vp, _c_float, _c_int, _vp, _vp, _c_int, _vp, _vp], _c_int),
    "ipc_collision_free_step": ([_vp, _vp, _vp, _c_float, _c_int, ctypes.POINTER(_c_double), _vp], _c_int),
    "ipc_ctx_last_ccd_counts": ([_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)], _c_int),
    "ipc_ctx_set_timing": ([_vp, _c_int], _c_int),
    "ipc_ctx_kernel_times": ([_vp, ctypes.POINTER(_c_double), ctypes.POINTER(ctypes.c_longlong)], _c_int),
    "ipc_mlp_forward": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ipc_mlp_forward_noalloc": ([_vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_mlp_backward": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ipc_mlp_mse": ([_vp, _vp, _vp, _c_int, _c_float, _vp, _vp, _vp], _c_int),
    "ipc_mlp_adam": ([_vp, _vp, _vp, _vp, ctypes.c_longlong, _c_float, _c_float, _c_float, _c_float, _c_int, _vp],
                     _c_int),
    "ipc_zero_f32": ([_vp, ctypes.c_longlong, _vp], _c_int),
    "ipc_ccd_intersections_ee": ([_vp, _vp, _vp, _c_float, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_ccd_intersections_pt": ([_vp, _vp, _vp, _c_float, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_ctx_init_fixed_bodies": ([_vp, _c_int, _vp], _c_int),
    "ipc_detect_collision_fixed": ([_vp, _vp, _vp, _c_float, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp],
                                   _c_int),
    "ipc_collision_free_step_fixed": ([_vp, _vp, _vp, _vp, _vp, _vp, _c_float, _vp, _vp, _vp], _c_int),
    "ipc_detect_collision": ([_vp, _vp, _c_float, _c_float, _c_int, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp],
                             _c_int),
    "ipc_count_intersections": ([_vp, _vp, _vp, _c_int, _vp, _vp], _c_int),
    "ipc_warmstart_features": ([_vp, _c_int, _vp, _vp, _c_int, _vp, _c_int, _c_float, _vp, _vp, _vp], _c_int),
    "ipc_contact_records_pack": ([_vp, _c_int, _c_int, _c_float, _vp, _vp, _vp], _c_int),
    "ipc_contact_records_unpack": ([_vp, _vp, _c_int, _c_float, _vp, _vp, _c_int, _vp], _c_int),
    "ipc_bsr_mv": ([_c_int, _vp, _vp, _vp, _vp, _vp, _c_float, _c_float, _vp], _c_int),
    "ipc_cg_create": ([ctypes.POINTER(_vp)], _c_int),
    "ipc_cg_destroy": ([_vp], _c_int),
    "ipc_bsr_cg": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_float, _c_int, _c_int, _c_int, ctypes.POINTER(_c_int),
                    ctypes.POINTER(_c_double), _vp], _c_int),
    "ipc_tet_rest": ([_vp, _c_int, _vp, _c_int, _vp, _vp, _vp], _c_int),
    "ipc_tet_hessian": ([_vp, _c_int, _vp, _c_int, _vp, _vp, _c_float, _c_float, _c_int, _vp, _vp, _vp, _vp, _vp],
                        _c_int),
    "ipc_tet_energy": ([_vp, _c_int, _vp, _c_int, _vp, _vp, _c_float, _c_float, _vp, _vp, _vp], _c_int),
    "ipc_abd_create": ([ctypes.POINTER(_vp)], _c_int),
    "ipc_abd_destroy": ([_vp], _c_int),
    "ipc_abd_reduce_contacts": ([_vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _c_double, _vp, _vp,
                                 _vp, _c_int, _vp, _vp], _c_int),
    "ipc_active_set_create": ([ctypes.POINTER(_vp)], _c_int),
    "ipc_active_set_destroy": ([_vp], _c_int),
    "ipc_al_vertex_toi": ([_vp, _c_int, _c_int, _vp, _c_int, _vp], _c_int),
    "ipc_al_create_hash": ([_vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp, _c_int, _vp], _c_int),
    "ipc_al_merge": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp, _c_int, _vp], _c_int),
    "ipc_al_linearize": ([_vp, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _c_double, _vp], _c_int),
    "ipc_al_update_slack": ([_vp, _c_int, _c_int, _vp, _c_int, _c_double, _vp], _c_int),
    "ipc_al_update_forces": ([_vp, _c_int, _c_int, _vp, _c_int, _c_double, _vp], _c_int),
    "ipc_bsr_create": ([ctypes.POINTER(_vp)], _c_int),
    "ipc_bsr_destroy": ([_vp], _c_int),
    "ipc_bsr_from_triplets": ([_vp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _c_int, _vp], _c_int),
    "ipc_bsr_refill": ([_vp, _vp, _c_int, _vp, _vp], _c_int),
    "ipc_bsr_axpy": ([_vp, _c_int, _vp, _vp, _vp, _c_int, _c_float, _vp, _vp, _vp, _c_int, _c_float, _vp, _vp, _vp,
                      _c_int, _vp], _c_int),
}

TIMER_KINDS = ("build_edge", "build_tri", "query_dcd_ee", "query_dcd_pt", "query_ccd_ee", "query_ccd_pt",
               "toi_ee", "toi_pt", "fixed_cull", "fixed_query_dcd", "fixed_query_ccd")

EXPORTED = tuple(_SIGNATURES)


class IPCError(RuntimeError):
    """A libipc_warp_amd entry point returned a failure status."""


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IPCError(
                f"{LIB_PATH} is not built; run `python -m ipc_warp_amd.build` (hipcc, gfx950). "
                "There is no fallback implementation."
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (argtypes, restype) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = L
    return _lib


IPC_EINVAL = -100000  # status of a refused argument (include/ipc_warp_amd.h)


def check(status: int, what: str) -> int:
    """Raise on a negative status; return a positive one (a required capacity)."""
    if status < 0:
        msg = lib().ipc_last_error().decode(errors="replace")
        raise IPCError(f"{what} failed with status {status}: {msg}")
    return status
