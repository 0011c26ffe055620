"""Loader of libflowmol_hip.so (built in-tree by flowmol_amd.build).

There is no fallback: if the library is missing or cannot be loaded, every call raises.
"""
import ctypes as C
import os

from ._abi import FmMacros, FmParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflowmol_hip.so")

# every symbol declared in include/flowmol_hip.h with its ctypes signature
_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
SIGNATURES = {
    "fm_init": (C.c_int, [C.POINTER(FmParams), _i32, _i32, C.POINTER(_vp)]),
    "fm_finalize": (None, [_vp]),
    "fm_last_error": (C.c_char_p, []),
    "fm_abi_version": (C.c_int, []),
    "fm_get_geometry": (C.c_int, [_vp, _ip, _dp, _dp, _dp, _dp]),
    "fm_get_capacity": (C.c_int, [_vp, _ip, _ip]),
    "fm_upload_state": (C.c_int, [_vp, _i32, _dp, _dp, _dp]),
    "fm_upload_moltype": (C.c_int, [_vp, _ip]),
    "fm_upload_tags": (C.c_int, [_vp, _ip, _dp]),
    "fm_upload_monomers": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "fm_download_state": (C.c_int, [_vp, _dp, _dp, _dp]),
    "fm_download_tags": (C.c_int, [_vp, _dp]),
    "fm_download_labels": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip, _ip, _ip]),
    "fm_download_potenergy_parts": (C.c_int, [_vp, _dp, _dp]),
    "fm_download_molecular": (C.c_int, [_vp, _dp, _dp]),
    "fm_download_rfmol": (C.c_int, [_vp, _dp]),
    "fm_get_np": (C.c_int, [_vp, _ip]),
    "fm_simtime": (C.c_double, [_vp]),
    "fm_set_simtime": (C.c_int, [_vp, C.c_double]),
    "fm_move_particles_vv": (C.c_int, [_vp, _i32]),
    "fm_move_particles_lfv": (C.c_int, [_vp]),
    "fm_checkrebuild": (C.c_int, [_vp, _i32, _ip]),
    "fm_linklist_deallocateall": (C.c_int, [_vp]),
    "fm_sendmols": (C.c_int, [_vp]),
    "fm_sort_mols": (C.c_int, [_vp]),
    "fm_assign_to_cell": (C.c_int, [_vp]),
    "fm_updateborders": (C.c_int, [_vp, _i32]),
    "fm_assign_to_neighbourlist": (C.c_int, [_vp]),
    "fm_compute_forces": (C.c_int, [_vp, _i32]),
    "fm_evaluate_macroscopic_properties": (C.c_int, [_vp, C.POINTER(FmMacros)]),
    "fm_get_mop": (C.c_int, [_vp, _dp, _i32]),
    "fm_reset_mop": (C.c_int, [_vp]),
    "fm_get_cv_bins": (C.c_int, [_vp, _ip, _ip, _dp]),
    "fm_get_cv_stress": (C.c_int, [_vp, _dp, _i64]),
    "fm_reset_cv_stress": (C.c_int, [_vp]),
    "fm_macros_sums_noalloc": (C.c_int, [_i32, _vp, _vp, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fm_derive_macros": (C.c_int, [_dp, _i32, C.c_double, _i32, C.POINTER(FmMacros)]),
    "fm_run_vv": (C.c_int, [_vp, _i32, _i32, C.POINTER(FmMacros), _i32, _ip]),
    "fm_run_lfv": (C.c_int, [_vp, _i32, _i32, C.POINTER(FmMacros), _i32, _ip]),
    "fm_download_cells": (C.c_int, [_vp, _ip]),
    "fm_download_neighbourlist": (C.c_int, [_vp, _ip, _ip, _ip, _i64, _lp]),
    "fm_download_bonds": (C.c_int, [_vp, _ip, _ip]),
    "fm_batch_create": (C.c_int, [C.POINTER(_vp)]),
    "fm_batch_destroy": (None, [_vp]),
    "fm_batch_use_graphs": (C.c_int, [_vp, _i32]),
    "fm_batch_add": (C.c_int, [_vp, _vp, _i32]),
    "fm_batch_remove": (C.c_int, [_vp, _vp]),
    "fm_batch_run": (C.c_int, [_vp, _i32, _ip]),
    "fm_batch_counters": (C.c_int, [_vp, _lp, _lp]),
    "fm_batch_size": (C.c_int, [_vp, _ip]),
    "fm_batch_get_iter": (C.c_int, [_vp, _vp, _ip]),
    "fm_memory_footprint": (C.c_int, [C.POINTER(FmParams), _i32, _lp]),
    "fm_synchronize": (C.c_int, [_vp]),
    "fm_kernel_timing": (C.c_int, [_vp, _i32]),
    "fm_get_kernel_timing": (C.c_int, [_vp, C.c_char_p, _i32, _dp, _lp, _i32, _ip]),
}

_lib = None


class FlowmolHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load():
    """Load the HIP library; raises if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m flowmol_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise FlowmolHipError(rc, load().fm_last_error().decode(errors="replace"))
