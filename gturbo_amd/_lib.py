"""ctypes binding of libgturbo_hip.so (the C ABI declared in include/gturbo.h).

The library is loaded from this package directory only.  There is no
fallback: if it is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgturbo_hip.so")

# Every symbol include/gturbo.h declares, with (restype, argtypes).
_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
SIGNATURES = {
    "gt_create": (C.c_int, [_I32, _P, C.POINTER(_P)]),
    "gt_destroy": (C.c_int, [_P]),
    "gt_last_error": (C.c_char_p, []),
    "gt_info": (C.c_int, [_P, _P, _P]),
    "gt_get_defk": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gt_get_forcing_modes": (C.c_int, [_P, _P, _P, _P]),
    "gt_plan_create": (C.c_int, [_P, _I32, C.POINTER(_I32)]),
    "gt_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "gt_enter_data": (C.c_int, [_P, _P, _I64, _I32]),
    "gt_exit_data": (C.c_int, [_P, _P, _I32]),
    "gt_update_host": (C.c_int, [_P, _P]),
    "gt_update_device": (C.c_int, [_P, _P]),
    "gt_map_data": (C.c_int, [_P, _P, _P, _I64]),
    "gt_unmap_data": (C.c_int, [_P, _P]),
    "gt_sync_ref": (C.c_int, [_P, _P]),
    "gt_copy": (C.c_int, [_P, _P, _P]),
    "gt_synchronize": (C.c_int, [_P]),
    "gt_fft_dir": (C.c_int, [_P, _P, _I32]),
    "gt_fft_inv": (C.c_int, [_P, _P, _I32]),
    "gt_produ": (C.c_int, [_P, _P, _P, _P]),
    "gt_trunc": (C.c_int, [_P, _P]),
    "gt_nlt": (C.c_int, [_P, _P, _P, _P]),
    "gt_step": (C.c_int, [_P, _P, _P, _P, _P]),
    "gt_step4": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "gt_forcing": (C.c_int, [_P, _P, C.POINTER(_I32)]),
    "gt_spectrum": (C.c_int, [_P, _P, _P, _P]),
    "gt_fluxes": (C.c_int, [_P, _P, _P, _P, _P]),
    "gt_glob": (C.c_int, [_P, _P, _P]),
    "gt_glob_noalloc": (C.c_int, [_P, _P, _P]),
    "gt_shell_index": (C.c_int, [_P, _P]),
    "gt_timing_enable": (C.c_int, [_P, _I32]),
    "gt_timing_classes": (C.c_int, [_P, _I32]),
    "gt_timing_read": (C.c_int, [_P, _P, _P]),
    "gt_adj_forward": (C.c_int, [_P, _P, _P, _I32, _I32, C.POINTER(_I32)]),
    "gt_adj_backward": (C.c_int, [_P, _P, _P]),
    "gt_ens_create": (C.c_int, [_I32, _P, _P, C.POINTER(_P)]),
    "gt_ens_destroy": (C.c_int, [_P]),
    "gt_ens_set_field": (C.c_int, [_P, _I32, _P, _I32]),
    "gt_ens_get_field": (C.c_int, [_P, _I32, _P, _I32]),
    "gt_ens_set_seed": (C.c_int, [_P, _I32, _I32]),
    "gt_ens_get_seed": (C.c_int, [_P, _I32, C.POINTER(_I32)]),
    "gt_ens_schedule": (C.c_int, [_P, _P, _I32]),
    "gt_ens_tune": (C.c_int, [_P, _I32, _P, C.POINTER(C.c_double)]),
    "gt_ens_run": (C.c_int, [_P, _I32]),
    "gt_ens_advance": (C.c_int, [_P, _I32, _I32]),
    "gt_ens_synchronize": (C.c_int, [_P]),
}

_lib = None
# Whether torch was imported before the library was loaded.  torch ships its
# own HIP runtime; the library binds to it only if it is already loaded, so
# device tensors can be shared with torch only in that order (otherwise the
# process holds two runtimes and torch finds no GPU).
TORCH_FIRST = False


class GturboError(RuntimeError):
    pass


def load():
    """Load the library (once).  Raises if it has not been built."""
    global _lib, TORCH_FIRST
    if _lib is None:
        TORCH_FIRST = "torch" in sys.modules
        if not os.path.exists(LIB_PATH):
            raise GturboError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
                "python -m gturbo_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().gt_last_error()
        raise GturboError(f"gturbo error {rc}: {msg.decode() if msg else ''}")
    return rc
