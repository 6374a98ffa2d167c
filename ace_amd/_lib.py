"""ctypes binding of libace_hip.so (the C-ABI declared in include/ace_hip.h).

There is no CPU fallback: if the library is missing or a GPU call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libace_hip.so")

_lib = None

P = C.c_void_p
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
lp = C.POINTER(C.c_longlong)

# name -> (restype, argtypes); must list every entry point of include/ace_hip.h
SIGNATURES = {
    "ace_create": (C.c_int, [C.POINTER(P), P]),
    "ace_destroy": (None, [P]),
    "ace_last_error": (C.c_char_p, [P]),
    "ace_input_width": (C.c_int, [P, C.c_int]),
    "ace_param_count": (C.c_int, [P, C.c_int]),
    "ace_set_params": (C.c_int, [P, C.c_int, dp, C.c_int]),
    "ace_get_params": (C.c_int, [P, C.c_int, dp, C.c_int]),
    "ace_set_input_scaling": (C.c_int, [P, C.c_int, dp, dp, C.c_int]),
    "ace_upload": (C.c_int, [P, C.c_int, ip, dp, ip, dp]),
    "ace_update_positions": (C.c_int, [P, dp]),
    "ace_build_neighbors": (C.c_int, [P, C.c_double]),
    "ace_neighbor_count": (C.c_int, [P, lp]),
    "ace_export_neighbors": (C.c_int, [P, lp, ip, ip]),
    "ace_set_precision": (C.c_int, [P, C.c_int]),
    "ace_symm": (C.c_int, [P, dp, C.c_int]),
    "ace_energy_forces": (C.c_int, [P, C.c_int, dp, dp, dp]),
    "ace_compute": (C.c_int, [P, C.c_int, C.c_double, dp, dp, dp]),
    "ace_energy_forces_noalloc": (C.c_int, [P, C.c_int, P, P, P, P]),
    "ace_md_init": (C.c_int, [P, dp, dp]),
    "ace_md_run": (C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, dp]),
    "ace_md_state": (C.c_int, [P, dp, dp]),
    "ace_autotune": (C.c_int, [P, C.c_int, dp]),
    "ace_set_path": (C.c_int, [P, C.c_int, C.c_int]),
    "ace_train_optimizer": (C.c_int, [P, C.c_int, C.c_double, C.c_double]),
    "ace_train_opt_state": (C.c_int, [P, C.c_int, dp, dp, dp, dp]),
    "ace_train_graph": (C.c_int, [P, C.c_int]),
    "ace_plain_energy_forces": (C.c_int, [P, dp, dp]),
    "ace_set_targets": (C.c_int, [P, dp, dp]),
    "ace_train_step": (C.c_int, [P, P, dp, dp]),
    "ace_forward": (C.c_int, [P, dp, dp, dp]),
    "ace_tile_shape": (C.c_int, [P, C.c_int, dp]),
    "ace_set_level_split": (C.c_int, [P, C.c_int]),
    "ace_set_force_gather": (C.c_int, [P, C.c_int]),
    "ace_set_tile_shape": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "ace_backward": (C.c_int, [P, dp, dp, dp, dp]),
    "ace_profile_enable": (C.c_int, [P, C.c_int]),
    "ace_profile_read": (C.c_int, [P, C.c_char_p, dp, lp]),
    "ace_sync": (C.c_int, [P]),
}


def load(path: str = LIB_PATH):
    """Load libace_hip.so and declare every entry point; raises if it is absent."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"ace_amd: {path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
