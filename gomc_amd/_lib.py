"""ctypes binding of libgomc_hip.so (the C ABI in include/gomc_hip.h).

The library is loaded from this package directory only.  If it is missing
the import fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

from .systems import SystemDesc


class BatchTuning(C.Structure):
    """gomc_batch_tuning (include/gomc_hip.h)."""
    _fields_ = [("team_min", C.c_int32 * 3), ("pair12_max_k", C.c_int32)]

LIB_PATH = Path(__file__).resolve().parent / "libgomc_hip.so"
HEADER = Path(__file__).resolve().parent.parent / "include" / "gomc_hip.h"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "gomc_box_create": (C.c_int, [C.POINTER(SystemDesc), C.c_int, C.POINTER(_vp)]),
    "gomc_box_destroy": (C.c_int, [_vp]),
    "gomc_last_error": (C.c_char_p, [_vp]),
    "gomc_box_info": (C.c_int, [_vp, _lp]),
    "gomc_box_dims": (C.c_int, [_vp, _dp]),
    "gomc_set_coords": (C.c_int, [_vp, _dp, _dp, _dp]),
    "gomc_get_coords": (C.c_int, [_vp, _dp, _dp, _dp]),
    "gomc_commit_mol": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "gomc_build_cells": (C.c_int, [_vp, _ip, _ip, _ip]),
    "gomc_trial_cells": (C.c_int, [_vp, _ip, _ip, _ip]),
    "gomc_box_inter": (C.c_int, [_vp, _dp, _dp]),
    "gomc_mol_inter": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip]),
    "gomc_particle_inter": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                      _dp, _dp, _ip]),
    "gomc_recip_init": (C.c_int, [_vp, _dp]),
    "gomc_recip_sums": (C.c_int, [_vp, _dp]),
    "gomc_recip_sums_noalloc": (C.c_int, [_vp, _dp]),
    "gomc_set_device_team": (C.c_int, [_vp, C.c_int]),
    "gomc_set_mol_box": (C.c_int, [_vp, C.c_int, C.c_int]),
    "gomc_mol_recip": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "gomc_swap_recip": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "gomc_recip_cache": (C.c_int, [_vp, C.c_int]),
    "gomc_recip_accept": (C.c_int, [_vp]),
    "gomc_recip_copy_ref_to_new": (C.c_int, [_vp]),
    "gomc_recip_get": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "gomc_mol_trial": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "gomc_system_total": (C.c_int, [_vp, _dp]),
    "gomc_box_force": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "gomc_box_force_recip": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "gomc_torque": (C.c_int, [_vp, _dp, _dp, _dp]),
    "gomc_box_virial": (C.c_int, [_vp, _dp]),
    "gomc_box_checkpoint": (C.c_int64, [_vp, C.c_int64, _vp, C.c_int64]),
    "gomc_box_restore": (C.c_int, [_vp, _vp, C.c_int64, _lp]),
    "gomc_mp_trial": (C.c_int, [_vp, C.c_int, C.c_uint64, C.c_double, _dp]),
    "gomc_mp_accept": (C.c_int, [_vp, C.c_int, _dp]),
    "gomc_mp_settings": (C.c_int, [_vp, _dp]),
    "gomc_run_moves": (C.c_int64, [_vp, C.c_int64, C.c_int64, _ip, _ip, _ip, _ip, _dp, _dp, _dp]),
    "gomc_run_moves_device": (C.c_int64, [_vp, C.c_int64, C.c_int64, _ip, _ip, _ip, _ip, _dp, _dp,
                                          _dp]),
    "gomc_run_moves_device_multi": (C.c_int, [C.POINTER(_vp), C.c_int, _lp, _lp, _ip, _lp, _lp, _dp,
                                              _lp, _ip]),
    "gomc_energies": (C.c_int, [_vp, _dp]),
    "gomc_mol_boxes": (C.c_int, [_vp, _ip]),
    "gomc_get_scales": (C.c_int, [_vp, _dp, _dp]),
    "gomc_set_precision": (C.c_int, [_vp, C.c_int]),
    "gomc_batch_create": (C.c_int, [C.POINTER(SystemDesc), C.c_int, C.c_int, C.c_int,
                                    C.POINTER(_vp)]),
    "gomc_batch_destroy": (C.c_int, [_vp]),
    "gomc_batch_last_error": (C.c_char_p, [_vp]),
    "gomc_batch_run": (C.c_int, [_vp, _lp]),
    "gomc_batch_stats": (C.c_int, [_vp, _dp, _lp, _lp]),
    "gomc_batch_exchange": (C.c_int, [_vp, _ip, C.c_int, C.c_int64, C.c_uint32, _ip]),
    "gomc_batch_get_coords": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "gomc_batch_box_layout": (C.c_int, [_vp, C.c_int, _ip]),
    "gomc_batch_box_times": (C.c_int, [_vp, _dp]),
    "gomc_batch_default_tuning": (C.c_int, [C.POINTER(BatchTuning)]),
    "gomc_batch_create_tuned": (C.c_int, [C.POINTER(SystemDesc), C.c_int, C.c_int, C.c_int,
                                          C.POINTER(BatchTuning), C.POINTER(_vp)]),
    "gomc_batch_set_trace": (C.c_int, [_vp, C.c_int, C.c_int64]),
    "gomc_batch_get_trace": (C.c_int, [_vp, _ip, _dp, _dp, _dp, _lp]),
    "gomc_server_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_double,
                                     C.POINTER(_vp)]),
    "gomc_server_destroy": (C.c_int, [_vp]),
    "gomc_server_last_error": (C.c_char_p, [_vp]),
    "gomc_server_submit": (C.c_int, [_vp, C.POINTER(SystemDesc), C.c_int64, _lp]),
    "gomc_server_poll": (C.c_int, [_vp, C.c_int64, _ip, _dp, _dp, _lp, _lp]),
    "gomc_server_poll_counts": (C.c_int, [_vp, C.c_int64, _dp]),
    "gomc_server_drain": (C.c_int, [_vp, C.c_double]),
    "gomc_server_stats": (C.c_int, [_vp, _dp]),
    "gomc_batch_run_cycles": (C.c_int, [_vp, C.c_int, C.c_int]),
    "gomc_batch_last_kernel_ms": (C.c_int, [_vp, _dp]),
    "gomc_place_molecules": (C.c_int, [C.c_int, _ip, _dp, C.c_int, _ip, C.c_double, C.c_uint64,
                                       C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]),
}


def header_symbols(path: Path = HEADER) -> list[str]:
    """Function names declared in include/gomc_hip.h."""
    text = re.sub(r"/\*.*?\*/", "", path.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(gomc_[a-z0-9_]+)\s*\(", text)))


_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load libgomc_hip.so and bind every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m gomc_amd.build` "
                          "(the HIP extension is required; there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    if a.dtype.name == "float64":
        return a.ctypes.data_as(_dp)
    if a.dtype.name == "int32":
        return a.ctypes.data_as(_ip)
    if a.dtype.name == "int64":
        return a.ctypes.data_as(_lp)
    raise TypeError(a.dtype)
