"""ctypes binding of libboxzy_hip.so (the C-ABI declared in include/boxzy.h).

The library is built in-tree (``python -m boxzyhydro_amd.build`` or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
# BOXZY_LIB selects another build of the same library (e.g. the host-sanitized
# one of `make -C boxzyhydro_amd/csrc host-asan`); a missing file still raises
LIB_PATH = os.environ.get("BOXZY_LIB") or os.path.join(HERE, "libboxzy_hip.so")

BX_OK = 0
BX_ERR_ARG = -1
BX_ERR_HIP = -2
BX_ERR_STATE = -3
BX_ERR_UNSUPPORTED = -4
BX_PREC_FP64 = 0
BX_PREC_MIXED = 1
BX_STEP_EXACT = 0
BX_STEP_FUSED = 1
BX_STEP_GRAPH = 2
BX_GRAV_POTENTIAL = 1
BX_GRAV_EXTERNAL = 2
BX_GRAV_COM_SEQUENTIAL = 4
BX_STATIC_MAX_CELLS = 1048576
BX_STATIC_BYTES_PER_CELL = 88


class BxParams(ctypes.Structure):
    """Mirror of ``bx_params`` (include/boxzy.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("gammafix", ctypes.c_double),
        ("small_rho", ctypes.c_double), ("small_eps", ctypes.c_double),
        ("vlimit", ctypes.c_double), ("cfl", ctypes.c_double),
        ("xabun", ctypes.c_double), ("yabun", ctypes.c_double),
        ("zabun", ctypes.c_double), ("mu_z", ctypes.c_double),
        ("avmag", ctypes.c_double * 3),
        ("no_outflow", ctypes.c_int32 * 6),
        ("precision", ctypes.c_int32),
        ("fluxangmom", ctypes.c_int32),
        ("yoffset", ctypes.c_double),
    ]


class BxEosParams(ctypes.Structure):
    """Mirror of ``bx_eos_params`` (include/boxzy.h)."""

    _fields_ = [
        ("H2STAT", ctypes.c_int32), ("NEOS_RHO", ctypes.c_int32),
        ("dTk_eos", ctypes.c_double), ("tk_bgrnd", ctypes.c_double),
        ("brot", ctypes.c_double), ("vib", ctypes.c_double),
        ("ac", ctypes.c_double), ("bc", ctypes.c_double),
        ("diss", ctypes.c_double), ("ionize", ctypes.c_double),
        ("rho_eos_low", ctypes.c_double), ("rho_eos_high", ctypes.c_double),
        ("tk_eos_cutoff", ctypes.c_double),
    ]


class BoxzyError(RuntimeError):
    pass


_lib = None

# name -> argtypes; every entry point returns int
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.c_int32
_SIGS = {
    "bx_create": [ctypes.POINTER(BxParams), _I32, ctypes.POINTER(_VP)],
    "bx_destroy": [_VP],
    "bx_ngrid": [_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "bx_ghost_table": [ctypes.POINTER(BxParams), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                       ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)],
    "bx_upload": [_VP, _DP, _DP, _DP, _DP],
    "bx_download": [_VP, _DP, _DP, _DP, _DP],
    "bx_download_cons_old": [_VP, _DP],
    "bx_state": [_VP],
    "bx_velocity": [_VP, _DP],
    "bx_courant": [_VP, _DP],
    "bx_source": [_VP, ctypes.c_double, _I32, ctypes.c_double],
    "bx_set_ghost_cells": [_VP],
    "bx_flux": [_VP, _I32, ctypes.c_double],
    "bx_copy_old": [_VP],
    "bx_avisc": [_VP, _DP],
    "bx_step": [_VP, _I32, _I32, ctypes.c_double, _DP, _DP, ctypes.POINTER(_I32), _DP,
                ctypes.POINTER(_I32)],
    "bx_step_static": [_VP, _VP, _VP, ctypes.c_double, _VP, _VP, _VP],
    "bx_fallback_count": [_VP, ctypes.POINTER(ctypes.c_int64)],
    "bx_graph_step_count": [_VP, ctypes.POINTER(ctypes.c_int64)],
    "bx_step_counts": [_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "bx_tune": [_VP, _I32, _I32, _I32],
    "bx_profile": [_VP, _I32, _DP, ctypes.POINTER(ctypes.c_int64)],
    "bx_sync": [_VP],
    "bx_batch_check": [_I32, ctypes.POINTER(BxParams), _DP, _DP, ctypes.POINTER(_I32)],
    "bx_batch_create": [_I32, ctypes.POINTER(BxParams), _DP, _DP, _DP, _DP, _I32, ctypes.POINTER(_VP)],
    "bx_batch_destroy": [_VP],
    "bx_batch_cells": [_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "bx_batch_order": [_VP, ctypes.POINTER(_I32)],
    "bx_batch_tune": [_VP, _I32, _I32],
    "bx_batch_gang": [_VP, _I32, _I32],
    "bx_batch_gang_fallbacks": [_VP, ctypes.POINTER(_I32)],
    "bx_batch_log_dt": [_VP, _I32],
    "bx_batch_run": [_VP, _I32, _I32],
    "bx_batch_sync": [_VP, _DP],
    "bx_batch_result": [_VP, _DP, _DP, _DP, _DP, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_uint32)],
    "bx_batch_dts": [_VP, _DP],
    "bx_boxes_create": [_I32, ctypes.POINTER(BxParams), _DP, _DP, _DP, _DP, _I32, ctypes.POINTER(_VP)],
    "bx_boxes_destroy": [_VP],
    "bx_boxes_cells": [_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "bx_boxes_order": [_VP, ctypes.POINTER(_I32)],
    "bx_boxes_tune": [_VP, _I32, _I32],
    "bx_boxes_log_dt": [_VP, _I32],
    "bx_boxes_run": [_VP, _I32, _I32],
    "bx_boxes_sync": [_VP, _DP],
    "bx_boxes_result": [_VP, _DP, _DP, _DP, _DP, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_uint32)],
    "bx_boxes_dts": [_VP, _DP],
    "bx_adjoint_forward": [_VP, _I32, _I32, _DP, _DP],
    "bx_adjoint_backward": [_VP, _VP, _VP],
    "bx_adjoint_tune": [_VP, _I32, _I32],
    "bx_adjoint_store": [_VP, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)],
    "bx_cons_device": [_VP, _VP, _I32],
    "bx_write_celldump": [_VP, ctypes.c_char_p, ctypes.c_double, _I32, _I32],
    "bx_read_celldump": [_VP, ctypes.c_char_p, _DP, ctypes.POINTER(_I32)],
    "bx_celldump_encode": [ctypes.POINTER(BxParams), ctypes.c_char_p, ctypes.c_double, _I32, _I32, _DP, _DP, _DP,
                           _DP],
    "bx_celldump_decode": [ctypes.POINTER(BxParams), ctypes.c_char_p, _DP, ctypes.POINTER(_I32), _DP, _DP, _DP,
                           _DP],
    "bx_eos_size": [ctypes.POINTER(BxEosParams), ctypes.POINTER(_I32)],
    "bx_eos_tables": [ctypes.POINTER(BxParams), ctypes.POINTER(BxEosParams)] + [_DP] * 11,
    "bx_eos_init": [_VP, ctypes.POINTER(BxEosParams)],
    "bx_download_muc": [_VP, _DP],
    "bx_gravity_config": [_VP, _I32, ctypes.c_double, ctypes.c_double],
    "bx_set_phi": [_VP, _DP],
    "bx_calc_gforce": [_VP],
    "bx_external_grav": [_VP, _DP],
    "bx_download_gforce": [_VP, _DP],
}

EXPORTED = tuple(_SIGS) + ("bx_last_error", "bx_eos_defaults")


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime libraries; when both are used in one
        # process they work if torch's are loaded first, so load them now
        if "torch" not in sys.modules and importlib.util.find_spec("torch") is not None:
            import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise BoxzyError(f"{LIB_PATH} is missing: build it with `python -m boxzyhydro_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.bx_last_error.argtypes = []
        L.bx_last_error.restype = ctypes.c_char_p
        L.bx_eos_defaults.argtypes = [ctypes.POINTER(BxEosParams)]
        L.bx_eos_defaults.restype = None
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != BX_OK:
        msg = lib().bx_last_error().decode(errors="replace")
        raise BoxzyError(f"{what} failed ({rc}): {msg}")


def dptr(a):
    """double* of a contiguous float64 numpy array, or NULL for None."""
    if a is None:
        return ctypes.cast(None, _DP)
    return a.ctypes.data_as(_DP)
