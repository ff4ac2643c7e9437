"""ctypes binding of libsarf.so (the C ABI declared in include/sarf.h).

There is no CPU fallback: importing this module fails loudly if the HIP
library has not been built, and every call that returns a non-zero status
raises :class:`SarfError` with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libsarf.so")

# status codes (sarf.h)
SARF_OK = 0
SARF_ERR_INVALID = -1
SARF_ERR_UNSUPPORTED = -2
SARF_ERR_HIP = -3
SARF_ERR_NOMEM = -4
SARF_ERR_NODEVICE = -5

SARF_BP_MIXED = 0
SARF_BP_F64 = 1
SARF_BP_ACCUMULATE = 1
SARF_BATCH_DIR_PER_BASIS = 2
SARF_SCENES_TILE_MIXED = 100
SARF_RC_TAPER = 1
SARF_RC_RAMP = 2
SARF_RC_F64 = 4


class SarfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libsarf error {code}: {msg}")
        self.code = code


class RcDesc(C.Structure):
    _fields_ = [("K", C.c_int), ("Np", C.c_int), ("Nfft", C.c_int), ("flags", C.c_int)]


class BpDesc(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nfft", C.c_int), ("Np", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("Wx", C.c_double), ("Wy", C.c_double),
        ("maxWr", C.c_double),
        ("precision", C.c_int), ("flags", C.c_int), ("tile", C.c_int), ("pulse_split", C.c_int),
        ("los_x", C.c_double), ("los_y", C.c_double),
    ]


class FocusDesc(C.Structure):
    _fields_ = [
        ("K", C.c_int), ("Np", C.c_int), ("Nfft", C.c_int), ("Nx", C.c_int), ("Ny", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("Wx", C.c_double), ("Wy", C.c_double),
        ("maxWr", C.c_double), ("precision", C.c_int),
    ]


_vp = C.c_void_p
_i = C.c_int
_d = C.c_double

# (name, restype, argtypes) for every symbol sarf.h declares
SIGNATURES = {
    "sarf_version": (_i, []),
    "sarf_last_error": (C.c_char_p, []),
    "sarf_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "sarf_ctx_destroy": (_i, [_vp]),
    "sarf_ctx_trim": (_i, [_vp]),
    "sarf_range_compress": (_i, [_vp, C.POINTER(RcDesc), _vp, _vp, _vp, _vp, _vp, _vp]),
    "sarf_range_compress_scenes": (_i, [_vp, _i, C.c_void_p, _vp]),
    "sarf_backproject": (_i, [_vp, C.POINTER(BpDesc), _vp, _vp, _vp, _vp, _vp, _vp]),
    "sarf_focus": (_i, [_vp, C.POINTER(FocusDesc), _vp, _vp, _vp, _vp, _vp]),
    "sarf_focus_mf": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sarf_synth_phase_history": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp]),
    "sarf_image_entropy": (_i, [_vp, _vp, C.c_long, _i, _vp, _vp]),
    "sarf_bp_batch_entropy": (_i, [_vp, C.POINTER(BpDesc), _vp, _vp, _vp, _vp, _i, _i, _vp, _vp,
                                   _vp, _vp, _vp]),
    "sarf_bp_entropy_grad": (_i, [_vp, C.POINTER(BpDesc), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sarf_pga": (_i, [_vp, C.c_void_p, _vp, _vp, _vp, _vp]),
    "sarf_bp_scenes": (_i, [_vp, _i, C.c_void_p, _i, _vp]),
    "sarf_synth_phase_history_nufft": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp,
                                            _vp]),
    "sarf_autotune": (_i, [_vp, C.POINTER(BpDesc), _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "sarf_tune_table_get": (_i, [_vp, C.c_void_p, _i, C.POINTER(_i)]),
    "sarf_tune_table_set": (_i, [_vp, C.c_void_p, _i]),
    "sarf_matched_filter": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sarf_image_db8": (_i, [_vp, _vp, _i, _i, C.c_float, _vp, _vp]),
    "sarf_db8_levels": (_i, [C.c_float, _vp]),
    "sarf_autotune_scenes": (_i, [_vp, _i, C.c_void_p, _i, _vp, _vp, _vp]),
    "sarf_pga_scenes": (_i, [_vp, _i, C.c_void_p, _i, _d, _d, _vp]),
    "sarf_grid_search": (_i, [_vp, C.POINTER(BpDesc), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                              _vp, _vp, _vp, _vp]),
    "sarf_backproject_noalloc": (_i, [_vp, C.POINTER(BpDesc), _vp, _vp, _vp, _vp, _vp, _vp]),
    "sarf_grid_covariance": (_i, [_vp, _vp, C.c_long, _vp]),
    "sarf_fit_linear": (_i, [_vp, _i, _i, _vp]),
    "sarf_fit_quadratic": (_i, [_vp, _i, _i, _vp]),
}


class MfDesc(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("K", C.c_int), ("Np", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("Wx", C.c_double), ("Wy", C.c_double),
        ("flags", C.c_int), ("pulse_split", C.c_int),
    ]


class TuneResult(C.Structure):
    _fields_ = [("tile", C.c_int), ("pulse_split", C.c_int), ("ms", C.c_float),
                ("default_ms", C.c_float), ("candidates", C.c_int)]


class TuneEntry(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Np", C.c_int), ("Nfft", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("maxWr", C.c_double),
                ("los_x", C.c_double), ("los_y", C.c_double),
                ("tile", C.c_int), ("pulse_split", C.c_int)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("prof", _vp), ("ant", _vp), ("r0", _vp), ("f0", _vp), ("img", _vp),
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nfft", C.c_int), ("Np", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("Wx", C.c_double), ("Wy", C.c_double),
        ("maxWr", C.c_double), ("los_x", C.c_double), ("los_y", C.c_double),
    ]


class PgaDesc(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("iters", C.c_int), ("db_down", C.c_double),
                ("rms_stop", C.c_double)]


class RcScene(C.Structure):
    _fields_ = [("ph", _vp), ("prof", _vp), ("taper", _vp), ("f0", _vp), ("r0", _vp),
                ("K", C.c_int), ("Np", C.c_int), ("Nfft", C.c_int), ("flags", C.c_int)]


SARF_GRID_MAX_DIMS = 10
SARF_GRID_MODEL_COEF = 0
SARF_GRID_MODEL_REF_D7 = 7
SARF_GRID_MODEL_REF_D10 = 10
SARF_GRID_SHRINK_REFERENCE = 0
SARF_GRID_SHRINK_CENTRED = 1


class GridDesc(C.Structure):
    _fields_ = [("dims", C.c_int), ("model", C.c_int), ("levels", C.c_int), ("shrink", C.c_int),
                ("start", C.c_double * SARF_GRID_MAX_DIMS), ("end", C.c_double * SARF_GRID_MAX_DIMS),
                ("num", C.c_int * SARF_GRID_MAX_DIMS)]


class GridLevel(C.Structure):
    _fields_ = [("start", C.c_double * SARF_GRID_MAX_DIMS), ("end", C.c_double * SARF_GRID_MAX_DIMS),
                ("best", C.c_double * SARF_GRID_MAX_DIMS), ("entropy", C.c_double),
                ("gpu_ms", C.c_float), ("ncand", C.c_int)]


class PgaScene(C.Structure):
    _fields_ = [("img", _vp), ("af_ph", _vp), ("rms", _vp), ("Nx", C.c_int), ("Ny", C.c_int)]


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"libsarf.so not found at {path}; build it with __graft_entry__.build() "
            "(there is no CPU fallback)"
        )
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def check(code: int) -> None:
    if code != SARF_OK:
        msg = lib().sarf_last_error()
        raise SarfError(code, msg.decode() if msg else "")
