"""ctypes binding of libtsphere.so (include/tsphere.h).

The library is loaded from the package directory (built in-tree by
``turing-sphere_amd/build.py``).  There is no fallback: if it is missing or
does not load, every entry point raises, so nothing can silently run on the
CPU instead of the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libtsphere.so")

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_COMPILE = 3
ERR_UNSUPPORTED = 4
ERR_NODEVICE = 5

FOURIER_AUTO, FOURIER_FFT, FOURIER_DFT = 0, 1, 2
LEG_FP32, LEG_BF16X3, LEG_F64REC = 0, 1, 2


class TsphereError(RuntimeError):
    """A non-zero status from the C ABI, with the library's message."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class ShtConfigC(C.Structure):
    _fields_ = [("lmax", C.c_int), ("mmax", C.c_int), ("nlat", C.c_int), ("nphi", C.c_int),
                ("fourier", C.c_int), ("legendre", C.c_int), ("max_fields", C.c_int)]


class PlanInfoC(C.Structure):
    _fields_ = [("lmax", C.c_int), ("mmax", C.c_int), ("nlat", C.c_int), ("nphi", C.c_int),
                ("nlm", C.c_longlong), ("fourier", C.c_int), ("legendre", C.c_int),
                ("leg_lat", C.c_int), ("max_fields", C.c_int), ("scratch_bytes", C.c_size_t)]


class LegTileC(C.Structure):
    _fields_ = [("nf", C.c_int), ("nwaves", C.c_int), ("k", C.c_int), ("ctc", C.c_int),
                ("impl", C.c_int), ("rows", C.c_int)]


_vp = C.c_void_p
_ll = C.c_longlong
_i = C.c_int
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)

# name -> argtypes (every function returns int status unless listed in _RESTYPE)
SIGNATURES = {
    "tsphere_version": [],
    "tsphere_last_error": [],
    "tsphere_device_count": [C.POINTER(C.c_int)],
    "tsphere_stream_create": [C.POINTER(_vp)],
    "tsphere_stream_destroy": [_vp],
    "tsphere_stream_sync": [_vp],
    "tsphere_device_sync": [],
    "tsphere_malloc": [C.c_size_t, C.POINTER(_vp)],
    "tsphere_free": [_vp],
    "tsphere_memcpy_htod": [_vp, _vp, C.c_size_t, _vp],
    "tsphere_memcpy_dtoh": [_vp, _vp, C.c_size_t, _vp],
    "tsphere_memcpy_dtod": [_vp, _vp, C.c_size_t, _vp],
    "tsphere_fill_f32": [_vp, C.c_float, _ll, _vp],
    "tsphere_event_create": [C.POINTER(_vp)],
    "tsphere_event_destroy": [_vp],
    "tsphere_event_record": [_vp, _vp],
    "tsphere_event_elapsed": [_vp, _vp, C.POINTER(C.c_float)],
    "tsphere_stream_wait": [_vp, _vp],
    "tsphere_plan_create": [C.POINTER(ShtConfigC), C.POINTER(_vp)],
    "tsphere_plan_destroy": [_vp],
    "tsphere_plan_query": [_vp, C.POINTER(PlanInfoC)],
    "tsphere_plan_grid": [_vp, _vp, _vp, _vp],
    "tsphere_plan_tables": [C.POINTER(ShtConfigC), _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tsphere_synth": [_vp, _vp, _vp, _vp],
    "tsphere_analys": [_vp, _vp, _vp, _vp],
    "tsphere_synth_batch": [_vp, _i, _vp, _ll, _vp, _ll, _vp],
    "tsphere_synth_batch_noalloc": [_vp, _i, _vp, _ll, _vp, _ll, _vp],
    "tsphere_analys_batch": [_vp, _i, _vp, _ll, _vp, _ll, _vp],
    "tsphere_synth_multi": [_vp, _i, C.POINTER(_vp), C.POINTER(_vp), _vp],
    "tsphere_analys_multi": [_vp, _i, C.POINTER(_vp), C.POINTER(_vp), _vp],
    "tsphere_synth_fields": [_vp, _i, _i, C.POINTER(_vp), _ll, C.POINTER(_vp), _ll, _vp],
    "tsphere_analys_fields": [_vp, _i, _i, C.POINTER(_vp), _ll, C.POINTER(_vp), _ll, _vp],
    "tsphere_synth_adjoint": [_vp, _i, _i, C.POINTER(_vp), _ll, C.POINTER(_vp), _ll, _vp],
    "tsphere_analys_adjoint": [_vp, _i, _i, C.POINTER(_vp), _ll, C.POINTER(_vp), _ll, _vp],
    "tsphere_accumulate": [_vp, _vp, _ll, _vp],
    "tsphere_plan_set_tile": [_vp, _i, C.POINTER(LegTileC)],
    "tsphere_plan_get_tile": [_vp, _i, _i, C.POINTER(LegTileC)],
    "tsphere_plan_tune": [_vp, _i, _i, _vp, C.POINTER(LegTileC), _fp],
    "tsphere_leg_synth": [_vp, _i, _vp, _ll, _vp, _vp],
    "tsphere_fourier_synth": [_vp, _i, _vp, _vp, _ll, _vp],
    "tsphere_fourier_analys": [_vp, _i, _vp, _ll, _vp, _vp],
    "tsphere_leg_analys": [_vp, _i, _vp, _vp, _ll, _vp],
    "tsphere_leg_synth_multi": [_vp, _i, C.POINTER(_vp), _vp, _vp],
    "tsphere_leg_analys_multi": [_vp, _i, _vp, C.POINTER(_vp), _vp],
    "tsphere_leg_imex_synth": [_vp, _i, C.POINTER(_vp), C.POINTER(_vp), _vp,
                               C.POINTER(C.c_float), C.POINTER(C.c_float), _vp, _vp],
    "tsphere_leg_spectral_imex": [_vp, _i, _vp, C.POINTER(_vp), _vp, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_int), _vp],
    "tsphere_leg_synth_fields": [_vp, _i, _i, C.POINTER(_vp), _ll, _vp, _vp],
    "tsphere_leg_analys_fields": [_vp, _i, _i, _vp, C.POINTER(_vp), _ll, _vp],
    "tsphere_leg_imex_synth_fields": [_vp, _i, _i, C.POINTER(_vp), C.POINTER(_vp), _ll, _vp, _vp,
                                      _vp, _vp],
    "tsphere_leg_analys_imex_fields": [_vp, _i, _i, _vp, C.POINTER(_vp), C.POINTER(_vp), _ll, _vp,
                                       _vp, _vp],
    "tsphere_plan_fourier_table":[_vp, C.POINTER(_vp)],
    "tsphere_module_compile": [C.c_char_p, C.c_char_p, C.POINTER(_vp)],
    "tsphere_module_destroy": [_vp],
    "tsphere_module_kernel": [_vp, C.c_char_p, C.POINTER(_vp)],
    "tsphere_kernel_destroy": [_vp],
    "tsphere_rtc_check": [C.c_char_p, C.c_char_p],
    "tsphere_kernel_launch": [_vp, _ll, _i, C.POINTER(_vp), _i, _fp, _vp],
    "tsphere_kernel_launch_grid": [_vp, C.c_uint, C.c_uint, C.c_uint, _i, C.POINTER(_vp), _i,
                                   _fp, _vp],
    "tsphere_graph_begin": [_vp],
    "tsphere_graph_end": [_vp, C.POINTER(_vp)],
    "tsphere_graph_launch": [_vp, _vp],
    "tsphere_graph_destroy": [_vp],
}
_RESTYPE = {"tsphere_last_error": C.c_char_p}

_lib = None


def load():
    """Load libtsphere.so; raise (never fall back) if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libtsphere.so not found at {LIB_PATH}: build it with "
            f"`python turing-sphere_amd/build.py` (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = L
    return L


def last_error() -> str:
    return load().tsphere_last_error().decode("utf-8", "replace")


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = last_error()
        raise TsphereError(rc, f"{what}: {msg}" if what else msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
