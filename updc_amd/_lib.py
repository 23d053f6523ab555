"""ctypes binding of libupdc_hip.so (include/updc_hip.h).

The library is the only compute path: there is no CPU fallback.  If the built
library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libupdc_hip.so"
EXE_PATH = LIB_PATH.parent / "updc_hip_core"
PROJECTOR_PATH = LIB_PATH.parent / "updc_hip_psf_projector"

UPDC_OK = 0
ERR_NAMES = {1: "ARG", 2: "HIP", 3: "IO", 4: "NUMERIC", 5: "STATE", 6: "NOMEM", 7: "FORMAT"}
PREC = {"f32": 0, "f64": 1, "mixed": 2}
FLAG_GRAPH = 1
FLAG_QUIET = 2
FLAG_PER_EXPOSURE = 4
FLAG_AUTOTUNE = 8
FLAG_FUSED_ADJOINT = 16
FLAG_COLUMN_SPECTRA = 32


class UpdcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"UPDC_{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class FreezeRow(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("exposure", C.c_int32), ("n_new_frozen", C.c_uint64),
                ("n_total_frozen", C.c_uint64), ("min_raw_mu", C.c_double)]


class SetupTimes(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("fft_autotune", "fft_plan", "psf_spectra", "linkmaps", "weight_scale",
                                         "csc_transpose", "initial_drizzle", "sensitivity")]


class IterTimes(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("eager", C.c_int32)] + [
        (n, C.c_double) for n in ("estimate_fft", "fused_exposure_wall", "forward_ifft_critical",
                                  "fused_detector_critical", "scatter_fft_critical", "adjoint_reduce_ifft", "update",
                                  "total")]


class IterInfo(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("pad", C.c_int32), ("signed_sum", C.c_double),
                ("n_negative", C.c_uint64), ("n_new_frozen", C.c_uint64), ("n_total_frozen", C.c_uint64),
                ("min_raw_mu", C.c_double), ("ms_total", C.c_double)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
SIGNATURES = {
    "updc_all": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                           C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "updc_last_error": (C.c_char_p, []),
    "updc_self_test": (C.c_int, [C.c_int]),
    "updc_polygon_area": (C.c_float, [_D, C.c_int]),
    "updc_overlap_area": (C.c_float, [_D, C.c_int, C.c_int]),
    "updc_ctx_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_int]),
    "updc_ctx_destroy": (C.c_int, [_P]),
    "updc_ctx_set_flags": (C.c_int, [_P, C.c_uint]),
    "updc_ctx_set_preprocess_workers": (C.c_int, [_P, C.c_int]),
    "updc_fft_shape_candidates": (C.c_int, [C.c_int, C.c_int, _I, _I, C.c_int, _I]),
    "updc_ctx_set_fft_cache": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "updc_ctx_set_column_kernel": (C.c_int, [_P, C.c_int]),
    "updc_ctx_set_setup_cache": (C.c_int, [_P, C.c_char_p]),
    "updc_get_setup_cache_hits": (C.c_int, [_P, _I]),
    "updc_get_spectrum_layout": (C.c_int, [_P, _I]),
    "updc_get_autotune": (C.c_int, [_P, _I, _I, _I, _D, C.c_int, _I, _I]),
    "updc_psf_project": (C.c_int, [C.c_int, _F, C.c_int, C.c_int, _D, _F, C.c_size_t, _I, _I, _D, _D,
                                   C.POINTER(C.c_uint64)]),
    "updc_op_layout": (C.c_int, [_P, _I, _I, _I, _I, _I]),
    "updc_op_r2c": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "updc_op_c2r": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_double, _P]),
    "updc_op_c2r_noalloc": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_double, _P]),
    "updc_op_gather": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "updc_op_scatter": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "updc_add_exposure": (C.c_int, [_P, _F, _F, C.c_int, C.c_int, _P, C.c_size_t, _F, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.c_double]),
    "updc_add_exposure_vertices": (C.c_int, [_P, _F, _F, C.c_int, C.c_int, _D, _F, C.c_int, C.c_int, C.c_int,
                                             C.c_double, C.c_double]),
    "updc_linkmap_size": (C.c_int, [_P, C.c_int, C.POINTER(C.c_size_t)]),
    "updc_get_linkmap": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "updc_prepare": (C.c_int, [_P]),
    "updc_iterate": (C.c_int, [_P, C.c_int, C.POINTER(FreezeRow), C.POINTER(IterInfo)]),
    "updc_get_iterate_progress": (C.c_int, [_P, _I, _I]),
    "updc_get_setup_times": (C.c_int, [_P, C.POINTER(SetupTimes)]),
    "updc_get_iteration_times": (C.c_int, [_P, C.POINTER(IterTimes), C.c_int, _I]),
    "updc_get_estimate": (C.c_int, [_P, _F]),
    "updc_get_estimate_f64": (C.c_int, [_P, _D]),
    "updc_set_estimate_f64": (C.c_int, [_P, _D]),
    "updc_get_wht": (C.c_int, [_P, _F]),
    "updc_get_denominator_f64": (C.c_int, [_P, _D]),
    "updc_get_initial_f64": (C.c_int, [_P, _D, _D]),
    "updc_get_psf_spectrum_f64": (C.c_int, [_P, C.c_int, _D]),
    "updc_get_scalars": (C.c_int, [_P, _D, _D, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "updc_get_frozen": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint8)]),
    "updc_get_exposure_info": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "updc_get_csc_layout": (C.c_int, [_P, C.c_int, _I, C.POINTER(C.c_uint64)]),
    "updc_stage_forward": (C.c_int, [_P, _D, _D]),
    "updc_stage_model": (C.c_int, [_P, _D, _D]),
    "updc_stage_detector": (C.c_int, [_P, C.c_int, _D, _D, _D, _D]),
    "updc_stage_adjoint": (C.c_int, [_P, _D, _D]),
    "updc_stage_scatter": (C.c_int, [_P, C.c_int, _D, _D]),
    "updc_stage_adjoint_images": (C.c_int, [_P, _D, _D]),
    "updc_reset_frozen": (C.c_int, [_P]),
    "updc_fft_r2c_f64": (C.c_int, [_P, _D, _D]),
    "updc_fft_c2r_f64": (C.c_int, [_P, _D, _D]),
    "updc_get_stage_times": (C.c_int, [_P, C.c_void_p, _D, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "updc_copy_calibration": (C.c_int, [C.c_size_t, C.c_int, C.c_int, _D]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m updc_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != UPDC_OK:
        raise UpdcError(rc, load().updc_last_error().decode(errors="replace"))


def exported_symbols() -> list[str]:
    return list(SIGNATURES)


if os.environ.get("UPDC_EAGER_LOAD"):
    load()
