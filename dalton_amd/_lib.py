"""ctypes binding of libccgpu.so (the C ABI declared in include/ccgpu.h).

There is no fallback: if the library is missing or fails to load, every
entry point raises."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libccgpu.so")

_dp = C.POINTER(C.c_double)

# (name, restype, argtypes) for every symbol of include/ccgpu.h
SIGNATURES = {
    "ccgpu_version": (C.c_int, []),
    "ccgpu_last_error": (C.c_char_p, []),
    "ccgpu_init": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_void_p)]),
    "ccgpu_finalize": (C.c_int, [C.c_void_p]),
    "ccgpu_df_factor_ao": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_long, _dp, C.POINTER(C.c_long)]),
    "ccgpu_df_factor_ao_cols": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_void_p, C.c_void_p, C.c_double,
                                          C.c_long, _dp, C.POINTER(C.c_long)]),
    "ccgpu_init_ao": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int,
                                C.POINTER(C.c_void_p), C.POINTER(C.c_long)]),
    "ccgpu_rhs": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int]),
    "ccgpu_lhtr": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "ccgpu_eta": (C.c_int, [C.c_void_p, _dp, _dp]),
    "ccgpu_energy": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "ccgpu_diis": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "ccgpu_ccsd_resume": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                    C.POINTER(C.c_int), _dp]),
    "ccgpu_rsp_filename": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_char_p]),
    "ccgpu_rsp_write": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_long, C.c_long, C.c_long, _dp, _dp, _dp]),
    "ccgpu_rsp_read": (C.c_int, [C.c_char_p, C.c_long, C.c_long, C.c_long, _dp, _dp, _dp, C.c_char_p,
                                 C.POINTER(C.c_int)]),
    "ccgpu_save_amplitudes": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "ccgpu_write_e_intermediates": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, _dp, _dp]),
    "ccgpu_write_gamma_intermediate": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, _dp]),
    "ccgpu_restart_amplitudes": (C.c_int, [C.c_void_p, C.c_char_p, _dp]),
    "ccgpu_lambda_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                     C.POINTER(C.c_int), _dp]),
    "ccgpu_onedens": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "ccgpu_lambda_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ccgpu_ccsd_guess": (C.c_int, [C.c_void_p, _dp]),
    "ccgpu_ccsd_iterate": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "ccgpu_ccsd_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                    C.POINTER(C.c_int), _dp]),
    "ccgpu_batch_init": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(_dp), C.POINTER(_dp), C.c_int, C.c_double,
                                   C.POINTER(C.c_void_p)]),
    "ccgpu_batch_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                     C.POINTER(C.c_int)]),
    "ccgpu_batch_get_amplitudes": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "ccgpu_batch_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ccgpu_batch_finalize": (C.c_int, [C.c_void_p]),
    "ccgpu_get_amplitudes": (C.c_int, [C.c_void_p, _dp, _dp]),
    "ccgpu_set_amplitudes": (C.c_int, [C.c_void_p, _dp, _dp]),
    "ccgpu_pt": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "ccgpu_pt_partial": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong, C.c_longlong, C.c_longlong, _dp, _dp, _dp,
                                   C.POINTER(C.c_longlong)]),
    "ccgpu_profile_start": (C.c_int, []),
    "ccgpu_profile_stop": (C.c_int, [C.c_char_p, C.c_int]),
    "ccgpu_mfma_ceiling": (C.c_int, [C.c_int, _dp]),
    "ccgpu_dgemm": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                              _dp, C.c_int, C.c_double, _dp, C.c_int]),
}
SIGNATURES["ccgpu_dgemm_mixed"] = SIGNATURES["ccgpu_dgemm"]
# device arrays and a stream handle (raw addresses)
SIGNATURES["ccgpu_dgemm_mixed_noalloc"] = (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                                                     C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double,
                                                     C.c_void_p, C.c_int, C.c_void_p])

_lib = None


class CCGPUError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ccgpu error {code}: {msg}")
        self.code = code


def lib():
    """Load libccgpu.so once; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built (python -m dalton_amd.build)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise CCGPUError(rc, lib().ccgpu_last_error().decode())


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)
