"""ctypes loader of libtaco_tcl4.so (the HIP engine).

The product has no CPU fallback: if the shared library is missing or cannot
be loaded, every entry point raises.  Build it with ``python -m taco_amd.build``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_size_t, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtaco_tcl4.so")

TACO_OK = 0
TACO_ERR_INVALID = 1
TACO_ERR_RANGE = 2
TACO_ERR_RUNTIME = 3
TACO_ERR_OVERFLOW = 4

_dp = POINTER(c_double)
_ip = POINTER(c_int)
_sp = POINTER(c_size_t)



class BatchSystem(ctypes.Structure):
    """taco_batch_system of include/taco_tcl4.h."""
    _fields_ = [
        ("gamma", _dp), ("Nt", c_size_t), ("nf", c_int), ("omegas", _dp), ("mirror", _ip),
        ("pair_to_freq", _ip), ("N", c_int), ("ops", _dp), ("num_ops", c_int), ("dt", c_double),
        ("tidx", _sp), ("ntau", c_size_t), ("steps", c_size_t), ("order", c_int), ("eps", _dp), ("r0", _dp),
    ]


class TapeSet(ctypes.Structure):
    """taco_tape_set of include/taco_tcl4.h."""
    _fields_ = [
        ("gamma", _dp), ("Nt", c_size_t), ("nf", c_int), ("omegas", _dp), ("mirror", _ip),
        ("pair_to_freq", _ip), ("N", c_int), ("eps", _dp), ("ops", _dp), ("num_ops", c_int), ("dt", c_double),
        ("steps", c_size_t), ("order", c_int), ("r0", _dp),
    ]


class BathParams(ctypes.Structure):
    """taco_bath_params of include/taco_tcl4.h."""
    _fields_ = [("kind", c_int), ("p0", c_double), ("p1", c_double), ("beta", c_double)]


_dpp = POINTER(_dp)

# name -> (restype, argtypes); every symbol declared in include/taco_tcl4.h.
SIGNATURES = {
    "taco_last_error": (c_char_p, []),
    "taco_version": (c_char_p, []),
    "taco_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "taco_ctx_destroy": (c_int, [c_void_p]),
    "taco_ctx_stream": (c_void_p, [c_void_p]),
    "taco_set_fcr_fft_pad_factor": (None, [c_size_t]),
    "taco_cr_sincos": (c_int, [c_void_p, _dp, c_size_t, _dp, _dp]),
    "taco_cr_exp": (c_int, [c_void_p, _dp, c_size_t, _dp]),
    "taco_bath_gamma_hip": (c_int, [c_void_p, c_size_t, c_double, c_int, c_double, c_double, c_double, _dp, c_int,
                                    _dp, _dp]),
    "taco_gamma_trapz_hip": (c_int, [c_void_p, _dp, c_size_t, c_double, _dp, c_int, _dp]),
    "taco_profile_enable": (c_int, [c_void_p, c_int]),
    "taco_fft_tile_choice": (c_int, [c_void_p, c_int, _ip]),
    "taco_profile_reset": (c_int, [c_void_p]),
    "taco_profile_get": (c_int, [c_void_p, c_int, _dp, POINTER(ctypes.c_long), _dp]),
    "taco_get_fcr_fft_pad_factor": (c_size_t, []),
    "taco_set_phase_mode": (c_int, [c_int]),
    "taco_get_phase_mode": (c_int, []),
    "taco_set_fft_mode": (c_int, [c_int]),
    "taco_get_fft_mode": (c_int, []),
    "taco_set_slice_budget": (c_int, [c_void_p, c_size_t]),
    "taco_get_slice_budget": (c_size_t, [c_void_p]),
    "taco_system_build": (c_int, [_dp, c_int, _dp, c_int, c_double, _dp, _dp, _dp, _ip, _dp, _ip, _ip, _ip]),
    "taco_system_build_hip": (c_int, [c_void_p, c_int, _dp, c_int, _dp, c_int, c_double, _dp, _dp, _dp, _ip, _dp,
                                      _ip, _ip, _ip]),
    "taco_bcf_fft": (c_int, [c_size_t, c_double, c_int, c_double, c_double, c_double, _dp]),
    "taco_gamma_trapz": (c_int, [_dp, c_size_t, c_double, _dp, c_int, _dp]),
    "taco_spectral_create": (c_int, [c_char_p, POINTER(c_char_p), _dp, c_int, POINTER(c_void_p)]),
    "taco_spectral_eval": (c_int, [c_void_p, _dp, c_size_t, _dp]),
    "taco_spectral_destroy": (c_int, [c_void_p]),
    "taco_bcf_fft_spectral": (c_int, [c_size_t, c_double, c_void_p, c_double, _dp]),
    "taco_gamma_rect": (c_int, [_dp, c_size_t, c_double, _dp, c_int, _dp]),
    "taco_triple_kernels": (c_int, [c_void_p, _dp, c_size_t, c_int, _dp, _ip, c_double, c_int, c_size_t,
                                    _dp, _dp, _dp]),
    "taco_l4_batch": (c_int, [c_void_p, _dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, c_int, c_double,
                              c_int, _sp, c_size_t, _dp, _dp]),
    "taco_l4_batch_dev": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                  c_void_p, c_int, c_double, c_int, _sp, c_size_t, c_void_p, _dp]),
    "taco_l4_single": (c_int, [c_void_p, _dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, c_int, c_double,
                               c_int, c_size_t, _dp]),
    "taco_l4_from_slices": (c_int, [c_void_p, _dp, _dp, _dp, _ip, c_int, c_int, _dp, c_int, _dp, _dp]),
    "taco_tcl2_generator": (c_int, [c_void_p, _dp, c_size_t, c_int, c_int, _dp, _dp, _ip, c_int, _sp, c_size_t,
                                    _dp]),
    "taco_mikx_device": (c_int, [c_void_p] * 4 + [c_int, c_int] + [c_void_p] * 5),
    "taco_assemble_gw_device": (c_int, [c_void_p] * 5 + [c_int, c_int, c_void_p, c_int, c_void_p]),
    "taco_fcr_batched_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_int, c_double, c_int, c_int,
                                     c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "taco_l4_from_slices_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                                            c_size_t, c_void_p, c_void_p, c_void_p]),
    "taco_rk4_propagate": (c_int, [c_void_p, _dp, _dp, c_size_t, c_int, _dp, c_double, c_size_t, _dp, _sp]),
    "taco_tcl4_propagate": (c_int, [c_void_p, _dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, _dp, c_int,
                                    c_double, c_size_t, c_int, _dp, _dp]),
    "taco_batch_create": (c_int, [c_void_p, POINTER(BatchSystem), c_int, POINTER(c_void_p)]),
    "taco_batch_run": (c_int, [c_void_p, c_int]),
    "taco_batch_l4": (c_int, [c_void_p, c_int, _dp]),
    "taco_batch_rho": (c_int, [c_void_p, c_int, _dp, _dp]),
    "taco_batch_destroy": (c_int, [c_void_p]),
    "taco_batched_run": (c_int, [c_void_p, POINTER(BatchSystem), c_int, POINTER(_dp), POINTER(_dp)]),
    "taco_tape_forward": (c_int, [c_void_p, _dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, _dp, c_int, c_double,
                                  c_size_t, c_int, _dp, c_size_t, _dp, POINTER(c_void_p)]),
    "taco_tape_adjoint": (c_int, [c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]),
    "taco_tape_destroy": (c_int, [c_void_p]),
    "taco_tape_forward_bath": (c_int, [c_void_p, c_int, c_double, c_double, c_double, c_size_t, c_int, _dp, _ip, _ip,
                                       c_int, _dp, _dp, c_int, c_double, c_size_t, c_int, _dp, c_size_t, _dp,
                                       POINTER(c_void_p)]),
    "taco_tape_adjoint_bath": (c_int, [c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]),
    "taco_mtape_forward": (c_int, [c_void_p, POINTER(TapeSet), c_int, c_size_t, _dpp, POINTER(c_void_p)]),
    "taco_mtape_adjoint": (c_int, [c_void_p, _dpp, _dpp, _dpp, _dpp, _dpp, _dpp]),
    "taco_mtape_forward_bath": (c_int, [c_void_p, POINTER(TapeSet), POINTER(BathParams), c_int, c_size_t, _dpp,
                                        POINTER(c_void_p)]),
    "taco_mtape_adjoint_bath": (c_int, [c_void_p, _dpp, _dpp, _dpp, _dpp, _dpp, _dpp]),
    "taco_mtape_keeps_slices": (c_int, [c_void_p]),
    "taco_mtape_set_slices": (c_int, [c_void_p, c_size_t, POINTER(c_size_t)]),
    "taco_mtape_destroy": (c_int, [c_void_p]),
    "taco_adjoint": (c_int, [c_void_p, _dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, _dp, c_int, c_double,
                             c_size_t, c_int, _dp, c_size_t, _dp, _dp, _dp, _dp, _dp, _dp]),
}

_lib = None


class TacoError(RuntimeError):
    """Base of the engine's errors; ``code`` is the C-ABI status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class TacoInvalidArgument(TacoError, ValueError):
    """std::invalid_argument in the reference."""


class TacoOutOfRange(TacoError, IndexError):
    """std::out_of_range in the reference."""


class TacoOverflow(TacoError, OverflowError):
    """std::overflow_error in the reference."""


_ERR = {TACO_ERR_INVALID: TacoInvalidArgument, TACO_ERR_RANGE: TacoOutOfRange, TACO_ERR_OVERFLOW: TacoOverflow}


def lib() -> ctypes.CDLL:
    """The loaded engine library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"taco_amd: {LIB_PATH} is not built; run `python -m taco_amd.build` "
                              "(the engine has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != TACO_OK:
        msg = lib().taco_last_error().decode()
        raise _ERR.get(rc, TacoError)(rc, msg)


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def sptr(a):
    return a.ctypes.data_as(_sp)
