"""ctypes binding of libldpc_mi355x.so (the C-ABI in include/ldpc_mi355x.h).

Loading fails loudly: there is no CPU fallback for any decoding entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libldpc_mi355x.so"

_c_int_p = C.POINTER(C.c_int)
_u32_p = C.POINTER(C.c_uint32)
_i8_p = C.POINTER(C.c_int8)
_u8_p = C.POINTER(C.c_uint8)


class Tuning(C.Structure):
    _fields_ = [("words_per_block", C.c_int), ("threads", C.c_int),
                ("early_term", C.c_int), ("post_in_lds", C.c_int), ("uniform", C.c_int),
                ("batch", C.c_int)]


class SimStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("bit_errors", C.c_uint64), ("frame_errors", C.c_uint64),
                ("syndrome_failures", C.c_uint64), ("avg_iterations", C.c_double), ("decode_ms", C.c_double)]


class LearnConfig(C.Structure):
    _fields_ = [("steps", C.c_int), ("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("ebn0_lo", C.c_double), ("ebn0_hi", C.c_double), ("rate", C.c_double),
                ("seed", C.c_uint64), ("first_frame", C.c_uint64), ("quant", C.c_int), ("pmin", C.c_double),
                ("pmax", C.c_double)]


_f_p = C.POINTER(C.c_float)

# name -> (restype, argtypes); mirrors include/ldpc_mi355x.h one to one
SIGNATURES = {
    "ldpc_last_error": (C.c_char_p, []),
    "ldpc_version": (C.c_char_p, []),
    "ldpc_code_create": (C.c_void_p, [C.c_int, C.c_int, _c_int_p, _u32_p, C.c_int]),
    "ldpc_code_create_grouped": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _u32_p]),
    "ldpc_code_load_table": (C.c_void_p, [C.c_char_p, C.c_int]),
    "ldpc_code_regular": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "ldpc_code_qc": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64]),
    "ldpc_code_ira": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "ldpc_code_reorder_by_colour": (C.c_void_p, [C.c_void_p]),
    "ldpc_code_destroy": (None, [C.c_void_p]),
    "ldpc_code_info": (C.c_int, [C.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "ldpc_code_export": (C.c_int, [C.c_void_p, _c_int_p, _u32_p]),
    "ldpc_code_levels": (C.c_int, [C.c_void_p, C.c_int, _c_int_p]),
    "ldpc_decoder_create": (C.c_void_p, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(Tuning)]),
    "ldpc_decoder_initialize": (C.c_int, [C.c_void_p]),
    "ldpc_decoder_destroy": (None, [C.c_void_p]),
    "ldpc_decoder_decode": (C.c_int, [C.c_void_p, _i8_p, _u8_p, C.c_int]),
    "ldpc_decoder_decode_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ldpc_decoder_stream": (C.c_void_p, [C.c_void_p]),
    "ldpc_decoder_config": (C.c_int, [C.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "ldpc_decoder_set_table": (C.c_int, [C.c_void_p, _i8_p, C.c_int, C.c_int]),
    "ldpc_decoder_load": (C.c_int, [C.c_void_p, _i8_p]),
    "ldpc_decoder_iterate": (C.c_int, [C.c_void_p, C.c_int]),
    "ldpc_decoder_get_state": (C.c_int, [C.c_void_p, _u32_p, _u32_p]),
    "ldpc_decoder_get_hard": (C.c_int, [C.c_void_p, _u8_p]),
    "ldpc_decoder_get_stop_info": (C.c_int, [C.c_void_p, _u8_p, _c_int_p]),
    "ldpc_interleave_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]),
    "ldpc_deinterleave_hd_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]),
    "ldpc_channel_generate": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_uint64, C.c_uint64,
                                        C.c_uint32, C.c_double, C.c_int, C.c_void_p]),
    "ldpc_channel_sigma": (C.c_double, [C.c_double, C.c_double, C.c_int]),
    "ldpc_decoder_count_errors": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int, C.c_void_p]),
    "ldpc_channel_generate_range": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint32,
                                              C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p]),
    "ldpc_simulate": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int,
                                C.POINTER(SimStats)]),
    "ldpc_trainer_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_size_t]),
    "ldpc_trainer_destroy": (None, [C.c_void_p]),
    "ldpc_trainer_loss_grad": (C.c_int, [C.c_void_p, _f_p, C.c_void_p, _f_p, _f_p, C.c_void_p]),
    "ldpc_trainer_fit": (C.c_int, [C.c_void_p, C.POINTER(LearnConfig), _f_p, _f_p]),
    "ldpc_pool_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(Tuning)]),
    "ldpc_pool_set_table": (C.c_int, [C.c_void_p, _i8_p, C.c_int, C.c_int]),
    "ldpc_pool_set_order": (C.c_int, [C.c_void_p, C.c_int]),
    "ldpc_pool_decode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ldpc_pool_decode_graph": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ldpc_pool_simulate": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64,
                                     C.POINTER(SimStats)]),
    "ldpc_pool_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64,
                                     C.POINTER(Tuning), C.POINTER(C.c_double), _c_int_p]),
    "ldpc_mixed_create": (C.c_void_p, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(Tuning), C.c_uint64]),
    "ldpc_mixed_set_table": (C.c_int, [C.c_void_p, C.c_int, _i8_p, C.c_int, C.c_int]),
    "ldpc_mixed_decode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                    C.c_void_p, C.c_int, C.c_void_p]),
    "ldpc_mixed_config": (C.c_int, [C.c_void_p, C.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                    _c_int_p]),
    "ldpc_mixed_destroy": (None, [C.c_void_p]),
    "ldpc_mixed_decode_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                            C.c_void_p, C.c_void_p]),
    "ldpc_pool_config": (C.c_int, [C.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "ldpc_pool_stream": (C.c_void_p, [C.c_void_p]),
    "ldpc_pool_destroy": (None, [C.c_void_p]),
    "ldpc_device_alloc": (C.c_void_p, [C.c_size_t]),
    "ldpc_device_free": (None, [C.c_void_p]),
    "ldpc_host_alloc_pinned": (C.c_void_p, [C.c_size_t]),
    "ldpc_host_free_pinned": (None, [C.c_void_p]),
    "ldpc_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "ldpc_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "ldpc_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "ldpc_device_synchronize": (C.c_int, []),
    "ldpc_device_count": (C.c_int, []),
    "ldpc_timer_create": (C.c_void_p, [C.c_void_p]),
    "ldpc_timer_start": (C.c_int, [C.c_void_p]),
    "ldpc_timer_stop": (C.c_int, [C.c_void_p]),
    "ldpc_timer_elapsed_ms": (C.c_double, [C.c_void_p]),
    "ldpc_timer_destroy": (None, [C.c_void_p]),
}

_lib = None


class LdpcError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LdpcError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m fast_ldpc_decoder_for_gpu_fixed_amd.build` (there is no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return load().ldpc_last_error().decode()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise LdpcError(f"{what} failed ({rc}): {last_error()}")


def ptr(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))
