"""ctypes binding of the C-ABI declared in include/morpheus_amd.h.

This is the reference-side binding a Python caller uses (the same shape a
cgo / JNI / N-API stub would have): plain pointers, sizes and enums.  The
shared library must have been built (``python -m morpheus_amd.build``);
there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmorpheus_amd.so"
HEADER = HERE.parent / "include" / "morpheus_amd.h"

# status codes (enum mph_status)
OK, ERR_NOT_IMPLEMENTED, ERR_INVALID_INPUT, ERR_FORMAT_CONVERSION, ERR_HIP, \
    ERR_OUT_OF_MEMORY, ERR_RUNTIME = range(7)
# analysis policy flags (mph_matrix_set_policy)
POLICY_NO_VALUE_COPIES = 1

# value types (enum mph_value_type)
F64, F32, BF16, I32 = 0, 1, 2, 3
VALUE_TYPES = {"f64": F64, "f32": F32, "bf16": BF16}


class MorpheusError(RuntimeError):
    """Base of the reference exception taxonomy (Morpheus::Exception)."""

    status = ERR_RUNTIME


class NotImplementedException(MorpheusError):
    status = ERR_NOT_IMPLEMENTED


class InvalidInputException(MorpheusError):
    status = ERR_INVALID_INPUT


class FormatConversionException(MorpheusError):
    status = ERR_FORMAT_CONVERSION


class RuntimeException(MorpheusError):
    status = ERR_RUNTIME


class HipError(MorpheusError):
    status = ERR_HIP


class IOException(MorpheusError):
    """Morpheus::IOException (core/src/Morpheus_Exceptions.hpp): MatrixMarket
    files that cannot be opened or parsed."""


_EXC = {
    ERR_NOT_IMPLEMENTED: NotImplementedException,
    ERR_INVALID_INPUT: InvalidInputException,
    ERR_FORMAT_CONVERSION: FormatConversionException,
    ERR_HIP: HipError,
    ERR_OUT_OF_MEMORY: HipError,
    ERR_RUNTIME: RuntimeException,
}

p = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
I = C.c_int
pi64 = C.POINTER(C.c_int64)
pI = C.POINTER(C.c_int)


class I32Stats(C.Structure):
    _fields_ = [("max", i64), ("min", i64), ("count", i64), ("sq_dev", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("rows_chunk", i32), ("lanes", i32), ("coo_items", i32),
                ("rows_per_thread", i32), ("xtile", i32)]


_SIGS = {
    "mph_version": (C.c_char_p, []),
    "mph_status_string": (C.c_char_p, [I]),
    "mph_last_error": (C.c_char_p, []),
    "mph_initialize": (I, [I]),
    "mph_finalize": (I, []),
    "mph_matrix_create_coo": (I, [C.POINTER(p), i64, i64, i64, p, p, p, I]),
    "mph_matrix_create_csr": (I, [C.POINTER(p), i64, i64, i64, p, p, p, I]),
    "mph_matrix_create_dia": (I, [C.POINTER(p), i64, i64, i64, i64, i64, p, p, I]),
    "mph_matrix_create_ell": (I, [C.POINTER(p), i64, i64, i64, i64, i64, i64, p, p, I]),
    "mph_matrix_create_hyb": (I, [C.POINTER(p), i64, i64, i64, i64, i64, i64, p, p,
                                  i64, p, p, p, I]),
    "mph_matrix_create_hdc": (I, [C.POINTER(p), i64, i64, i64, i64, i64, p, p,
                                  i64, p, p, p, I]),
    "mph_matrix_destroy": (I, [p]),
    "mph_matrix_analyse": (I, [p, p]),
    "mph_matrix_format": (I, [p]),
    "mph_matrix_set_tuning": (I, [p, C.POINTER(Tuning)]),
    "mph_matrix_get_tuning": (I, [p, C.POINTER(Tuning)]),
    "mph_matrix_set_policy": (I, [p, I]),
    "mph_index_narrow": (I, [p, i64, p, p]),
    "mph_index_widen": (I, [p, i64, p, p]),
    "mph_values_cast": (I, [p, I, p, I, i64, p]),
    "mph_multiply": (I, [p, p, p, I, p]),
    "mph_multiply_multi": (I, [p, i32, p, p, I, p]),
    "mph_csr_multiply": (I, [i64, i64, i64, p, p, p, I, p, p, I, p]),
    "mph_coo_multiply": (I, [i64, i64, i64, p, p, p, I, p, p, I, p]),
    "mph_coo_multiply_noalloc": (I, [i64, i64, i64, p, p, p, I, p, p, I, p]),
    "mph_matrix_plan_bytes": (I, [p, I, pi64]),
    "mph_probe_read": (I, [p, i64, p, p]),
    "mph_probe_gather": (I, [p, i64, p, p, p]),
    "mph_dia_multiply": (I, [i64, i64, i64, i64, i64, p, p, I, p, p, I, p]),
    "mph_ell_multiply": (I, [i64, i64, i64, i64, i64, i64, p, p, I, p, p, I, p]),
    "mph_convert_csr_to_coo": (I, [i64, i64, p, p, p, I, p, p, p, p]),
    "mph_convert_coo_to_csr": (I, [i64, i64, p, p, p, I, p, p, p, p]),
    "mph_convert_coo_to_dia_size": (I, [i64, i64, i64, p, p, pi64, p]),
    "mph_convert_coo_to_dia": (I, [i64, i64, i64, p, p, p, I, i64, i64, p, p, p]),
    "mph_convert_dia_to_coo": (I, [i64, i64, i64, i64, i64, p, p, I, p, p, p, p]),
    "mph_convert_coo_to_ell_size": (I, [i64, i64, p, pi64, p]),
    "mph_convert_coo_to_ell": (I, [i64, i64, p, p, p, I, i64, i64, i64, p, p, p]),
    "mph_convert_ell_to_coo": (I, [i64, i64, i64, i64, p, p, I, p, p, p, p]),
    "mph_convert_coo_to_hyb_size": (I, [i64, i64, p, pi64, pi64, pi64, p]),
    "mph_convert_coo_to_hyb": (I, [i64, i64, p, p, p, I, i64, i64, i64, p, p, p, p, p, p]),
    "mph_convert_hyb_to_coo": (I, [i64, i64, i64, i64, p, p, i64, p, p, p, I, p, p, p, p]),
    "mph_convert_coo_to_hdc_size": (I, [i64, i64, i64, p, p, pi64, pi64, pi64, p]),
    "mph_convert_coo_to_hdc": (I, [i64, i64, i64, p, p, p, I, i64, i64, p, p, p, p, p, p]),
    "mph_convert_hdc_to_coo": (I, [i64, i64, i64, i64, i64, p, p, i64, p, p, p, I, p, p, p, p]),
    "mph_coo_sort_by_row_and_column": (I, [i64, p, p, p, I, p]),
    "mph_coo_is_sorted": (I, [i64, p, p, pI, p]),
    "mph_coo_count_nnz_per_diagonal": (I, [i64, i64, i64, p, p, p, I, p]),
    "mph_matrix_update_diagonal": (I, [p, p, p]),
    "mph_matrix_get_diagonal": (I, [p, p, p]),
    "mph_dot": (I, [i64, p, p, I, C.POINTER(C.c_double), p]),
    "mph_waxpby": (I, [i64, C.c_double, p, C.c_double, p, p, I, p]),
    "mph_reduce": (I, [i64, p, I, C.POINTER(C.c_double), p]),
    "mph_cg_create": (I, [C.POINTER(p), p]),
    "mph_cg_destroy": (I, [p]),
    "mph_cg_start": (I, [p, p, p, p]),
    "mph_cg_iterate": (I, [p, p, i32, p]),
    "mph_cg_residual": (I, [p, C.POINTER(C.c_double), p]),
    "mph_matrix_count_nnz_per_row": (I, [p, p, i32, p]),
    "mph_matrix_count_nnz_per_diagonal": (I, [p, p, i32, p]),
    "mph_i32_stats": (I, [p, i64, C.c_double, i32, p, p]),
    "mph_exceeds_tolerance": (I, [i64, i64, i64]),
    "mph_batch_create": (I, [C.POINTER(p), i32, C.POINTER(p), i32, C.POINTER(p),
                             C.POINTER(p), p]),
    "mph_batch_destroy": (I, [p]),
    "mph_batch_run": (I, [p, I, p]),
    "mph_batch_run_graph": (I, [p, I, p]),
    "mph_batch_info": (I, [p, pi64, C.POINTER(i32)]),
    "mph_matrix_autotune": (I, [p, i32, p, p, i32, p, C.POINTER(C.c_float)]),
    "mph_autotune_format": (I, [i64, i64, i64, p, p, p, I, i32, i32, p, pI,
                                C.POINTER(C.c_float)]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(mph_[a-z0-9_]+)\s*\(", text)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m morpheus_amd.build` "
                "(there is no CPU fallback for the MI355X kernels)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status == OK:
        return
    msg = lib().mph_last_error().decode(errors="replace")
    exc = _EXC.get(status, MorpheusError)
    raise exc(f"{what}: {lib().mph_status_string(status).decode()}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
