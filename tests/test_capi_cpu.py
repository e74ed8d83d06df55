"""The C-ABI library loads and exports every symbol include/morpheus_amd.h
declares; host-only entry points behave; no GPU compute is called here."""
import ctypes as C

import numpy as np
import pytest

import oracle
from morpheus_amd import capi


def test_header_declares_entry_points():
    syms = capi.header_symbols()
    assert "mph_multiply" in syms and "mph_convert_coo_to_dia" in syms
    assert len(syms) >= 40


def test_library_exports_every_header_symbol():
    L = capi.lib()
    missing = [s for s in capi.header_symbols() if not hasattr(L, s)]
    assert missing == []


def test_binding_covers_header():
    assert set(capi.header_symbols()) == set(capi._SIGS)


def test_status_strings_and_version():
    L = capi.lib()
    assert b"gfx950" in L.mph_version()
    for s in range(7):
        assert L.mph_status_string(s)


@pytest.mark.parametrize("case", [(10 ** 6, 16 * 10 ** 6, 20000), (2_097_152, 14_581_760, 7),
                                  (4_096_000, 109_215_352, 27), (10, 32, 9), (10 ** 8, 1, 2),
                                  (10 ** 7, 0, 11), (3 * 10 ** 7, 10 ** 6, 4)])
def test_exceeds_tolerance_matches_oracle(case):
    assert bool(capi.lib().mph_exceeds_tolerance(*case)) == oracle.exceeds_tolerance(*case)


def test_invalid_arguments_rejected_before_any_device_call():
    L = capi.lib()
    h = C.c_void_p()
    assert L.mph_matrix_create_csr(C.byref(h), -1, 4, 0, None, None, None, 0) == capi.ERR_INVALID_INPUT
    assert L.mph_matrix_create_csr(C.byref(h), 4, 4, 3, None, None, None, 0) == capi.ERR_INVALID_INPUT
    assert L.mph_matrix_create_csr(C.byref(h), 4, 4, 0, C.c_void_p(8), None, None, 7) == \
        capi.ERR_INVALID_INPUT
    assert L.mph_matrix_create_ell(C.byref(h), 4, 4, 4, 5, 4, 4, C.c_void_p(8), C.c_void_p(8), 0) \
        == capi.ERR_INVALID_INPUT
    assert b"entries_per_row" in L.mph_last_error()
    with pytest.raises(capi.InvalidInputException):
        capi.call("mph_matrix_create_dia", C.byref(h), 10, 10, 5, 2, 4, C.c_void_p(8),
                  C.c_void_p(8), 0)


def test_multiply_rejects_null_handle():
    assert capi.lib().mph_multiply(None, None, None, 1, None) == capi.ERR_INVALID_INPUT


def test_host_only_import_does_not_need_torch():
    import importlib
    import sys
    mod = importlib.import_module("morpheus_amd")
    assert "morpheus_amd.device" not in sys.modules or mod is not None
    assert np is not None
