"""CPU oracle for the Morpheus SpMV / Convert hot path.

TEST INFRASTRUCTURE ONLY.  A ctypes wrapper over ``libmph_oracle.so`` (the
plain-C restatement of the reference Serial backend in ``mph_oracle.c`` and
of its OpenMP multiply in ``mph_omp_baseline.c``).  Only tests/, the smoke
check and the benchmark's cpu_baseline leg may import it, and only to check
or to time the CPU baseline; the product package never does.

The reference itself (C++ on Kokkos 3.5.00, googletest) cannot be built in
this image -- Kokkos is not vendored and cannot be fetched -- so there is no
``oracle/_ref`` build; the restatement is pinned instead by the golden
fixtures, closed forms and brute-force checks in tests/test_oracle_golden.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from morpheus_amd.formats import (ALIGNMENT, COO_FORMAT, CSR_FORMAT, DIA_FORMAT,
                                  ELL_FORMAT, HDC_FORMAT, HYB_FORMAT, Coo, Csr, Dia,
                                  Ell, Hdc, Hyb, pad_size)

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmph_oracle.so"

OK, INVALID_INPUT, FORMAT_CONVERSION = 0, 2, 3


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: oracle status {code}")
        self.code = code


class FormatConversion(OracleError):
    pass


_lib = None

p = C.c_void_p
i64 = C.c_int64
I = C.c_int
pi64 = C.POINTER(C.c_int64)

_SIGS = {
    "orc_csr_spmv": (None, [i64, p, p, p, p, p, I]),
    "orc_coo_spmv": (None, [i64, i64, p, p, p, p, p, I]),
    "orc_dia_spmv": (None, [i64, i64, i64, i64, p, p, p, p, I]),
    "orc_ell_spmv": (None, [i64, i64, i64, p, p, p, p, I]),
    "orc_csr_absrow": (None, [i64, p, p, p, p, p]),
    "orc_csr_spmv_compensated": (None, [i64, p, p, p, p, p]),
    "orc_csr_spmv_f32": (None, [i64, p, p, p, p, p, I]),
    "orc_exceeds_tolerance": (I, [i64, i64, i64]),
    "orc_pad_size": (i64, [i64, i64]),
    "orc_coo_is_sorted": (I, [i64, p, p]),
    "orc_coo_sort_by_row_and_column": (I, [i64, p, p, p]),
    "orc_count_nnz_per_diagonal": (None, [i64, i64, i64, p, p, p]),
    "orc_csr_update_diagonal": (None, [i64, p, p, p, p]),
    "orc_coo_update_diagonal": (None, [i64, p, p, p, p]),
    "orc_dia_update_diagonal": (None, [i64, i64, i64, i64, p, p, p]),
    "orc_ell_update_diagonal": (None, [i64, i64, i64, p, p, p]),
    "orc_dot": (C.c_double, [i64, p, p]),
    "orc_waxpby": (None, [i64, C.c_double, p, C.c_double, p, p]),
    "orc_reduce": (C.c_double, [i64, p]),
    "orc_cg_csr": (C.c_double, [i64, p, p, p, p, p, i64]),
    "orc_csr_nnz_per_row": (None, [i64, p, p, I]),
    "orc_coo_nnz_per_row": (None, [i64, i64, p, p, I]),
    "orc_dia_nnz_per_row": (None, [i64, i64, i64, p, p, I]),
    "orc_ell_nnz_per_row": (None, [i64, i64, i64, p, p, I]),
    "orc_csr_nnz_per_diagonal": (None, [i64, i64, p, p, p, I]),
    "orc_dia_nnz_per_diagonal": (None, [i64, i64, i64, i64, p, p, p, I]),
    "orc_ell_nnz_per_diagonal": (None, [i64, i64, i64, i64, p, p, p, I]),
    "orc_i32_stats": (None, [i64, p, C.c_double, C.c_int32, pi64, pi64,
                             C.POINTER(C.c_double), pi64]),
    "orc_csr_to_coo": (None, [i64, i64, p, p, p, p, p, p]),
    "orc_coo_to_csr": (None, [i64, i64, p, p, p, p, p, p]),
    "orc_coo_dia_ndiag": (i64, [i64, p, p]),
    "orc_coo_to_dia": (I, [i64, i64, i64, p, p, p, i64, i64, p, p]),
    "orc_dia_to_coo": (I, [i64, i64, i64, i64, i64, p, p, p, p, p]),
    "orc_coo_ell_entries_per_row": (i64, [i64, i64, p]),
    "orc_coo_to_ell": (I, [i64, i64, p, p, p, i64, i64, i64, p, p]),
    "orc_ell_to_coo": (I, [i64, i64, i64, i64, p, p, p, p, p]),
    "orc_coo_hyb_sizes": (I, [i64, i64, p, pi64, pi64, pi64]),
    "orc_coo_to_hyb": (I, [i64, i64, p, p, p, i64, i64, i64, p, p, p, p, p]),
    "orc_hyb_to_coo": (I, [i64, i64, i64, i64, p, p, i64, p, p, p, p, p, p]),
    "orc_coo_hdc_sizes": (I, [i64, i64, i64, p, p, pi64, pi64, pi64]),
    "orc_coo_to_hdc": (I, [i64, i64, i64, p, p, p, i64, i64, p, p, p, p, p]),
    "orc_hdc_to_coo": (I, [i64, i64, i64, i64, i64, p, p, i64, p, p, p, p, p, p]),
    "orc_omp_max_threads": (I, []),
    "orc_omp_csr_spmv": (None, [i64, p, p, p, p, p, I]),
    "orc_omp_coo_spmv": (None, [i64, i64, p, p, p, p, p, I]),
    "orc_omp_dia_spmv": (None, [i64, i64, i64, i64, p, p, p, p, I]),
    "orc_omp_ell_spmv": (None, [i64, i64, i64, p, p, p, p, I]),
    "orc_omp_copy_f64": (None, [i64, p, p]),
    "orc_omp_copy_i32": (None, [i64, p, p]),
}


def build() -> Path:
    """Compile the oracle (gcc, via oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a.size else None


def _chk(code, what):
    if code == FORMAT_CONVERSION:
        raise FormatConversion(code, what)
    if code != OK:
        raise OracleError(code, what)


def _i32(n):
    return np.zeros(max(int(n), 0), dtype=np.int32)


def _f64(n):
    return np.zeros(max(int(n), 0), dtype=np.float64)


# ------------------------------------------------------------------ multiply
def spmv(A, x, y=None, init=True):
    """Serial multiply of any host container; returns y."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = _f64(A.nrows) if y is None else np.array(y, dtype=np.float64, copy=True)
    if isinstance(A, Csr):
        L.orc_csr_spmv(A.nrows, _ptr(A.row_offsets), _ptr(A.column_indices),
                       _ptr(A.values), _ptr(x), _ptr(y), int(init))
    elif isinstance(A, Coo):
        L.orc_coo_spmv(A.nrows, A.nnz, _ptr(A.row_indices), _ptr(A.column_indices),
                       _ptr(A.values), _ptr(x), _ptr(y), int(init))
    elif isinstance(A, Dia):
        L.orc_dia_spmv(A.nrows, A.ncols, A.ndiag, A.pitch, _ptr(A.offsets),
                       _ptr(A.values), _ptr(x), _ptr(y), int(init))
    elif isinstance(A, Ell):
        L.orc_ell_spmv(A.nrows, A.width, A.pitch, _ptr(A.column_indices),
                       _ptr(A.values), _ptr(x), _ptr(y), int(init))
    elif isinstance(A, Hyb):
        # core/src/impl/Hyb/Serial/Morpheus_Multiply_Impl.hpp:51-52
        y = spmv(A.ell, x, y, init)
        y = spmv(A.coo, x, y, False)
    elif isinstance(A, Hdc):
        # core/src/impl/Hdc/Serial/Morpheus_Multiply_Impl.hpp:51-52
        y = spmv(A.dia, x, y, init)
        y = spmv(A.csr, x, y, False)
    else:
        raise TypeError(type(A))
    return y


def absrow(A: Csr, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = _f64(A.nrows)
    lib().orc_csr_absrow(A.nrows, _ptr(A.row_offsets), _ptr(A.column_indices),
                         _ptr(A.values), _ptr(x), _ptr(out))
    return out


def spmv_compensated(A: Csr, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = _f64(A.nrows)
    lib().orc_csr_spmv_compensated(A.nrows, _ptr(A.row_offsets), _ptr(A.column_indices),
                                   _ptr(A.values), _ptr(x), _ptr(out))
    return out


F32_PLAIN, F32_LANES, F32_TWOSUM, F32_DOT2 = 0, 1, 2, 3


def spmv_f32(A: Csr, x, mode: int) -> np.ndarray:
    """FP32 row sums of FP32-rounded values and x in one of four schemes
    (plain serial, 64 lanes + tree, TwoSum only, Dot2): the witnesses of the
    FP32 accuracy check, not a reference restatement."""
    v = np.ascontiguousarray(A.values, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(A.nrows, dtype=np.float32)
    lib().orc_csr_spmv_f32(A.nrows, _ptr(A.row_offsets), _ptr(A.column_indices),
                           _ptr(v), _ptr(x), _ptr(out), int(mode))
    return out


# ------------------------------------------------------------------ helpers
def exceeds_tolerance(nrows, nnz, ncols) -> bool:
    return bool(lib().orc_exceeds_tolerance(nrows, nnz, ncols))


def is_sorted(A: Coo) -> bool:
    return bool(lib().orc_coo_is_sorted(A.nnz, _ptr(A.row_indices), _ptr(A.column_indices)))


def sort_by_row_and_column(A: Coo) -> Coo:
    B = Coo(A.nrows, A.ncols, A.row_indices.copy(), A.column_indices.copy(), A.values.copy())
    _chk(lib().orc_coo_sort_by_row_and_column(B.nnz, _ptr(B.row_indices),
                                              _ptr(B.column_indices), _ptr(B.values)),
         "sort_by_row_and_column")
    return B


def count_nnz_per_diagonal(A: Coo) -> np.ndarray:
    out = _i32(A.nrows + A.ncols - 1)
    lib().orc_count_nnz_per_diagonal(A.nrows, A.ncols, A.nnz, _ptr(A.row_indices),
                                     _ptr(A.column_indices), _ptr(out))
    return out


# ------------------------------------------------------------ matrix operations
def update_diagonal(A, diag):
    """Morpheus::update_diagonal (Serial) on a copy of the host container;
    returns the updated copy."""
    import copy
    B = copy.deepcopy(A)
    d = np.ascontiguousarray(diag, dtype=np.float64)
    L = lib()
    if isinstance(B, Csr):
        L.orc_csr_update_diagonal(B.nrows, _ptr(B.row_offsets), _ptr(B.column_indices),
                                  _ptr(B.values), _ptr(d))
    elif isinstance(B, Coo):
        L.orc_coo_update_diagonal(B.nnz, _ptr(B.row_indices), _ptr(B.column_indices),
                                  _ptr(B.values), _ptr(d))
    elif isinstance(B, Dia):
        L.orc_dia_update_diagonal(B.nrows, B.ncols, B.ndiag, B.pitch, _ptr(B.offsets),
                                  _ptr(B.values), _ptr(d))
    elif isinstance(B, Ell):
        L.orc_ell_update_diagonal(B.nrows, B.width, B.pitch, _ptr(B.column_indices),
                                  _ptr(B.values), _ptr(d))
    elif isinstance(B, Hyb):
        # core/src/impl/Hyb/Serial/Morpheus_MatrixOperations_Impl.hpp: ELL, then COO
        B.ell = update_diagonal(B.ell, d)
        B.coo = update_diagonal(B.coo, d)
    elif isinstance(B, Hdc):
        B.dia = update_diagonal(B.dia, d)
        B.csr = update_diagonal(B.csr, d)
    else:
        raise TypeError(type(B))
    return B


def get_diagonal(A) -> np.ndarray:
    """The diagonal the product sees: the sum of every stored (i, i) entry
    (the reference's get_diagonal throws NotImplementedException)."""
    C_ = to_coo(A)
    d = np.zeros(A.nrows)
    m = C_.row_indices == C_.column_indices
    np.add.at(d, C_.row_indices[m], C_.values[m])
    return d


# ------------------------------------------------------------ vector algorithms
def dot(x, y=None) -> float:
    """Morpheus::dot (Serial): sequential sum of x_i y_i (x_i^2 when y is x)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = x if y is None else np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_dot(x.size, _ptr(x), _ptr(y))


def waxpby(alpha, x, beta, y) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    w = _f64(x.size)
    lib().orc_waxpby(x.size, float(alpha), _ptr(x), float(beta), _ptr(y), _ptr(w))
    return w


def reduce(v) -> float:
    """Morpheus::reduce (Serial): Kahan summation."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_reduce(v.size, _ptr(v))


def cg(A: Csr, b, x0, iters):
    """`iters` textbook CG iterations from x0; returns (x, r.r)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    rr = lib().orc_cg_csr(A.nrows, _ptr(A.row_offsets), _ptr(A.column_indices),
                          _ptr(A.values), _ptr(b), _ptr(x), int(iters))
    return x, rr


# ---------------------------------------------------------------- analytics
def nnz_per_row(A, out=None, init=True) -> np.ndarray:
    """Impl::count_nnz_per_row of any host container (Serial semantics;
    core/src/impl/*/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp)."""
    L = lib()
    out = _i32(A.nrows) if out is None else np.array(out, dtype=np.int32, copy=True)
    if isinstance(A, Csr):
        L.orc_csr_nnz_per_row(A.nrows, _ptr(A.row_offsets), _ptr(out), int(init))
    elif isinstance(A, Coo):
        L.orc_coo_nnz_per_row(A.nrows, A.nnz, _ptr(A.row_indices), _ptr(out), int(init))
    elif isinstance(A, Dia):
        L.orc_dia_nnz_per_row(A.nrows, A.ncols, A.ndiag, _ptr(A.offsets), _ptr(out), int(init))
    elif isinstance(A, Ell):
        L.orc_ell_nnz_per_row(A.nrows, A.width, A.pitch, _ptr(A.column_indices), _ptr(out),
                              int(init))
    elif isinstance(A, Hyb):
        # core/src/impl/Hyb/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp: ELL, then COO
        out = nnz_per_row(A.coo, nnz_per_row(A.ell, out, init), False)
    elif isinstance(A, Hdc):
        # core/src/impl/Hdc/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp: DIA, then CSR
        out = nnz_per_row(A.csr, nnz_per_row(A.dia, out, init), False)
    else:
        raise TypeError(type(A))
    return out


def nnz_per_diagonal(A, out=None, init=True) -> np.ndarray:
    """Impl::count_nnz_per_diagonal of any host container; entry d is
    diagonal d - (nrows - 1)."""
    L = lib()
    n = A.nrows + A.ncols - 1
    out = _i32(n) if out is None else np.array(out, dtype=np.int32, copy=True)
    if isinstance(A, Csr):
        L.orc_csr_nnz_per_diagonal(A.nrows, A.ncols, _ptr(A.row_offsets),
                                   _ptr(A.column_indices), _ptr(out), int(init))
    elif isinstance(A, Coo):
        if init:
            out[:] = 0
        L.orc_count_nnz_per_diagonal(A.nrows, A.ncols, A.nnz, _ptr(A.row_indices),
                                     _ptr(A.column_indices), _ptr(out))
    elif isinstance(A, Dia):
        L.orc_dia_nnz_per_diagonal(A.nrows, A.ncols, A.ndiag, A.pitch, _ptr(A.offsets),
                                   _ptr(A.values), _ptr(out), int(init))
    elif isinstance(A, Ell):
        L.orc_ell_nnz_per_diagonal(A.nrows, A.ncols, A.width, A.pitch,
                                   _ptr(A.column_indices), _ptr(A.values), _ptr(out), int(init))
    elif isinstance(A, Hyb):
        out = nnz_per_diagonal(A.coo, nnz_per_diagonal(A.ell, out, init), False)
    elif isinstance(A, Hdc):
        out = nnz_per_diagonal(A.csr, nnz_per_diagonal(A.dia, out, init), False)
    else:
        raise TypeError(type(A))
    return out


def i32_stats(v, mean=0.0, threshold=0):
    """(max, min, sum of squared deviations from mean, count of entries >
    threshold and != 0) of an int32 vector."""
    v = np.ascontiguousarray(v, dtype=np.int32)
    mx, mn, c = C.c_int64(), C.c_int64(), C.c_int64()
    sq = C.c_double()
    lib().orc_i32_stats(v.size, _ptr(v), float(mean), int(threshold), C.byref(mx),
                        C.byref(mn), C.byref(sq), C.byref(c))
    return mx.value, mn.value, sq.value, c.value


def nnz_of(A) -> int:
    """A.nnnz() of the reference containers."""
    if isinstance(A, Hyb):
        return A.ell.nnz + A.coo.nnz
    if isinstance(A, Hdc):
        return A.dia.nnz + A.csr.nnz
    return int(A.nnz) if not isinstance(A, Csr) else int(A.row_offsets[-1])


def row_stats(A):
    """max_nnnz, min_nnnz and std_nnnz as core/src/Morpheus_MatrixAnalytics.hpp
    :126-204 computes them: std around the integer mean nnnz / nrows,
    truncated to the size type."""
    counts = nnz_per_row(A)
    mean = float(nnz_of(A) // A.nrows)
    mx, mn, sq, _ = i32_stats(counts, mean)
    return mx, mn, int(np.sqrt(sq / A.nrows))


def count_diagonals(A, threshold=0) -> int:
    """core/src/Morpheus_MatrixAnalytics.hpp:220-259 (count_diagonals with
    threshold 0, count_true_diagonals otherwise)."""
    return i32_stats(nnz_per_diagonal(A), 0.0, threshold)[3]


# ------------------------------------------------------------------ convert
def to_coo(A) -> Coo:
    """Serial X -> COO."""
    L = lib()
    if isinstance(A, Coo):
        return Coo(A.nrows, A.ncols, A.row_indices.copy(), A.column_indices.copy(),
                   A.values.copy())
    n = A.nnz
    ri, ci, v = _i32(n), _i32(n), _f64(n)
    if isinstance(A, Csr):
        L.orc_csr_to_coo(A.nrows, n, _ptr(A.row_offsets), _ptr(A.column_indices),
                         _ptr(A.values), _ptr(ri), _ptr(ci), _ptr(v))
    elif isinstance(A, Dia):
        _chk(L.orc_dia_to_coo(A.nrows, A.ncols, n, A.ndiag, A.pitch, _ptr(A.offsets),
                              _ptr(A.values), _ptr(ri), _ptr(ci), _ptr(v)), "dia_to_coo")
    elif isinstance(A, Ell):
        _chk(L.orc_ell_to_coo(A.nrows, n, A.width, A.pitch, _ptr(A.column_indices),
                              _ptr(A.values), _ptr(ri), _ptr(ci), _ptr(v)), "ell_to_coo")
    elif isinstance(A, Hyb):
        e, c = A.ell, A.coo
        _chk(L.orc_hyb_to_coo(e.nrows, n, e.width, e.pitch, _ptr(e.column_indices),
                              _ptr(e.values), c.nnz, _ptr(c.row_indices),
                              _ptr(c.column_indices), _ptr(c.values), _ptr(ri), _ptr(ci),
                              _ptr(v)), "hyb_to_coo")
    elif isinstance(A, Hdc):
        d, c = A.dia, A.csr
        _chk(L.orc_hdc_to_coo(d.nrows, d.ncols, n, d.ndiag, d.pitch, _ptr(d.offsets),
                              _ptr(d.values), c.nnz, _ptr(c.row_offsets),
                              _ptr(c.column_indices), _ptr(c.values), _ptr(ri), _ptr(ci),
                              _ptr(v)), "hdc_to_coo")
    else:
        raise TypeError(type(A))
    return Coo(A.nrows, A.ncols, ri, ci, v)


def from_coo(A: Coo, fmt: int):
    """Serial COO -> X."""
    L = lib()
    n, nc, nnz = A.nrows, A.ncols, A.nnz
    ri, ci, vv = _ptr(A.row_indices), _ptr(A.column_indices), _ptr(A.values)
    if fmt == COO_FORMAT:
        return to_coo(A)
    if fmt == CSR_FORMAT:
        ro, c, v = _i32(n + 1), _i32(nnz), _f64(nnz)
        L.orc_coo_to_csr(n, nnz, ri, ci, vv, _ptr(ro), _ptr(c), _ptr(v))
        return Csr(n, nc, ro, c, v)
    if fmt == DIA_FORMAT:
        nd = L.orc_coo_dia_ndiag(nnz, ri, ci) if nnz else 0
        if nnz and exceeds_tolerance(n, nnz, nd):  # checked before allocating
            raise FormatConversion(FORMAT_CONVERSION, "coo_to_dia")
        pitch = pad_size(n, ALIGNMENT)
        off, v = _i32(nd), _f64(pitch * nd)
        _chk(L.orc_coo_to_dia(n, nc, nnz, ri, ci, vv, nd, pitch, _ptr(off), _ptr(v)),
             "coo_to_dia")
        return Dia(n, nc, nnz, off, v, pitch)
    if fmt == ELL_FORMAT:
        K = L.orc_coo_ell_entries_per_row(n, nnz, ri)
        if exceeds_tolerance(n, nnz, K):  # checked before allocating
            raise FormatConversion(FORMAT_CONVERSION, "coo_to_ell")
        width = pad_size(K, ALIGNMENT)
        c, v = _i32(width * n), _f64(width * n)
        _chk(L.orc_coo_to_ell(n, nnz, ri, ci, vv, K, width, n, _ptr(c), _ptr(v)),
             "coo_to_ell")
        return Ell(n, nc, nnz, K, width, c, v, n)
    if fmt == HYB_FORMAT:
        K, en, cn = C.c_int64(), C.c_int64(), C.c_int64()
        _chk(L.orc_coo_hyb_sizes(n, nnz, ri, C.byref(K), C.byref(en), C.byref(cn)),
             "coo_hyb_sizes")
        width = pad_size(K.value, ALIGNMENT)
        ec, ev = _i32(width * n), _f64(width * n)
        cr, cc, cv = _i32(cn.value), _i32(cn.value), _f64(cn.value)
        _chk(L.orc_coo_to_hyb(n, nnz, ri, ci, vv, K.value, width, n, _ptr(ec), _ptr(ev),
                              _ptr(cr), _ptr(cc), _ptr(cv)), "coo_to_hyb")
        return Hyb(Ell(n, nc, en.value, K.value, width, ec, ev, n), Coo(n, nc, cr, cc, cv))
    if fmt == HDC_FORMAT:
        nd, dn, cn = C.c_int64(), C.c_int64(), C.c_int64()
        _chk(L.orc_coo_hdc_sizes(n, nc, nnz, ri, ci, C.byref(nd), C.byref(dn), C.byref(cn)),
             "coo_hdc_sizes")
        pitch = pad_size(n, ALIGNMENT)
        off, dv = _i32(nd.value), _f64(pitch * nd.value)
        ro, c, v = _i32(n + 1), _i32(cn.value), _f64(cn.value)
        _chk(L.orc_coo_to_hdc(n, nc, nnz, ri, ci, vv, nd.value, pitch, _ptr(off), _ptr(dv),
                              _ptr(ro), _ptr(c), _ptr(v)), "coo_to_hdc")
        return Hdc(Dia(n, nc, dn.value, off, dv, pitch), Csr(n, nc, ro, c, v))
    raise ValueError(fmt)


def convert(A, fmt: int):
    """Serial convert through the COO hub (core/src/impl/Morpheus_Convert_Impl.hpp:83-106);
    same-format conversion is an element-wise copy."""
    if A.format == fmt:
        import copy
        return copy.deepcopy(A)
    return from_coo(to_coo(A), fmt)


# ------------------------------------------------------------------ OpenMP
def omp_threads() -> int:
    return int(lib().orc_omp_max_threads())


def omp_spmv_fn(A):
    """A zero-argument callable running the OpenMP restatement of the
    reference multiply on A (arrays first-touched by the worker threads),
    plus the output vector it writes."""
    L = lib()

    def ft_f(a):
        out = np.empty_like(a)
        L.orc_omp_copy_f64(a.size, _ptr(a), _ptr(out))
        return out

    def ft_i(a):
        out = np.empty_like(a)
        L.orc_omp_copy_i32(a.size, _ptr(a), _ptr(out))
        return out

    x = ft_f(np.random.default_rng(1).uniform(-1, 1, A.ncols))
    y = ft_f(np.zeros(A.nrows))
    if isinstance(A, Csr):
        ro, ci, v = ft_i(A.row_offsets), ft_i(A.column_indices), ft_f(A.values)
        return (lambda: L.orc_omp_csr_spmv(A.nrows, _ptr(ro), _ptr(ci), _ptr(v), _ptr(x),
                                           _ptr(y), 1)), x, y
    if isinstance(A, Coo):
        ri, ci, v = ft_i(A.row_indices), ft_i(A.column_indices), ft_f(A.values)
        return (lambda: L.orc_omp_coo_spmv(A.nrows, A.nnz, _ptr(ri), _ptr(ci), _ptr(v),
                                           _ptr(x), _ptr(y), 1)), x, y
    if isinstance(A, Dia):
        off, v = ft_i(A.offsets), ft_f(A.values)
        return (lambda: L.orc_omp_dia_spmv(A.nrows, A.ncols, A.ndiag, A.pitch, _ptr(off),
                                           _ptr(v), _ptr(x), _ptr(y), 1)), x, y
    if isinstance(A, Ell):
        ci, v = ft_i(A.column_indices), ft_f(A.values)
        return (lambda: L.orc_omp_ell_spmv(A.nrows, A.width, A.pitch, _ptr(ci), _ptr(v),
                                           _ptr(x), _ptr(y), 1)), x, y
    if isinstance(A, Hyb):
        f1, x1, y1 = omp_spmv_fn(A.ell)
        ri, ci, v = ft_i(A.coo.row_indices), ft_i(A.coo.column_indices), ft_f(A.coo.values)

        def run():
            f1()
            L.orc_omp_coo_spmv(A.nrows, A.coo.nnz, _ptr(ri), _ptr(ci), _ptr(v), _ptr(x1),
                               _ptr(y1), 0)
        return run, x1, y1
    if isinstance(A, Hdc):
        f1, x1, y1 = omp_spmv_fn(A.dia)
        ro, ci, v = ft_i(A.csr.row_offsets), ft_i(A.csr.column_indices), ft_f(A.csr.values)

        def run():
            f1()
            L.orc_omp_csr_spmv(A.nrows, _ptr(ro), _ptr(ci), _ptr(v), _ptr(x1), _ptr(y1), 0)
        return run, x1, y1
    raise TypeError(type(A))
