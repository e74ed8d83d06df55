"""Host-side (numpy) containers for the six sparse formats.

These mirror the storage of the reference containers as the HIP backend
sees them (LayoutLeft, int32 indices): ``core/src/Morpheus_CooMatrix.hpp``,
``Morpheus_CsrMatrix.hpp:534-535``, ``Morpheus_DiaMatrix.hpp:192-201``
(values ``pad(nrows, 32) x ndiag``, column-major), ``Morpheus_EllMatrix.hpp:
189-202`` (``nrows x pad(K, 32)`` column-major, padding index ``-1``),
``Morpheus_HybMatrix.hpp`` (ELL + COO) and ``Morpheus_HdcMatrix.hpp``
(DIA + CSR).  Values are kept as float64 on the host; the device copies may
be narrowed to float32 / bfloat16.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Morpheus::formats_e (core/src/Morpheus_FormatsRegistry.hpp:56-64)
COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT = range(6)
NFORMATS = 6
FORMAT_NAMES = ["COO", "CSR", "DIA", "ELL", "HYB", "HDC"]
ALIGNMENT = 32  # default alignment of DIA rows / ELL columns


def pad_size(length: int, alignment: int = ALIGNMENT) -> int:
    """core/src/impl/Dia/Morpheus_Utils_Impl.hpp:33-36"""
    return alignment * ((length + alignment - 1) // alignment)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class Coo:
    nrows: int
    ncols: int
    row_indices: np.ndarray
    column_indices: np.ndarray
    values: np.ndarray
    format: int = field(default=COO_FORMAT, init=False)

    def __post_init__(self):
        self.row_indices = _i32(self.row_indices)
        self.column_indices = _i32(self.column_indices)
        self.values = _f64(self.values)

    @property
    def nnz(self) -> int:
        return int(self.values.size)


@dataclass
class Csr:
    nrows: int
    ncols: int
    row_offsets: np.ndarray
    column_indices: np.ndarray
    values: np.ndarray
    format: int = field(default=CSR_FORMAT, init=False)

    def __post_init__(self):
        self.row_offsets = _i32(self.row_offsets)
        self.column_indices = _i32(self.column_indices)
        self.values = _f64(self.values)

    @property
    def nnz(self) -> int:
        return int(self.values.size)


@dataclass
class Dia:
    """values is column-major: values[r + pitch*d] = A(r, r + offsets[d])."""

    nrows: int
    ncols: int
    nnz: int
    offsets: np.ndarray
    values: np.ndarray  # flat, pitch * ndiag
    pitch: int
    format: int = field(default=DIA_FORMAT, init=False)

    def __post_init__(self):
        self.offsets = _i32(self.offsets)
        self.values = _f64(self.values)

    @property
    def ndiag(self) -> int:
        return int(self.offsets.size)

    def value(self, r: int, d: int) -> float:
        return float(self.values[r + self.pitch * d])


@dataclass
class Ell:
    """slot (r, n) at [r + pitch*n]; width = pad(K, 32) stored columns."""

    nrows: int
    ncols: int
    nnz: int
    entries_per_row: int
    width: int
    column_indices: np.ndarray  # flat, pitch * width
    values: np.ndarray
    pitch: int
    format: int = field(default=ELL_FORMAT, init=False)

    def __post_init__(self):
        self.column_indices = _i32(self.column_indices)
        self.values = _f64(self.values)


@dataclass
class Hyb:
    ell: Ell
    coo: Coo
    format: int = field(default=HYB_FORMAT, init=False)

    @property
    def nrows(self):
        return self.ell.nrows

    @property
    def ncols(self):
        return self.ell.ncols

    @property
    def nnz(self):
        return self.ell.nnz + self.coo.nnz


@dataclass
class Hdc:
    dia: Dia
    csr: Csr
    format: int = field(default=HDC_FORMAT, init=False)

    @property
    def nrows(self):
        return self.dia.nrows

    @property
    def ncols(self):
        return self.dia.ncols

    @property
    def nnz(self):
        return self.dia.nnz + self.csr.nnz


def csr_from_triplets(nrows, ncols, rows, cols, vals) -> Csr:
    """Row-sorted, column-sorted CSR from unique (row, col, val) triplets."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=nrows)
    ro = np.zeros(nrows + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    return Csr(nrows, ncols, ro, cols, vals)


def csr_to_coo_host(A: Csr) -> Coo:
    rows = np.repeat(np.arange(A.nrows, dtype=np.int32), np.diff(A.row_offsets))
    return Coo(A.nrows, A.ncols, rows, A.column_indices.copy(), A.values.copy())


def to_dense(A) -> np.ndarray:
    """Dense float64 matrix of any host container (test helper)."""
    D = np.zeros((A.nrows, A.ncols))
    if isinstance(A, Coo):
        np.add.at(D, (A.row_indices, A.column_indices), A.values)
    elif isinstance(A, Csr):
        c = csr_to_coo_host(A)
        np.add.at(D, (c.row_indices, c.column_indices), c.values)
    elif isinstance(A, Dia):
        for d, off in enumerate(A.offsets):
            for r in range(A.nrows):
                col = r + off
                if 0 <= col < A.ncols:
                    D[r, col] += A.values[r + A.pitch * d]
    elif isinstance(A, Ell):
        for n in range(A.width):
            for r in range(A.nrows):
                c = A.column_indices[r + A.pitch * n]
                if c != -1:
                    D[r, c] += A.values[r + A.pitch * n]
    elif isinstance(A, Hyb):
        D = to_dense(A.ell) + to_dense(A.coo)
    elif isinstance(A, Hdc):
        D = to_dense(A.dia) + to_dense(A.csr)
    else:
        raise TypeError(type(A))
    return D
