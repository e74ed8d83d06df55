"""MatrixMarket I/O for the Python side: ``read_matrix_market`` and
``write_matrix_market`` with the behaviour of Morpheus::IO
(core/src/Morpheus_MatrixMarket.hpp:62-209,
core/src/impl/Morpheus_MatrixMarket_Impl.hpp:42-496), the same as the C++
drop-in's morpheus_amd/include/Morpheus_AMD_MatrixMarket.hpp:

* a coordinate file gives a host ``Coo``: 1-based indices checked against
  the size line, "symmetric" entries mirrored right after each off-diagonal
  one, "pattern" values 1, entries stably sorted by (row, column);
* an array file gives a 2-D numpy array filled row by row (as the reference
  reads it; a vector file is its single column);
* complex data, hermitian / skew-symmetric matrices, pattern arrays and
  non-general arrays raise NotImplementedException; malformed files raise
  IOException, with the reference's messages;
* "unsymmetric" is read as "general" (the reference's reader turns it into
  a zero-filled matrix).

Parsing is host work; the matrix goes to the GPU with ``device.to_device``.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .capi import IOException, NotImplementedException
from .formats import Coo

_STORAGE = ("array", "coordinate")
_TYPES = ("complex", "real", "integer", "pattern")
_SYMMETRY = ("general", "symmetric", "unsymmetric", "hermitian", "skew-symmetric")


def _banner(line: str):
    t = line.split()
    if len(t) != 5 or t[0] != "%%MatrixMarket" or t[1] != "matrix":
        raise IOException("invalid MatrixMarket banner")
    storage, typ, sym = t[2], t[3], t[4]
    if storage not in _STORAGE:
        raise IOException(f"invalid MatrixMarket storage format [{storage}]")
    if typ not in _TYPES:
        raise IOException(f"invalid MatrixMarket data type [{typ}]")
    if sym not in _SYMMETRY:
        raise IOException(f"invalid MatrixMarket symmetry [{sym}]")
    return storage, typ, sym


def read_matrix_market(path):
    """Reads a MatrixMarket file into a host ``Coo`` (coordinate) or a 2-D
    numpy array (array)."""
    try:
        text = Path(path).read_text()
    except OSError as e:
        raise IOException(f'unable to open file "{path}" for reading') from e
    lines = text.splitlines()
    if not lines:
        raise IOException("invalid MatrixMarket banner")
    storage, typ, sym = _banner(lines[0])
    k = 1
    while k < len(lines) and lines[k].startswith("%"):
        k += 1
    size = lines[k].split() if k < len(lines) else []
    body = " ".join(lines[k + 1:]).split()
    if storage == "coordinate":
        return _coordinate(size, body, typ, sym)
    return _array(size, body, typ, sym)


def _coordinate(size, body, typ, sym):
    if len(size) != 3:
        raise IOException("invalid MatrixMarket coordinate format")
    nrows, ncols, nnz = (int(s) for s in size)
    if typ == "complex":
        raise NotImplementedException("Morpheus does not currently support complex matrices")
    per = 2 if typ == "pattern" else 3
    if len(body) < per * nnz:
        raise IOException("unexpected EOF while reading MatrixMarket entries")
    vals = np.array(body[:per * nnz], dtype=object).reshape(nnz, per) if nnz else \
        np.zeros((0, per), dtype=object)
    r = vals[:, 0].astype(np.int64)
    c = vals[:, 1].astype(np.int64)
    v = np.ones(nnz) if typ == "pattern" else vals[:, 2].astype(np.float64)
    if nnz:
        if r.min() < 1:
            raise IOException("found invalid row index (index < 1)")
        if c.min() < 1:
            raise IOException("found invalid column index (index < 1)")
        if r.max() > nrows:
            raise IOException("found invalid row index (index > num_rows)")
        if c.max() > ncols:
            raise IOException("found invalid column index (index > num_columns)")
    r, c = r - 1, c - 1
    if sym == "hermitian":
        raise NotImplementedException(
            "MatrixMarket I/O does not currently support hermitian matrices")
    if sym == "skew-symmetric":
        raise NotImplementedException(
            "MatrixMarket I/O does not currently support skew-symmetric matrices")
    if sym == "symmetric":
        off = r != c
        # entry n, then its mirror when off the diagonal
        rr = np.stack([r, np.where(off, c, -1)], axis=1).ravel()
        cc = np.stack([c, np.where(off, r, -1)], axis=1).ravel()
        vv = np.stack([v, v], axis=1).ravel()
        keep = rr >= 0
        r, c, v = rr[keep], cc[keep], vv[keep]
    order = np.lexsort((c, r))  # stable
    return Coo(nrows, ncols, r[order], c[order], v[order])


def _array(size, body, typ, sym):
    if len(size) != 2:
        raise IOException("invalid MatrixMarket array format")
    nrows, ncols = (int(s) for s in size)
    if typ == "pattern":
        raise NotImplementedException("pattern array MatrixMarket format is not supported")
    if typ == "complex":
        raise NotImplementedException("complex type is not supported")
    if len(body) < nrows * ncols:
        raise IOException("unexpected EOF while reading MatrixMarket entries")
    dense = np.array(body[:nrows * ncols], dtype=np.float64).reshape(nrows, ncols)
    if sym != "general":
        raise NotImplementedException(
            "only general array symmetric MatrixMarket format is supported")
    return dense


def write_matrix_market(A, path, precision: int = 6):
    """Writes a host ``Coo`` (coordinate real general) or a 1-D / 2-D numpy
    array (array real general), 1-based, values with `precision` significant
    digits (6 is the C++ stream default the reference writes with)."""
    path = Path(path)
    try:
        f = path.open("w")
    except OSError as e:
        raise IOException(f'unable to open file "{path}" for writing') from e
    with f:
        if isinstance(A, Coo):
            f.write("%%MatrixMarket matrix coordinate real general\n")
            f.write(f"{A.nrows}\t{A.ncols}\t{A.nnz}\n")
            for r, c, v in zip(A.row_indices, A.column_indices, A.values):
                f.write(f"{r + 1} {c + 1} {float(v):.{precision}g}\n")
        else:
            a = np.asarray(A)
            f.write("%%MatrixMarket matrix array real general\n")
            if a.ndim == 1:
                f.write(f"{a.size}\t1\n")
            else:
                f.write(f"{a.shape[0]}\t{a.shape[1]}\n")
            for v in a.ravel():
                f.write(f"{float(v):.{precision}g}\n")
