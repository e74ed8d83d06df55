"""Seeded synthetic matrices for the SpMV workloads (numpy, int32 indices).

Every generator returns a row-sorted, column-sorted ``formats.Csr`` with
float64 values; the device side narrows values when a reduced-precision
storage type is requested.

* ``poisson2d(g)``            2D 5-point Poisson, diagonal 4, neighbours -1
* ``stencil3d_7(g)``          3D 7-point stencil, diagonal 6, neighbours -1
* ``stencil3d_27(g)``         3D 27-point stencil, diagonal 26, neighbours -1
* ``powerlaw_rows(...)``      power-law ("Zipf") row lengths, uniform columns
* ``banded(...)``             band of width 2h+1 plus random off-band entries
* ``diagonal_test_matrix(n)`` the reference test fixture's pattern: random
  values in [-2, 2] on the main diagonal and every 5th diagonal either side
  (core/tests/Test_Multiply.hpp:161-186, core/tests/utils/MatrixGenerator.hpp:53-65)
* ``batch_set(...)``          512 mixed matrices for the batched SpMM
"""
from __future__ import annotations

import math

import numpy as np

from .formats import Csr


def _stencil(dims, offsets, diag_value, off_value) -> Csr:
    """CSR of a structured-grid stencil; offsets are (dz, dy, dx) tuples."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    strides = []
    acc = 1
    for d in reversed(dims):
        strides.append(acc)
        acc *= d
    strides = list(reversed(strides))
    coords = np.indices(dims, dtype=np.int32).reshape(len(dims), -1)
    lin = [sum(o[k] * strides[k] for k in range(len(dims))) for o in offsets]
    order = np.argsort(lin, kind="stable")
    offsets = [offsets[i] for i in order]
    cols = np.empty((n, len(offsets)), dtype=np.int32)
    valid = np.ones((n, len(offsets)), dtype=bool)
    base = np.arange(n, dtype=np.int64)
    for j, o in enumerate(offsets):
        ok = np.ones(n, dtype=bool)
        for k in range(len(dims)):
            c = coords[k] + o[k]
            ok &= (c >= 0) & (c < dims[k])
        valid[:, j] = ok
        cols[:, j] = (base + sum(o[k] * strides[k] for k in range(len(dims)))).astype(np.int32)
    vals = np.where(np.array([all(v == 0 for v in o) for o in offsets])[None, :],
                    diag_value, off_value).astype(np.float64)
    vals = np.broadcast_to(vals, (n, len(offsets)))
    counts = valid.sum(axis=1)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    return Csr(n, n, ro, cols[valid], vals[valid])


def poisson2d(g: int) -> Csr:
    offs = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)]
    return _stencil((g, g), offs, 4.0, -1.0)


def stencil3d_7(g: int) -> Csr:
    offs = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    return _stencil((g, g, g), offs, 6.0, -1.0)


def stencil3d_27(g: int) -> Csr:
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    return _stencil((g, g, g), offs, 26.0, -1.0)


# ------------------------------------------------------------ power-law rows
def _powerlaw_survival(t, s, R):
    """P(r >= t) for a continuous density proportional to r^-s on [1, R]."""
    a = 1.0 - s
    return (np.power(t, a) - R ** a) / (1.0 - R ** a)


def _powerlaw_mean_len(s, R, cap):
    """Exact mean of L = 1 + floor((r-1) c), c = (cap-1)/(R-1)."""
    c = (cap - 1) / (R - 1)
    j = np.arange(1, cap, dtype=np.float64)
    t = np.minimum(1.0 + j / c, R)
    return 1.0 + float(np.sum(_powerlaw_survival(t, s, R)))


def powerlaw_parameters(s: float, mean: float, cap: int):
    """Range R of the power law such that the rescaled lengths
    L = 1 + floor((r-1)(cap-1)/(R-1)) have the requested mean and reach cap."""
    lo, hi = math.log(cap + 1.0), math.log(1e300)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        m = _powerlaw_mean_len(s, math.exp(mid), cap)
        if m > mean:
            lo = mid
        else:
            hi = mid
    return math.exp(0.5 * (lo + hi))


def powerlaw_row_lengths(nrows: int, s: float, mean: float, cap: int, rng):
    R = powerlaw_parameters(s, mean, cap)
    c = (cap - 1) / (R - 1)
    a = 1.0 - s
    u = rng.random(nrows)
    # inverse CDF of the density r^-s on [1, R]
    r = np.power(1.0 - u * (1.0 - R ** a), 1.0 / a)
    L = 1 + np.floor((r - 1.0) * c).astype(np.int64)
    return np.minimum(L, cap)


def _unique_columns(rows: np.ndarray, ncols: int, rng) -> np.ndarray:
    """Uniform random columns, distinct within each row (redraw collisions)."""
    cols = rng.integers(0, ncols, size=rows.size, dtype=np.int64)
    while True:
        key = rows * ncols + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(ks.size, dtype=bool)
        dup[1:] = ks[1:] == ks[:-1]
        if not dup.any():
            return cols[order], order
        idx = order[dup]
        cols[idx] = rng.integers(0, ncols, size=idx.size, dtype=np.int64)


def powerlaw_rows(nrows: int, ncols: int, s: float, mean: float, cap: int,
                  seed: int) -> Csr:
    rng = np.random.default_rng(seed)
    L = powerlaw_row_lengths(nrows, s, mean, min(cap, ncols), rng)
    rows = np.repeat(np.arange(nrows, dtype=np.int64), L)
    cols, order = _unique_columns(rows, ncols, rng)
    rows = rows[order]
    vals = rng.standard_normal(rows.size)
    ro = np.zeros(nrows + 1, dtype=np.int64)
    np.cumsum(L, out=ro[1:])
    return Csr(nrows, ncols, ro, cols, vals)


def banded(nrows: int, half_width: int, offband_frac: float, seed: int) -> Csr:
    """Band |col-row| <= half_width plus offband_frac * (band nnz) random
    entries outside the band (deduplicated); values N(0, 1)."""
    rng = np.random.default_rng(seed)
    r = np.arange(nrows, dtype=np.int64)
    rows_l, cols_l = [], []
    for o in range(-half_width, half_width + 1):
        c = r + o
        ok = (c >= 0) & (c < nrows)
        rows_l.append(r[ok])
        cols_l.append(c[ok])
    band_rows = np.concatenate(rows_l)
    band_cols = np.concatenate(cols_l)
    n_off = int(round(offband_frac * band_rows.size))
    orow = rng.integers(0, nrows, size=n_off, dtype=np.int64)
    ocol = rng.integers(0, nrows, size=n_off, dtype=np.int64)
    keep = np.abs(ocol - orow) > half_width
    rows = np.concatenate([band_rows, orow[keep]])
    cols = np.concatenate([band_cols, ocol[keep]])
    key = np.unique(rows * nrows + cols)
    rows, cols = key // nrows, key % nrows
    vals = rng.standard_normal(rows.size)
    counts = np.bincount(rows, minlength=nrows)
    ro = np.zeros(nrows + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    return Csr(nrows, nrows, ro, cols, vals)


def diagonal_test_matrix(n: int, seed: int, freq: int = 5) -> Csr:
    rng = np.random.default_rng(seed)
    offs = [0] + [s * d for d in range(freq, n, freq) for s in (1, -1)]
    rows_l, cols_l = [], []
    r = np.arange(n, dtype=np.int64)
    for o in sorted(offs):
        c = r + o
        ok = (c >= 0) & (c < n)
        rows_l.append(r[ok])
        cols_l.append(c[ok])
    rows = np.concatenate(rows_l)
    cols = np.concatenate(cols_l)
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    vals = rng.uniform(-2.0, 2.0, size=rows.size)
    counts = np.bincount(rows, minlength=n)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    return Csr(n, n, ro, cols, vals)


def random_vector(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=n)


# ------------------------------------------------------------ named workloads
def P2(seed: int = 5374857):
    """2D 5-point Poisson on a 256x256 grid (65,536 rows, 326,656 nnz)."""
    return poisson2d(256)


def S7(g: int = 128):
    """3D 7-point stencil on a 128^3 grid (2,097,152 rows, 14,581,760 nnz)."""
    return stencil3d_7(g)


def S27(g: int = 160):
    """3D 27-point stencil on a 160^3 grid (4,096,000 rows, 109,215,352 nnz)."""
    return stencil3d_27(g)


def Z15(nrows: int = 1_000_000, seed: int = 15):
    """1M rows, power-law(1.5) row lengths with mean 16 capped at 20k."""
    return powerlaw_rows(nrows, nrows, 1.5, 16.0, 20_000, seed)


def Z12(nrows: int = 2_000_000, seed: int = 12):
    """2M rows, power-law(1.2) row lengths with mean 30 capped at 200k."""
    return powerlaw_rows(nrows, nrows, 1.2, 30.0, 200_000, seed)


def B9(nrows: int = 4_000_000, seed: int = 9):
    """4M rows, bandwidth 9, plus 2 % random off-band entries."""
    return banded(nrows, 4, 0.02, seed)


def batch_set(nmat: int = 512, seed: int = 2024, nmin: int = 1000, nmax: int = 65536):
    """Mixed batch: log-uniform sizes; a third each of 2D 5-point stencils,
    power-law(1.5) rows with mean 12, and banded with bandwidth 5..33."""
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(nmat):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        kind = i % 3
        if kind == 0:
            g = max(2, int(round(math.sqrt(n))))
            mats.append(("stencil", poisson2d(g)))
        elif kind == 1:
            mats.append(("powerlaw", powerlaw_rows(n, n, 1.5, 12.0, min(n, 2000),
                                                   int(rng.integers(1 << 30)))))
        else:
            bw = int(rng.integers(2, 17)) * 2 + 1  # odd bandwidth in [5, 33]
            mats.append(("banded", banded(n, bw // 2, 0.0, int(rng.integers(1 << 30)))))
    return mats
