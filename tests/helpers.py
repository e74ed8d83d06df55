"""Shared test helpers: the reference's 10x10 fixture in all six formats,
the FP error bound, and exact container comparison."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from morpheus_amd.formats import Coo, Csr, Dia, Ell, Hdc, Hyb, pad_size

GOLDEN = Path(__file__).resolve().parent / "golden"
U64 = 2.0 ** -53


def _small():
    return json.loads((GOLDEN / "small_matrix.json").read_text())


def small_fixture():
    """The reference fixture (core/tests/utils/Macros_*Matrix.hpp
    reset_small_container) as host containers, keyed by format name."""
    d = _small()
    k = d["constants"]
    n, m = k["SMALL_MATRIX_NROWS"], k["SMALL_MATRIX_NCOLS"]
    al = k["SMALL_MATRIX_ALIGNMENT"]
    F = d["formats"]

    def vec(entries, size, dtype):
        out = np.zeros(size, dtype=dtype)
        for idx, v in entries:
            out[idx[0]] = v
        return out

    nnz = k["SMALL_MATRIX_NNZ"]
    coo = Coo(n, m, vec(F["Coo"]["row_indices"], nnz, np.int32),
              vec(F["Coo"]["column_indices"], nnz, np.int32),
              vec(F["Coo"]["values"], nnz, np.float64))
    csr = Csr(n, m, vec(F["Csr"]["row_offsets"], n + 1, np.int32),
              vec(F["Csr"]["column_indices"], nnz, np.int32),
              vec(F["Csr"]["values"], nnz, np.float64))

    def dia(entries_off, entries_val, nd, nnz_):
        pitch = pad_size(n, al)
        off = vec(entries_off, nd, np.int32)
        vals = np.zeros(pitch * nd)
        for (r, c), v in entries_val:
            vals[r + pitch * c] = v
        return Dia(n, m, nnz_, off, vals, pitch)

    def ell(entries_ci, entries_v, K, nnz_):
        width = pad_size(K, al)
        ci = np.full(width * n, -1, dtype=np.int32)
        vals = np.zeros(width * n)
        for (r, c), v in entries_ci:
            ci[r + n * c] = v
        for (r, c), v in entries_v:
            vals[r + n * c] = v
        return Ell(n, m, nnz_, K, width, ci, vals, n)

    D = F["Dia"]
    dia_c = dia(D["diagonal_offsets"], D["values"], k["SMALL_DIA_MATRIX_NDIAGS"], nnz)
    E = F["Ell"]
    ell_c = ell(E["column_indices"], E["values"], k["SMALL_ELL_ENTRIES_PER_ROW"], nnz)
    H = F["Hyb"]
    hyb_ell = ell(H["ell.column_indices"], H["ell.values"], k["SMALL_HYB_ENTRIES_PER_ROW"],
                  k["SMALL_HYB_ELL_NNZ"])
    cn = k["SMALL_HYB_COO_NNZ"]
    hyb_coo = Coo(n, m, vec(H["coo.row_indices"], cn, np.int32),
                  vec(H["coo.column_indices"], cn, np.int32), vec(H["coo.values"], cn, np.float64))
    hyb = Hyb(hyb_ell, hyb_coo)
    G = F["Hdc"]
    hdc_dia = dia(G["dia.diagonal_offsets"], G["dia.values"], k["SMALL_HDC_DIA_NDIAG"],
                  k["SMALL_HDC_DIA_NNZ"])
    hc = k["SMALL_HDC_CSR_NNZ"]
    hdc_csr = Csr(n, m, vec(G["csr.row_offsets"], n + 1, np.int32),
                  vec(G["csr.column_indices"], hc, np.int32), vec(G["csr.values"], hc, np.float64))
    hdc = Hdc(hdc_dia, hdc_csr)
    return {"COO": coo, "CSR": csr, "DIA": dia_c, "ELL": ell_c, "HYB": hyb, "HDC": hdc}


def same_container(a, b) -> bool:
    """Bit-exact equality of two host containers (values compared as bits)."""
    if type(a) is not type(b):
        return False
    if isinstance(a, Hyb):
        return same_container(a.ell, b.ell) and same_container(a.coo, b.coo)
    if isinstance(a, Hdc):
        return same_container(a.dia, b.dia) and same_container(a.csr, b.csr)
    if (a.nrows, a.ncols) != (b.nrows, b.ncols):
        return False
    fields = {Coo: ["row_indices", "column_indices", "values"],
              Csr: ["row_offsets", "column_indices", "values"],
              Dia: ["offsets", "values"],
              Ell: ["column_indices", "values"]}[type(a)]
    if isinstance(a, Dia) and (a.pitch != b.pitch or a.nnz != b.nnz):
        return False
    if isinstance(a, Ell) and (a.entries_per_row, a.width, a.pitch, a.nnz) != \
            (b.entries_per_row, b.width, b.pitch, b.nnz):
        return False
    for f in fields:
        x, y = getattr(a, f), getattr(b, f)
        if x.shape != y.shape:
            return False
        if x.dtype.kind == "f":
            if not np.array_equal(x.view(np.uint64), y.view(np.uint64)):
                return False
        elif not np.array_equal(x, y):
            return False
    return True


def fp64_bound(nnz_row, absrow, y0abs=None):
    """|y - y_ref| <= 4 * n_row * u * sum_j |a_ij x_j| (y0 counted as one more
    term when accumulating)."""
    n = np.asarray(nnz_row, dtype=np.float64)
    s = np.asarray(absrow, dtype=np.float64)
    if y0abs is not None:
        n = n + 1
        s = s + np.abs(y0abs)
    return 4.0 * np.maximum(n, 1) * U64 * s + 1e-300


def max_rel_err(y, yref, absrow):
    """max_i |y_i - yref_i| / sum_j |a_ij x_j| (rows with zero scale skipped)."""
    y = np.asarray(y, dtype=np.float64)
    s = np.asarray(absrow)
    mask = s > 0
    if not mask.any():
        return 0.0
    return float(np.max(np.abs(y[mask] - yref[mask]) / s[mask]))


def anti_diagonal_matrix(n: int, m: int, diags, seed: int) -> Csr:
    """The reference's AntiDiagonalMatrixGenerator
    (core/tests/utils/MatrixGenerator.hpp:147-160): A(i, m - j - 1) is set
    wherever |j - i| equals |d| for a listed d, with values U(-2, 2) (drawn
    here from a seeded generator)."""
    rng = np.random.default_rng(seed)
    dense = np.zeros((n, m))
    i = np.arange(n)[:, None]
    j = np.arange(m)[None, :]
    mask = np.zeros((n, m), dtype=bool)
    for d in diags:
        mask |= np.abs(j - i) == abs(int(d))
    # column j of the generator's loop lands at m - j - 1
    mask = mask[:, ::-1]
    vals = rng.uniform(-2.0, 2.0, size=int(mask.sum()))
    vals[vals == 0] = 1.0
    dense[mask] = vals
    rows, cols = np.nonzero(dense)
    counts = np.bincount(rows, minlength=n)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    return Csr(n, m, ro, cols, dense[rows, cols])


def reference_diags(n: int, freq: int = 5):
    """The diagonal list of the reference's analytics tests
    (core/tests/Test_MatrixAnalytics.hpp:145-160): 0 and +-k*freq."""
    return [0] + [s * d for d in range(freq, n, freq) for s in (1, -1)]


def dense_of(A: Csr) -> np.ndarray:
    D = np.zeros((A.nrows, A.ncols))
    r = np.repeat(np.arange(A.nrows), np.diff(A.row_offsets))
    D[r, A.column_indices] = A.values
    return D


U32 = 2.0 ** -24


def round_sig(v, bits: int):
    """Round to `bits` significant bits (24: FP32, 8: BF16), to nearest."""
    m, e = np.frexp(np.asarray(v, dtype=np.float64))
    return np.ldexp(np.round(m * 2.0 ** bits) / 2.0 ** bits, e)


def cancellation_csr(nrows: int, ncols: int, long_lengths, seed: int, bits: int = 24,
                     rho: float = 1e-3, short_max: int = 8) -> tuple:
    """A CSR matrix and x whose long rows cancel heavily, with every value
    exact in `bits` significant bits (24: FP32, 8: BF16) and x exact in FP32,
    so every product is exact in FP64.

    Each row holds its diagonal entry.  Rows listed in `long_lengths` (placed
    at random rows) have that many entries: in column order the first half
    of the products are positive and the second half negative, scaled so the
    row sum is about rho * (sum of the positive products); sum |a x| is then
    about 2 / rho times |sum a x|, and every partial sum in column order is
    large.  Other rows hold 1..short_max entries with values in [-1, 1].
    Returns (A, x, long_rows)."""
    from morpheus_amd.formats import Csr
    rng = np.random.default_rng(seed)
    x = round_sig(rng.uniform(0.5, 2.0, ncols), 24)
    long_rows = rng.choice(nrows, size=len(long_lengths), replace=False)
    lens = rng.integers(1, short_max + 1, nrows)
    lens[long_rows] = long_lengths
    cols, vals = [], []
    for r in range(nrows):
        n = int(lens[r])
        c = rng.choice(ncols - 1, size=n - 1, replace=False)
        c = np.sort(np.concatenate([c + (c >= r), [r]])) if r < ncols else np.sort(
            rng.choice(ncols, size=n, replace=False))
        if r in long_rows:
            a = round_sig(rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-6, 6, n), bits)
            h = n // 2
            p = a * x[c]
            scale = p[:h].sum() * (1.0 - rho) / p[h:].sum()
            a[h:] = -round_sig(a[h:] * scale, bits)
        else:
            a = round_sig(rng.uniform(-1.0, 1.0, n), bits)
            a[a == 0] = 0.5
        cols.append(c)
        vals.append(a)
    ro = np.zeros(nrows + 1, dtype=np.int64)
    np.cumsum(lens, out=ro[1:])
    A = Csr(nrows, ncols, ro, np.concatenate(cols), np.concatenate(vals))
    return A, x, np.sort(long_rows)


def compensated_bound(nnz_row, absrow, y_exact):
    """The accuracy of a Dot2-compensated FP32 row sum over exact inputs:
    |y - y_exact| <= 4 u |y_exact| + n u^2 sum_j |a_ij x_j| (u = 2^-24)."""
    n = np.asarray(nnz_row, dtype=np.float64)
    return 4.0 * U32 * np.abs(y_exact) + n * U32 * U32 * np.asarray(absrow) + 1e-300


def plain_fp32_bound(nnz_row, absrow):
    """What any FP32 summation order guarantees: (n + 2) u sum_j |a_ij x_j|."""
    n = np.asarray(nnz_row, dtype=np.float64)
    return (n + 2.0) * U32 * np.asarray(absrow) + 1e-300
