"""Pins the CPU oracle (oracle/mph_oracle.c) before it is trusted as the
checker: against the reference's own 10x10 fixture in all six formats,
closed forms, scipy, numpy and brute force.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  HDC_FORMAT, HYB_FORMAT, Coo, Csr, Dia, Ell,
                                  csr_to_coo_host, pad_size, to_dense)
from tests.helpers import fp64_bound, same_container, small_fixture

FMT = {"CSR": CSR_FORMAT, "DIA": DIA_FORMAT, "ELL": ELL_FORMAT, "HYB": HYB_FORMAT,
       "HDC": HDC_FORMAT, "COO": COO_FORMAT}


@pytest.fixture(scope="module")
def small():
    return small_fixture()


# ------------------------------------------------------------ golden fixture
def test_fixture_formats_describe_one_matrix(small):
    D = to_dense(small["COO"])
    assert np.count_nonzero(D) == 32
    for name, A in small.items():
        assert np.array_equal(to_dense(A), D), name


@pytest.mark.parametrize("name", ["CSR", "DIA", "ELL", "HYB", "HDC"])
def test_coo_to_format_matches_reference_fixture(small, name):
    """COO -> X reproduces the reference's hand-written X fixture bit for bit
    (HYB K=3 split, HDC threshold floor(0.2*10)=2, ascending DIA offsets,
    ELL -1 padding)."""
    got = oracle.from_coo(small["COO"], FMT[name])
    assert same_container(got, small[name]), name


@pytest.mark.parametrize("name", ["CSR", "DIA", "ELL", "HYB", "HDC"])
def test_format_to_coo_matches_reference_fixture(small, name):
    got = oracle.to_coo(small[name])
    assert same_container(got, small["COO"]), name


@pytest.mark.parametrize("src", ["COO", "CSR", "DIA", "ELL", "HYB", "HDC"])
@pytest.mark.parametrize("dst", ["COO", "CSR", "DIA", "ELL", "HYB", "HDC"])
def test_any_to_any_through_coo(small, src, dst):
    got = oracle.convert(small[src], FMT[dst])
    assert same_container(got, small[dst])


@pytest.mark.parametrize("name", ["COO", "CSR", "DIA", "ELL", "HYB", "HDC"])
@pytest.mark.parametrize("init", [True, False])
def test_multiply_small_vs_dense(small, name, init):
    """core/tests/Test_Multiply.hpp:193-271: y vs the dense loop, 1e-13."""
    A = small[name]
    x = gen.random_vector(10, 5374857)
    y0 = np.ones(10)
    D = to_dense(small["COO"])
    ref = D @ x + (0 if init else y0)
    got = oracle.spmv(A, x, y0, init)
    assert np.max(np.abs(got - ref)) < 1e-13


@pytest.mark.parametrize("name", ["COO", "CSR", "DIA", "ELL", "HYB", "HDC"])
def test_multiply_integer_values_exact(small, name):
    """Integer-valued data: the reference tests int values with tolerance 0."""
    A = oracle.convert(small["COO"], FMT[name])
    for attr in ("values",):
        pass
    coo = small["COO"]
    ints = Coo(10, 10, coo.row_indices, coo.column_indices, np.arange(1, 33, dtype=np.float64))
    Ai = oracle.convert(ints, FMT[name])
    x = np.arange(-5, 5, dtype=np.float64)
    ref = to_dense(ints) @ x
    assert np.array_equal(oracle.spmv(Ai, x), ref)
    assert A is not None


# ------------------------------------------------------------ closed forms
def _boundary_count_2d(g):
    i, j = np.divmod(np.arange(g * g), g)
    return (i == 0).astype(int) + (i == g - 1) + (j == 0) + (j == g - 1)


@pytest.mark.parametrize("fmt", ["CSR", "COO", "DIA", "ELL", "HYB", "HDC"])
def test_poisson2d_rowsums_closed_form(fmt):
    """2D 5-point Poisson, x = 1: row sum = number of missing neighbours
    (0 interior, 1 edge, 2 corner)."""
    g = 16
    A = oracle.convert(gen.poisson2d(g), FMT[fmt])
    y = oracle.spmv(A, np.ones(g * g))
    assert np.array_equal(y, _boundary_count_2d(g).astype(float))


def test_stencil3d_rowsums_closed_form():
    g = 12
    A = gen.stencil3d_7(g)
    assert A.nnz == 7 * g ** 3 - 6 * g ** 2
    c = np.indices((g, g, g)).reshape(3, -1)
    expect = sum(((c[k] == 0).astype(int) + (c[k] == g - 1)) for k in range(3))
    for fmt in ("CSR", "DIA", "ELL", "HDC"):
        y = oracle.spmv(oracle.convert(A, FMT[fmt]), np.ones(g ** 3))
        assert np.array_equal(y, expect.astype(float)), fmt


def test_stencil_nnz_closed_forms():
    assert gen.poisson2d(256).nnz == 326_656
    assert gen.stencil3d_7(16).nnz == 7 * 16 ** 3 - 6 * 16 ** 2
    g = 10
    assert gen.stencil3d_27(g).nnz == (3 * g - 2) ** 3


# ------------------------------------------------------------ scipy / brute
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_csr_spmv_vs_scipy(seed):
    A = gen.powerlaw_rows(3000, 2500, 1.5, 8.0, 1500, seed)
    x = gen.random_vector(A.ncols, seed)
    S = sp.csr_matrix((A.values, A.column_indices, A.row_offsets), shape=(A.nrows, A.ncols))
    ref = S @ x
    got = oracle.spmv(A, x)
    absrow = oracle.absrow(A, x)
    assert np.all(np.abs(got - ref) <= fp64_bound(np.diff(A.row_offsets), absrow))


def test_compensated_is_closer_than_plain():
    rng = np.random.default_rng(7)
    n = 20000
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    A = Csr(1, n, [0, n], np.arange(n), v)
    x = np.ones(n)
    exact = float(np.sum(np.array([np.longdouble(t) for t in v], dtype=np.longdouble)))
    assert abs(oracle.spmv_compensated(A, x)[0] - exact) <= abs(oracle.spmv(A, x)[0] - exact)


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_all_formats(seed):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(1, 9)), int(rng.integers(1, 9))
    D = np.where(rng.random((n, m)) < 0.4, rng.integers(-4, 5, (n, m)), 0).astype(float)
    r, c = np.nonzero(D)
    coo = Coo(n, m, r, c, D[r, c])
    x = rng.integers(-3, 4, m).astype(float)
    for f in FMT.values():
        try:
            A = oracle.from_coo(coo, f)
        except oracle.FormatConversion:
            continue
        except oracle.OracleError as e:
            # non-square HDC: the reference records the midpoint bin as
            # offset 0 and then overflows its CSR part (undefined there);
            # the restatement reports INVALID_INPUT instead
            assert f == HDC_FORMAT and n != m and e.code == oracle.INVALID_INPUT
            continue
        assert np.array_equal(to_dense(A), D)
        assert np.array_equal(oracle.spmv(A, x), D @ x)


# ------------------------------------------------------------ helpers
def test_sort_is_stable_lexicographic():
    rng = np.random.default_rng(3)
    n = 5000
    r = rng.integers(0, 50, n)
    c = rng.integers(0, 40, n)
    v = np.arange(n, dtype=float)  # identifies the original position
    got = oracle.sort_by_row_and_column(Coo(50, 40, r, c, v))
    order = np.lexsort((c, r))  # numpy lexsort is stable
    assert np.array_equal(got.row_indices, r[order])
    assert np.array_equal(got.column_indices, c[order])
    assert np.array_equal(got.values, v[order])
    assert oracle.is_sorted(got)
    assert not oracle.is_sorted(Coo(50, 40, r, c, v))


def test_exceeds_tolerance_float_decision():
    def ref(nr, ne, nc):
        size = np.float32(nc) * np.float32(nr)
        fill = size / max(np.float32(1.0), np.float32(ne))
        return bool(fill > np.float32(10.0) and size > np.float32(100e6))
    cases = [(10 ** 6, 16 * 10 ** 6, 20000), (2_097_152, 14_581_760, 7),
             (4_096_000, 109_215_352, 27), (10, 32, 9), (10 ** 8, 1, 2), (10 ** 7, 0, 11),
             (50_000_000, 20_000_000, 3)]
    for cse in cases:
        assert oracle.exceeds_tolerance(*cse) == ref(*cse), cse
    assert oracle.exceeds_tolerance(10 ** 6, 16 * 10 ** 6, 20000)  # Z15 ELL throws


def test_count_nnz_per_diagonal_vs_bincount():
    A = csr_to_coo_host(gen.powerlaw_rows(500, 700, 1.5, 6.0, 300, 11))
    got = oracle.count_nnz_per_diagonal(A)
    ref = np.bincount(A.column_indices.astype(np.int64) - A.row_indices + A.nrows - 1,
                      minlength=A.nrows + A.ncols - 1)
    assert np.array_equal(got, ref)


def test_dia_to_coo_drops_zeros_including_negative_zero():
    """core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:106: stored zeros
    (+0.0 and -0.0) are dropped; the unfilled slots stay (0, 0, 0.0) and the
    stable (row, col) sort places them after the real entries at (0, 0)."""
    pitch = 32
    vals = np.zeros(pitch * 1)
    vals[0], vals[1], vals[2] = 1.5, -0.0, 2.5
    A = Dia(3, 3, 3, np.array([0]), vals, pitch)
    got = oracle.to_coo(A)
    assert list(got.row_indices) == [0, 0, 2]
    assert list(got.column_indices) == [0, 0, 2]
    assert list(got.values) == [1.5, 0.0, 2.5]


def test_ell_to_coo_keeps_explicit_zeros():
    ci = np.full(32 * 2, -1, dtype=np.int32)
    v = np.zeros(32 * 2)
    ci[0], v[0] = 1, 0.0   # row 0 slot 0: explicit zero
    ci[1], v[1] = 0, 3.0   # row 1 slot 0
    A = Ell(2, 2, 2, 1, 32, ci, v, 2)
    got = oracle.to_coo(A)
    assert list(got.row_indices) == [0, 1] and list(got.values) == [0.0, 3.0]


def test_ell_hyb_tolerance_exceptions():
    """Z15-shaped data: ELL conversion throws, HYB does not."""
    n = 2_000_000
    rows = np.concatenate([np.arange(n), np.zeros(19_999, dtype=np.int64)])
    cols = np.concatenate([np.zeros(n, dtype=np.int64), np.arange(1, 20_000)])
    order = np.lexsort((cols, rows))
    coo = Coo(n, n, rows[order], cols[order], np.ones(rows.size))
    with pytest.raises(oracle.FormatConversion):
        oracle.from_coo(coo, ELL_FORMAT)
    with pytest.raises(oracle.FormatConversion):
        oracle.from_coo(coo, DIA_FORMAT)


def test_p2_roundtrip_csr_coo_dia_csr_bitexact():
    """CSR -> COO -> DIA -> CSR is the identity on P2 (2D 5-point Poisson 256^2)."""
    A = gen.P2()
    coo = oracle.to_coo(A)
    dia = oracle.from_coo(coo, DIA_FORMAT)
    assert list(dia.offsets) == [-256, -1, 0, 1, 256]
    back = oracle.from_coo(oracle.to_coo(dia), CSR_FORMAT)
    assert same_container(back, A)


def test_hdc_threshold_small_rows_selects_all_diagonals():
    """With fewer than 5 rows the threshold (int)(0.2*n) is 0 and every
    diagonal, populated or not, goes to the DIA part."""
    coo = Coo(4, 4, [0, 1, 3], [0, 3, 0], [1.0, 2.0, 3.0])
    h = oracle.from_coo(coo, HDC_FORMAT)
    assert h.dia.ndiag == 7 and h.csr.nnz == 0
    assert list(h.dia.offsets) == [-3, -2, -1, 0, 1, 2, 3]


def test_pad_size():
    L = oracle.lib()
    for n, a in [(0, 32), (1, 32), (32, 32), (33, 32), (65_536, 32), (7, 4)]:
        assert L.orc_pad_size(n, a) == pad_size(n, a) == a * ((n + a - 1) // a)


@pytest.mark.parametrize("fmt", ["csr", "coo", "dia", "ell"])
def test_openmp_baseline_matches_serial(fmt):
    """The OpenMP restatement timed as the CPU baseline
    (core/src/impl/{Csr,Coo,Dia,Ell}/OpenMP/Morpheus_Multiply_Impl.hpp)
    computes the same product as the Serial loops, within the FP64 bound
    (COO boundary rows are combined with atomics, so their order differs)."""
    from morpheus_amd import generators as gen
    if fmt in ("csr", "coo"):
        A = gen.powerlaw_rows(4000, 3000, 1.5, 10.0, 2000, seed=5)
        if fmt == "coo":
            A = oracle.to_coo(A)
    else:
        A = oracle.from_coo(oracle.to_coo(gen.stencil3d_7(12)), FMT[fmt.upper()])
    run, x, y = oracle.omp_spmv_fn(A)
    run()
    ref = oracle.spmv(A, x)
    csr = oracle.convert(A, FMT["CSR"])
    bound = fp64_bound(np.diff(csr.row_offsets), oracle.absrow(csr, x))
    assert np.all(np.abs(y - ref) <= bound)
