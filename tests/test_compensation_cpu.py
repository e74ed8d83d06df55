"""Pins of the oracle's accuracy checkers, and of the FP32 compensation check.

* ``absrow`` (sum_j |a_ij x_j|, the scale of every FP64 bound and of every
  FP32 / BF16 relative error) equals scipy's |A| @ |x| exactly on
  integer-valued data and to rounding elsewhere, so neither a missing abs()
  nor a scale that is too large or too small passes;
* the compensated FP64 row sum (the reference of the reduced-precision
  checks) is strictly closer to the exact sum than the plain Serial sum
  whenever the plain sum rounds, and within one ulp of the exact sum -- a
  compensated sum that dropped its error term would equal the plain sum and
  fail;
* on the cancellation fixture (rows of 1.5k-60k entries, sum |a x| about
  2000 |sum a x|, inputs exact in FP32 or BF16) the bound
  |y - y_exact| <= 4 u |y_exact| + n u^2 sum |a x| (u = 2^-24) that the GPU
  tests apply rejects every plain FP32 accumulation order (serial; 64 lanes
  then a tree) and TwoSum on rounded products, and admits Dot2 -- so a GPU
  kernel that drops any part of its compensation fails the GPU test.
The FP32 witnesses are ``orc_csr_spmv_f32`` in oracle/mph_oracle.c.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import Csr
from tests.helpers import cancellation_csr, compensated_bound


def _scipy(A):
    return sp.csr_matrix((A.values, A.column_indices, A.row_offsets), shape=(A.nrows, A.ncols))


def _exact_rows(A, x, rows):
    """Exact row sums (rational arithmetic) of the listed rows."""
    out = []
    for r in rows:
        b, e = A.row_offsets[r], A.row_offsets[r + 1]
        out.append(sum(Fraction(float(a)) * Fraction(float(x[c]))
                       for a, c in zip(A.values[b:e], A.column_indices[b:e])))
    return out


# ------------------------------------------------------------------ absrow
@pytest.mark.parametrize("seed", [1, 2])
def test_absrow_exact_on_integer_data(seed):
    rng = np.random.default_rng(seed)
    A = gen.powerlaw_rows(3000, 2500, 1.5, 8.0, 1500, seed)
    v = rng.integers(-9, 10, A.nnz).astype(np.float64)
    v[v == 0] = 3.0
    A = Csr(A.nrows, A.ncols, A.row_offsets, A.column_indices, v)
    x = rng.integers(-5, 6, A.ncols).astype(np.float64)
    expect = abs(_scipy(A)) @ np.abs(x)  # integers below 2^53: exact
    got = oracle.absrow(A, x)
    assert np.array_equal(got, expect)
    # the scale is not |sum a x| and not sum a x: rows with both signs differ
    y = oracle.spmv(A, x)
    assert np.all(got >= np.abs(y)) and np.any(got > np.abs(y))


def test_absrow_float_data_to_rounding():
    A = gen.powerlaw_rows(2000, 2000, 1.5, 10.0, 1000, 4)
    x = gen.random_vector(A.ncols, 4)
    expect = abs(_scipy(A)) @ np.abs(x)
    got = oracle.absrow(A, x)
    n = np.diff(A.row_offsets)
    assert np.all(np.abs(got - expect) <= 2.0 * n * 2.0 ** -53 * expect)


# ------------------------------------------------------- compensated FP64 sum
def test_compensated_strictly_closer_and_within_one_ulp():
    rng = np.random.default_rng(7)
    n = 20000
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    A = Csr(1, n, [0, n], np.arange(n), v)
    x = np.ones(n)
    exact = sum(Fraction(float(t)) for t in v)
    ex = float(exact)
    plain = oracle.spmv(A, x)[0]
    comp = oracle.spmv_compensated(A, x)[0]
    assert abs(Fraction(plain) - exact) > 0, "the plain sum must round for this check"
    assert abs(Fraction(comp) - exact) < abs(Fraction(plain) - exact)
    assert abs(comp - ex) <= np.spacing(abs(ex))


def test_compensated_products_exact_to_one_ulp():
    """Values and x of 26 significant bits: every product is exact in FP64,
    so the compensated row sum must be within one ulp of the rational sum
    (Neumaier on exact products), on rows that cancel."""
    rng = np.random.default_rng(11)
    rows, n = 8, 4000
    ro = np.arange(0, rows * n + 1, n)
    ci = np.concatenate([rng.choice(n, n, replace=False) for _ in range(rows)])
    m = rng.integers(1 << 25, 1 << 26, rows * n) * rng.choice([-1, 1], rows * n)
    v = np.ldexp(m.astype(np.float64), rng.integers(-30, -20, rows * n))
    x = np.ldexp(rng.integers(1 << 25, 1 << 26, n).astype(np.float64), -26)
    A = Csr(rows, n, ro, ci, v)
    comp = oracle.spmv_compensated(A, x)
    plain = oracle.spmv(A, x)
    for r, e in enumerate(_exact_rows(A, x, range(rows))):
        assert abs(comp[r] - float(e)) <= np.spacing(abs(float(e)))
        assert abs(Fraction(comp[r]) - e) <= abs(Fraction(plain[r]) - e)


# ---------------------------------------------- FP32 compensation witnesses
@pytest.fixture(scope="module", params=[24, 8], ids=["fp32", "bf16"])
def cancel(request):
    A, x, long_rows = cancellation_csr(3000, 80000, [1500, 2048, 3000, 5000, 20000, 60000], 5,
                                       bits=request.param)
    return A, x, long_rows


def test_cancellation_fixture_exact_reference(cancel):
    A, x, long_rows = cancel
    # the products are exact in FP64: the compensated oracle is the exact
    # sum to one FP64 ulp (checked in rationals on the long rows)
    ref = oracle.spmv_compensated(A, x)
    for r, e in zip(long_rows, _exact_rows(A, x, long_rows)):
        assert abs(ref[r] - float(e)) <= np.spacing(abs(float(e)))
    # the cancellation is heavy: sum |a x| >= 10^3 |sum a x| on every long row
    scale = oracle.absrow(A, x)
    assert np.all(scale[long_rows] >= 1e3 * np.abs(ref[long_rows]))


@pytest.mark.parametrize("mode,name", [(oracle.F32_PLAIN, "serial"), (oracle.F32_LANES, "lanes"),
                                       (oracle.F32_TWOSUM, "twosum")])
def test_bound_rejects_uncompensated_fp32(cancel, mode, name):
    A, x, long_rows = cancel
    ref = oracle.spmv_compensated(A, x)
    bound = compensated_bound(np.diff(A.row_offsets), oracle.absrow(A, x), ref)
    y = oracle.spmv_f32(A, x, mode).astype(np.float64)
    ratio = np.abs(y[long_rows] - ref[long_rows]) / bound[long_rows]
    assert ratio.max() > 2.0, f"{name}: the FP32 bound does not reject it ({ratio.max():.2f})"


def test_bound_admits_dot2(cancel):
    A, x, long_rows = cancel
    ref = oracle.spmv_compensated(A, x)
    bound = compensated_bound(np.diff(A.row_offsets), oracle.absrow(A, x), ref)
    y = oracle.spmv_f32(A, x, oracle.F32_DOT2).astype(np.float64)
    ratio = np.abs(y[long_rows] - ref[long_rows]) / bound[long_rows]
    assert ratio.max() <= 0.5, f"Dot2 at {ratio.max():.2f} of the bound"
