"""The oracle's dense-vector algorithms and its CG, pinned without the
reference build: dot against math.fsum (exact) within the textbook bound,
waxpby bit for bit against numpy (no FMA in either), the Kahan reduce
against math.fsum, and CG by its closed forms and convergence."""
import math

import numpy as np
import pytest

import oracle
from morpheus_amd import generators as gen

U = 2.0 ** -53


@pytest.mark.parametrize("n", [1, 7, 1000, 100_000])
def test_dot_within_bound(n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    exact = math.fsum(x * y)  # exact sum of the rounded products
    assert abs(oracle.dot(x, y) - exact) <= 2 * n * U * np.sum(np.abs(x * y)) + 1e-300
    assert oracle.dot(x) == oracle.dot(x, x.copy())


def test_waxpby_bitwise_and_special_cases():
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(5000), rng.standard_normal(5000)
    for a, b in [(2.5, -0.75), (1.0, 3.0), (-4.0, 1.0), (0.0, 0.0)]:
        assert np.array_equal(oracle.waxpby(a, x, b, y), a * x + b * y)


@pytest.mark.parametrize("n", [1, 10, 12345])
def test_kahan_reduce(n):
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    exact = math.fsum(v)
    assert abs(oracle.reduce(v) - exact) <= 2 * U * abs(exact) + 4 * n * U * U * np.sum(np.abs(v))


def test_cg_scaled_identity_converges_in_one_step():
    I = gen.stencil3d_7(1)  # the 1x1 matrix [6]
    x, rr = oracle.cg(I, np.array([3.0]), np.zeros(1), 1)
    assert x[0] == 0.5 and rr == 0.0


def test_cg_converges_on_laplacian():
    A = gen.stencil3d_7(12)          # SPD 3-D Laplacian, 1728 rows
    rng = np.random.default_rng(5)
    b = rng.standard_normal(A.nrows)
    x, rr = oracle.cg(A, b, np.zeros(A.nrows), 200)
    res = b - oracle.spmv(A, x)
    assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(b)
    # the recurrence residual tracks the true one
    assert math.sqrt(rr) <= 1e-8 * np.linalg.norm(b)


def test_cg_first_step_closed_form():
    A = gen.stencil3d_7(6)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(A.nrows)
    x, _ = oracle.cg(A, b, np.zeros(A.nrows), 1)
    Ab = oracle.spmv(A, b)
    alpha = math.fsum(b * b) / math.fsum(b * Ab)
    assert np.allclose(x, alpha * b, rtol=1e-13, atol=0)
