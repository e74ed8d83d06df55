"""GPU parity of the matrix analytics (core/src/Morpheus_MatrixAnalytics.hpp
:112-261) through the C-ABI, against the oracle, bit for bit: the counts
are integers.  Mirrors core/tests/Test_MatrixAnalytics.hpp (the diagonal
and anti-diagonal generators at 10 / 100 / 1000 rows, init and accumulate,
min / max / std, count_diagonals) plus large and skewed matrices, the three
value types and the per-format quirks."""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi, device
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  FORMAT_NAMES, HDC_FORMAT, HYB_FORMAT, Coo, Dia, Ell)
from tests.helpers import anti_diagonal_matrix, reference_diags

pytestmark = pytest.mark.gpu

FORMATS = [COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT]


def _torch():
    import torch
    return torch


def _host_or_skip(A, fmt):
    try:
        return oracle.convert(A, fmt)
    except oracle.FormatConversion:
        pytest.skip("the reference refuses this conversion too")


def check_counts(H, dtype="f64"):
    torch = _torch()
    D = device.to_device(H, dtype)
    n = H.nrows
    got = device.count_nnz_per_row(D).cpu().numpy()
    assert np.array_equal(got, oracle.nnz_per_row(H)), "per row"
    acc = device.count_nnz_per_row(D, torch.ones(n, dtype=torch.int32, device="cuda"),
                                   init=False).cpu().numpy()
    assert np.array_equal(acc, oracle.nnz_per_row(H) + 1), "per row, accumulate"
    nd = H.nrows + H.ncols - 1
    got = device.count_nnz_per_diagonal(D).cpu().numpy()
    assert np.array_equal(got, oracle.nnz_per_diagonal(H)), "per diagonal"
    acc = device.count_nnz_per_diagonal(D, torch.full((nd,), 2, dtype=torch.int32, device="cuda"),
                                        init=False).cpu().numpy()
    assert np.array_equal(acc, oracle.nnz_per_diagonal(H) + 2), "per diagonal, accumulate"
    mx, mn, sd = oracle.row_stats(H)
    assert device.max_nnnz(D) == mx
    assert device.min_nnnz(D) == mn
    assert device.std_nnnz(D) == sd
    assert device.count_diagonals(D) == oracle.count_diagonals(H)
    assert device.count_true_diagonals(D, 5) == oracle.count_diagonals(H, 5)
    D.close()


@pytest.mark.parametrize("n", [10, 100, 1000])
@pytest.mark.parametrize("fmt", FORMATS, ids=FORMAT_NAMES)
def test_reference_diagonal_matrices(gpu, n, fmt):
    check_counts(_host_or_skip(gen.diagonal_test_matrix(n, seed=n), fmt))


@pytest.mark.parametrize("n", [10, 100, 1000])
@pytest.mark.parametrize("fmt", FORMATS, ids=FORMAT_NAMES)
def test_reference_anti_diagonal_matrices(gpu, n, fmt):
    check_counts(_host_or_skip(anti_diagonal_matrix(n, n, reference_diags(n), seed=n), fmt))


@pytest.mark.parametrize("fmt", [COO_FORMAT, CSR_FORMAT, HYB_FORMAT, HDC_FORMAT],
                         ids=["COO", "CSR", "HYB", "HDC"])
def test_powerlaw_and_rectangular(gpu, fmt):
    check_counts(oracle.convert(gen.powerlaw_rows(30000, 50000, 1.5, 16.0, 12000, seed=5), fmt))


@pytest.mark.parametrize("fmt", FORMATS, ids=FORMAT_NAMES)
@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_stencil_value_types(gpu, fmt, dtype):
    check_counts(oracle.convert(gen.stencil3d_7(24), fmt), dtype)


def test_dia_stored_zeros_and_padding(gpu):
    off = np.array([-1, 0, 2], np.int32)
    pitch = 8
    v = np.zeros(pitch * 3)
    for d, o in enumerate(off):
        for r in range(4):
            if 0 <= r + o < 4:
                v[r + pitch * d] = 1.0 + r + d
    v[1 + pitch * 1] = 0.0
    v[5 + pitch * 1] = 7.0
    check_counts(Dia(4, 4, 8, off, v, pitch))


def test_ell_first_slots_quirk(gpu):
    ci = np.array([0, 1, 2, -1], np.int32)
    v = np.array([0.0, 5.0, 3.0, 0.0])
    check_counts(Ell(2, 3, 3, 2, 2, ci, v, 2))


def test_coo_explicit_zeros(gpu):
    check_counts(Coo(3, 3, np.array([0, 1, 2]), np.array([0, 2, 1]), np.array([0.0, 1.0, 0.0])))


def test_large_stencil(gpu):
    check_counts(oracle.convert(gen.stencil3d_7(96), CSR_FORMAT))


def test_i32_stats_large(gpu):
    torch = _torch()
    rng = np.random.default_rng(3)
    v = rng.integers(-1000, 5000, 3_000_001).astype(np.int32)
    t = torch.from_numpy(v).cuda()
    mx, mn, sq, c = device.i32_stats(t, 123.5, 40)
    omx, omn, osq, oc = oracle.i32_stats(v, 123.5, 40)
    assert (mx, mn, c) == (omx, omn, oc)
    assert sq == pytest.approx(osq, rel=1e-12)
    # a fixed reduction order: the same bits every time
    assert device.i32_stats(t, 123.5, 40)[2] == sq


def test_out_of_range_indices_are_refused(gpu):
    torch = _torch()
    D = device.to_device(Coo(3, 3, np.array([0, 1, 2]), np.array([0, 2, 1]),
                             np.array([1.0, 1.0, 1.0])), "f64")
    D.arrays["ci"][1] = 7
    torch.cuda.synchronize()
    with pytest.raises(capi.InvalidInputException):
        device.count_nnz_per_diagonal(D)
    D.close()
