"""On-GPU Convert, bit-exact against the CPU oracle (reference Serial
converters) and the reference's hand-written fixture in all six formats.

The reference HIP backend throws NotImplementedException for every pair
(core/tests/HIP/Test_Convert.hpp:164-165,178-179); the library converts on the
GPU instead, so parity is against the Serial output
(core/tests/Serial/Test_Convert.hpp, core/tests/Serial/Test_Convert_Dynamic.hpp)."""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  HDC_FORMAT, HYB_FORMAT, Coo, Csr, Dia, Ell, Hdc, Hyb,
                                  csr_to_coo_host)
from tests.helpers import same_container, small_fixture

pytestmark = pytest.mark.gpu

FMT = {"COO": COO_FORMAT, "CSR": CSR_FORMAT, "DIA": DIA_FORMAT, "ELL": ELL_FORMAT,
       "HYB": HYB_FORMAT, "HDC": HDC_FORMAT}
ALL = list(FMT)


def dev():
    from morpheus_amd import device
    return device


def gpu_convert(Ah, fmt, dtype="f64"):
    import torch
    D = dev().to_device(Ah, dtype)
    out = dev().convert(D, fmt)
    torch.cuda.synchronize()
    return out.to_host()


def round_values(A, dtype):
    """Host container with values rounded to the device storage type."""
    import torch
    if dtype == "f64":
        return A

    def r(v):
        t = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64))
        t = t.to(torch.float32) if dtype == "f32" else t.to(torch.bfloat16)
        return t.to(torch.float64).numpy()
    if isinstance(A, Hyb):
        return Hyb(round_values(A.ell, dtype), round_values(A.coo, dtype))
    if isinstance(A, Hdc):
        return Hdc(round_values(A.dia, dtype), round_values(A.csr, dtype))
    import copy
    B = copy.deepcopy(A)
    B.values = r(B.values)
    return B


# ------------------------------------------------------------ reference fixture
@pytest.mark.parametrize("src", ALL)
@pytest.mark.parametrize("dst", ALL)
def test_fixture_any_to_any(gpu, src, dst):
    fx = small_fixture()
    got = gpu_convert(fx[src], FMT[dst])
    assert same_container(got, fx[dst]), (src, dst)


# ------------------------------------------------------------ workloads
@pytest.fixture(scope="module")
def mats():
    return {
        "p2": gen.P2(),
        "s7": gen.stencil3d_7(24),
        "pl": gen.powerlaw_rows(20000, 20000, 1.5, 12.0, 3000, seed=21),
        "band": gen.banded(30000, 4, 0.02, seed=22),
        "diag": gen.diagonal_test_matrix(1000, seed=5),
    }


@pytest.mark.parametrize("name", ["p2", "s7", "pl", "band", "diag"])
@pytest.mark.parametrize("fmt", ["COO", "DIA", "ELL", "HYB", "HDC", "CSR"])
@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_csr_to_format_and_back(gpu, mats, name, fmt, dtype):
    A = round_values(mats[name], dtype)
    try:
        ref = oracle.convert(A, FMT[fmt])
    except oracle.FormatConversion:
        with pytest.raises(capi.FormatConversionException):
            gpu_convert(A, FMT[fmt], dtype)
        return
    got = gpu_convert(A, FMT[fmt], dtype)
    assert same_container(got, ref)
    back = gpu_convert(ref, CSR_FORMAT, dtype)
    assert same_container(back, oracle.convert(ref, CSR_FORMAT))


def test_p2_roundtrip_csr_coo_dia_csr(gpu, mats):
    """P2 (2D 5-point Poisson 256^2): CSR -> COO -> DIA -> CSR bit-exact, on the GPU."""
    import torch
    device = dev()
    A = mats["p2"]
    D = device.to_device(A)
    coo = device.convert(D, COO_FORMAT)
    dia = device.convert(coo, DIA_FORMAT)
    back = device.convert(dia, CSR_FORMAT)
    torch.cuda.synchronize()
    assert list(dia.to_host().offsets) == [-256, -1, 0, 1, 256]
    assert same_container(back.to_host(), A)


# ------------------------------------------------------------ reference edge semantics
def _coo(n, m, r, c, v):
    return Coo(n, m, np.asarray(r), np.asarray(c), np.asarray(v, dtype=float))


def test_unsorted_coo_to_csr_is_stable(gpu):
    rng = np.random.default_rng(1)
    n = 3000
    r = rng.integers(0, 500, n)
    c = rng.integers(0, 400, n)
    A = _coo(500, 400, r, c, np.arange(n, dtype=float))
    assert same_container(gpu_convert(A, CSR_FORMAT), oracle.from_coo(A, CSR_FORMAT))


def test_sort_by_row_and_column(gpu):
    import ctypes as C
    import torch
    rng = np.random.default_rng(2)
    n = 50000
    r = rng.integers(0, 3000, n).astype(np.int32)
    c = rng.integers(0, 70000, n).astype(np.int32)
    v = np.arange(n, dtype=float)
    ref = oracle.sort_by_row_and_column(_coo(3000, 70000, r, c, v))
    tr, tc, tv = (torch.from_numpy(a).cuda() for a in (r, c, v))
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    flag = C.c_int(7)
    capi.call("mph_coo_is_sorted", n, p(tr), p(tc), C.byref(flag), s)
    assert flag.value == 0
    capi.call("mph_coo_sort_by_row_and_column", n, p(tr), p(tc), p(tv), capi.F64, s)
    capi.call("mph_coo_is_sorted", n, p(tr), p(tc), C.byref(flag), s)
    assert flag.value == 1
    assert np.array_equal(tr.cpu().numpy(), ref.row_indices)
    assert np.array_equal(tc.cpu().numpy(), ref.column_indices)
    assert np.array_equal(tv.cpu().numpy(), ref.values)


def test_repeated_entries_later_wins(gpu):
    """Repeated (row, col): the Serial scatter keeps the later entry."""
    A = _coo(6, 6, [0, 0, 2, 2, 2, 5], [1, 1, 2, 2, 2, 0], [1, 2, 3, 4, 5, 6])
    for f in (DIA_FORMAT, HDC_FORMAT):
        assert same_container(gpu_convert(A, f), oracle.from_coo(A, f))


def test_rows_in_several_runs_later_run_wins(gpu):
    """COO -> ELL places entries by their position in their run of equal
    rows; a row met again in a later run overwrites (Ell/Serial:160-172)."""
    A = _coo(4, 5, [1, 1, 0, 1, 3], [0, 1, 4, 3, 2], [1, 2, 3, 4, 5])
    assert same_container(gpu_convert(A, ELL_FORMAT), oracle.from_coo(A, ELL_FORMAT))


def test_dia_stored_zeros_dropped_with_phantom_slots(gpu):
    pitch = 32
    vals = np.zeros(pitch * 2)
    vals[0], vals[1], vals[2] = 1.5, -0.0, 2.5
    vals[pitch + 0] = 7.0
    A = Dia(3, 3, 5, np.array([0, 1]), vals, pitch)
    assert same_container(gpu_convert(A, COO_FORMAT), oracle.to_coo(A))


def test_ell_explicit_zeros_kept(gpu):
    ci = np.full(32 * 3, -1, dtype=np.int32)
    v = np.zeros(32 * 3)
    ci[0], v[0] = 2, 0.0
    ci[1], v[1] = 0, 3.0
    ci[3 + 1], v[3 + 1] = 1, -2.0
    A = Ell(3, 3, 3, 2, 32, ci, v, 3)
    assert same_container(gpu_convert(A, COO_FORMAT), oracle.to_coo(A))


def test_tolerance_exceptions(gpu):
    """Z15-shaped: a 20k-entry row among 2M rows makes ELL and DIA
    inadmissible (FormatConversionException), HYB admissible."""
    n = 2_000_000
    rows = np.concatenate([np.arange(n), np.zeros(19_999, dtype=np.int64)])
    cols = np.concatenate([np.zeros(n, dtype=np.int64), np.arange(1, 20_000)])
    order = np.lexsort((cols, rows))
    A = _coo(n, n, rows[order], cols[order], np.ones(rows.size))
    with pytest.raises(capi.FormatConversionException):
        gpu_convert(A, ELL_FORMAT)
    with pytest.raises(capi.FormatConversionException):
        gpu_convert(A, DIA_FORMAT)
    got = gpu_convert(A, HYB_FORMAT)
    assert same_container(got, oracle.from_coo(A, HYB_FORMAT))


def test_count_nnz_per_diagonal(gpu, mats):
    import ctypes as C
    import torch
    A = csr_to_coo_host(mats["pl"])
    ri, ci = torch.from_numpy(A.row_indices).cuda(), torch.from_numpy(A.column_indices).cuda()
    out = torch.full((A.nrows + A.ncols - 1,), 5, dtype=torch.int32, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    capi.call("mph_coo_count_nnz_per_diagonal", A.nrows, A.ncols, A.nnz,
              C.c_void_p(ri.data_ptr()), C.c_void_p(ci.data_ptr()), C.c_void_p(out.data_ptr()),
              1, s)
    assert np.array_equal(out.cpu().numpy(), oracle.count_nnz_per_diagonal(A))


@pytest.mark.parametrize("seed", range(12))
def test_random_shapes_all_pairs(gpu, seed):
    rng = np.random.default_rng(100 + seed)
    n, m = int(rng.integers(1, 40)), int(rng.integers(1, 40))
    D = np.where(rng.random((n, m)) < 0.25, rng.standard_normal((n, m)), 0.0)
    r, c = np.nonzero(D)
    base = _coo(n, m, r, c, D[r, c])
    for f in ALL:
        try:
            ref = oracle.from_coo(base, FMT[f])
        except oracle.FormatConversion:
            continue
        except oracle.OracleError as e:
            assert f == "HDC" and n != m and e.code == oracle.INVALID_INPUT
            with pytest.raises(capi.InvalidInputException):
                gpu_convert(base, FMT[f])
            continue
        got = gpu_convert(base, FMT[f])
        assert same_container(got, ref), f
        assert same_container(gpu_convert(ref, COO_FORMAT), oracle.to_coo(ref)), f


def test_hdc_fixture_threshold_cases(gpu):
    coo = _coo(4, 4, [0, 1, 3], [0, 3, 0], [1.0, 2.0, 3.0])
    got = gpu_convert(coo, HDC_FORMAT)
    ref = oracle.from_coo(coo, HDC_FORMAT)
    assert same_container(got, ref)
    assert got.dia.ndiag == 7
