"""Parity at the full sizes of the larger workloads.

* Precision sweep: the 3D 7-point stencil (128^3), the Zipf(1.5) matrix
  (1M rows) and the banded matrix (4M rows, bandwidth 9, 2 % off-band), every
  format the reference admits for the matrix.  FP64 (the drop-in multiply):
  every row within |y - y_ref| <= 4 n_row u sum_j |a_ij x_j| of the Serial
  restatement.  FP32 and BF16 storage (int32 indices, FP32 accumulation,
  Dot2 compensation on rows longer than 1k): the max-row relative error
  max_i |y_i - y_ref_i| / sum_j |a_ij x_j| against the compensated FP64
  oracle on the original FP64 matrix and x stays within 1e-5 (FP32) / 1e-2
  (BF16).
* Large DynamicMatrix switching: the 3D 27-point stencil (160^3, 109M nnz) and
  the 2M-row Zipf(1.2) matrix (mean 30, max 200k): on-GPU Convert from CSR to
  DIA / ELL / HYB / HDC is bit-exact (against the Serial restatement, and by
  converting back to CSR), formats the reference refuses are refused the same
  way, the autotuner picks a format per matrix, and its SpMV with k in
  {1, 4, 16} right-hand sides in FP64 / FP32 / BF16 is checked against the FP64
  oracle (the FP64 bound, or the FP32 / BF16 tolerances).

x is drawn from a seeded generator; all sizes are the workloads' own.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi, device
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  FORMAT_NAMES, HDC_FORMAT, HYB_FORMAT)
from tests.helpers import fp64_bound, max_rel_err, same_container

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "bf16": 1e-2}


def _torch():
    import torch
    return torch


def _spmv(D, x, k=1):
    torch = _torch()
    vdt = device.vector_dtype(D.dtype)
    xt = torch.from_numpy(np.ascontiguousarray(x)).cuda().to(vdt)
    if k == 1:
        y = torch.full((D.nrows,), float("nan"), dtype=vdt, device="cuda")
        D.multiply(xt, y)
    else:
        y = torch.full((D.nrows, k), float("nan"), dtype=vdt, device="cuda")
        D.multiply_multi(xt, y)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


# ------------------------------------------------------------ precision sweep
@pytest.fixture(scope="module", params=["S7", "Z15", "B9"])
def sweep_matrix(request):
    return request.param, getattr(gen, request.param)()


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_precision_sweep(gpu, sweep_matrix, dtype):
    name, A = sweep_matrix
    x = gen.random_vector(A.ncols, 21)
    yref = oracle.spmv_compensated(A, x) if dtype != "f64" else oracle.spmv(A, x)
    scale = oracle.absrow(A, x)
    nrow = np.diff(A.row_offsets)
    base = device.to_device(A, dtype)
    checked = []
    for f, fname in enumerate(FORMAT_NAMES):
        try:
            D = base if f == CSR_FORMAT else device.convert(base, f)
        except capi.FormatConversionException:
            # the reference refuses the same conversions (fill tolerance)
            assert f in (DIA_FORMAT, ELL_FORMAT), f"{name}: {fname} refused"
            continue
        y = _spmv(D, x)
        err = max_rel_err(y, yref, scale)
        print(f"{name} {fname} {dtype}: max rel err {err:.3e}", flush=True)
        if dtype == "f64":
            # the drop-in bound, per row: 4 n_row u sum_j |a_ij x_j|
            bad = np.count_nonzero(~(np.abs(y - yref) <= fp64_bound(nrow, scale)))
            assert bad == 0, f"{name} {fname}: {bad} rows outside the FP64 bound"
        else:
            assert err <= TOL[dtype], f"{name} {fname} {dtype}: {err:.3e}"
        checked.append(fname)
        if D is not base:
            D.close()
    base.close()
    assert {"CSR", "COO", "HYB", "HDC"} <= set(checked)


# ------------------------------------------------------------ large switching
@pytest.fixture(scope="module", params=["S27", "Z12"])
def large(request):
    A = getattr(gen, request.param)()
    return request.param, A


def test_large_convert_bit_exact(gpu, large):
    name, A = large
    D = device.to_device(A, "f64")
    admitted = []
    for f in (DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT):
        try:
            M = device.convert(D, f)
        except capi.FormatConversionException:
            # the Serial restatement refuses it too
            with pytest.raises(oracle.FormatConversion):
                oracle.from_coo(oracle.to_coo(A), f)
            continue
        admitted.append(FORMAT_NAMES[f])
        # back to CSR: identity (no stored zeros, sorted columns)
        back = device.convert(M, CSR_FORMAT).to_host()
        assert same_container(back, A), f"{name}: {FORMAT_NAMES[f]} -> CSR differs"
        M.close()
    # one conversion per matrix against the Serial restatement, bit for bit
    f = DIA_FORMAT if name == "S27" else HYB_FORMAT
    M = device.convert(D, f)
    assert same_container(M.to_host(), oracle.from_coo(oracle.to_coo(A), f))
    M.close()
    D.close()
    if name == "S27":
        assert set(admitted) == {"DIA", "ELL", "HYB", "HDC"}
    else:
        # Zipf(1.2) rows up to 200k: ELL and DIA exceed the fill tolerance
        assert "HYB" in admitted and "ELL" not in admitted and "DIA" not in admitted


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_large_autotuned_spmm(gpu, large, dtype):
    name, A = large
    D = device.to_device(A, dtype)
    best, times = device.autotune_format(D, k=4, reps=2)
    M = D if best == D.format else device.convert(D, best)
    print(f"{name} {dtype}: autotuned format {FORMAT_NAMES[best]} "
          f"({', '.join(f'{FORMAT_NAMES[i]} {t:.3f} ms' for i, t in times.items())})",
          flush=True)
    rng = np.random.default_rng(31)
    nrow = np.diff(A.row_offsets)
    for k in (1, 4, 16):
        X = rng.uniform(-1.0, 1.0, (A.ncols, k)) if k > 1 else rng.uniform(-1.0, 1.0, A.ncols)
        Y = _spmv(M, X, k)
        cols = [0] if k == 1 else sorted({0, k // 2, k - 1})
        for j in cols:
            xj = X if k == 1 else np.ascontiguousarray(X[:, j])
            yj = Y if k == 1 else Y[:, j]
            scale = oracle.absrow(A, xj)
            if dtype == "f64":
                ref = oracle.spmv(A, xj)
                bad = np.count_nonzero(~(np.abs(yj - ref) <= fp64_bound(nrow, scale)))
                assert bad == 0, f"{name} k={k} col {j}: {bad} rows outside the FP64 bound"
            else:
                err = max_rel_err(yj, oracle.spmv_compensated(A, xj), scale)
                assert err <= TOL[dtype], f"{name} {dtype} k={k} col {j}: {err:.3e}"
    if M is not D:
        M.close()
    D.close()
