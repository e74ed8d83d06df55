"""A HYB's own storage at scale: the split plan's ELL pass bounded per row.

Reference: core/src/impl/Hyb/HIP/Morpheus_Multiply_Impl.hpp:42-53 (an ELL
pass that writes y, then the COO part accumulated onto it) and the Serial
ELL loop core/src/impl/Ell/Serial/Morpheus_Multiply_Impl.hpp, which skips
padding (-1) slots.

With the policy MPH_POLICY_NO_VALUE_COPIES a large HYB keeps no merged CSR
copy: its ELL part runs as its own pass and the COO tail is added over the
rows that have one.  When the ELL slots are mostly padding, the analysis
records per row 1 + its last filled slot and the ELL pass loads no slot past
it.  These tests run such HYBs -- slots filled in any position, rows without
any ELL entry, explicit zeros, a 3000-entry tail row -- in every value type,
initialising and accumulating, k = 1 and k > 1, against the Serial HYB
product (ELL pass, then COO) of the oracle.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi
from morpheus_amd import generators as gen
from morpheus_amd.formats import Coo, Ell, Hyb
from tests.helpers import fp64_bound, max_rel_err
from tests.test_multiply_gpu import holey_hyb

pytestmark = pytest.mark.gpu


def _abs_hyb(H):
    e, c = H.ell, H.coo
    return Hyb(Ell(e.nrows, e.ncols, e.nnz, e.entries_per_row, e.width, e.column_indices,
                   np.abs(e.values), e.pitch),
               Coo(c.nrows, c.ncols, c.row_indices, c.column_indices, np.abs(c.values)))


def _nnz_row(H):
    e, c = H.ell, H.coo
    filled = (e.column_indices.reshape(e.width, e.pitch)[:, :e.nrows] >= 0).sum(axis=0)
    return filled + np.bincount(c.row_indices, minlength=H.nrows)


def _run(H, dtype, X, y0, init, xtile=0):
    import torch
    from morpheus_amd import device
    D = device.to_device(H, dtype)
    D.set_policy(capi.POLICY_NO_VALUE_COPIES)
    if xtile:
        D.set_tuning(xtile=xtile)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(np.ascontiguousarray(X)).cuda().to(vdt)
    yt = torch.from_numpy(np.ascontiguousarray(y0)).cuda().to(vdt)
    outs = []
    for _ in range(2):
        y = yt.clone()
        if X.ndim == 1:
            D.multiply(xt, y, init)
        else:
            D.multiply_multi(xt, y, init)
        torch.cuda.synchronize()
        outs.append(y.cpu().numpy().astype(np.float64))
    D.close()
    return outs


@pytest.mark.parametrize("fill", [0.05, 0.2, 0.8])
@pytest.mark.parametrize("init", [True, False])
def test_split_hyb_fp64(gpu, fill, init):
    """70k rows: fill 0.05 / 0.2 visit per-row counts (at most half the
    slots), 0.8 visits every slot; each within the per-row FP64 bound and
    the same bits on a second call."""
    H = holey_hyb(70000, fill, 11)
    x = gen.random_vector(H.ncols, 21)
    y0 = gen.random_vector(H.nrows, 22)
    y, y2 = _run(H, "f64", x, y0, init)
    assert np.array_equal(y, y2)
    ref = oracle.spmv(H, x, y0, init)
    bound = fp64_bound(_nnz_row(H), oracle.spmv(_abs_hyb(H), np.abs(x)), None if init else y0)
    bad = np.count_nonzero(~(np.abs(y - ref) <= bound))
    assert bad == 0, f"{bad} rows outside the FP64 bound"


def test_split_hyb_small_xtiled(gpu):
    """5,000 rows with x-tiling forced on also take the split plan (and its
    per-row counts)."""
    H = holey_hyb(5000, 0.1, 12, tail_every=3)
    x = gen.random_vector(H.ncols, 23)
    y0 = np.zeros(H.nrows)
    y, _ = _run(H, "f64", x, y0, True, xtile=2)
    ref = oracle.spmv(H, x)
    bound = fp64_bound(_nnz_row(H), oracle.spmv(_abs_hyb(H), np.abs(x)))
    assert np.all(np.abs(y - ref) <= bound)


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("fill", [0.05, 0.8])
def test_split_hyb_reduced_precision(gpu, dtype, tol, fill):
    """FP32 / BF16 storage: the counted one-row ELL kernel replaces the row
    pair kernel; max_i |y_i - y_ref_i| / sum_j |a_ij x_j| against the FP64
    product of the original matrix."""
    H = holey_hyb(70000, fill, 13)
    x = gen.random_vector(H.ncols, 24)
    y, y2 = _run(H, dtype, x, np.zeros(H.nrows), True)
    assert np.array_equal(y, y2)
    err = max_rel_err(y, oracle.spmv(H, x), oracle.spmv(_abs_hyb(H), np.abs(x)))
    assert err <= tol, err


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("init", [True, False])
def test_split_hyb_spmm(gpu, k, init):
    """k right-hand sides over the counted split plan: every column within
    the FP64 bound of the single-vector Serial product."""
    H = holey_hyb(70000, 0.1, 14)
    rng = np.random.default_rng(25)
    X = rng.uniform(-1, 1, (H.ncols, k))
    Y0 = rng.uniform(-1, 1, (H.nrows, k))
    Y, _ = _run(H, "f64", X, Y0, init)
    absy = _abs_hyb(H)
    for j in range(k):
        ref = oracle.spmv(H, X[:, j], Y0[:, j], init)
        bound = fp64_bound(_nnz_row(H), oracle.spmv(absy, np.abs(X[:, j])),
                           None if init else Y0[:, j])
        assert np.all(np.abs(Y[:, j] - ref) <= bound), j
