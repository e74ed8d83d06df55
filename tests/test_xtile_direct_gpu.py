"""The x-tiled pass without a value copy (MPH_POLICY_NO_VALUE_COPIES).

Reference: core/src/impl/Csr/Serial/Morpheus_Multiply_Impl.hpp:51-57 (one sum
per row), and the containers' shallow value arrays
(core/src/Morpheus_CsrMatrix.hpp:148-163), which the policy keeps as the
only source of values.

Under the policy the x-tiled plan copies the long rows' columns (structure)
but not their values: the pass reads each value from the caller's array at
its run's position, its lanes taking the items of a sub-block in stride
order.  It sums in another order than the plan with the value copy, so each
product is held to the FP64 per-row bound (FP32 / BF16: the compensated
oracle within 1e-5 / 1e-2) and to the same bits on a second call.
Matrices: wide power-law rows (x-tiled by default), runs that fill a
sub-block, rows dense enough to halve the tile, tiled rows at both ends;
CSR, row-sorted COO, HYB and HDC; k = 1 and the per-column SpMM pass.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi
from morpheus_amd import generators as gen
from morpheus_amd.formats import Csr
from tests.helpers import fp64_bound, max_rel_err
from tests.test_multiply_gpu import as_format

pytestmark = pytest.mark.gpu


def _wide():
    return gen.powerlaw_rows(20000, 400000, 1.5, 40.0, 30000, seed=31)


def _full_runs():
    # runs of 512 items in one tile (a whole sub-block) and rows of 450 items
    # per tile over many tiles
    rng = np.random.default_rng(12)
    m, n = 300000, 400
    rows = []
    for r in range(n):
        if r == 5:
            cols = np.concatenate([np.arange(512), 8192 + rng.choice(200000, 100, replace=False)])
        elif r % 50 == 7:
            cols = np.concatenate([t * 4096 + rng.choice(4096, 450, replace=False) for t in range(0, 70, 3)])
        else:
            cols = rng.choice(m, int(rng.integers(0, 12)), replace=False)
        rows.append(np.unique(cols))
    ro = np.concatenate([[0], np.cumsum([c.size for c in rows])]).astype(np.int32)
    return Csr(n, m, ro, np.concatenate(rows).astype(np.int32), rng.standard_normal(int(ro[-1])))


def _dense_rows():
    # 1,500 items in each of 15 windows of 4,096 columns: the tile is halved
    rng = np.random.default_rng(21)
    m, n = 300000, 600
    rows = []
    for r in range(n):
        if r % 15 == 3:
            cols = np.concatenate([t * 4096 + rng.choice(4096, 1500, replace=False) for t in range(0, 60, 4)])
        else:
            cols = rng.choice(m, int(rng.integers(0, 30)), replace=False)
        rows.append(np.unique(cols))
    ro = np.concatenate([[0], np.cumsum([c.size for c in rows])]).astype(np.int32)
    return Csr(n, m, ro, np.concatenate(rows).astype(np.int32), rng.standard_normal(int(ro[-1])))


def _ends():
    rng = np.random.default_rng(3)
    n, m = 3000, 300000
    lens = np.zeros(n, dtype=np.int64)
    lens[[0, 1, 1500, n - 1]] = [5000, 3000, 4000, 2500]
    lens[rng.choice(np.arange(10, n - 10), 800, replace=False)] = rng.integers(1, 20, 800)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=ro[1:])
    ci = np.concatenate([np.sort(rng.choice(m, L, replace=False)) for L in lens if L > 0])
    return Csr(n, m, ro, ci, rng.standard_normal(ro[-1]))


MATS = {"wide": _wide, "full runs": _full_runs, "dense rows": _dense_rows, "ends": _ends}


def _run(H, dtype, X, y0, init, policy, xtile=0):
    import torch
    from morpheus_amd import device
    D = device.to_device(H, dtype)
    D.set_policy(policy)
    if xtile:
        D.set_tuning(xtile=xtile)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(np.ascontiguousarray(X)).cuda().to(vdt)
    y = torch.from_numpy(np.ascontiguousarray(y0)).cuda().to(vdt)
    if X.ndim == 1:
        D.multiply(xt, y, init)
    else:
        D.multiply_multi(xt, y, init)
    torch.cuda.synchronize()
    out = y.cpu().numpy().astype(np.float64)
    D.close()
    return out


@pytest.mark.parametrize("mat", list(MATS))
@pytest.mark.parametrize("name", ["CSR", "COO", "HYB", "HDC"])
@pytest.mark.parametrize("init", [True, False])
def test_direct_fp64(gpu, mat, name, init):
    A = MATS[mat]()
    H = as_format(A, name)
    x = gen.random_vector(A.ncols, 41)
    y0 = gen.random_vector(A.nrows, 42)
    yd = _run(H, "f64", x, y0, init, capi.POLICY_NO_VALUE_COPIES, xtile=2)
    ref = oracle.spmv(A, x, y0, init)
    bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, x), None if init else y0)
    bad = np.count_nonzero(~(np.abs(yd - ref) <= bound))
    assert bad == 0, f"{bad} rows outside the FP64 bound"
    yd2 = _run(H, "f64", x, y0, init, capi.POLICY_NO_VALUE_COPIES, xtile=2)
    assert np.array_equal(yd, yd2)


@pytest.mark.parametrize("mat", ["wide", "dense rows"])
@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
def test_direct_reduced_precision(gpu, mat, dtype, tol):
    """FP32 / BF16: the direct pass (Dot2 pairs through the scans) within the
    tolerance of the compensated oracle."""
    A = MATS[mat]()
    x = gen.random_vector(A.ncols, 43)
    y0 = np.zeros(A.nrows)
    yd = _run(A, dtype, x, y0, True, capi.POLICY_NO_VALUE_COPIES)
    ref = oracle.spmv_compensated(A, x)
    assert max_rel_err(yd, ref, oracle.absrow(A, x)) <= tol


@pytest.mark.parametrize("k", [2, 4])
def test_direct_spmm_columns(gpu, k):
    """SpMM on the wide matrix: the per-column x-tiled pass over the direct
    plan, every column within the FP64 bound."""
    A = _wide()
    X = np.random.default_rng(44).uniform(-1, 1, (A.ncols, k))
    Y0 = np.zeros((A.nrows, k))
    Yd = _run(A, "f64", X, Y0, True, capi.POLICY_NO_VALUE_COPIES)
    for j in range(k):
        ref = oracle.spmv(A, X[:, j])
        bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, X[:, j]))
        assert np.all(np.abs(Yd[:, j] - ref) <= bound)
