"""SpMM over x-tiled rows with the k-wide pass.

Reference: core/src/impl/Csr/Serial/Morpheus_Multiply_Impl.hpp:51-57 (one sum
per row, per right-hand side).

With few right-hand sides, the x-tiled long rows of a CSR plan (with its
value copy) run through a pass that stages 4 or 2 columns of the x tile in
LDS at once, so each copied item is read once for those columns; the leftover
columns take the single-column pass.  Each column is reduced exactly as the
single-column pass reduces it, so the product must equal, bit for bit, the
same product with MPH_XTILE_K=0 (one pass per column) wherever that also
takes the x-tiled route (k <= 4 for runs of 32 items or more per tile, k <=
2 otherwise; beyond, the k-wide pass serves k <= 16 / 4 and the comparison
is with the oracle alone), and stay within the FP64 per-row bound of the
oracle.  Cases: long rows whose runs per tile average hundreds of items and
short runs, k = 2, 3, 4, 8 and 16, accumulation, FP32 and BF16, a plan
without the value copy (the per-column pass), COO and HYB.
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


def _long_runs():
    # 50,000 rows over 300,000 columns (x of 2.4 MB: x-tiled): 40 rows of
    # 8,000-30,000 sorted columns, the others 0-10 entries
    rng = np.random.default_rng(17)
    n, m = 50000, 300000
    lens = rng.integers(0, 11, n)
    lens[rng.choice(n, 40, replace=False)] = rng.integers(8000, 30000, 40)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=ro[1:])
    ci = np.concatenate([np.sort(rng.choice(m, L, replace=False)) for L in lens if L > 0])
    return Csr(n, m, ro, ci, rng.standard_normal(ro[-1]))


def _short_runs():
    return gen.powerlaw_rows(300000, 300000, 1.5, 16.0, 20000, seed=23)


MATS = {"long runs": _long_runs, "short runs": _short_runs}


def _run(monkeypatch, H, dtype, X, Y0, init, kwide, policy=0):
    import torch
    from morpheus_amd import device
    if kwide:
        monkeypatch.delenv("MPH_XTILE_K", raising=False)
    else:
        monkeypatch.setenv("MPH_XTILE_K", "0")
    D = device.to_device(H, dtype)
    if policy:
        D.set_policy(policy)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(np.ascontiguousarray(X)).cuda().to(vdt)
    y = torch.from_numpy(np.ascontiguousarray(Y0)).cuda().to(vdt)
    D.multiply_multi(xt, y, init)
    torch.cuda.synchronize()
    out = y.cpu().numpy().astype(np.float64)
    D.close()
    monkeypatch.delenv("MPH_XTILE_K", raising=False)
    return out


def test_device_lds_limit(gpu):
    """The k-wide pass needs about 148 KiB of LDS per workgroup for 4 FP64
    columns; print what the device allows (the pass falls back to 2 columns
    or to one pass per column below that)."""
    import torch
    p = torch.cuda.get_device_properties(0)
    print("LDS per block", getattr(p, "shared_memory_per_block", None),
          "opt-in", getattr(p, "shared_memory_per_block_optin", None),
          "per CU", getattr(p, "shared_memory_per_multiprocessor", None))


@pytest.mark.parametrize("mat,k,same_route", [("long runs", 2, True), ("long runs", 3, True),
                                              ("long runs", 4, True), ("long runs", 8, False),
                                              ("long runs", 16, False),
                                              ("short runs", 2, True), ("short runs", 4, False)])
@pytest.mark.parametrize("init", [True, False])
def test_kwide_equals_columns_fp64(gpu, monkeypatch, mat, k, same_route, init):
    A = MATS[mat]()
    rng = np.random.default_rng(61)
    X = rng.uniform(-1, 1, (A.ncols, k))
    Y0 = rng.uniform(-1, 1, (A.nrows, k))
    yk = _run(monkeypatch, A, "f64", X, Y0, init, True)
    if same_route:
        yc = _run(monkeypatch, A, "f64", X, Y0, init, False)
        assert np.array_equal(yk, yc), f"{np.count_nonzero(yk != yc)} entries differ"
    for j in range(k):
        y0 = None if init else Y0[:, j]
        ref = oracle.spmv(A, np.ascontiguousarray(X[:, j]), Y0[:, j], init)
        bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, np.ascontiguousarray(X[:, j])), y0)
        assert np.all(np.abs(yk[:, j] - ref) <= bound)


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("k", [2, 4])
def test_kwide_reduced_precision(gpu, monkeypatch, dtype, tol, k):
    A = _long_runs()
    rng = np.random.default_rng(62)
    X = rng.uniform(-1, 1, (A.ncols, k))
    Y0 = np.zeros((A.nrows, k))
    yk = _run(monkeypatch, A, dtype, X, Y0, True, True)
    yc = _run(monkeypatch, A, dtype, X, Y0, True, False)
    assert np.array_equal(yk, yc)
    for j in range(k):
        xj = np.ascontiguousarray(X[:, j])
        assert max_rel_err(yk[:, j], oracle.spmv_compensated(A, xj), oracle.absrow(A, xj)) <= tol


@pytest.mark.parametrize("name", ["COO", "HYB"])
def test_kwide_other_formats(gpu, monkeypatch, name):
    A = _long_runs()
    H = as_format(A, name)
    rng = np.random.default_rng(63)
    X = rng.uniform(-1, 1, (A.ncols, 4))
    Y0 = np.zeros((A.nrows, 4))
    yk = _run(monkeypatch, H, "f64", X, Y0, True, True)
    for j in range(4):
        xj = np.ascontiguousarray(X[:, j])
        ref = oracle.spmv(A, xj)
        assert np.all(np.abs(yk[:, j] - ref) <= fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, xj)))


def test_kwide_without_value_copy(gpu, monkeypatch):
    """A plan without the value copy keeps one pass per column."""
    A = _long_runs()
    rng = np.random.default_rng(64)
    X = rng.uniform(-1, 1, (A.ncols, 4))
    Y0 = np.zeros((A.nrows, 4))
    yd = _run(monkeypatch, A, "f64", X, Y0, True, True, capi.POLICY_NO_VALUE_COPIES)
    for j in range(4):
        xj = np.ascontiguousarray(X[:, j])
        ref = oracle.spmv(A, xj)
        assert np.all(np.abs(yd[:, j] - ref) <= fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, xj)))
