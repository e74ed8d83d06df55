"""Batched SpMM (one persistent launch over many matrices) and the
autotuner, through the C-ABI, against the CPU oracle."""
import numpy as np
import pytest

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  HDC_FORMAT, HYB_FORMAT, Csr)
from tests.helpers import fp64_bound, max_rel_err

pytestmark = pytest.mark.gpu


def small_batch(n=30, seed=7):
    rng = np.random.default_rng(seed)
    mats = []
    fmts = [CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT, COO_FORMAT]
    for i in range(n):
        rows = int(rng.integers(300, 3000))
        kind = i % 3
        if kind == 0:
            A = gen.poisson2d(max(2, int(np.sqrt(rows))))
        elif kind == 1:
            A = gen.powerlaw_rows(rows, rows, 1.5, 10.0, min(rows, 1500), int(rng.integers(1 << 30)))
        else:
            A = gen.banded(rows, int(rng.integers(2, 9)), 0.0, int(rng.integers(1 << 30)))
        f = fmts[i % len(fmts)]
        try:
            H = oracle.convert(A, f)
        except oracle.FormatConversion:
            f, H = CSR_FORMAT, A
        mats.append((A, H))
    return mats


def run_batch(mats, k, dtype, init=True, graph=False, reps=1, Y0=None):
    import torch
    from morpheus_amd import device
    D = [device.to_device(H, dtype) for _, H in mats]
    vdt = device.vector_dtype(dtype)
    rng = np.random.default_rng(k)
    Xs = [rng.uniform(-1, 1, (A.ncols, k)) for A, _ in mats]
    X = [torch.from_numpy(x).cuda().to(vdt) for x in Xs]
    if Y0 is None:
        Y = [torch.full((A.nrows, k), float("nan"), dtype=vdt, device="cuda") for A, _ in mats]
    else:
        Y = [torch.from_numpy(y).cuda().to(vdt) for y in Y0]
    B = device.Batch(D, k, X, Y)
    outs = []
    for _ in range(reps):
        if Y0 is not None:
            for y, y0 in zip(Y, Y0):
                y.copy_(torch.from_numpy(y0).to(vdt))
        (B.run_graph if graph else B.run)(init)
        torch.cuda.synchronize()
        outs.append([y.cpu().double().numpy() for y in Y])
    B.close()
    return Xs, outs


@pytest.mark.parametrize("k", [1, 8])
def test_batch_fp64_within_bound(gpu, k):
    mats = small_batch()
    Xs, outs = run_batch(mats, k, "f64")
    for (A, _), X, Y in zip(mats, Xs, outs[0]):
        for j in range(k):
            ref = oracle.spmv(A, X[:, j])
            bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, X[:, j]))
            assert np.all(np.abs(Y[:, j] - ref) <= bound)


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
def test_batch_reduced_precision(gpu, dtype, tol):
    mats = small_batch(seed=11)
    Xs, outs = run_batch(mats, 8, dtype)
    worst = 0.0
    for (A, _), X, Y in zip(mats, Xs, outs[0]):
        for j in range(8):
            ref = oracle.spmv_compensated(A, X[:, j])
            worst = max(worst, max_rel_err(Y[:, j], ref, oracle.absrow(A, X[:, j])))
    assert worst <= tol, worst


def test_batch_accumulate(gpu):
    mats = small_batch(n=12, seed=3)
    Y0 = [np.full((A.nrows, 4), 0.25) for A, _ in mats]
    Xs, outs = run_batch(mats, 4, "f64", init=False, Y0=Y0)
    for (A, _), X, Y in zip(mats, Xs, outs[0]):
        for j in range(4):
            ref = oracle.spmv(A, X[:, j], np.full(A.nrows, 0.25), init=False)
            bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, X[:, j]),
                               np.full(A.nrows, 0.25))
            assert np.all(np.abs(Y[:, j] - ref) <= bound)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_batch_rows_per_group_agree(gpu, k, dtype, monkeypatch):
    """Row tiles whose lane groups take 1 or 4 rows (8 or 2 entries of each
    per batch), for DIA / ELL members and for CSR-shaped ones, with and
    without the LDS-staged X rows of narrow DIA members, add every row's
    entries in the same order: identical bits, and within the FP64 bound /
    the FP32 tolerance of the oracle."""
    mats = small_batch(n=18, seed=13)
    outs = {}
    for rows, csr_rows, stage in (("4", "1", "1"), ("1", "1", "1"), ("4", "4", "1"),
                                  ("1", "4", "1"), ("4", "1", "0"), ("1", "1", "0")):
        monkeypatch.setenv("MPH_BATCH_ROWS_PER_GROUP", rows)
        monkeypatch.setenv("MPH_BATCH_CSR_ROWS_PER_GROUP", csr_rows)
        monkeypatch.setenv("MPH_BATCH_STAGE", stage)
        Xs, o = run_batch(mats, k, dtype)
        outs[(rows, csr_rows, stage)] = o[0]
    ref = outs[("4", "1", "1")]
    for key, o in outs.items():
        for a, b in zip(ref, o):
            assert np.array_equal(a, b), key
    for (A, _), X, Y in zip(mats, Xs, ref):
        for j in range(k):
            scale = oracle.absrow(A, X[:, j])
            if dtype == "f64":
                bound = fp64_bound(np.diff(A.row_offsets), scale)
                assert np.all(np.abs(Y[:, j] - oracle.spmv(A, X[:, j])) <= bound)
            else:
                assert max_rel_err(Y[:, j], oracle.spmv_compensated(A, X[:, j]), scale) <= 1e-5


def test_batch_graph_replays_match_eager(gpu):
    """The captured graph is replayed several times with eager work and
    synchronisations in between; every replay equals the eager launch."""
    import torch
    mats = small_batch(n=16, seed=5)
    _, eager = run_batch(mats, 8, "f32")
    _, replays = run_batch(mats, 8, "f32", graph=True, reps=4, Y0=[np.zeros((A.nrows, 8)) for A, _ in mats])
    for rep in replays:
        torch.ones(1000, device="cuda").sum().item()  # unrelated eager work
        for a, b in zip(eager[0], rep):
            assert np.array_equal(a, b)


def test_batch_coo_members(gpu):
    """COO members (row-sorted) run through their derived row offsets; every
    member is COO, rows of length 0 included."""
    mats = []
    for i, rows in enumerate([1, 7, 513, 2000, 4100]):
        A = gen.powerlaw_rows(rows, rows, 1.5, 6.0, min(rows, 900), 40 + i)
        mats.append((A, oracle.convert(A, COO_FORMAT)))
    # every third row empty
    P = gen.powerlaw_rows(3000, 3000, 1.5, 6.0, 900, 5)
    keep = np.repeat(np.arange(P.nrows) % 3 != 0, np.diff(P.row_offsets))
    lens = np.where(np.arange(P.nrows) % 3 != 0, np.diff(P.row_offsets), 0)
    E = Csr(P.nrows, P.ncols, np.concatenate([[0], np.cumsum(lens)]),
            P.column_indices[keep], P.values[keep])
    mats.append((E, oracle.convert(E, COO_FORMAT)))
    Xs, outs = run_batch(mats, 4, "f64")
    for (A, _), X, Y in zip(mats, Xs, outs[0]):
        for j in range(4):
            ref = oracle.spmv(A, X[:, j])
            bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, X[:, j]))
            assert np.all(np.abs(Y[:, j] - ref) <= bound)


def test_batch_refuses_unsorted_coo(gpu):
    import torch
    from morpheus_amd import device
    from morpheus_amd.capi import InvalidInputException
    from morpheus_amd.formats import Coo
    A = gen.powerlaw_rows(500, 500, 1.5, 6.0, 100, 3)
    C = oracle.convert(A, COO_FORMAT)
    perm = np.random.default_rng(1).permutation(C.nnz)
    U = Coo(C.nrows, C.ncols, C.row_indices[perm], C.column_indices[perm], C.values[perm])
    D = device.to_device(U, "f64")
    X = [torch.zeros((500, 2), dtype=torch.float64, device="cuda")]
    Y = [torch.zeros((500, 2), dtype=torch.float64, device="cuda")]
    with pytest.raises(InvalidInputException):
        device.Batch([D], 2, X, Y)


def test_autotune_format_picks_fastest_admissible(gpu):
    import torch
    from morpheus_amd import device
    A = gen.poisson2d(200)
    D = device.to_device(A, "f64")
    best, ms = device.autotune_format(D, k=1, reps=3)
    torch.cuda.synchronize()
    assert best in ms and ms[best] == min(ms.values())
    assert set(ms) >= {CSR_FORMAT, DIA_FORMAT, ELL_FORMAT}
    # and the chosen format computes the product correctly
    M = device.convert(D, best)
    x = gen.random_vector(A.ncols, 4)
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    M.multiply(xt, y)
    torch.cuda.synchronize()
    ref = oracle.spmv(A, x)
    assert np.all(np.abs(y.cpu().numpy() - ref) <= fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, x)))


def test_autotune_format_skips_inadmissible(gpu):
    from morpheus_amd import device
    n = 200_000
    A = gen.powerlaw_rows(n, n, 1.5, 16.0, 20_000, seed=1)
    best, ms = device.autotune_format(device.to_device(A, "f64"), k=1, reps=2)
    assert ELL_FORMAT not in ms and DIA_FORMAT not in ms  # fill tolerance exceeded
    assert COO_FORMAT in ms and CSR_FORMAT in ms


@pytest.mark.parametrize("fmt", [CSR_FORMAT, COO_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT])
def test_autotune_launch_keeps_results_correct(gpu, fmt):
    import torch
    from morpheus_amd import device
    A = gen.powerlaw_rows(20000, 20000, 1.5, 10.0, 5000, 9) if fmt not in (DIA_FORMAT, ELL_FORMAT) \
        else gen.banded(20000, 3, 0.0, 2)
    D = device.convert(device.to_device(A, "f64"), fmt)
    device.autotune_launch(D, k=1, reps=2)
    x = gen.random_vector(A.ncols, 4)
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    D.multiply(xt, y)
    torch.cuda.synchronize()
    ref = oracle.spmv(A, x)
    assert np.all(np.abs(y.cpu().numpy() - ref) <= fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, x)))
