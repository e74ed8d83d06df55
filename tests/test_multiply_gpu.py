"""GPU parity of Morpheus::multiply for all six formats, through the C-ABI,
against the CPU oracle (reference Serial restatement).

FP64: every row must satisfy |y - y_ref| <= 4 * nnz_row * u * sum_j |a_ij x_j|
(u = 2^-53), integer-valued cases must match exactly.  FP32 / BF16 storage:
max-row relative error vs the compensated FP64 oracle <= 1e-5 / 1e-2.
Mirrors core/tests/Test_Multiply.hpp:193-271 (init and accumulate, reference
fixture sizes 10/100/1000) plus the benchmark workloads."""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  HDC_FORMAT, HYB_FORMAT, Coo, Csr, csr_to_coo_host)
from tests.helpers import fp64_bound, max_rel_err, small_fixture

pytestmark = pytest.mark.gpu

FMT = {"COO": COO_FORMAT, "CSR": CSR_FORMAT, "DIA": DIA_FORMAT, "ELL": ELL_FORMAT,
       "HYB": HYB_FORMAT, "HDC": HDC_FORMAT}
ALL = list(FMT)


def _torch():
    import torch
    return torch


def gpu_spmv(Ah, x, dtype="f64", y0=None, init=True, tuning=None, k=1):
    torch = _torch()
    from morpheus_amd import device
    D = device.to_device(Ah, dtype)
    if tuning:
        D.set_tuning(**tuning)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda").to(vdt)
    if y0 is None:
        yt = torch.full((Ah.nrows,) if k == 1 else (Ah.nrows, k), float("nan"),
                        dtype=vdt, device="cuda")
    else:
        yt = torch.from_numpy(np.ascontiguousarray(y0)).to("cuda").to(vdt)
    if k == 1:
        D.multiply(xt, yt, init)
    else:
        D.multiply_multi(xt, yt, init)
    torch.cuda.synchronize()
    out = yt.cpu().numpy().astype(np.float64)
    D.close()
    return out


def check_fp64(A_csr, fmt_host, x, y0=None, init=True, tuning=None):
    y = gpu_spmv(fmt_host, x, "f64", y0, init, tuning)
    ref = oracle.spmv(A_csr, x, y0, init)
    bound = fp64_bound(np.diff(A_csr.row_offsets), oracle.absrow(A_csr, x),
                       None if init else y0)
    bad = np.nonzero(~(np.abs(y - ref) <= bound))[0]
    assert bad.size == 0, f"{bad.size} rows out of bound, first {bad[:5]}: {y[bad[:5]]} vs {ref[bad[:5]]}"
    return y


def as_format(A_csr, name):
    return oracle.convert(A_csr, FMT[name])


# ------------------------------------------------------------ reference fixture
@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("init", [True, False])
def test_small_fixture(gpu, name, init):
    fx = small_fixture()
    x = gen.random_vector(10, 5374857)
    y0 = np.ones(10)
    check_fp64(fx["CSR"], fx[name], x, y0, init)


@pytest.mark.parametrize("n", [10, 100, 1000])
@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("init", [True, False])
def test_reference_sizes(gpu, n, name, init):
    A = gen.diagonal_test_matrix(n, seed=n)
    x = gen.random_vector(n, 5374857)
    y0 = np.linspace(-1, 1, n)
    check_fp64(A, as_format(A, name), x, y0, init)


@pytest.mark.parametrize("name", ALL)
def test_integer_values_exact(gpu, name):
    A = gen.diagonal_test_matrix(300, seed=3)
    A = Csr(A.nrows, A.ncols, A.row_offsets, A.column_indices,
            np.round(A.values * 4.0))
    x = np.arange(A.ncols, dtype=np.float64) % 7 - 3
    y = gpu_spmv(as_format(A, name), x)
    assert np.array_equal(y, oracle.spmv(A, x))


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("name", ALL)
def test_stencil7_closed_form(gpu, name):
    g = 40
    A = gen.stencil3d_7(g)
    c = np.indices((g, g, g)).reshape(3, -1)
    expect = sum(((c[k] == 0).astype(int) + (c[k] == g - 1)) for k in range(3)).astype(float)
    y = gpu_spmv(as_format(A, name), np.ones(g ** 3))
    assert np.array_equal(y, expect)


# ------------------------------------------------------------ long rows
@pytest.fixture(scope="module")
def powerlaw():
    return gen.powerlaw_rows(30000, 30000, 1.5, 16.0, 12000, seed=77)


@pytest.mark.parametrize("name", ["CSR", "COO", "HYB", "HDC"])
def test_powerlaw_long_rows(gpu, powerlaw, name):
    A = powerlaw
    x = gen.random_vector(A.ncols, 1)
    check_fp64(A, as_format(A, name), x)


@pytest.mark.parametrize("name", ["CSR", "COO", "HYB"])
def test_deterministic_bitwise(gpu, powerlaw, name):
    Ah = as_format(powerlaw, name)
    x = gen.random_vector(powerlaw.ncols, 2)
    a = gpu_spmv(Ah, x)
    b = gpu_spmv(Ah, x)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("tuning", [dict(lanes=1), dict(lanes=8), dict(lanes=64),
                                    dict(rows_chunk=256), dict(rows_chunk=2048),
                                    dict(coo_items=4), dict(coo_items=16),
                                    dict(rows_per_thread=2), dict(xtile=1), dict(xtile=2),
                                    dict(xtile=2, rows_chunk=256)])
@pytest.mark.parametrize("name", ["CSR", "COO", "HYB", "HDC", "DIA", "ELL"])
def test_tuning_variants(gpu, name, tuning):
    A = gen.powerlaw_rows(5000, 5000, 1.5, 10.0, 3000, seed=5) if name in ("CSR", "COO", "HYB", "HDC") \
        else gen.stencil3d_7(16)
    x = gen.random_vector(A.ncols, 3)
    check_fp64(A, as_format(A, name), x, tuning=tuning)


# ------------------------------------------------------------ x-tiled long rows
@pytest.fixture(scope="module")
def wide():
    """x of 3.2 MB (FP64): the long rows take the x-tiled pass by default."""
    return gen.powerlaw_rows(20000, 400000, 1.5, 40.0, 30000, seed=31)


@pytest.mark.parametrize("name", ["CSR", "COO", "HYB", "HDC"])
@pytest.mark.parametrize("init", [True, False])
def test_xtile_long_rows(gpu, wide, name, init):
    A = wide
    assert np.diff(A.row_offsets).max() > 2048
    x = gen.random_vector(A.ncols, 4)
    y0 = gen.random_vector(A.nrows, 5)
    check_fp64(A, as_format(A, name), x, y0, init)


@pytest.mark.parametrize("name", ["CSR", "COO", "HYB", "HDC"])
def test_xtile_forced_and_deterministic(gpu, powerlaw, name):
    Ah = as_format(powerlaw, name)
    x = gen.random_vector(powerlaw.ncols, 6)
    a = check_fp64(powerlaw, Ah, x, tuning=dict(xtile=2))
    b = gpu_spmv(Ah, x, tuning=dict(xtile=2))
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
def test_xtile_reduced_precision(gpu, wide, dtype, tol):
    A = wide
    x = gen.random_vector(A.ncols, 7)
    y = gpu_spmv(A, x, dtype)
    ref = oracle.spmv_compensated(A, x)
    assert max_rel_err(y, ref, oracle.absrow(A, x)) <= tol


@pytest.mark.parametrize("name", ["CSR", "COO", "HYB", "HDC"])
@pytest.mark.parametrize("init", [True, False])
def test_xtile_long_rows_at_the_ends(gpu, name, init):
    """Tiled long rows first and last, empty rows around them: the rows the
    segmented / short-block passes skip are still cleared or accumulated."""
    rng = np.random.default_rng(3)
    n, m = 3000, 300000
    lens = np.zeros(n, dtype=np.int64)
    lens[[0, 1, 1500, n - 1]] = [5000, 3000, 4000, 2500]
    lens[rng.choice(np.arange(10, n - 10), 800, replace=False)] = rng.integers(1, 20, 800)
    lens[[5, 6, 7, 2000, 2001]] = 0
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=ro[1:])
    ci = np.concatenate([np.sort(rng.choice(m, L, replace=False)) for L in lens if L > 0])
    A = Csr(n, m, ro, ci, rng.standard_normal(ro[-1]))
    x = gen.random_vector(m, 12)
    y0 = gen.random_vector(n, 13)
    check_fp64(A, as_format(A, name), x, y0, init, tuning=dict(xtile=2))


@pytest.mark.parametrize("xtile", [0, 2])
@pytest.mark.parametrize("init", [True, False])
def test_xtile_run_limit(gpu, xtile, init):
    """Runs of a long row inside one 4096-column tile: 512 items (a whole
    sub-block: the row is x-tiled) and 513 (the row keeps the position
    chunks), beside rows of 450 items per tile over many tiles."""
    rng = np.random.default_rng(12)
    m, n = 300000, 400
    rows = []
    for r in range(n):
        if r == 5:
            cols = np.concatenate([np.arange(512), 8192 + rng.choice(200000, 100, replace=False)])
        elif r == 9:
            cols = np.concatenate([4096 + np.arange(513), 12288 + rng.choice(200000, 100, replace=False)])
        elif r % 50 == 7:
            cols = np.concatenate([t * 4096 + rng.choice(4096, 450, replace=False) for t in range(0, 70, 3)])
        else:
            cols = rng.choice(m, int(rng.integers(0, 12)), replace=False)
        rows.append(np.unique(cols))
    ro = np.concatenate([[0], np.cumsum([c.size for c in rows])]).astype(np.int32)
    A = Csr(n, m, ro, np.concatenate(rows).astype(np.int32),
            rng.standard_normal(int(ro[-1])))
    x = gen.random_vector(m, 13)
    y0 = gen.random_vector(n, 14)
    for name in ("CSR", "COO", "HYB"):
        check_fp64(A, as_format(A, name), x, y0, init, tuning=dict(xtile=xtile))


@pytest.mark.parametrize("xtile", [0, 2])
def test_xtile_dense_rows_halve_the_tile(gpu, xtile):
    """Rows with 1,500 items in each of 15 windows of 4,096 columns: every
    run exceeds a sub-block in the full tile (4,096 FP64 / 8,192 FP32
    columns), so the plan halves the tile until the runs fit (1,024
    columns); FP64 within the per-row bound, FP32 / BF16 against the
    compensated oracle."""
    rng = np.random.default_rng(21)
    m, n = 300000, 600
    rows = []
    for r in range(n):
        if r % 15 == 3:
            cols = np.concatenate([t * 4096 + rng.choice(4096, 1500, replace=False)
                                   for t in range(0, 60, 4)])
        else:
            cols = rng.choice(m, int(rng.integers(0, 30)), replace=False)
        rows.append(np.unique(cols))
    ro = np.concatenate([[0], np.cumsum([c.size for c in rows])]).astype(np.int32)
    A = Csr(n, m, ro, np.concatenate(rows).astype(np.int32),
            rng.standard_normal(int(ro[-1])))
    x = gen.random_vector(m, 22)
    y0 = gen.random_vector(n, 23)
    for name in ("CSR", "COO", "HYB"):
        check_fp64(A, as_format(A, name), x, y0, False, tuning=dict(xtile=xtile))
    ref = oracle.spmv_compensated(A, x)
    for dtype, tol in (("f32", 1e-5), ("bf16", 1e-2)):
        y = gpu_spmv(A, x, dtype, tuning=dict(xtile=xtile))
        assert max_rel_err(y, ref, oracle.absrow(A, x)) <= tol


def test_xtile_table_cap(gpu):
    """32M columns (8,192 FP64 tiles) and 9,000 rows longer than 512: the
    per-row tile-bound table would exceed its cap (2^26 entries), so only
    the 8,191 longest rows are x-tiled and the rest keep the position
    chunks."""
    rng = np.random.default_rng(17)
    m, n = 32 * 1024 * 1024, 9000
    lens = rng.integers(513, 700, n)
    ro = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(m, L, replace=False)) for L in lens]).astype(np.int32)
    A = Csr(n, m, ro, ci, rng.standard_normal(int(ro[-1])))
    x = gen.random_vector(m, 18)
    check_fp64(A, A, x)


def test_xtile_unsorted_columns(gpu):
    """Long rows whose columns are not sorted keep the position chunks."""
    A = gen.powerlaw_rows(4000, 300000, 1.5, 30.0, 6000, seed=8)
    rng = np.random.default_rng(1)
    ci = A.column_indices.copy()
    v = A.values.copy()
    for r in range(A.nrows):
        a, b = A.row_offsets[r], A.row_offsets[r + 1]
        perm = rng.permutation(b - a)
        ci[a:b] = ci[a:b][perm]
        v[a:b] = v[a:b][perm]
    U = Csr(A.nrows, A.ncols, A.row_offsets, ci, v)
    x = gen.random_vector(U.ncols, 2)
    check_fp64(U, U, x, tuning=dict(xtile=2))


@pytest.mark.parametrize("name", ["CSR", "COO", "HYB"])
@pytest.mark.parametrize("k", [2, 3, 8, 12])
@pytest.mark.parametrize("init", [True, False])
def test_xtile_plan_spmm(gpu, wide, name, k, init):
    """SpMM on a handle with x-tiled rows: up to 8 right-hand sides the
    x-tiled pass runs once per column (its columns equal the single-vector
    product of the tiled rows bit for bit) and the other rows take the k-wide
    kernel over the split plan; 12 keeps the position chunks.  A k = 1
    product on the same handle before and after.  FP64 within the per-row
    bound, FP32 within 1e-5."""
    torch = _torch()
    from morpheus_amd import device
    A = wide
    X = np.stack([gen.random_vector(A.ncols, 10 + j) for j in range(k)], axis=1)
    Y0 = np.stack([gen.random_vector(A.nrows, 40 + j) for j in range(k)], axis=1)
    nrow = np.diff(A.row_offsets)
    for dtype in ("f64", "f32"):
        D = device.to_device(as_format(A, name), dtype)
        vdt = device.vector_dtype(dtype)
        Xt = torch.from_numpy(np.ascontiguousarray(X)).cuda().to(vdt)
        Y = torch.from_numpy(np.ascontiguousarray(Y0)).cuda().to(vdt)
        y1 = torch.full((A.nrows,), float("nan"), dtype=vdt, device="cuda")
        D.multiply(Xt[:, 0].contiguous(), y1)
        D.multiply_multi(Xt, Y, init)
        y2 = torch.full((A.nrows,), float("nan"), dtype=vdt, device="cuda")
        D.multiply(Xt[:, 0].contiguous(), y2)
        torch.cuda.synchronize()
        assert torch.equal(y1, y2)
        Yh = Y.double().cpu().numpy()
        y0s = Y.new_tensor(Y0).double().cpu().numpy()  # Y0 as stored
        for j in range(k):
            if dtype == "f64":
                ref = oracle.spmv(A, X[:, j], Y0[:, j].copy(), init)
                bound = fp64_bound(nrow, oracle.absrow(A, X[:, j]), None if init else Y0[:, j])
                assert np.all(np.abs(Yh[:, j] - ref) <= bound), f"column {j}"
            else:
                ref = oracle.spmv_compensated(A, X[:, j]) + (0 if init else y0s[:, j])
                scale = oracle.absrow(A, X[:, j]) + (0 if init else np.abs(y0s[:, j]))
                assert max_rel_err(Yh[:, j], ref, scale) <= 1e-5, f"column {j}"
        D.close()


# ------------------------------------------------------------ split two-pass HYB / HDC
@pytest.fixture(scope="module")
def tall():
    """100k rows: HYB / HDC run the ELL / DIA part as its own pass and then
    the rows that have a tail / CSR part (a compacted row list), with the
    long rows x-tiled or not."""
    mats = {"powerlaw": gen.powerlaw_rows(100000, 100000, 1.5, 8.0, 3000, seed=41),
            "banded": gen.banded(100000, 4, 0.02, seed=42)}
    return {(m, f): (A, as_format(A, f)) for m, A in mats.items() for f in ("HYB", "HDC")}


@pytest.mark.parametrize("mat", ["powerlaw", "banded"])
@pytest.mark.parametrize("name", ["HYB", "HDC"])
@pytest.mark.parametrize("xtile", [0, 1, 2])
@pytest.mark.parametrize("init", [True, False])
def test_split_two_pass(gpu, tall, mat, name, xtile, init):
    A, Ah = tall[(mat, name)]
    x = gen.random_vector(A.ncols, 14)
    y0 = gen.random_vector(A.nrows, 15)
    check_fp64(A, Ah, x, y0, init, tuning=dict(xtile=xtile))


def test_split_two_pass_deterministic(gpu, tall):
    _, Ah = tall[("powerlaw", "HYB")]
    x = gen.random_vector(Ah.ncols, 16)
    a = gpu_spmv(Ah, x, tuning=dict(xtile=2))
    b = gpu_spmv(Ah, x, tuning=dict(xtile=2))
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def holey_hyb(n, fill, seed, tail_every=10):
    """A HYB built by hand: K = 4 ELL slots per row, each filled with
    probability `fill` in any position (holes of -1 between entries, some
    explicit zeros), and a row-sorted COO tail on every `tail_every`-th row
    plus one row of 3000 entries."""
    from morpheus_amd.formats import Ell, Hyb
    rng = np.random.default_rng(seed)
    K, width, pitch = 4, 32, n
    ci = np.full(pitch * width, -1, np.int32)
    v = np.zeros(pitch * width)
    for k in range(K):
        on = rng.random(n) < fill
        ci[k * pitch: k * pitch + n] = np.where(on, rng.integers(0, n, n), -1)
        v[k * pitch: k * pitch + n] = np.where(on, rng.standard_normal(n) * (rng.random(n) > 0.05), 0.0)
    ell = Ell(n, n, int((ci >= 0).sum()), K, width, ci, v, pitch)
    rows = [np.repeat(r, rng.integers(1, 4)) for r in range(0, n, tail_every)]
    rows.insert(len(rows) // 2, np.repeat(n // 2 + 1, 3000))
    ri = np.sort(np.concatenate(rows), kind="stable")
    cc = rng.integers(0, n, ri.size)
    coo = Coo(n, n, ri, cc, rng.standard_normal(ri.size))
    return Hyb(ell, coo)


@pytest.mark.parametrize("fill,tail_every", [(0.2, 10), (0.8, 10), (0.8, 4)])
@pytest.mark.parametrize("xtile", [0, 2])
@pytest.mark.parametrize("init", [True, False])
def test_hyb_merged_and_split_plans(gpu, fill, tail_every, xtile, init):
    """70k rows: a HYB whose ELL slots are a fifth filled runs over the merged
    CSR copy, one four-fifths filled with a tail on every tenth row runs the
    ELL pass then the tail, and one with a tail on every fourth row the merged
    copy again; all against the Serial HYB product (ELL, then COO)."""
    H = holey_hyb(70000, fill, 3, tail_every)
    x = gen.random_vector(H.ncols, 17)
    y0 = gen.random_vector(H.nrows, 18)
    y = gpu_spmv(H, x, y0=y0, init=init, tuning=dict(xtile=xtile))
    ref = oracle.spmv(H, x, y0, init)
    from morpheus_amd.formats import Ell, Hyb
    e, c = H.ell, H.coo
    absH = Hyb(Ell(e.nrows, e.ncols, e.nnz, e.entries_per_row, e.width, e.column_indices,
                   np.abs(e.values), e.pitch),
               Coo(c.nrows, c.ncols, c.row_indices, c.column_indices, np.abs(c.values)))
    absrow = oracle.spmv(absH, np.abs(x))
    nnz_row = (e.column_indices.reshape(e.width, e.pitch)[:, :e.nrows] >= 0).sum(axis=0) + \
        np.bincount(c.row_indices, minlength=H.nrows)
    bound = fp64_bound(nnz_row, absrow, None if init else y0)
    assert np.all(np.abs(y - ref) <= bound)


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("fill", [0.2, 0.8])
def test_hyb_plans_reduced_precision(gpu, dtype, tol, fill):
    """The merged and split HYB plans in FP32 / BF16 storage against the FP64
    product of the original matrix, max_i |y_i - y_ref_i| / sum_j |a_ij x_j|."""
    from morpheus_amd.formats import Ell, Hyb
    H = holey_hyb(70000, fill, 6)
    x = gen.random_vector(H.ncols, 19)
    y = gpu_spmv(H, x, dtype)
    ref = oracle.spmv(H, x)
    e, c = H.ell, H.coo
    absH = Hyb(Ell(e.nrows, e.ncols, e.nnz, e.entries_per_row, e.width, e.column_indices,
                   np.abs(e.values), e.pitch),
               Coo(c.nrows, c.ncols, c.row_indices, c.column_indices, np.abs(c.values)))
    err = max_rel_err(y, ref, oracle.spmv(absH, np.abs(x)))
    assert err <= tol, err


def test_hyb_merged_plan_spmm(gpu):
    H = holey_hyb(70000, 0.2, 4)
    X = np.random.default_rng(5).uniform(-1, 1, (H.ncols, 3))
    Y = gpu_spmv(H, X, k=3)
    for j in range(3):
        ref = oracle.spmv(H, X[:, j])
        assert np.allclose(Y[:, j], ref, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ hipGraph capture
@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("xtile", [1, 2])
def test_graph_replays_match_eager(gpu, name, xtile):
    """mph_multiply captured in a hipGraph: four replays, with x rewritten
    eagerly between them, each bit-identical to an eager product on the
    same x (the captured launches hold no memset nodes)."""
    torch = _torch()
    from morpheus_amd import device
    A = gen.powerlaw_rows(70000, 300000, 1.5, 8.0, 4000, seed=9) \
        if name in ("CSR", "COO", "HYB", "HDC") else gen.stencil3d_7(24)
    D = device.to_device(as_format(A, name), "f64")
    D.set_tuning(xtile=xtile)
    D.analyse()
    x = torch.zeros(A.ncols, dtype=torch.float64, device="cuda")
    y = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    ye = torch.empty_like(y)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        D.multiply(x, y, True)
    for rep in range(4):
        xh = gen.random_vector(A.ncols, 40 + rep)
        x.copy_(torch.from_numpy(xh))
        g.replay()
        torch.cuda.synchronize()
        got = y.cpu().numpy().copy()
        D.multiply(x, ye, True)
        torch.cuda.synchronize()
        assert np.array_equal(got.view(np.uint64), ye.cpu().numpy().view(np.uint64)), f"replay {rep}"
        if rep == 0:
            ref = oracle.spmv(A, xh)
            bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, xh))
            assert np.all(np.abs(got - ref) <= bound)
    D.close()


# ------------------------------------------------------------ edge cases
def _with_empty_rows(n=5000, seed=4):
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(np.arange(300, n - 300), size=400, replace=False))
    rows = np.concatenate([rows, np.full(50, rows[10])])  # one longer row
    cols = rng.integers(0, n, rows.size)
    key = np.unique(rows.astype(np.int64) * n + cols)
    rows, cols = key // n, key % n
    counts = np.bincount(rows, minlength=n)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ro[1:])
    return Csr(n, n, ro, cols, rng.standard_normal(rows.size))


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("init", [True, False])
def test_empty_rows_and_long_gaps(gpu, name, init):
    A = _with_empty_rows()
    x = gen.random_vector(A.ncols, 9)
    y0 = np.full(A.nrows, 0.5)
    try:
        Ah = as_format(A, name)
    except oracle.FormatConversion:
        pytest.skip("inadmissible for this format")
    check_fp64(A, Ah, x, y0, init)


@pytest.mark.parametrize("name", ALL)
def test_no_entries(gpu, name):
    A = Csr(64, 64, np.zeros(65), [], [])
    y = gpu_spmv(oracle.convert(A, FMT[name]) if name != "CSR" else A, np.ones(64))
    assert np.array_equal(y, np.zeros(64))


def test_unsorted_coo_falls_back(gpu, powerlaw):
    coo = csr_to_coo_host(powerlaw)
    perm = np.random.default_rng(0).permutation(coo.nnz)
    shuffled = Coo(coo.nrows, coo.ncols, coo.row_indices[perm], coo.column_indices[perm],
                   coo.values[perm])
    x = gen.random_vector(coo.ncols, 5)
    y = gpu_spmv(shuffled, x)
    ref = oracle.spmv(powerlaw, x)
    # atomics: order-dependent rounding, still inside the bound
    bound = fp64_bound(np.diff(powerlaw.row_offsets), oracle.absrow(powerlaw, x))
    assert np.all(np.abs(y - ref) <= bound)


# ------------------------------------------------------------ reduced precision
@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("name", ALL)
def test_reduced_precision_error(gpu, powerlaw, dtype, tol, name):
    A = powerlaw if name not in ("DIA", "ELL") else gen.banded(20000, 4, 0.0, 3)
    x = gen.random_vector(A.ncols, 11)
    y = gpu_spmv(as_format(A, name), x, dtype)
    ref = oracle.spmv_compensated(A, x)
    err = max_rel_err(y, ref, oracle.absrow(A, x))
    assert err <= tol, err


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("name", ["DIA", "ELL"])
@pytest.mark.parametrize("n", [20000, 20001, 31])
@pytest.mark.parametrize("init", [True, False])
def test_row_pair_kernels(gpu, dtype, tol, name, n, init):
    """FP32 / BF16 DIA and ELL SpMV with two adjacent rows per thread (the
    default launch shape; an odd row count, where ELL's pitch is odd and the
    one-row kernel runs, and DIA's even pitch holds the pair's second row):
    against the FP64 product, and equal to the one-row kernel."""
    A = gen.banded(n, 4, 0.0, 7)
    Ah = as_format(A, name)
    x = gen.random_vector(A.ncols, 21)
    y0 = gen.random_vector(A.nrows, 22)
    y = gpu_spmv(Ah, x, dtype, y0=None if init else y0, init=init)
    y1 = gpu_spmv(Ah, x, dtype, y0=None if init else y0, init=init, tuning=dict(rows_per_thread=1))
    assert np.array_equal(y, y1)
    ref = oracle.spmv_compensated(A, x) + (0 if init else y0)
    scale = oracle.absrow(A, x) + (0 if init else np.abs(y0))
    assert max_rel_err(y, ref, scale) <= tol


# ------------------------------------------------------------ SpMM
@pytest.mark.parametrize("name", ["CSR", "COO", "DIA", "ELL", "HYB", "HDC"])
@pytest.mark.parametrize("k", [2, 8, 11])
def test_spmm_matches_columnwise(gpu, name, k):
    A = gen.powerlaw_rows(4000, 4000, 1.5, 8.0, 2500, seed=8) if name not in ("DIA", "ELL") \
        else gen.poisson2d(60)
    Ah = as_format(A, name)
    X = np.random.default_rng(k).uniform(-1, 1, (A.ncols, k))
    Y = gpu_spmv(Ah, X, k=k)
    absrow = np.stack([oracle.absrow(A, X[:, j]) for j in range(k)], 1)
    ref = np.stack([oracle.spmv(A, X[:, j]) for j in range(k)], 1)
    bound = fp64_bound(np.diff(A.row_offsets)[:, None], absrow)
    assert np.all(np.abs(Y - ref) <= bound)


@pytest.mark.parametrize("name", ["DIA", "ELL"])
@pytest.mark.parametrize("dtype,k", [("f64", 2), ("f64", 6), ("f64", 16), ("f64", 40),
                                     ("f32", 4), ("f32", 12), ("f32", 20), ("bf16", 8),
                                     ("f32", 5)])
@pytest.mark.parametrize("init", [True, False])
def test_spmm_lane_groups(gpu, name, dtype, k, init):
    """DIA / ELL SpMM whose X / Y rows start on 16 bytes runs lane groups
    (16 bytes of a row per lane; column passes of up to 32 FP64 / 64 FP32
    columns, the rest in narrower passes); FP32 k = 5 keeps one thread per
    row.  FP64: each column equals the single-vector product bit for bit
    (same summation order); FP32 / BF16 within the tolerance of the
    compensated FP64 oracle; init = false accumulates into Y."""
    torch = _torch()
    from morpheus_amd import device
    # row counts that leave a partial last workgroup for every lane-group width
    A = gen.stencil3d_7(23) if name == "DIA" else gen.banded(20001, 9, 0.02, seed=4)
    Ah = as_format(A, name)
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, (A.ncols, k))
    Y0 = rng.uniform(-1, 1, (A.nrows, k))
    D = device.to_device(Ah, dtype)
    vdt = device.vector_dtype(dtype)
    Xt = torch.from_numpy(X).cuda().to(vdt)
    Yt = torch.from_numpy(Y0).cuda().to(vdt)
    D.multiply_multi(Xt, Yt, init)
    cols = []
    for j in range(k):
        yj = torch.from_numpy(Y0[:, j].copy()).cuda().to(vdt)
        D.multiply(Xt[:, j].contiguous(), yj, init)
        cols.append(yj)
    torch.cuda.synchronize()
    Y = Yt.cpu().numpy().astype(np.float64)
    D.close()
    if dtype == "f64":
        Yc = torch.stack(cols, 1).cpu().numpy()
        assert np.array_equal(Y.view(np.uint64), Yc.view(np.uint64))
        nrow = np.diff(A.row_offsets)
        for j in range(k):
            ref = oracle.spmv(A, X[:, j], Y0[:, j].copy(), init)
            bound = fp64_bound(nrow, oracle.absrow(A, X[:, j]), None if init else Y0[:, j])
            assert np.all(np.abs(Y[:, j] - ref) <= bound)
    else:
        tol = 1e-5 if dtype == "f32" else 1e-2
        y0c = Yt.new_tensor(Y0).cpu().numpy().astype(np.float64)  # Y0 as stored
        for j in range(k):
            ref = oracle.spmv_compensated(A, X[:, j]) + (0 if init else y0c[:, j])
            scale = oracle.absrow(A, X[:, j]) + (0 if init else np.abs(y0c[:, j]))
            assert max_rel_err(Y[:, j], ref, scale) <= tol


def test_one_shot_entry_points(gpu):
    torch = _torch()
    import ctypes as C
    A = gen.poisson2d(50)
    x = gen.random_vector(A.ncols, 1)
    ro = torch.from_numpy(A.row_offsets).cuda()
    ci = torch.from_numpy(A.column_indices).cuda()
    v = torch.from_numpy(A.values).cuda()
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    capi.call("mph_csr_multiply", A.nrows, A.ncols, A.nnz, p(ro), p(ci), p(v), capi.F64,
              p(xt), p(yt), 1, s)
    torch.cuda.synchronize()
    ref = oracle.spmv(A, x)
    assert np.allclose(yt.cpu().numpy(), ref, rtol=0, atol=1e-13)
