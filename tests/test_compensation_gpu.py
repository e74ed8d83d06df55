"""FP32 / BF16 accumulation with Dot2 compensation on rows longer than 1024
entries, through the C-ABI on the GPU.

The cancellation fixture (tests/helpers.py: rows of 1.5k-60k entries whose
products cancel to about 1/2000 of sum |a x|, values exact in FP32 / BF16, x
exact in FP32) runs through every path that sums such a row: CSR long-row
chunks and the x-tiled pass, COO's row plan and its segmented kernel, HYB
and HDC (fused and split plans), DIA and ELL with more than 1024 diagonals /
slots per row, SpMM with k = 2 / 4 / 8, and the batched SpMM.  A long row
must satisfy |y - y_exact| <= 4 u |y_exact| + n u^2 sum |a x| (u = 2^-24),
y_exact being the oracle's compensated FP64 sum of the exact products; the
CPU tests show that this bound rejects plain FP32 accumulation in any order
and TwoSum on rounded products, and admits Dot2.  Rows of at most 1024
entries are held to what any FP32 order guarantees, (n + 2) u sum |a x|.
In the split plans of HYB / HDC the ELL / DIA part runs as its own plain
FP32 pass (K <= 9 slots, one diagonal here) before the compensated tail
joins it, so those rows also get that part's plain bound,
(K + 2) u sum |a x| over the part.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi, device
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HDC_FORMAT,
                                  HYB_FORMAT, FORMAT_NAMES)
from tests.helpers import U32, cancellation_csr, compensated_bound, plain_fp32_bound

pytestmark = pytest.mark.gpu

LONG = [1500, 2048, 3000, 5000, 20000, 60000]
DT = {24: "f32", 8: "bf16"}


@pytest.fixture(scope="module", params=[24, 8], ids=["fp32", "bf16"])
def fixture(request):
    # 2^19 columns: x is 2 MiB in FP32, where the x-tiled pass turns on
    A, x, long_rows = cancellation_csr(20000, 1 << 19, LONG, 5, bits=request.param)
    return request.param, A, x, long_rows


def _run(D, x, k=1, init=True, y0=None):
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x)).float().cuda()
    shape = (D.nrows,) if k == 1 else (D.nrows, k)
    if y0 is None:
        y = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
    else:
        y = torch.from_numpy(np.ascontiguousarray(y0)).float().cuda()
    if k == 1:
        D.multiply(xt, y, init=init)
    else:
        D.multiply_multi(xt, y, init=init)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


def _prefix_bound(A, x, fmt, K):
    """Per row, the plain FP32 bound (m + 2) u sum |a x| of the part a split
    plan sums in its own pass: the first K entries (HYB's ELL part) or the
    diagonal entry (HDC's DIA part, the main diagonal in this fixture)."""
    out = np.zeros(A.nrows)
    for r in range(A.nrows):
        b, e = A.row_offsets[r], A.row_offsets[r + 1]
        c, v = A.column_indices[b:e], A.values[b:e]
        if fmt == HYB_FORMAT:
            out[r] = (min(K, e - b) + 2) * np.sum(np.abs(v[:K] * x[c[:K]]))
        else:
            out[r] = 3 * np.sum(np.abs(v[c == r] * x[r]))
    return out


def _check(A, x, y, long_rows, what, extra=None, scale=1.0):
    ref = oracle.spmv_compensated(A, x) * scale
    absrow = oracle.absrow(A, x) * scale
    n = np.diff(A.row_offsets)
    bound = plain_fp32_bound(n, absrow)
    bound[long_rows] = compensated_bound(n[long_rows], absrow[long_rows], ref[long_rows])
    if extra is not None:
        bound[long_rows] += U32 * extra[long_rows] * scale
    err = np.abs(y - ref)
    bad = np.nonzero(~(err <= bound))[0]
    worst = float(np.max(err[long_rows] / bound[long_rows]))
    print(f"{what}: long rows at {worst:.3f} of the compensated bound", flush=True)
    assert bad.size == 0, f"{what}: rows {bad[:8]} outside the bound (long rows {long_rows})"


def test_csr_paths(gpu, fixture):
    bits, A, x, long_rows = fixture
    D = device.to_device(A, DT[bits])
    for xt, name in ((1, "chunks"), (2, "x-tiled"), (0, "default")):
        D.set_tuning(xtile=xt)
        _check(A, x, _run(D, x), long_rows, f"CSR {name} {DT[bits]}")
    # accumulate: the old y joins the compensated sum
    D.set_tuning(xtile=2)
    y0 = np.zeros(A.nrows)
    _check(A, x, _run(D, x, init=False, y0=y0), long_rows, f"CSR x-tiled accumulate {DT[bits]}")
    D.close()


def test_coo_paths(gpu, fixture):
    bits, A, x, long_rows = fixture
    D = device.to_device(A, DT[bits])
    C = device.convert(D, COO_FORMAT)
    for items, xt, name in ((0, 1, "row plan"), (0, 2, "row plan x-tiled"),
                            (4, 1, "segmented T=4"), (16, 1, "segmented T=16")):
        C.set_tuning(coo_items=items, xtile=xt)
        _check(A, x, _run(C, x), long_rows, f"COO {name} {DT[bits]}")
    C.close()
    D.close()


@pytest.mark.parametrize("fmt", [HYB_FORMAT, HDC_FORMAT], ids=["HYB", "HDC"])
def test_hyb_hdc_paths(gpu, fixture, fmt):
    bits, A, x, long_rows = fixture
    D = device.to_device(A, DT[bits])
    M = device.convert(D, fmt)
    H = oracle.convert(A, fmt)
    K = H.ell.entries_per_row if fmt == HYB_FORMAT else 0
    extra = _prefix_bound(A, x, fmt, K)
    # fused (one pass, fewer than 65,536 rows, x-tiling off) and split plans
    M.set_tuning(xtile=1)
    _check(A, x, _run(M, x), long_rows, f"{FORMAT_NAMES[fmt]} fused {DT[bits]}")
    M.set_tuning(xtile=2)
    _check(A, x, _run(M, x), long_rows, f"{FORMAT_NAMES[fmt]} split {DT[bits]}", extra=extra)
    M.close()
    D.close()


@pytest.mark.parametrize("k", [2, 4, 8])
def test_spmm_paths(gpu, fixture, k):
    bits, A, x, long_rows = fixture
    # column j of X is x * 2^(j - 2): exact, and every column cancels alike
    X = np.stack([x * 2.0 ** (j - 2) for j in range(k)], axis=1)
    D = device.to_device(A, DT[bits])
    for fmt in (CSR_FORMAT, HYB_FORMAT, HDC_FORMAT):
        M = D if fmt == CSR_FORMAT else device.convert(D, fmt)
        for xt in (1, 2):
            M.set_tuning(xtile=xt)
            Y = _run(M, X, k=k)
            # (the SpMM of HYB / HDC always fuses the ELL / DIA part)
            for j in (0, k - 1):
                _check(A, x, Y[:, j], long_rows,
                       f"{FORMAT_NAMES[fmt]} SpMM k={k} xtile={xt} col {j} {DT[bits]}",
                       scale=2.0 ** (j - 2))
        if M is not D:
            M.close()
    D.close()


@pytest.fixture(scope="module", params=[24, 8], ids=["fp32", "bf16"])
def dense_rows(request):
    # every row long: DIA (about 2300 diagonals) and ELL (K up to 1500)
    # are within the fill tolerance and take the compensated kernel
    A, x, long_rows = cancellation_csr(300, 2000, [1100 + 2 * i for i in range(300)], 9,
                                       bits=request.param)
    return request.param, A, x, long_rows


@pytest.mark.parametrize("fmt", [DIA_FORMAT, ELL_FORMAT], ids=["DIA", "ELL"])
def test_dia_ell_long_rows(gpu, dense_rows, fmt):
    bits, A, x, long_rows = dense_rows
    D = device.to_device(A, DT[bits])
    M = device.convert(D, fmt)
    _check(A, x, _run(M, x), long_rows, f"{FORMAT_NAMES[fmt]} {DT[bits]}")
    X = np.stack([x * 2.0 ** (j - 1) for j in range(4)], axis=1)
    Y = _run(M, X, k=4)
    for j in (0, 3):
        _check(A, x, Y[:, j], long_rows, f"{FORMAT_NAMES[fmt]} k=4 col {j} {DT[bits]}",
               scale=2.0 ** (j - 1))
    M.close()
    D.close()


def test_batch(gpu, fixture, dense_rows):
    import torch
    bits, A, x, long_rows = fixture
    _, B, xb, long_b = dense_rows
    if bits != dense_rows[0]:
        pytest.skip("one value type per batch")
    k = 8
    D = device.to_device(A, DT[bits])
    E = device.to_device(B, DT[bits])
    members = [(A, x, long_rows, D), (A, x, long_rows, device.convert(D, COO_FORMAT)),
               (A, x, long_rows, device.convert(D, HDC_FORMAT)),
               (B, xb, long_b, device.convert(E, ELL_FORMAT)),
               (B, xb, long_b, device.convert(E, DIA_FORMAT))]
    Xs = [np.stack([xx * 2.0 ** (j - 4) for j in range(k)], axis=1) for _, xx, _, _ in members]
    X = [torch.from_numpy(v).float().cuda() for v in Xs]
    Y = [torch.full((M.nrows, k), float("nan"), dtype=torch.float32, device="cuda")
         for _, _, _, M in members]
    bt = device.Batch([M for _, _, _, M in members], k, X, Y)
    for run in (bt.run, bt.run_graph):
        run(True)
        torch.cuda.synchronize()
        for (AA, xx, lr, M), y in zip(members, Y):
            Yh = y.double().cpu().numpy()
            for j in (0, k - 1):
                _check(AA, xx, Yh[:, j], lr, f"batch {FORMAT_NAMES[M.format]} col {j} {DT[bits]}",
                       scale=2.0 ** (j - 4))
    bt.close()
    for _, _, _, M in members:
        M.close()
    E.close()
