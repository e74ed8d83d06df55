"""DIA SpMM with consecutive rows per lane group and X reuse along runs of
diagonals.

Reference: core/src/impl/Dia/Serial/Morpheus_Multiply_Impl.hpp (one sum per
row over the diagonals in storage order, entries whose column falls outside
the matrix skipped), per right-hand side.

The lane-group DIA SpMM kernel takes one row per lane group; the run form
(by default with 4 or more lanes per row; forced on / off here with
MPH_DIA_SPMM_RUNS=1 / 0) takes 4 consecutive rows and, for each run of
diagonals whose offsets step by one, loads the X rows the run covers once
(the run's values through LDS).  Each row sums its diagonals in the same
order with the same skips, so the
two must agree bit for bit.  Cases: 27- and 5-point stencils, a band,
unsorted and repeated offsets, offsets past the matrix edge, rectangular
shapes, rows not a multiple of the group's rows, k = 4, 8, 16 and 32,
accumulation, FP32 and BF16; FP64 also against the oracle's bound.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import CSR_FORMAT, DIA_FORMAT, Dia
from tests.helpers import fp64_bound

pytestmark = pytest.mark.gpu


def _dia(nrows, ncols, offsets, seed):
    rng = np.random.default_rng(seed)
    offsets = np.asarray(offsets, np.int32)
    pitch = ((nrows + 31) // 32) * 32
    v = np.zeros(pitch * offsets.size)
    for d, o in enumerate(offsets):
        r = np.arange(nrows)
        ok = (r + o >= 0) & (r + o < ncols)
        v[d * pitch + r[ok]] = rng.uniform(-1, 1, ok.sum())
    return Dia(nrows, ncols, int(np.count_nonzero(v)), offsets, v, pitch)


def _cases():
    return {
        "S27": oracle.convert(gen.stencil3d_27(20), DIA_FORMAT),
        "5-point": oracle.convert(gen.poisson2d(70), DIA_FORMAT),
        "band 9": _dia(5003, 5003, np.arange(-4, 5), 1),
        "unsorted, repeated": _dia(3001, 3001, [3, -2, 0, 1, 3, 700, -1, 701, 702, -900, 2], 2),
        "past the edge": _dia(2000, 2000, [-3000, -1999, -1, 0, 1, 1999, 2500], 3),
        "rectangular wide": _dia(1500, 4000, [0, 1, 2, 2000, 2001], 4),
        "rectangular tall": _dia(4000, 1500, [-3000, -2999, -1, 0, 1], 5),
    }


NAMES = ["S27", "5-point", "band 9", "unsorted, repeated", "past the edge",
         "rectangular wide", "rectangular tall"]


def _run(monkeypatch, D, dtype, X, Y0, init, runs):
    import torch
    from morpheus_amd import device
    monkeypatch.setenv("MPH_DIA_SPMM_RUNS", "1" if runs else "0")
    M = device.to_device(D, dtype)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(np.ascontiguousarray(X)).cuda().to(vdt)
    y = torch.from_numpy(np.ascontiguousarray(Y0)).cuda().to(vdt)
    M.multiply_multi(xt, y, init)
    torch.cuda.synchronize()
    out = y.cpu().numpy().astype(np.float64)
    M.close()
    monkeypatch.delenv("MPH_DIA_SPMM_RUNS", raising=False)
    return out


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("k", [4, 16])
@pytest.mark.parametrize("init", [True, False])
def test_runs_equal_groups_fp64(gpu, monkeypatch, name, k, init):
    D = _cases()[name]
    rng = np.random.default_rng(71)
    X = rng.uniform(-1, 1, (D.ncols, k))
    Y0 = rng.uniform(-1, 1, (D.nrows, k))
    yr = _run(monkeypatch, D, "f64", X, Y0, init, True)
    yg = _run(monkeypatch, D, "f64", X, Y0, init, False)
    assert np.array_equal(yr, yg), f"{np.count_nonzero(yr != yg)} entries differ"
    C = oracle.convert(D, CSR_FORMAT)
    for j in (0, k - 1):
        xj = np.ascontiguousarray(X[:, j])
        ref = oracle.spmv(C, xj, Y0[:, j], init)
        bound = fp64_bound(np.diff(C.row_offsets), oracle.absrow(C, xj), None if init else Y0[:, j])
        assert np.all(np.abs(yr[:, j] - ref) <= bound)


@pytest.mark.parametrize("name", ["S27", "unsorted, repeated", "rectangular tall"])
@pytest.mark.parametrize("dtype,k", [("f64", 8), ("f64", 32), ("f32", 8), ("f32", 16), ("bf16", 8),
                                     ("bf16", 16)])
def test_runs_equal_groups_widths(gpu, monkeypatch, name, dtype, k):
    D = _cases()[name]
    rng = np.random.default_rng(72)
    X = rng.uniform(-1, 1, (D.ncols, k))
    Y0 = np.zeros((D.nrows, k))
    yr = _run(monkeypatch, D, dtype, X, Y0, True, True)
    yg = _run(monkeypatch, D, dtype, X, Y0, True, False)
    assert np.array_equal(yr, yg)
