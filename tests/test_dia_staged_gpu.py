"""DIA SpMV with x staged in LDS per cluster of nearby diagonals.

Reference: core/src/impl/Dia/HIP/Morpheus_Multiply_Impl.hpp:41-85 and the
Serial loop core/src/impl/Dia/Serial/Morpheus_Multiply_Impl.hpp (one sum per
row over the diagonals in storage order, entries whose column falls outside
the matrix skipped).

The analysis cuts the diagonals (storage order) into clusters whose offsets
span at most 512; a workgroup stages the x entries of each cluster of two or
more diagonals once.  The staged kernel sums the diagonals in the same order
as the plain one, so the two must agree bit for bit; the plain kernel is
forced with the launch shape rows_per_thread = 1.  Cases: stencils (one and
three clusters), unsorted and duplicate offsets, offsets past the matrix
edge, rectangular matrices, more clusters than staged (the plain kernel
then), windows beyond the LDS budget, rows not a multiple of 256, an odd
pitch, and accumulation; in FP64 (within the per-row bound of the oracle)
and in FP32 / BF16 storage (the row-pair kernels, staged the same way).
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import CSR_FORMAT, DIA_FORMAT, Dia
from tests.helpers import fp64_bound

pytestmark = pytest.mark.gpu


def _dia(nrows, ncols, offsets, seed, pitch=None):
    rng = np.random.default_rng(seed)
    offsets = np.asarray(offsets, np.int32)
    pitch = pitch or ((nrows + 31) // 32) * 32
    v = np.zeros(pitch * offsets.size)
    for d, o in enumerate(offsets):
        r = np.arange(nrows)
        ok = (r + o >= 0) & (r + o < ncols)
        v[d * pitch + r[ok]] = rng.uniform(-1, 1, ok.sum())
    return Dia(nrows, ncols, int(np.count_nonzero(v)), offsets, v, pitch)


def _product(D, x, y0, init, plain, dtype="f64"):
    import torch
    from morpheus_amd import device
    M = device.to_device(D, dtype)
    if plain:
        M.set_tuning(rows_per_thread=1)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(x).cuda().to(vdt)
    yt = torch.from_numpy(y0.copy()).cuda().to(vdt)
    M.multiply(xt, yt, init)
    torch.cuda.synchronize()
    out = yt.cpu().numpy().astype(np.float64)
    M.close()
    return out


def _cases():
    s27 = oracle.convert(gen.stencil3d_27(24), DIA_FORMAT)
    s7 = oracle.convert(gen.S7(20), DIA_FORMAT)
    return {
        "S27 (3 clusters)": s27,
        "S7 (1 staged, 2 direct)": s7,
        "band 33 (1 cluster)": _dia(10007, 10007, np.arange(-16, 17), 1),
        "unsorted, repeated": _dia(5000, 5000, [3, -2, 0, 3, 700, -1, 701, -900, 2], 2),
        "past the edge": _dia(3000, 3000, [-5000, -2999, -1, 0, 1, 2999, 4000], 3),
        "rectangular wide": _dia(4000, 9000, [0, 1, 2, 4500, 4501, 8000], 4),
        "rectangular tall": _dia(9000, 4000, [-8000, -7999, -5, 0, 5], 5),
        "9 clusters (plain)": _dia(20000, 20000, [k * 1000 + j for k in range(-4, 5) for j in (0, 1)], 6),
        "over the LDS budget (plain)": _dia(20000, 20000, [-1800, -1300, -800, -300, 200, 700, 1200] +
                                           [-1799, -1299, -799, -299, 201, 701, 1201], 7),
        "two diagonals": _dia(300, 300, [0, 1], 8),
        "odd pitch": _dia(10007, 10007, [-300, -1, 0, 1, 300], 9, pitch=10007),
    }


CASES = ["S27 (3 clusters)", "S7 (1 staged, 2 direct)", "band 33 (1 cluster)", "unsorted, repeated",
         "past the edge", "rectangular wide", "rectangular tall", "9 clusters (plain)",
         "over the LDS budget (plain)", "two diagonals", "odd pitch"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("init", [True, False])
def test_staged_equals_plain(gpu, name, init):
    D = _cases()[name]
    x = gen.random_vector(D.ncols, 31)
    y0 = gen.random_vector(D.nrows, 32)
    ys = _product(D, x, y0, init, plain=False)
    yp = _product(D, x, y0, init, plain=True)
    assert np.array_equal(ys, yp), f"{np.count_nonzero(ys != yp)} rows differ"
    C = oracle.convert(D, CSR_FORMAT)
    ref = oracle.spmv(C, x, y0, init)
    bound = fp64_bound(np.diff(C.row_offsets), oracle.absrow(C, x), None if init else y0)
    assert np.all(np.abs(ys - ref) <= bound)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 1e-2)])
def test_staged_reduced_precision(gpu, name, dtype, tol):
    """FP32 / BF16 storage: the staged row-pair kernel (even pitch) or the
    staged one-row kernel (odd pitch) against the plain one-row kernel, bit
    for bit, and against the FP64 product of the original matrix,
    max_i |y_i - y_ref_i| / sum_j |a_ij x_j|."""
    from tests.helpers import max_rel_err
    D = _cases()[name]
    x = gen.random_vector(D.ncols, 33)
    y0 = np.zeros(D.nrows)
    ys = _product(D, x, y0, True, plain=False, dtype=dtype)
    yp = _product(D, x, y0, True, plain=True, dtype=dtype)
    assert np.array_equal(ys, yp), f"{np.count_nonzero(ys != yp)} rows differ"
    C = oracle.convert(D, CSR_FORMAT)
    err = max_rel_err(ys, oracle.spmv(C, x), oracle.absrow(C, x))
    assert err <= tol, err
