"""The batched SpMM workload at its full size, against the CPU oracle.

512 matrices (log-uniform 1k-64k rows; a third each of 2D 5-point stencils,
power-law(1.5) rows with mean 12 and banded matrices of bandwidth 5-33; 7.6M
rows, 89M entries in all), k = 8 row-major right-hand sides U[-1, 1], one
launch per kind of member over the batch's tile list (direct and replayed
from a hipGraph), in FP64, FP32 and BF16 storage.  Two format assignments: the autotuner's pick per matrix, and
every admissible format in turn (so that all six appear).  Two columns of
every matrix's Y are checked against the oracle: FP64 per row within
4 n_row u sum |a x|; FP32 / BF16 as max_i |y_i - y_ref_i| / sum_j |a_ij x_j|
against the compensated FP64 oracle on the original FP64 matrix, within
1e-5 / 1e-2.  Reference semantics per matrix: core/src/Morpheus_Multiply.hpp:
57-78; the reference's own multiply tests (core/tests/Test_Multiply.hpp:
193-271) check one matrix and one vector at a time.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi, device
from morpheus_amd import generators as gen
from morpheus_amd.formats import CSR_FORMAT, FORMAT_NAMES
from tests.helpers import fp64_bound, max_rel_err

pytestmark = pytest.mark.gpu

K = 8
TOL = {"f32": 1e-5, "bf16": 1e-2}


@pytest.fixture(scope="module")
def workload(gpu):
    import torch
    mats = [A for _, A in gen.batch_set(512)]
    rng = np.random.default_rng(88)
    Xh = [rng.uniform(-1.0, 1.0, (A.ncols, K)) for A in mats]
    # the autotuner's format per matrix (FP32 trials with k = 8, as the
    # benchmark), and every admissible format in turn
    D32 = [device.to_device(A, "f32") for A in mats]
    tuned = [device.autotune_format(D, k=K, reps=1)[0] for D in D32]
    for D in D32:
        D.close()
    turn = [(i % 6) for i in range(len(mats))]
    torch.cuda.synchronize()
    return mats, Xh, {"autotuned": tuned, "every format": turn}


def _batch_members(mats, fmts, dtype):
    out, used = [], []
    for A, f in zip(mats, fmts):
        D = device.to_device(A, dtype)
        if f != CSR_FORMAT:
            try:
                M = device.convert(D, f)
                D.close()
                D = M
            except capi.FormatConversionException:
                pass  # inadmissible for this matrix: it stays CSR
        out.append(D)
        used.append(D.format)
    return out, used


@pytest.mark.parametrize("assign", ["autotuned", "every format"])
@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_batch_512_k8(gpu, workload, assign, dtype):
    import torch
    mats, Xh, assignments = workload
    F, used = _batch_members(mats, assignments[assign], dtype)
    vdt = device.vector_dtype(dtype)
    X = [torch.from_numpy(x).cuda().to(vdt) for x in Xh]
    Y = [torch.full((D.nrows, K), float("nan"), dtype=vdt, device="cuda") for D in F]
    B = device.Batch(F, K, X, Y)
    hist = {FORMAT_NAMES[f]: used.count(f) for f in sorted(set(used))}
    print(f"{dtype} {assign}: formats {hist}", flush=True)
    if assign == "every format":
        assert len(hist) == 6, hist
    worst = 0.0
    for run in (B.run, B.run_graph):
        for y in Y:
            y.fill_(float("nan"))
        run(True)
        torch.cuda.synchronize()
        for A, x, y, f in zip(mats, Xh, Y, used):
            Yh = y.double().cpu().numpy()
            nrow = np.diff(A.row_offsets)
            for j in (0, K - 1):
                xj = np.ascontiguousarray(x[:, j])
                if dtype == "f64":
                    ref = oracle.spmv(A, xj)
                    bad = np.count_nonzero(~(np.abs(Yh[:, j] - ref) <=
                                             fp64_bound(nrow, oracle.absrow(A, xj))))
                    assert bad == 0, f"{FORMAT_NAMES[f]} member: {bad} rows outside the FP64 bound"
                else:
                    err = max_rel_err(Yh[:, j], oracle.spmv_compensated(A, xj),
                                      oracle.absrow(A, xj))
                    worst = max(worst, err)
                    assert err <= TOL[dtype], f"{FORMAT_NAMES[f]} member: {err:.3e}"
    print(f"{dtype} {assign}: max rel err {worst:.3e}", flush=True)
    B.close()
    for D in F:
        D.close()
