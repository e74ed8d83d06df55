"""Seeded random matrices through every format's plan, against the oracle.

Each case draws a shape (1 to 150k rows, x from a few entries to 2.4 MB so
that the x-tiled pass both runs and does not), a row-length law (uniform
short rows, power law with long rows, banded with off-band entries, rows
left empty), explicit zeros, a launch shape (default, fixed lanes, chunk,
rows per thread, x-tiling forced on or off) and init / accumulate; the
product of each of the six formats the reference admits for the matrix is
checked in FP64 against the per-row bound 4 (nnz_row + 1) u sum |a x| and
in FP32 / BF16 against max_i |y_i - y_ref_i| / sum_j |a_ij x_j| <= 1e-5 /
1e-2 (the compensated FP64 oracle on the original matrix)."""
import numpy as np
import pytest

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  HDC_FORMAT, HYB_FORMAT, Csr)
from tests.helpers import fp64_bound, max_rel_err

pytestmark = pytest.mark.gpu

FORMATS = [CSR_FORMAT, COO_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT]
TUNINGS = [dict(), dict(lanes=1), dict(lanes=16), dict(rows_chunk=512),
           dict(rows_per_thread=1), dict(xtile=1), dict(xtile=2)]


def random_case(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([1, 7, 300, 5000, 70000, 150000]))
    m = int(rng.choice([n, max(1, n // 3), 2 * n + 5, 300000]))
    kind = seed % 4
    if kind == 0:
        lens = rng.integers(0, 6, n)
    elif kind == 1:
        lens = gen.powerlaw_row_lengths(n, 1.5, 6.0, min(m, 3000), rng)
    elif kind == 2:
        lens = np.full(n, 9)
    else:
        lens = np.where(rng.random(n) < 0.3, 0, rng.integers(1, 40, n))
    lens = np.minimum(lens, m).astype(np.int64)
    rows = []
    for r in range(n):
        if lens[r] == 0:
            rows.append(np.empty(0, np.int64))
        elif kind == 2:  # a band around the diagonal, sometimes one off-band entry
            band = np.clip(r + np.arange(-4, 5), 0, m - 1)
            extra = rng.integers(0, m, 1) if rng.random() < 0.1 else np.empty(0, np.int64)
            rows.append(np.unique(np.concatenate([band, extra])))
        else:
            rows.append(np.sort(rng.choice(m, int(lens[r]), replace=False)))
    ro = np.concatenate([[0], np.cumsum([c.size for c in rows])]).astype(np.int32)
    cols = (np.concatenate(rows) if n else np.empty(0)).astype(np.int32)
    vals = rng.standard_normal(cols.size)
    vals[rng.random(cols.size) < 0.02] = 0.0  # explicit zeros
    return Csr(n, m, ro, cols, vals), rng


def gpu_product(Ah, x, dtype, y0, init, tuning):
    import torch
    from morpheus_amd import device
    D = device.to_device(Ah, dtype)
    if tuning:
        D.set_tuning(**tuning)
    vdt = device.vector_dtype(dtype)
    xt = torch.from_numpy(x).to("cuda").to(vdt)
    yt = (torch.full((Ah.nrows,), float("nan"), dtype=vdt, device="cuda") if init
          else torch.from_numpy(y0).to("cuda").to(vdt))
    D.multiply(xt, yt, init)
    torch.cuda.synchronize()
    out = yt.cpu().double().numpy()
    D.close()
    return out


@pytest.mark.parametrize("seed", range(40))
def test_random_matrix_all_formats(gpu, seed):
    A, rng = random_case(seed)
    x = rng.uniform(-1, 1, A.ncols)
    y0 = rng.uniform(-1, 1, A.nrows)
    init = bool(seed % 3)
    tuning = TUNINGS[seed % len(TUNINGS)]
    print(f"seed {seed}: {A.nrows} x {A.ncols}, nnz {A.nnz}, init {init}, tuning {tuning}",
          flush=True)
    ref = oracle.spmv(A, x, None if init else y0, init)
    refc = oracle.spmv_compensated(A, x) + (0 if init else y0)
    absrow = oracle.absrow(A, x)
    bound = fp64_bound(np.diff(A.row_offsets), absrow, None if init else y0)
    scale = absrow + (0 if init else np.abs(y0))
    for f in FORMATS:
        try:
            H = oracle.convert(A, f)
        except (oracle.FormatConversion, oracle.OracleError):
            # inadmissible, or (HDC of a non-square matrix under 5 rows) a
            # case where the reference's converter is undefined
            continue
        y = gpu_product(H, x, "f64", y0, init, tuning)
        assert np.all(np.abs(y - ref) <= bound), (f, seed)
        for dtype, tol in (("f32", 1e-5), ("bf16", 1e-2)):
            y = gpu_product(H, x, dtype, y0, init, tuning)
            assert max_rel_err(y, refc, scale) <= tol, (f, dtype, seed)
