"""HDC from 65,536 rows with the CSR part added by the DIA pass.

Reference: core/src/impl/Hdc/Serial/Morpheus_Multiply_Impl.hpp (the DIA
product, then the CSR product accumulating onto it).

A split FP64 HDC whose CSR rows all hold at most 8 entries runs one pass: a
thread sums its row's diagonals (x staged in LDS per cluster) and then its
row's CSR entries in order, instead of the DIA pass followed by a list pass
that read-modify-writes the scattered y entries of the rows with CSR
entries.  MPH_HDC_TAIL=0 keeps the two passes.  Both must stay within the
FP64 per-row bound of the oracle, for y = A x and y += A x; a CSR row of more
than 8 entries keeps the two passes.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import generators as gen
from morpheus_amd.formats import HDC_FORMAT, Csr
from tests.helpers import fp64_bound

pytestmark = pytest.mark.gpu


def _band():
    return gen.banded(120000, 4, 0.02, 5)


def _band_long_row():
    A = gen.banded(120000, 4, 0.02, 6)
    # row 777 gains 40 entries far from the band
    rows = [A.column_indices[A.row_offsets[r]:A.row_offsets[r + 1]] for r in range(A.nrows)]
    vals = [A.values[A.row_offsets[r]:A.row_offsets[r + 1]] for r in range(A.nrows)]
    extra = np.arange(50000, 50080, 2)
    rows[777] = np.union1d(rows[777], extra)
    vals[777] = np.random.default_rng(7).uniform(-1, 1, rows[777].size)
    ro = np.concatenate([[0], np.cumsum([r.size for r in rows])])
    return Csr(A.nrows, A.ncols, ro, np.concatenate(rows), np.concatenate(vals))


def _run(monkeypatch, H, x, y0, init, tail):
    import torch
    from morpheus_amd import device
    if tail:
        monkeypatch.delenv("MPH_HDC_TAIL", raising=False)
    else:
        monkeypatch.setenv("MPH_HDC_TAIL", "0")
    D = device.to_device(H, "f64")
    xt = torch.from_numpy(x).cuda()
    yt = torch.from_numpy(y0.copy()).cuda()
    D.multiply(xt, yt, init)
    torch.cuda.synchronize()
    out = yt.cpu().numpy()
    D.close()
    monkeypatch.delenv("MPH_HDC_TAIL", raising=False)
    return out


@pytest.mark.parametrize("mat", ["band", "band with a long row"])
@pytest.mark.parametrize("init", [True, False])
def test_hdc_tail(gpu, monkeypatch, mat, init):
    A = _band() if mat == "band" else _band_long_row()
    H = oracle.convert(A, HDC_FORMAT)
    assert H.csr.nnz > 0 and H.dia.ndiag > 1
    x = gen.random_vector(A.ncols, 81)
    y0 = gen.random_vector(A.nrows, 82)
    ref = oracle.spmv(A, x, y0, init)
    bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, x), None if init else y0)
    for tail in (True, False):
        y = _run(monkeypatch, H, x, y0, init, tail)
        bad = np.count_nonzero(~(np.abs(y - ref) <= bound))
        assert bad == 0, f"tail={tail}: {bad} rows outside the FP64 bound"
