"""Values written in place and the analysis' value copies.

Reference containers hold no derived state
(core/src/Morpheus_CsrMatrix.hpp:148-163): a product always reads the
current values.  Through the C-ABI a handle's analysis may keep copies of
values (the x-tiled long rows' tile-major copy, a HYB's merged CSR copy);
with the policy MPH_POLICY_NO_VALUE_COPIES (which the C++ drop-in sets) it
keeps none (the x-tiled pass then reads the values in place), so values written in place are seen by the next product without
a new analysis.  With the default policy the same product is right after
mph_matrix_analyse.  Each product is checked against the Serial oracle on
the rewritten matrix with the FP64 per-row bound.
"""
import numpy as np
import pytest

import oracle
from morpheus_amd import capi
from morpheus_amd import generators as gen
from morpheus_amd.formats import CSR_FORMAT, HYB_FORMAT, Coo, Ell, Hyb
from tests.helpers import fp64_bound
from tests.test_multiply_gpu import holey_hyb

pytestmark = pytest.mark.gpu


def _wide_csr():
    # x of 3.2 MB (FP64) and rows of up to 30k entries: x-tiled by default
    return gen.powerlaw_rows(20000, 400000, 1.5, 40.0, 30000, seed=31)


def _scaled(H, s):
    """The host container with every value multiplied by s."""
    if isinstance(H, Hyb):
        e, c = H.ell, H.coo
        return Hyb(Ell(e.nrows, e.ncols, e.nnz, e.entries_per_row, e.width, e.column_indices,
                       e.values * s, e.pitch),
                   Coo(c.nrows, c.ncols, c.row_indices, c.column_indices, c.values * s))
    return type(H)(H.nrows, H.ncols, H.row_offsets, H.column_indices, H.values * s)


def _check(H, x, y):
    C = oracle.convert(H, CSR_FORMAT)
    ref = oracle.spmv(C, x)
    bound = fp64_bound(np.diff(C.row_offsets), oracle.absrow(C, x))
    bad = np.count_nonzero(~(np.abs(y - ref) <= bound))
    assert bad == 0, f"{bad} rows outside the FP64 bound"


def _value_tensors(D):
    return [D.arrays[k] for k in ("v", "ev", "cv") if k in D.arrays]


@pytest.mark.parametrize("case", ["CSR x-tiled", "HYB merged"])
@pytest.mark.parametrize("policy", ["no copies", "default + analyse"])
def test_values_written_in_place(gpu, case, policy):
    import torch
    from morpheus_amd import device
    H = _wide_csr() if case == "CSR x-tiled" else holey_hyb(70000, 0.2, 3)
    D = device.to_device(H, "f64")
    if policy == "no copies":
        D.set_policy(capi.POLICY_NO_VALUE_COPIES)
    D.analyse()
    x = gen.random_vector(H.ncols, 7)
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(H.nrows, dtype=torch.float64, device="cuda")
    D.multiply(xt, y)
    torch.cuda.synchronize()
    _check(H, x, y.cpu().numpy())
    # rewrite every value in place (same arrays, no new handle)
    for t in _value_tensors(D):
        t.mul_(-3.0)
    if policy != "no copies":
        D.analyse()
    D.multiply(xt, y)
    torch.cuda.synchronize()
    _check(_scaled(H, -3.0), x, y.cpu().numpy())
    D.close()


@pytest.mark.parametrize("case", ["CSR x-tiled", "HYB merged"])
def test_policy_reads_no_copy(gpu, case):
    """The no-copies policy replaces the x-tiled copy by the direct pass and
    gives the merged HYB plan up, and stays within the bound; switching it
    back restores the default plan."""
    import torch
    from morpheus_amd import device
    H = _wide_csr() if case == "CSR x-tiled" else holey_hyb(70000, 0.2, 3)
    D = device.to_device(H, "f64")
    x = gen.random_vector(H.ncols, 8)
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(H.nrows, dtype=torch.float64, device="cuda")
    out = {}
    for pol in (0, capi.POLICY_NO_VALUE_COPIES, 0):
        D.set_policy(pol)
        D.multiply(xt, y)
        torch.cuda.synchronize()
        _check(H, x, y.cpu().numpy())
        out.setdefault(pol, []).append(y.cpu().numpy().copy())
    # the default plan is rebuilt identically: same bits
    assert np.array_equal(out[0][0], out[0][1])
    D.close()
