"""GPU parity of update_diagonal / get_diagonal through the C-ABI: the
updated containers equal the oracle's bit for bit in every format and
value type, the quirks included, and a product after an update uses the
new values (the x-tiled plan, which copies values, is rebuilt)."""
import numpy as np
import pytest

import oracle
from morpheus_amd import device
from morpheus_amd import generators as gen
from morpheus_amd.formats import (COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                                  FORMAT_NAMES, HDC_FORMAT, HYB_FORMAT, Coo, Csr, Dia, Ell)
from tests.helpers import fp64_bound, same_container

pytestmark = pytest.mark.gpu
FORMATS = [COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT]


def _torch():
    import torch
    return torch


def check_update(H, d, dtype="f64"):
    torch = _torch()
    D = device.to_device(H, dtype)
    dd = torch.from_numpy(np.ascontiguousarray(d)).cuda().to(device.vector_dtype(dtype))
    device.update_diagonal(D, dd)
    torch.cuda.synchronize()
    got = D.to_host()
    # the oracle works in FP64: feed it the stored (rounded) values
    ref = oracle.update_diagonal(device.to_device(H, dtype).to_host(),
                                 dd.double().cpu().numpy() if dtype == "f64" else
                                 dd.to(torch.bfloat16 if dtype == "bf16" else torch.float32)
                                 .double().cpu().numpy())
    assert same_container(got, ref)
    g = device.get_diagonal(D).double().cpu().numpy()
    assert np.allclose(g, oracle.get_diagonal(ref), rtol=1e-6 if dtype != "f64" else 0, atol=0)
    D.close()


@pytest.mark.parametrize("n", [10, 100, 1000])
@pytest.mark.parametrize("fmt", FORMATS, ids=FORMAT_NAMES)
def test_reference_matrices(gpu, n, fmt):
    A = gen.diagonal_test_matrix(n, seed=n)
    check_update(oracle.convert(A, fmt), np.random.default_rng(n).uniform(-1, 1, n))


@pytest.mark.parametrize("fmt", FORMATS, ids=FORMAT_NAMES)
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_value_types(gpu, fmt, dtype):
    A = gen.stencil3d_7(16)
    check_update(oracle.convert(A, fmt), np.random.default_rng(7).uniform(-1, 1, A.nrows), dtype)


def test_quirk_containers(gpu):
    check_update(Csr(2, 2, [0, 3, 4], [0, 0, 1, 1], [1.0, 2.0, 3.0, 4.0]), [9.0, 8.0])
    check_update(Coo(2, 2, [0, 0, 1], [0, 0, 1], [1.0, 2.0, 4.0]), [9.0, 8.0])
    off = np.array([0, 1], np.int32)
    check_update(Dia(3, 3, 5, off, np.array([1.0, 0.0, 3.0, 0.0, 5.0, 6.0, 7.0, 0.0]), 4),
                 [9.0, 8.0, 7.0])
    check_update(Ell(2, 2, 3, 2, 2, np.array([0, 1, 0, -1], np.int32),
                     np.array([1.0, 2.0, 3.0, 0.0]), 2), [9.0, 8.0])


@pytest.mark.parametrize("fmt", [CSR_FORMAT, HYB_FORMAT], ids=["CSR", "HYB"])
def test_product_after_update_uses_new_values(gpu, fmt):
    """Long rows with sorted columns and a wide x take the x-tiled pass,
    whose copy of the values must be rebuilt after the update."""
    torch = _torch()
    P = gen.powerlaw_rows(3000, 300000, 1.5, 40.0, 20000, seed=4)
    # a diagonal entry in every row that has none, so the update touches
    # every row (long ones included)
    r = np.repeat(np.arange(P.nrows), np.diff(P.row_offsets))
    c = P.column_indices.astype(np.int64)
    has = np.zeros(P.nrows, bool)
    has[r[r == c]] = True
    add = np.nonzero(~has)[0]
    r = np.concatenate([r, add])
    c = np.concatenate([c, add])
    v = np.concatenate([P.values, np.ones(add.size)])
    o = np.lexsort((c, r))
    ro = np.zeros(P.nrows + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=P.nrows), out=ro[1:])
    A = Csr(P.nrows, P.ncols, ro, c[o], v[o])
    H = oracle.convert(A, fmt)
    D = device.to_device(H, "f64")
    D.set_tuning(xtile=2)
    x = gen.random_vector(A.ncols, 3)
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    D.multiply(xt, y)
    d = np.random.default_rng(1).uniform(5, 6, A.nrows)
    device.update_diagonal(D, torch.from_numpy(d).cuda())
    D.multiply(xt, y)
    torch.cuda.synchronize()
    Hn = oracle.update_diagonal(H, d)
    An = oracle.convert(oracle.to_coo(Hn), CSR_FORMAT)
    ref = oracle.spmv(An, x)
    bound = fp64_bound(np.diff(An.row_offsets), oracle.absrow(An, x))
    assert np.all(np.abs(y.cpu().numpy() - ref) <= bound)
    D.close()
