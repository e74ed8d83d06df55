"""GPU parity of the dense-vector algorithms (Morpheus::dot / waxpby /
reduce) and of the fused CG iteration through the C-ABI, against the
oracle: waxpby bit for bit, dot and reduce within the summation bounds,
CG iterates within a tight relative tolerance of the textbook iteration
(they differ only in summation order) and to convergence; repeat calls and
hipGraph replays give the same bits."""
import math

import numpy as np
import pytest

import oracle
from morpheus_amd import device
from morpheus_amd import generators as gen
from morpheus_amd.formats import CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HDC_FORMAT, HYB_FORMAT

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def _torch():
    import torch
    return torch


def _dev(a, dtype=None):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda().to(dtype or torch.float64)


@pytest.mark.parametrize("n", [0, 1, 1000, 3_000_001])
def test_dot_reduce_waxpby_f64(gpu, n):
    torch = _torch()
    rng = np.random.default_rng(n + 1)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    xd, yd = _dev(x), _dev(y)
    d = device.dot(xd, yd)
    assert abs(d - math.fsum(x * y)) <= 2 * max(n, 1) * U * np.sum(np.abs(x * y)) + 1e-300
    assert device.dot(xd, yd) == d                       # same bits every call
    s = device.reduce(xd)
    assert abs(s - math.fsum(x)) <= 4 * U * np.sum(np.abs(x)) + 1e-300
    w = torch.empty_like(xd)
    device.waxpby(2.5, xd, -0.75, yd, w)
    assert np.array_equal(w.cpu().numpy(), oracle.waxpby(2.5, x, -0.75, y))
    device.waxpby(1.0, xd, 3.0, yd, w)
    assert np.array_equal(w.cpu().numpy(), oracle.waxpby(1.0, x, 3.0, y))


def test_dot_reduce_f32(gpu):
    torch = _torch()
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(100_000).astype(np.float32), rng.standard_normal(100_000).astype(np.float32)
    xd, yd = _dev(x, torch.float32), _dev(y, torch.float32)
    exact = math.fsum(x.astype(np.float64) * y.astype(np.float64))
    assert abs(device.dot(xd, yd) - exact) <= 1e-12 * np.sum(np.abs(x * y))
    w = torch.empty_like(xd)
    device.waxpby(0.5, xd, 2.0, yd, w)
    assert np.array_equal(w.cpu().numpy(), np.float32(0.5) * x + np.float32(2.0) * y)


@pytest.fixture(scope="module")
def laplacian():
    return gen.stencil3d_7(32)   # SPD, 32768 rows


@pytest.mark.parametrize("fmt", [CSR_FORMAT, DIA_FORMAT, ELL_FORMAT, HYB_FORMAT, HDC_FORMAT],
                         ids=["CSR", "DIA", "ELL", "HYB", "HDC"])
def test_cg_matches_textbook_iteration(gpu, laplacian, fmt):
    torch = _torch()
    A = laplacian
    rng = np.random.default_rng(12)
    b = rng.standard_normal(A.nrows)
    x0 = rng.standard_normal(A.nrows) * 0.1
    D = device.to_device(oracle.convert(A, fmt), "f64")
    cg = device.CG(D)
    bd, xd = _dev(b), _dev(x0)
    cg.start(bd, xd)
    cg.iterate(xd, 25)
    rr = cg.residual()
    xr, rr_ref = oracle.cg(A, b, x0, 25)
    got = xd.cpu().numpy()
    assert np.linalg.norm(got - xr) <= 1e-10 * np.linalg.norm(xr)
    assert abs(rr - rr_ref) <= 1e-8 * rr_ref
    cg.close()
    D.close()


def test_cg_converges(gpu, laplacian):
    torch = _torch()
    A = laplacian
    b = np.random.default_rng(2).standard_normal(A.nrows)
    D = device.to_device(A, "f64")
    cg = device.CG(D)
    bd = _dev(b)
    xd = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
    cg.start(bd, xd)
    for _ in range(20):
        cg.iterate(xd, 10)
        if math.sqrt(cg.residual()) <= 1e-10 * np.linalg.norm(b):
            break
    res = b - oracle.spmv(A, xd.cpu().numpy())
    assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(b)
    cg.close()
    D.close()


def test_cg_graph_replays_continue_the_iteration(gpu, laplacian):
    """One captured iteration replayed 12 times equals 12 eager iterations,
    bit for bit (the scalars live on the device; no memset nodes)."""
    torch = _torch()
    A = laplacian
    b = np.random.default_rng(8).standard_normal(A.nrows)
    D = device.to_device(A, "f64")
    D.analyse()
    eager, graphed = device.CG(D), device.CG(D)
    bd = _dev(b)
    xe = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
    xg = torch.zeros_like(xe)
    eager.start(bd, xe)
    eager.iterate(xe, 12)
    graphed.start(bd, xg)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        graphed.iterate(xg, 1)
    # capture records without running: x and the state are still at the start
    for _ in range(12):
        g.replay()
        torch.cuda.synchronize()
    assert np.array_equal(xe.cpu().numpy().view(np.uint64), xg.cpu().numpy().view(np.uint64))
    assert eager.residual() == graphed.residual()
    eager.close()
    graphed.close()
    D.close()


def test_cg_f32(gpu, laplacian):
    torch = _torch()
    A = laplacian
    b = np.random.default_rng(3).standard_normal(A.nrows)
    D = device.to_device(A, "f32")
    cg = device.CG(D)
    bd = _dev(b, torch.float32)
    xd = torch.zeros(A.nrows, dtype=torch.float32, device="cuda")
    cg.start(bd, xd)
    cg.iterate(xd, 150)
    # FP32 vectors: the attainable residual is about u_f32 * cond(A) ~ 3e-5
    res = b - oracle.spmv(A, xd.double().cpu().numpy())
    assert np.linalg.norm(res) <= 1e-4 * np.linalg.norm(b)
    cg.close()
    D.close()
