"""mph_coo_multiply_noalloc (no counterpart in the reference; for callers
that cannot allocate) against the existing one-shot mph_coo_multiply, both
through the C-ABI: two calls in a row, the second on a smaller matrix, in
FP64 (integer-valued data: equal bit for bit; random data: both within the
per-row FP64 bound of the oracle), FP32 and BF16 (within the FP32 bounds of
the oracle); accumulation (init = 0); the size limit rejected with a clear
error before any device work; and a capture into a hipGraph under torch's
global capture mode, which fails if the call allocates or synchronises,
replayed three times with eager work between the replays.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from morpheus_amd import capi, device
from morpheus_amd import generators as gen
from morpheus_amd.formats import Coo
from tests.helpers import U32, fp64_bound, plain_fp32_bound

pytestmark = pytest.mark.gpu

NOALLOC_MAX_NNZ = 268_435_456


def _coo(A):
    return oracle.to_coo(A)


def _arrays(Co, dtype):
    import torch
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[dtype]
    ri = torch.from_numpy(np.ascontiguousarray(Co.row_indices, dtype=np.int32)).cuda()
    ci = torch.from_numpy(np.ascontiguousarray(Co.column_indices, dtype=np.int32)).cuda()
    v = torch.from_numpy(np.ascontiguousarray(Co.values)).cuda().to(tdt)
    return ri, ci, v


def _call(name, Co, dtype, x, y, init, stream=None):
    ri, ci, v = _arrays(Co, dtype)
    capi.call(name, Co.nrows, Co.ncols, int(ri.numel()), C.c_void_p(ri.data_ptr()),
              C.c_void_p(ci.data_ptr()), C.c_void_p(v.data_ptr()), capi.VALUE_TYPES[dtype],
              C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), int(init),
              device.stream_handle(stream))
    return ri, ci, v  # keep the arrays alive until the caller synchronises


def _pair(A, dtype, x_host, init=True, y0=None):
    """(no-alloc result, one-shot result) as float64 host arrays."""
    import torch
    vdt = device.vector_dtype(dtype)
    Co = _coo(A)
    x = torch.from_numpy(x_host).cuda().to(vdt)
    out = []
    for name in ("mph_coo_multiply_noalloc", "mph_coo_multiply"):
        y = (torch.full((A.nrows,), float("nan"), dtype=vdt, device="cuda") if y0 is None
             else torch.from_numpy(y0).cuda().to(vdt))
        keep = _call(name, Co, dtype, x, y, init)
        torch.cuda.synchronize()
        del keep
        out.append(y.double().cpu().numpy())
    return out


def _int_valued(A, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(-4, 5, A.nnz).astype(np.float64)
    v[v == 0] = 1.0
    return type(A)(A.nrows, A.ncols, A.row_offsets, A.column_indices, v)


def test_noalloc_equals_one_shot_fp64(gpu):
    big = gen.powerlaw_rows(60000, 60000, 1.5, 16.0, 5000, 3)   # long rows cut by chunks
    small = gen.poisson2d(40)                                    # the smaller second call
    for A, seed in ((big, 1), (small, 2)):
        Ai = _int_valued(A, seed)
        x = np.random.default_rng(seed).integers(-3, 4, A.ncols).astype(np.float64)
        na, os_ = _pair(Ai, "f64", x)
        assert np.array_equal(na, os_) and np.array_equal(na, oracle.spmv(Ai, x))
        xr = gen.random_vector(A.ncols, seed)
        na, os_ = _pair(A, "f64", xr)
        bound = fp64_bound(np.diff(A.row_offsets), oracle.absrow(A, xr))
        ref = oracle.spmv(A, xr)
        assert np.all(np.abs(na - ref) <= bound) and np.all(np.abs(os_ - ref) <= bound)
        assert np.all(np.abs(na - os_) <= 2 * bound)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_noalloc_equals_one_shot_reduced(gpu, dtype):
    for A, seed in ((gen.powerlaw_rows(40000, 50000, 1.5, 12.0, 3000, 5), 5),
                    (gen.banded(3000, 4, 0.02, 6), 6)):
        x = gen.random_vector(A.ncols, seed)
        na, os_ = _pair(A, dtype, x)
        # the product of the stored (rounded) values and x
        vq = device.to_device(A, dtype).to_host().values
        Aq = type(A)(A.nrows, A.ncols, A.row_offsets, A.column_indices, vq)
        xq = x.astype(np.float32).astype(np.float64)
        ref = oracle.spmv_compensated(Aq, xq)
        bound = plain_fp32_bound(np.diff(A.row_offsets), oracle.absrow(Aq, xq))
        assert np.all(np.abs(na - ref) <= bound) and np.all(np.abs(os_ - ref) <= bound)


def test_noalloc_accumulate(gpu):
    A = gen.powerlaw_rows(20000, 20000, 1.5, 10.0, 2000, 8)
    Ai = _int_valued(A, 8)
    x = np.random.default_rng(8).integers(-3, 4, A.ncols).astype(np.float64)
    y0 = np.random.default_rng(9).integers(-5, 6, A.nrows).astype(np.float64)
    na, os_ = _pair(Ai, "f64", x, init=False, y0=y0)
    assert np.array_equal(na, os_) and np.array_equal(na, y0 + oracle.spmv(Ai, x))


def test_noalloc_rejects_beyond_limit(gpu):
    st = capi.lib().mph_coo_multiply_noalloc(1000, 1000, NOALLOC_MAX_NNZ + 1, None, None, None,
                                             capi.VALUE_TYPES["f64"], None, None, 1, None)
    assert st == 2  # MPH_ERR_INVALID_INPUT, before any device work
    msg = capi.lib().mph_last_error()
    assert b"exceed" in msg and b"268435456" in msg


def test_noalloc_graph_capture(gpu):
    import torch
    A = gen.powerlaw_rows(30000, 30000, 1.5, 12.0, 3000, 12)
    Ai = _int_valued(A, 12)
    Co = _coo(Ai)
    ri, ci, v = _arrays(Co, "f64")
    x = torch.from_numpy(np.random.default_rng(12).integers(-3, 4, A.ncols).astype(np.float64)).cuda()
    y = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    ref = oracle.spmv(Ai, x.cpu().numpy())
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="global"):
        capi.call("mph_coo_multiply_noalloc", Co.nrows, Co.ncols, Co.nnz,
                  C.c_void_p(ri.data_ptr()), C.c_void_p(ci.data_ptr()), C.c_void_p(v.data_ptr()),
                  capi.VALUE_TYPES["f64"], C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), 1,
                  device.stream_handle(torch.cuda.current_stream()))
    for rep in range(3):
        y.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), ref), f"replay {rep}"
        # eager work between the replays: the one-shot call on the same data
        y2 = torch.empty_like(y)
        capi.call("mph_coo_multiply", Co.nrows, Co.ncols, Co.nnz, C.c_void_p(ri.data_ptr()),
                  C.c_void_p(ci.data_ptr()), C.c_void_p(v.data_ptr()), capi.VALUE_TYPES["f64"],
                  C.c_void_p(x.data_ptr()), C.c_void_p(y2.data_ptr()), 1, None)
        torch.cuda.synchronize()
        assert np.array_equal(y2.cpu().numpy(), ref)
    del g
