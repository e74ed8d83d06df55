"""Device containers and the Python mirror of the reference front doors.

Device memory is held in torch tensors (plumbing only); every computation
goes through the C-ABI of ``libmorpheus_amd.so``:

* ``DeviceMatrix``  -- a matrix in one of the six formats resident in HBM,
  with its analysis handle (reference containers over ``HIPSpace``);
* ``multiply(A, x, y, init=True)`` -- ``Morpheus::multiply``
  (core/src/Morpheus_Multiply.hpp:57-78);
* ``convert(A, fmt)`` -- ``Morpheus::convert`` on the GPU, through a
  temporary COO exactly like the reference hub
  (core/src/impl/Morpheus_Convert_Impl.hpp:83-106).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi
from .formats import (ALIGNMENT, COO_FORMAT, CSR_FORMAT, DIA_FORMAT, ELL_FORMAT,
                      HDC_FORMAT, HYB_FORMAT, Coo, Csr, Dia, Ell, Hdc, Hyb, pad_size)

DTYPES = {
    "f64": (torch.float64, torch.float64),
    "f32": (torch.float32, torch.float32),
    "bf16": (torch.bfloat16, torch.float32),
}


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else None


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def vector_dtype(dtype: str):
    return DTYPES[dtype][1]


class DeviceMatrix:
    """A sparse matrix resident on the GPU plus its C-ABI handle."""

    def __init__(self, fmt: int, nrows: int, ncols: int, nnz: int, dtype: str,
                 arrays: dict, meta: dict):
        self.format = fmt
        self.nrows, self.ncols, self.nnz = int(nrows), int(ncols), int(nnz)
        self.dtype = dtype
        self.vt = capi.VALUE_TYPES[dtype]
        self.arrays = arrays
        self.meta = meta
        self.handle = C.c_void_p()
        self._create()

    # -------------------------------------------------------------- handle
    def _create(self):
        a, m, vt = self.arrays, self.meta, self.vt
        h = C.byref(self.handle)
        if self.format == COO_FORMAT:
            capi.call("mph_matrix_create_coo", h, self.nrows, self.ncols, self.nnz,
                      _p(a["ri"]), _p(a["ci"]), _p(a["v"]), vt)
        elif self.format == CSR_FORMAT:
            capi.call("mph_matrix_create_csr", h, self.nrows, self.ncols, self.nnz,
                      _p(a["ro"]), _p(a["ci"]), _p(a["v"]), vt)
        elif self.format == DIA_FORMAT:
            capi.call("mph_matrix_create_dia", h, self.nrows, self.ncols, self.nnz,
                      m["ndiag"], m["pitch"], _p(a["off"]), _p(a["v"]), vt)
        elif self.format == ELL_FORMAT:
            capi.call("mph_matrix_create_ell", h, self.nrows, self.ncols, self.nnz,
                      m["K"], m["width"], m["pitch"], _p(a["ci"]), _p(a["v"]), vt)
        elif self.format == HYB_FORMAT:
            capi.call("mph_matrix_create_hyb", h, self.nrows, self.ncols, m["ell_nnz"],
                      m["K"], m["width"], m["pitch"], _p(a["eci"]), _p(a["ev"]),
                      m["coo_nnz"], _p(a["cri"]), _p(a["cci"]), _p(a["cv"]), vt)
        elif self.format == HDC_FORMAT:
            capi.call("mph_matrix_create_hdc", h, self.nrows, self.ncols, m["dia_nnz"],
                      m["ndiag"], m["pitch"], _p(a["off"]), _p(a["dv"]),
                      m["csr_nnz"], _p(a["ro"]), _p(a["ci"]), _p(a["cv"]), vt)
        else:
            raise capi.InvalidInputException(f"unknown format {self.format}")

    def close(self):
        if self.handle:
            capi.lib().mph_matrix_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def analyse(self, stream=None):
        capi.call("mph_matrix_analyse", self.handle, stream_handle(stream))

    def set_tuning(self, rows_chunk=0, lanes=0, coo_items=0, rows_per_thread=0, xtile=0):
        t = capi.Tuning(rows_chunk, lanes, coo_items, rows_per_thread, xtile)
        capi.call("mph_matrix_set_tuning", self.handle, C.byref(t))

    def set_policy(self, flags: int):
        """mph_matrix_set_policy: capi.POLICY_NO_VALUE_COPIES keeps every
        product on the caller's value arrays (the x-tiled pass reads them in
        place; no HYB merged copy)."""
        capi.call("mph_matrix_set_policy", self.handle, int(flags))

    def get_tuning(self):
        t = capi.Tuning()
        capi.call("mph_matrix_get_tuning", self.handle, C.byref(t))
        return dict(rows_chunk=t.rows_chunk, lanes=t.lanes, coo_items=t.coo_items,
                    rows_per_thread=t.rows_per_thread, xtile=t.xtile)

    # -------------------------------------------------------------- products
    def multiply(self, x, y, init=True, stream=None):
        capi.call("mph_multiply", self.handle, _p(x), _p(y), int(bool(init)),
                  stream_handle(stream))

    def multiply_multi(self, X, Y, init=True, stream=None):
        k = X.shape[1] if X.dim() == 2 else 1
        capi.call("mph_multiply_multi", self.handle, int(k), _p(X), _p(Y),
                  int(bool(init)), stream_handle(stream))

    def new_vector(self, n, k=1, fill=None):
        shape = (n,) if k == 1 else (n, k)
        t = torch.empty(shape, dtype=vector_dtype(self.dtype), device="cuda")
        if fill is not None:
            t.fill_(fill)
        return t

    # -------------------------------------------------------------- bytes
    def algorithmic_bytes(self, k: int = 1, init: bool = True) -> int:
        """Minimal stored representation read once + x read once + y written
        once (k right-hand sides)."""
        si = 4
        sv = {"f64": 8, "f32": 4, "bf16": 2}[self.dtype]
        sx = 8 if self.dtype == "f64" else 4
        m = self.meta
        if self.format == COO_FORMAT:
            mat = self.nnz * (2 * si + sv)
        elif self.format == CSR_FORMAT:
            mat = (self.nrows + 1) * si + self.nnz * (si + sv)
        elif self.format == DIA_FORMAT:
            mat = m["ndiag"] * si + dia_inrange(self.nrows, self.ncols, m["offsets"]) * sv
        elif self.format == ELL_FORMAT:
            mat = self.nrows * m["K"] * (si + sv)
        elif self.format == HYB_FORMAT:
            mat = self.nrows * m["K"] * (si + sv) + m["coo_nnz"] * (2 * si + sv)
        else:
            mat = (m["ndiag"] * si + dia_inrange(self.nrows, self.ncols, m["offsets"]) * sv
                   + (self.nrows + 1) * si + m["csr_nnz"] * (si + sv))
        vec = k * (self.ncols * sx + self.nrows * sx * (1 if init else 2))
        return int(mat + vec)

    def plan_bytes(self, init: bool = True) -> int:
        """Bytes the analysed plan of one k = 1 product moves (mph_matrix_plan_bytes):
        what its kernels read of the stored arrays or of the analysis' copies
        (the tiled copy of x-tiled long rows, HYB's merged copy), the
        partials written and read back, x once, y once (twice when
        accumulating); ELL padding past the last filled slot is not read."""
        b = C.c_int64(0)
        capi.call("mph_matrix_plan_bytes", self.handle, int(bool(init)), C.byref(b))
        return int(b.value)

    # -------------------------------------------------------------- host
    def to_host(self):
        a, m = self.arrays, self.meta

        def i(t):
            return t.cpu().numpy().astype(np.int32) if t is not None else np.zeros(0, np.int32)

        def f(t):
            return t.float().cpu().numpy().astype(np.float64) if self.dtype == "bf16" \
                else t.cpu().numpy().astype(np.float64)

        if self.format == COO_FORMAT:
            return Coo(self.nrows, self.ncols, i(a["ri"]), i(a["ci"]), f(a["v"]))
        if self.format == CSR_FORMAT:
            return Csr(self.nrows, self.ncols, i(a["ro"]), i(a["ci"]), f(a["v"]))
        if self.format == DIA_FORMAT:
            return Dia(self.nrows, self.ncols, self.nnz, i(a["off"]), f(a["v"]), m["pitch"])
        if self.format == ELL_FORMAT:
            return Ell(self.nrows, self.ncols, self.nnz, m["K"], m["width"], i(a["ci"]),
                       f(a["v"]), m["pitch"])
        if self.format == HYB_FORMAT:
            return Hyb(Ell(self.nrows, self.ncols, m["ell_nnz"], m["K"], m["width"],
                           i(a["eci"]), f(a["ev"]), m["pitch"]),
                       Coo(self.nrows, self.ncols, i(a["cri"]), i(a["cci"]), f(a["cv"])))
        return Hdc(Dia(self.nrows, self.ncols, m["dia_nnz"], i(a["off"]), f(a["dv"]), m["pitch"]),
                   Csr(self.nrows, self.ncols, i(a["ro"]), i(a["ci"]), f(a["cv"])))


def dia_inrange(nrows: int, ncols: int, offsets) -> int:
    tot = 0
    for o in np.asarray(offsets, dtype=np.int64):
        tot += max(0, min(nrows - max(0, -o), ncols - max(0, o)))
    return int(tot)


# ------------------------------------------------------------------ upload
def _i(a, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(device)


def _v(a, dtype, device="cuda"):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)
    return t.to(DTYPES[dtype][0])


def to_device(A, dtype: str = "f64") -> DeviceMatrix:
    """Morpheus::copy of a host container to HIPSpace (same format)."""
    if isinstance(A, Coo):
        return DeviceMatrix(COO_FORMAT, A.nrows, A.ncols, A.nnz, dtype,
                            dict(ri=_i(A.row_indices), ci=_i(A.column_indices),
                                 v=_v(A.values, dtype)), {})
    if isinstance(A, Csr):
        return DeviceMatrix(CSR_FORMAT, A.nrows, A.ncols, A.nnz, dtype,
                            dict(ro=_i(A.row_offsets), ci=_i(A.column_indices),
                                 v=_v(A.values, dtype)), {})
    if isinstance(A, Dia):
        return DeviceMatrix(DIA_FORMAT, A.nrows, A.ncols, A.nnz, dtype,
                            dict(off=_i(A.offsets), v=_v(A.values, dtype)),
                            dict(ndiag=A.ndiag, pitch=A.pitch, offsets=A.offsets.copy()))
    if isinstance(A, Ell):
        return DeviceMatrix(ELL_FORMAT, A.nrows, A.ncols, A.nnz, dtype,
                            dict(ci=_i(A.column_indices), v=_v(A.values, dtype)),
                            dict(K=A.entries_per_row, width=A.width, pitch=A.pitch))
    if isinstance(A, Hyb):
        e, c = A.ell, A.coo
        return DeviceMatrix(HYB_FORMAT, e.nrows, e.ncols, e.nnz + c.nnz, dtype,
                            dict(eci=_i(e.column_indices), ev=_v(e.values, dtype),
                                 cri=_i(c.row_indices), cci=_i(c.column_indices),
                                 cv=_v(c.values, dtype)),
                            dict(K=e.entries_per_row, width=e.width, pitch=e.pitch,
                                 ell_nnz=e.nnz, coo_nnz=c.nnz))
    if isinstance(A, Hdc):
        d, c = A.dia, A.csr
        return DeviceMatrix(HDC_FORMAT, d.nrows, d.ncols, d.nnz + c.nnz, dtype,
                            dict(off=_i(d.offsets), dv=_v(d.values, dtype),
                                 ro=_i(c.row_offsets), ci=_i(c.column_indices),
                                 cv=_v(c.values, dtype)),
                            dict(ndiag=d.ndiag, pitch=d.pitch, offsets=d.offsets.copy(),
                                 dia_nnz=d.nnz, csr_nnz=c.nnz))
    raise TypeError(type(A))


def multiply(A: DeviceMatrix, x, y, init: bool = True, stream=None) -> None:
    """Morpheus::multiply<Morpheus::HIP>(A, x, y, init)."""
    if x.dim() == 2 and x.shape[1] > 1:
        A.multiply_multi(x, y, init, stream)
    else:
        A.multiply(x, y, init, stream)


# ------------------------------------------------------------------ convert
_VSIZE = {"f64": 8, "f32": 4, "bf16": 2}


def _empty_i(n):
    return torch.empty(max(int(n), 0), dtype=torch.int32, device="cuda")


def _zeros_i(n):
    return torch.zeros(max(int(n), 0), dtype=torch.int32, device="cuda")


def _zeros_v(n, dtype):
    return torch.zeros(max(int(n), 0), dtype=DTYPES[dtype][0], device="cuda")


def to_coo(A: DeviceMatrix, stream=None) -> DeviceMatrix:
    """Any format -> COO on the GPU (the reference Serial semantics)."""
    s = stream_handle(stream)
    a, m, vt = A.arrays, A.meta, A.vt
    n, nnz = A.nrows, A.nnz
    ri, ci, v = _zeros_i(nnz), _zeros_i(nnz), _zeros_v(nnz, A.dtype)
    if A.format == COO_FORMAT:
        ri.copy_(a["ri"])
        ci.copy_(a["ci"])
        v.copy_(a["v"])
    elif A.format == CSR_FORMAT:
        capi.call("mph_convert_csr_to_coo", n, nnz, _p(a["ro"]), _p(a["ci"]), _p(a["v"]),
                  vt, _p(ri), _p(ci), _p(v), s)
    elif A.format == DIA_FORMAT:
        capi.call("mph_convert_dia_to_coo", n, A.ncols, nnz, m["ndiag"], m["pitch"],
                  _p(a["off"]), _p(a["v"]), vt, _p(ri), _p(ci), _p(v), s)
    elif A.format == ELL_FORMAT:
        capi.call("mph_convert_ell_to_coo", n, nnz, m["width"], m["pitch"], _p(a["ci"]),
                  _p(a["v"]), vt, _p(ri), _p(ci), _p(v), s)
    elif A.format == HYB_FORMAT:
        capi.call("mph_convert_hyb_to_coo", n, nnz, m["width"], m["pitch"], _p(a["eci"]),
                  _p(a["ev"]), m["coo_nnz"], _p(a["cri"]), _p(a["cci"]), _p(a["cv"]), vt,
                  _p(ri), _p(ci), _p(v), s)
    elif A.format == HDC_FORMAT:
        capi.call("mph_convert_hdc_to_coo", n, A.ncols, nnz, m["ndiag"], m["pitch"],
                  _p(a["off"]), _p(a["dv"]), m["csr_nnz"], _p(a["ro"]), _p(a["ci"]),
                  _p(a["cv"]), vt, _p(ri), _p(ci), _p(v), s)
    return DeviceMatrix(COO_FORMAT, n, A.ncols, nnz, A.dtype, dict(ri=ri, ci=ci, v=v), {})


def from_coo(A: DeviceMatrix, fmt: int, stream=None) -> DeviceMatrix:
    """COO -> any format on the GPU (the reference Serial semantics)."""
    assert A.format == COO_FORMAT
    s = stream_handle(stream)
    a, vt, dt = A.arrays, A.vt, A.dtype
    n, nc, nnz = A.nrows, A.ncols, A.nnz
    if fmt == COO_FORMAT:
        return to_coo(A, stream)
    if fmt == CSR_FORMAT:
        ro, ci, v = _zeros_i(n + 1), _zeros_i(nnz), _zeros_v(nnz, dt)
        capi.call("mph_convert_coo_to_csr", n, nnz, _p(a["ri"]), _p(a["ci"]), _p(a["v"]), vt,
                  _p(ro), _p(ci), _p(v), s)
        return DeviceMatrix(CSR_FORMAT, n, nc, nnz, dt, dict(ro=ro, ci=ci, v=v), {})
    if fmt == DIA_FORMAT:
        nd = C.c_int64()
        capi.call("mph_convert_coo_to_dia_size", n, nc, nnz, _p(a["ri"]), _p(a["ci"]),
                  C.byref(nd), s)
        pitch = pad_size(n, ALIGNMENT)
        off, v = _zeros_i(nd.value), _zeros_v(pitch * nd.value, dt)
        capi.call("mph_convert_coo_to_dia", n, nc, nnz, _p(a["ri"]), _p(a["ci"]), _p(a["v"]),
                  vt, nd.value, pitch, _p(off), _p(v), s)
        return DeviceMatrix(DIA_FORMAT, n, nc, nnz, dt, dict(off=off, v=v),
                            dict(ndiag=nd.value, pitch=pitch, offsets=off.cpu().numpy()))
    if fmt == ELL_FORMAT:
        K = C.c_int64()
        capi.call("mph_convert_coo_to_ell_size", n, nnz, _p(a["ri"]), C.byref(K), s)
        width, pitch = pad_size(K.value, ALIGNMENT), n
        ci, v = _zeros_i(width * pitch), _zeros_v(width * pitch, dt)
        capi.call("mph_convert_coo_to_ell", n, nnz, _p(a["ri"]), _p(a["ci"]), _p(a["v"]), vt,
                  K.value, width, pitch, _p(ci), _p(v), s)
        return DeviceMatrix(ELL_FORMAT, n, nc, nnz, dt, dict(ci=ci, v=v),
                            dict(K=K.value, width=width, pitch=pitch))
    if fmt == HYB_FORMAT:
        K, en, cn = C.c_int64(), C.c_int64(), C.c_int64()
        capi.call("mph_convert_coo_to_hyb_size", n, nnz, _p(a["ri"]), C.byref(K), C.byref(en),
                  C.byref(cn), s)
        width, pitch = pad_size(K.value, ALIGNMENT), n
        eci, ev = _zeros_i(width * pitch), _zeros_v(width * pitch, dt)
        cri, cci, cv = _zeros_i(cn.value), _zeros_i(cn.value), _zeros_v(cn.value, dt)
        capi.call("mph_convert_coo_to_hyb", n, nnz, _p(a["ri"]), _p(a["ci"]), _p(a["v"]), vt,
                  K.value, width, pitch, _p(eci), _p(ev), _p(cri), _p(cci), _p(cv), s)
        return DeviceMatrix(HYB_FORMAT, n, nc, nnz, dt,
                            dict(eci=eci, ev=ev, cri=cri, cci=cci, cv=cv),
                            dict(K=K.value, width=width, pitch=pitch, ell_nnz=en.value,
                                 coo_nnz=cn.value))
    if fmt == HDC_FORMAT:
        nd, dn, cn = C.c_int64(), C.c_int64(), C.c_int64()
        capi.call("mph_convert_coo_to_hdc_size", n, nc, nnz, _p(a["ri"]), _p(a["ci"]),
                  C.byref(nd), C.byref(dn), C.byref(cn), s)
        pitch = pad_size(n, ALIGNMENT)
        off, dv = _zeros_i(nd.value), _zeros_v(pitch * nd.value, dt)
        ro, ci, cv = _zeros_i(n + 1), _zeros_i(cn.value), _zeros_v(cn.value, dt)
        capi.call("mph_convert_coo_to_hdc", n, nc, nnz, _p(a["ri"]), _p(a["ci"]), _p(a["v"]),
                  vt, nd.value, pitch, _p(off), _p(dv), _p(ro), _p(ci), _p(cv), s)
        return DeviceMatrix(HDC_FORMAT, n, nc, nnz, dt, dict(off=off, dv=dv, ro=ro, ci=ci, cv=cv),
                            dict(ndiag=nd.value, pitch=pitch, offsets=off.cpu().numpy(),
                                 dia_nnz=dn.value, csr_nnz=cn.value))
    raise capi.InvalidInputException(f"unknown format {fmt}")


# ------------------------------------------------------------ matrix operations
def update_diagonal(A: DeviceMatrix, diagonal, stream=None):
    """Morpheus::update_diagonal<HIP>(A, diagonal): A's values in place."""
    capi.call("mph_matrix_update_diagonal", A.handle, _p(diagonal), stream_handle(stream))


def get_diagonal(A: DeviceMatrix, out=None, stream=None):
    """The diagonal the product sees (sum of the stored (i, i) entries)."""
    if out is None:
        out = torch.zeros(min(A.nrows, A.ncols), dtype=vector_dtype(A.dtype), device="cuda")
    capi.call("mph_matrix_get_diagonal", A.handle, _p(out), stream_handle(stream))
    return out


# ------------------------------------------------------------ vector algorithms
def _vt_of(v) -> int:
    return capi.F64 if v.dtype == torch.float64 else capi.F32


def dot(x, y, stream=None) -> float:
    """Morpheus::dot<HIP>(n, x, y) of two device vectors."""
    out = C.c_double()
    capi.call("mph_dot", int(x.numel()), _p(x), _p(y), _vt_of(x), C.byref(out),
              stream_handle(stream))
    return out.value


def waxpby(alpha, x, beta, y, w, stream=None):
    """Morpheus::waxpby<HIP>(n, alpha, x, beta, y, w): w = alpha x + beta y."""
    capi.call("mph_waxpby", int(x.numel()), float(alpha), _p(x), float(beta), _p(y), _p(w),
              _vt_of(x), stream_handle(stream))
    return w


def reduce(x, stream=None) -> float:
    """Morpheus::reduce<HIP>(x, n)."""
    out = C.c_double()
    capi.call("mph_reduce", int(x.numel()), _p(x), _vt_of(x), C.byref(out), stream_handle(stream))
    return out.value


class CG:
    """Conjugate gradients on a square DeviceMatrix, one SpMV and three fused
    vector launches per iteration, scalars on the device."""

    def __init__(self, A: DeviceMatrix):
        self.A = A
        self.handle = C.c_void_p()
        capi.call("mph_cg_create", C.byref(self.handle), A.handle)

    def start(self, b, x, stream=None):
        capi.call("mph_cg_start", self.handle, _p(b), _p(x), stream_handle(stream))

    def iterate(self, x, iters, stream=None):
        capi.call("mph_cg_iterate", self.handle, _p(x), int(iters), stream_handle(stream))

    def residual(self, stream=None) -> float:
        rr = C.c_double()
        capi.call("mph_cg_residual", self.handle, C.byref(rr), stream_handle(stream))
        return rr.value

    def close(self):
        if self.handle:
            capi.lib().mph_cg_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ analytics
def count_nnz_per_row(A: DeviceMatrix, out=None, init=True, stream=None):
    """Morpheus::count_nnz_per_row<HIP>(A, out, init) into an int32 device
    vector (a new one when out is None)."""
    if out is None:
        out = torch.zeros(A.nrows, dtype=torch.int32, device="cuda")
    capi.call("mph_matrix_count_nnz_per_row", A.handle, _p(out), int(bool(init)),
              stream_handle(stream))
    return out


def count_nnz_per_diagonal(A: DeviceMatrix, out=None, init=True, stream=None):
    """Morpheus::count_nnz_per_diagonal<HIP>(A, out, init): nrows + ncols - 1
    int32 counters, entry col - row + nrows - 1."""
    if out is None:
        out = torch.zeros(A.nrows + A.ncols - 1, dtype=torch.int32, device="cuda")
    capi.call("mph_matrix_count_nnz_per_diagonal", A.handle, _p(out), int(bool(init)),
              stream_handle(stream))
    return out


def i32_stats(v, mean=0.0, threshold=0, stream=None):
    """(max, min, sum of squared deviations from mean, count of entries >
    threshold and != 0) of an int32 device vector."""
    st = capi.I32Stats()
    capi.call("mph_i32_stats", _p(v), int(v.numel()), float(mean), int(threshold),
              C.byref(st), stream_handle(stream))
    return st.max, st.min, st.sq_dev, st.count


def max_nnnz(A: DeviceMatrix) -> int:
    """core/src/Morpheus_MatrixAnalytics.hpp:126-140."""
    return i32_stats(count_nnz_per_row(A))[0]


def min_nnnz(A: DeviceMatrix) -> int:
    """core/src/Morpheus_MatrixAnalytics.hpp:151-165."""
    return i32_stats(count_nnz_per_row(A))[1]


def std_nnnz(A: DeviceMatrix) -> int:
    """core/src/Morpheus_MatrixAnalytics.hpp:177-191: the deviation around the
    integer mean nnnz / nrows, truncated to the size type."""
    sq = i32_stats(count_nnz_per_row(A), float(A.nnz // A.nrows))[2]
    return int((sq / A.nrows) ** 0.5)


def count_diagonals(A: DeviceMatrix) -> int:
    """core/src/Morpheus_MatrixAnalytics.hpp:220-234."""
    return i32_stats(count_nnz_per_diagonal(A))[3]


def count_true_diagonals(A: DeviceMatrix, threshold: int) -> int:
    """core/src/Morpheus_MatrixAnalytics.hpp:245-259."""
    return i32_stats(count_nnz_per_diagonal(A), 0.0, threshold)[3]


def convert(A: DeviceMatrix, fmt: int, stream=None) -> DeviceMatrix:
    """Morpheus::convert<Morpheus::HIP>(src, dst): via a temporary COO when
    neither side is COO (core/src/impl/Morpheus_Convert_Impl.hpp:83-106)."""
    if fmt == A.format:
        # same-format conversion is an element-wise copy
        # (e.g. core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:45-70)
        return DeviceMatrix(A.format, A.nrows, A.ncols, A.nnz, A.dtype,
                            {k: t.clone() for k, t in A.arrays.items()}, dict(A.meta))
    if A.format == COO_FORMAT:
        return from_coo(A, fmt, stream)
    coo = to_coo(A, stream)
    if fmt == COO_FORMAT:
        return coo
    return from_coo(coo, fmt, stream)


# ------------------------------------------------------------------ batched
class Batch:
    """Many device matrices x k right-hand sides in one persistent launch
    (mph_batch_*).  X[i] is (ncols_i, k) and Y[i] (nrows_i, k), row-major."""

    def __init__(self, mats, k, X, Y, stream=None):
        self.mats, self.X, self.Y, self.k = list(mats), list(X), list(Y), int(k)
        n = len(self.mats)
        hs = (C.c_void_p * n)(*[m.handle.value for m in self.mats])
        xs = (C.c_void_p * n)(*[x.data_ptr() for x in self.X])
        ys = (C.c_void_p * n)(*[y.data_ptr() for y in self.Y])
        self.handle = C.c_void_p()
        capi.call("mph_batch_create", C.byref(self.handle), n, hs, self.k, xs, ys,
                  stream_handle(stream))

    def run(self, init=True, stream=None):
        capi.call("mph_batch_run", self.handle, int(bool(init)), stream_handle(stream))

    def run_graph(self, init=True, stream=None):
        capi.call("mph_batch_run_graph", self.handle, int(bool(init)), stream_handle(stream))

    def info(self):
        """(tiles of the tile kernel, rows of the Dot2 kernel)"""
        nt, nw = C.c_int64(), C.c_int32()
        capi.call("mph_batch_info", self.handle, C.byref(nt), C.byref(nw))
        return int(nt.value), int(nw.value)

    def close(self):
        if self.handle:
            capi.lib().mph_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_format(A: DeviceMatrix, k: int = 1, reps: int = 5, stream=None):
    """Fastest format for the CSR matrix A with k right-hand sides, by timed
    trials on the GPU; returns (format, {format: ms}) with inadmissible
    formats omitted."""
    assert A.format == CSR_FORMAT
    best = C.c_int()
    ms = (C.c_float * 6)()
    a = A.arrays
    capi.call("mph_autotune_format", A.nrows, A.ncols, A.nnz, _p(a["ro"]), _p(a["ci"]),
              _p(a["v"]), A.vt, int(k), int(reps), stream_handle(stream), C.byref(best), ms)
    return int(best.value), {f: float(ms[f]) for f in range(6) if ms[f] >= 0}


def autotune_launch(A: DeviceMatrix, k: int = 1, reps: int = 5, stream=None) -> float:
    """Tunes A's launch shape in place by timed trials; returns the best ms."""
    x = torch.empty((A.ncols, k) if k > 1 else (A.ncols,), dtype=vector_dtype(A.dtype),
                    device="cuda").uniform_(-1, 1)
    y = torch.empty((A.nrows, k) if k > 1 else (A.nrows,), dtype=vector_dtype(A.dtype),
                    device="cuda")
    best = C.c_float()
    capi.call("mph_matrix_autotune", A.handle, int(k), _p(x), _p(y), int(reps),
              stream_handle(stream), C.byref(best))
    return float(best.value)
