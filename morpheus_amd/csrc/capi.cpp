// capi.cpp -- the C-ABI: library state, matrix handles, analysis and the
// per-format multiply dispatch.  Host C++ compiled into the same shared
// library as the kernels.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "mph_internal.hpp"

namespace mph {

namespace {
thread_local std::string g_last_error;
std::mutex g_dev_mutex;
int g_cu_count = 0;
int g_cu_device = -1;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int device_cu_count() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lk(g_dev_mutex);
  if (dev != g_cu_device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    g_cu_count = prop.multiProcessorCount;
    g_cu_device = dev;
  }
  return g_cu_count;
}

void release_plans(mph_matrix_impl* A) {
  RowBlockPlan& p = A->rb;
  dfree(p.blocks);
  if (p.partial) dfree(p.partial);
  if (p.tail_ro) dfree(p.tail_ro);
  dfree(p.m_ro);
  dfree(p.m_ci);
  dfree(p.m_val);
  if (p.sblocks) dfree(p.sblocks);
  dfree(p.krec);
  dfree(p.skrec);
  if (p.smap) dfree(p.smap);
  if (p.sro) dfree(p.sro);
  dfree(p.ell_cnt);
  xtile_release(p.xt);
  p = RowBlockPlan{};
  CooPlan& c = A->coo;
  if (c.first_val) dfree(c.first_val);
  if (c.last_val) dfree(c.last_val);
  if (c.chunks) dfree(c.chunks);
  xtile_release(c.xt);
  c = CooPlan{};
  A->dia_stage = DiaStage{};
  A->analysed = false;
}

namespace {

bool valid_vt(int vt) { return vt == MPH_F64 || vt == MPH_F32 || vt == MPH_BF16; }

int check_common(mph_matrix_t* out, int64_t nrows, int64_t ncols, int vt) {
  if (!out) return invalid("null output handle");
  if (nrows < 0 || ncols < 0) return invalid("negative dimensions");
  if (nrows > INT32_MAX || ncols > INT32_MAX)
    return invalid("dimensions exceed the int32 index range");
  if (!valid_vt(vt)) return invalid("unknown value type");
  return MPH_OK;
}

int check_array(const void* p, int64_t n, const char* what) {
  if (n < 0) return invalid(std::string("negative size for ") + what);
  if (n > 0 && !p) return invalid(std::string("null pointer for ") + what);
  return MPH_OK;
}

mph_matrix_t new_handle(int format, int64_t nrows, int64_t ncols, int vt) {
  mph_matrix_t A = new mph_matrix_s();
  A->format = format;
  A->nrows = nrows;
  A->ncols = ncols;
  A->vt = vt;
  hipGetDevice(&A->device);
  return A;
}

// HYB with a COO tail that is not row-sorted: ELL pass, then an atomic
// scatter of the tail (the tail is wrapped in its own COO handle).
struct HybFallback {
  mph_matrix_s tail;
};

// one pass over every index array before the first product
int check_inputs(mph_matrix_t A, hipStream_t s) {
  const int64_t n = A->nrows, m = A->ncols;
  switch (A->format) {
    case MPH_COO:
      MPH_CHECK_STATUS(check_index_range(A->coo_ri, A->coo_nnz, n, false, s, "COO row_indices"));
      return check_index_range(A->coo_ci, A->coo_nnz, m, false, s, "COO column_indices");
    case MPH_CSR:
    case MPH_HDC:
      MPH_CHECK_STATUS(check_row_offsets(A->csr_ro, n, A->csr_nnz, s, "CSR row_offsets"));
      return check_index_range(A->csr_ci, A->csr_nnz, m, false, s, "CSR column_indices");
    case MPH_ELL:
      return check_index_range(A->ell_ci, A->ell_width * A->ell_pitch, m, true, s,
                               "ELL column_indices");
    case MPH_HYB:
      MPH_CHECK_STATUS(check_index_range(A->ell_ci, A->ell_width * A->ell_pitch, m, true, s,
                                         "HYB ELL column_indices"));
      MPH_CHECK_STATUS(check_index_range(A->coo_ri, A->coo_nnz, n, false, s, "HYB COO row_indices"));
      return check_index_range(A->coo_ci, A->coo_nnz, m, false, s, "HYB COO column_indices");
  }
  return MPH_OK;
}

int analyse(mph_matrix_t A, hipStream_t s) {
  if (A->analysed) return MPH_OK;
  release_plans(A);  // what a failed earlier attempt left behind
  MPH_CHECK_STATUS(check_inputs(A, s));
  int st = MPH_OK;
  switch (A->format) {
    case MPH_COO:
      st = coo_analyse(A, s);
      break;
    case MPH_CSR:
      st = rowblock_analyse(A, s);
      break;
    case MPH_HDC:
      st = dia_analyse(A, s);
      if (st == MPH_OK) st = rowblock_analyse(A, s);
      break;
    case MPH_DIA:
      st = dia_analyse(A, s);
      break;
    case MPH_ELL:
    case MPH_HYB: {
      int pad_only = 1;
      st = ell_padding_only(A->ell_ci, A->ell_pitch, A->nrows, A->ell_K,
                            A->ell_width, s, &pad_only);
      if (st != MPH_OK) break;
      A->ell_loop = pad_only ? A->ell_K : A->ell_width;
      if (A->format == MPH_HYB) st = rowblock_analyse(A, s);
      break;
    }
  }
  if (st == MPH_OK) A->analysed = true;
  return st;
}

mph_matrix_impl hyb_tail_view(mph_matrix_t A) {
  mph_matrix_impl t{};
  t.format = MPH_COO;
  t.vt = A->vt;
  t.nrows = A->nrows;
  t.ncols = A->ncols;
  t.nnz = A->coo_nnz;
  t.coo_ri = A->coo_ri;
  t.coo_ci = A->coo_ci;
  t.coo_v = A->coo_v;
  t.coo_nnz = A->coo_nnz;
  return t;
}

int multiply_impl(mph_matrix_t A, int k, const void* x, void* y, int init,
                  hipStream_t s) {
  if (!A) return invalid("null matrix handle");
  if (k < 1) return invalid("k must be >= 1");
  if ((A->nrows > 0 && !y) || (A->ncols > 0 && !x))
    return invalid("null x or y");
  int st = analyse(A, s);
  if (st == MPH_ERR_NOT_IMPLEMENTED && A->format == MPH_HYB) {
    // unsorted tail: two passes
    MPH_CHECK_STATUS(ell_multiply(A, k, x, y, init, s));
    mph_matrix_impl t = hyb_tail_view(A);
    st = coo_analyse(&t, s);
    if (st == MPH_OK) st = coo_multiply(&t, k, x, y, 0, s);
    hipStreamSynchronize(s);
    release_plans(&t);
    return st;
  }
  if (st != MPH_OK) return st;
  switch (A->format) {
    case MPH_COO:
      return coo_multiply(A, k, x, y, init, s);
    case MPH_HDC:
      // an HDC whose CSR part is empty is its DIA part
      if (A->csr_nnz == 0) return dia_multiply(A, k, x, y, init, s);
      if (A->nrows == 0) return MPH_OK;
      return rowblock_multiply(A, k, x, y, init, s);
    case MPH_HYB:
      // a HYB whose COO part is empty is its ELL part
      if (A->coo_nnz == 0) return ell_multiply(A, k, x, y, init, s);
      if (A->nrows == 0) return MPH_OK;
      return rowblock_multiply(A, k, x, y, init, s);
    case MPH_CSR:
      if (A->nrows == 0) return MPH_OK;
      return rowblock_multiply(A, k, x, y, init, s);
    case MPH_DIA:
      return dia_multiply(A, k, x, y, init, s);
    case MPH_ELL:
      return ell_multiply(A, k, x, y, init, s);
  }
  return invalid("unknown format");
}

// Bytes of the arrays one k = 1 product streams under the analysed plan,
// each counted once (x once, y written once, or read and written when
// accumulating or when a second pass adds to it).
int64_t csr_plan_bytes(const mph_matrix_impl* A, const RowBlockPlan& p, int64_t list_rows,
                       int64_t items) {
  const int64_t si = 4, sv = (int64_t)value_size(A->vt);
  const int64_t acc = 8;  // a partial: an FP64 sum, or an FP32 Dot2 pair
  int64_t b = (list_rows + 1) * si;
  const int64_t tiled = p.xt.rows > 0 ? p.xt.items : 0;
  b += (items - tiled) * (si + sv);
  if (p.xt.rows > 0) {
    // tile-major copy (16-bit column + value per slot), the (tile, row)
    // partials written and read back, the tiled rows' ids
    b += tiled * (2 + sv) + 2 * (int64_t)p.xt.ntiles * p.xt.rows * acc + (int64_t)p.xt.rows * si;
    // a direct plan reads the values in place and each run's position
    if (p.xt.rpos) b += (int64_t)p.xt.groups * p.xt.ntiles * (p.xt.gr + 1) * si;
  }
  if (p.split) b += list_rows * si;  // the list's row map
  return b;
}

int plan_bytes(mph_matrix_t A, int init, int64_t* out) {
  const int64_t si = 4, sv = (int64_t)value_size(A->vt), sx = (int64_t)vector_size(A->vt);
  const int64_t x = A->ncols * sx, y = A->nrows * sx * (init ? 1 : 2);
  int64_t b = 0;
  auto dia_part = [&]() -> int64_t {
    std::vector<int32_t> off((size_t)A->ndiag);
    if (A->ndiag > 0 &&
        hipMemcpy(off.data(), A->dia_off, sizeof(int32_t) * A->ndiag, hipMemcpyDeviceToHost) !=
            hipSuccess)
      return -1;
    int64_t v = A->ndiag * si;
    for (int32_t o : off)
      v += std::max<int64_t>(0, std::min<int64_t>(A->nrows - std::max<int64_t>(0, -(int64_t)o),
                                                  A->ncols - std::max<int64_t>(0, (int64_t)o))) * sv;
    return v;
  };
  const RowBlockPlan& p = A->rb;
  switch (A->format) {
    case MPH_DIA: {
      const int64_t d = dia_part();
      if (d < 0) return runtime_error("DIA offsets unreadable");
      b = d;
      break;
    }
    case MPH_ELL:
      b = A->ell_loop * A->nrows * (si + sv);
      break;
    case MPH_COO:
      if (A->coo.rowplan) {
        b = csr_plan_bytes(A, p, A->nrows, A->nnz);  // derived row offsets, no row indices
      } else if (A->coo.sorted) {
        const int64_t tiled = A->coo.xt.rows > 0 ? A->coo.xt.items : 0;
        b = (A->nnz - tiled) * (2 * si + sv) + A->coo.nchunks * 2 * 8;
        if (tiled) b += tiled * (2 + sv) + 2 * (int64_t)A->coo.xt.ntiles * A->coo.xt.rows * 8;
        if (A->coo.prezero) b += A->nrows * sx;
      } else {
        b = A->nnz * (2 * si + sv) + A->nrows * sx;  // clear, then atomic adds
      }
      break;
    case MPH_CSR:
      b = csr_plan_bytes(A, p, A->nrows, A->nnz);
      break;
    case MPH_HYB:
    case MPH_HDC: {
      const bool hyb = A->format == MPH_HYB;
      const int64_t tail = hyb ? A->coo_nnz : A->csr_nnz;
      if (tail == 0) {  // the ELL / DIA part alone
        if (hyb) {
          b = A->ell_loop * A->nrows * (si + sv);
        } else {
          b = dia_part();
          if (b < 0) return runtime_error("DIA offsets unreadable");
        }
        break;
      }
      if (hyb && p.m_ro) {  // merged copy: the CSR plan over it
        int32_t mnnz = 0;
        if (hipMemcpy(&mnnz, p.m_ro + A->nrows, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
          return runtime_error("merged row offsets unreadable");
        b = csr_plan_bytes(A, p, A->nrows, mnnz);
        break;
      }
      int64_t pre = 0;
      if (hyb && p.ell_cnt) {
        pre = p.ell_visited * (si + sv) + A->nrows;  // the slots up to each row's count
      } else if (hyb) {
        pre = A->ell_loop * A->nrows * (si + sv);
      } else {
        pre = dia_part();
        if (pre < 0) return runtime_error("DIA offsets unreadable");
      }
      if (p.split) {
        // the ELL / DIA pass writes y; the tail rows' list reads and writes it again
        int64_t listed = 0;
        if (p.nsblocks > 0 &&
            hipMemcpy(&listed, p.sblocks + p.nsblocks, sizeof(int32_t), hipMemcpyDeviceToHost) !=
                hipSuccess)
          return runtime_error("split plan unreadable");
        listed = (int32_t)listed;
        b = pre + csr_plan_bytes(A, p, listed, tail) + 2 * listed * sx;
      } else {
        b = pre + csr_plan_bytes(A, p, A->nrows, tail);
      }
      break;
    }
    default:
      return invalid("unknown format");
  }
  *out = b + x + y;
  return MPH_OK;
}

}  // namespace

}  // namespace mph

using namespace mph;

extern "C" {

const char* mph_version(void) { return "morpheus_amd 0.1.0 (gfx950)"; }

const char* mph_status_string(int status) {
  switch (status) {
    case MPH_OK: return "ok";
    case MPH_ERR_NOT_IMPLEMENTED: return "not implemented";
    case MPH_ERR_INVALID_INPUT: return "invalid input";
    case MPH_ERR_FORMAT_CONVERSION: return "format conversion (fill tolerance exceeded)";
    case MPH_ERR_HIP: return "HIP runtime error";
    case MPH_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MPH_ERR_RUNTIME: return "runtime error";
  }
  return "unknown status";
}

const char* mph_last_error(void) { return g_last_error.c_str(); }

int mph_initialize(int device) {
  if (device >= 0) MPH_HIP_CHECK(hipSetDevice(device));
  int n = 0;
  MPH_HIP_CHECK(hipGetDeviceCount(&n));
  if (n == 0) {
    set_error("no HIP device");
    return MPH_ERR_HIP;
  }
  device_cu_count();
  void* scratch = nullptr;  // the no-allocation entry point's scratch address
  return noalloc_scratch(&scratch);
}

int mph_finalize(void) { return MPH_OK; }

int mph_matrix_create_coo(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, const int32_t* ri, const int32_t* ci,
                          const void* v, int vt) {
  MPH_CHECK_STATUS(check_common(out, nrows, ncols, vt));
  MPH_CHECK_STATUS(check_array(ri, nnz, "row_indices"));
  MPH_CHECK_STATUS(check_array(ci, nnz, "column_indices"));
  MPH_CHECK_STATUS(check_array(v, nnz, "values"));
  mph_matrix_t A = new_handle(MPH_COO, nrows, ncols, vt);
  A->nnz = A->coo_nnz = nnz;
  A->coo_ri = ri;
  A->coo_ci = ci;
  A->coo_v = v;
  *out = A;
  return MPH_OK;
}

int mph_matrix_create_csr(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, const int32_t* ro, const int32_t* ci,
                          const void* v, int vt) {
  MPH_CHECK_STATUS(check_common(out, nrows, ncols, vt));
  MPH_CHECK_STATUS(check_array(ro, nrows + 1, "row_offsets"));
  MPH_CHECK_STATUS(check_array(ci, nnz, "column_indices"));
  MPH_CHECK_STATUS(check_array(v, nnz, "values"));
  mph_matrix_t A = new_handle(MPH_CSR, nrows, ncols, vt);
  A->nnz = A->csr_nnz = nnz;
  A->csr_ro = ro;
  A->csr_ci = ci;
  A->csr_v = v;
  *out = A;
  return MPH_OK;
}

int mph_matrix_create_dia(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, int64_t ndiag, int64_t pitch,
                          const int32_t* off, const void* v, int vt) {
  MPH_CHECK_STATUS(check_common(out, nrows, ncols, vt));
  if (ndiag < 0 || (ndiag > 0 && pitch < nrows))
    return invalid("DIA pitch must be >= nrows");
  MPH_CHECK_STATUS(check_array(off, ndiag, "diagonal_offsets"));
  MPH_CHECK_STATUS(check_array(v, ndiag * pitch, "values"));
  mph_matrix_t A = new_handle(MPH_DIA, nrows, ncols, vt);
  A->nnz = A->dia_nnz = nnz;
  A->ndiag = ndiag;
  A->dia_pitch = pitch;
  A->dia_off = off;
  A->dia_v = v;
  *out = A;
  return MPH_OK;
}

int mph_matrix_create_ell(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, int64_t K, int64_t width, int64_t pitch,
                          const int32_t* ci, const void* v, int vt) {
  MPH_CHECK_STATUS(check_common(out, nrows, ncols, vt));
  if (K < 0 || width < K) return invalid("ELL width must be >= entries_per_row");
  if (width > 0 && pitch < nrows) return invalid("ELL pitch must be >= nrows");
  MPH_CHECK_STATUS(check_array(ci, width * pitch, "column_indices"));
  MPH_CHECK_STATUS(check_array(v, width * pitch, "values"));
  mph_matrix_t A = new_handle(MPH_ELL, nrows, ncols, vt);
  A->nnz = A->ell_nnz = nnz;
  A->ell_K = K;
  A->ell_width = width;
  A->ell_pitch = pitch;
  A->ell_ci = ci;
  A->ell_v = v;
  *out = A;
  return MPH_OK;
}

int mph_matrix_create_hyb(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t ell_nnz, int64_t K, int64_t width,
                          int64_t pitch, const int32_t* eci, const void* ev,
                          int64_t coo_nnz, const int32_t* cri,
                          const int32_t* cci, const void* cv, int vt) {
  MPH_CHECK_STATUS(check_common(out, nrows, ncols, vt));
  if (K < 0 || width < K) return invalid("ELL width must be >= entries_per_row");
  if (width > 0 && pitch < nrows) return invalid("ELL pitch must be >= nrows");
  MPH_CHECK_STATUS(check_array(eci, width * pitch, "ell column_indices"));
  MPH_CHECK_STATUS(check_array(ev, width * pitch, "ell values"));
  MPH_CHECK_STATUS(check_array(cri, coo_nnz, "coo row_indices"));
  MPH_CHECK_STATUS(check_array(cci, coo_nnz, "coo column_indices"));
  MPH_CHECK_STATUS(check_array(cv, coo_nnz, "coo values"));
  mph_matrix_t A = new_handle(MPH_HYB, nrows, ncols, vt);
  A->ell_nnz = ell_nnz;
  A->coo_nnz = coo_nnz;
  A->nnz = ell_nnz + coo_nnz;
  A->ell_K = K;
  A->ell_width = width;
  A->ell_pitch = pitch;
  A->ell_ci = eci;
  A->ell_v = ev;
  A->coo_ri = cri;
  A->coo_ci = cci;
  A->coo_v = cv;
  *out = A;
  return MPH_OK;
}

int mph_matrix_create_hdc(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t dia_nnz, int64_t ndiag, int64_t pitch,
                          const int32_t* off, const void* dv, int64_t csr_nnz,
                          const int32_t* ro, const int32_t* ci, const void* cv,
                          int vt) {
  MPH_CHECK_STATUS(check_common(out, nrows, ncols, vt));
  if (ndiag < 0 || (ndiag > 0 && pitch < nrows))
    return invalid("DIA pitch must be >= nrows");
  MPH_CHECK_STATUS(check_array(off, ndiag, "dia diagonal_offsets"));
  MPH_CHECK_STATUS(check_array(dv, ndiag * pitch, "dia values"));
  MPH_CHECK_STATUS(check_array(ro, nrows + 1, "csr row_offsets"));
  MPH_CHECK_STATUS(check_array(ci, csr_nnz, "csr column_indices"));
  MPH_CHECK_STATUS(check_array(cv, csr_nnz, "csr values"));
  mph_matrix_t A = new_handle(MPH_HDC, nrows, ncols, vt);
  A->dia_nnz = dia_nnz;
  A->csr_nnz = csr_nnz;
  A->nnz = dia_nnz + csr_nnz;
  A->ndiag = ndiag;
  A->dia_pitch = pitch;
  A->dia_off = off;
  A->dia_v = dv;
  A->csr_ro = ro;
  A->csr_ci = ci;
  A->csr_v = cv;
  *out = A;
  return MPH_OK;
}

int mph_matrix_destroy(mph_matrix_t A) {
  if (!A) return MPH_OK;
  release_plans(A);
  delete A;
  return MPH_OK;
}

int mph_matrix_analyse(mph_matrix_t A, mph_stream_t stream) {
  if (!A) return invalid("null matrix handle");
  release_plans(A);
  int st = analyse(A, reinterpret_cast<hipStream_t>(stream));
  // an unsorted HYB tail is handled at multiply time by the two-pass form
  if (st == MPH_ERR_NOT_IMPLEMENTED && A->format == MPH_HYB) return MPH_OK;
  return st;
}

int mph_matrix_format(mph_matrix_t A) { return A ? A->format : -1; }

int mph_matrix_plan_bytes(mph_matrix_t A, int init, int64_t* bytes) {
  if (!A || !bytes) return invalid("null argument");
  int st = analyse(A, nullptr);
  if (st == MPH_ERR_NOT_IMPLEMENTED && A->format == MPH_HYB) {
    // unsorted tail: the ELL pass, then the COO tail by atomics
    const int64_t si = 4, sv = (int64_t)value_size(A->vt), sx = (int64_t)vector_size(A->vt);
    *bytes = A->ell_loop * A->nrows * (si + sv) + A->coo_nnz * (2 * si + sv) +
             A->ncols * sx + A->nrows * sx * (init ? 2 : 3);
    return MPH_OK;
  }
  MPH_CHECK_STATUS(st);
  return plan_bytes(A, init ? 1 : 0, bytes);
}

int mph_matrix_set_tuning(mph_matrix_t A, const mph_tuning* t) {
  if (!A || !t) return invalid("null argument");
  if (t->lanes != 0 && (t->lanes < 1 || t->lanes > 64 || (t->lanes & (t->lanes - 1))))
    return invalid("lanes must be a power of two in [1, 64]");
  if (t->rows_chunk != 0 && (t->rows_chunk < 256 || t->rows_chunk > 4096 ||
                             (A->vt == MPH_F64 && t->rows_chunk > 2048)))
    return invalid("rows_chunk out of range");
  if (t->coo_items != 0 && t->coo_items != 4 && t->coo_items != 8 && t->coo_items != 16)
    return invalid("coo_items must be 4, 8 or 16");
  if (t->rows_per_thread != 0 && t->rows_per_thread != 1 && t->rows_per_thread != 2)
    return invalid("rows_per_thread must be 1 or 2");
  if (t->xtile < 0 || t->xtile > 2) return invalid("xtile must be 0, 1 or 2");
  const bool replan = t->rows_chunk != A->tuning.rows_chunk ||
                      t->coo_items != A->tuning.coo_items || t->xtile != A->tuning.xtile;
  A->tuning = *t;
  if (replan) release_plans(A);
  return MPH_OK;
}

int mph_matrix_set_policy(mph_matrix_t A, int flags) {
  if (!A) return invalid("null matrix");
  if (flags & ~MPH_POLICY_NO_VALUE_COPIES) return invalid("unknown policy flag");
  if (flags != A->policy) {
    A->policy = flags;
    release_plans(A);
  }
  return MPH_OK;
}

int mph_matrix_get_tuning(mph_matrix_t A, mph_tuning* t) {
  if (!A || !t) return invalid("null argument");
  *t = A->tuning;
  return MPH_OK;
}

int mph_multiply(mph_matrix_t A, const void* x, void* y, int init,
                 mph_stream_t stream) {
  return multiply_impl(A, 1, x, y, init, reinterpret_cast<hipStream_t>(stream));
}

int mph_multiply_multi(mph_matrix_t A, int32_t k, const void* X, void* Y,
                       int init, mph_stream_t stream) {
  return multiply_impl(A, k, X, Y, init, reinterpret_cast<hipStream_t>(stream));
}

static int one_shot(mph_matrix_t A, const void* x, void* y, int init,
                    mph_stream_t stream) {
  int st = mph_multiply(A, x, y, init, stream);
  // the transient analysis is released after the product has run
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  mph_matrix_destroy(A);
  if (st == MPH_OK && e != hipSuccess) {
    set_error(hipGetErrorString(e));
    st = MPH_ERR_HIP;
  }
  return st;
}

int mph_csr_multiply(int64_t nrows, int64_t ncols, int64_t nnz,
                     const int32_t* ro, const int32_t* ci, const void* v,
                     int vt, const void* x, void* y, int init,
                     mph_stream_t stream) {
  mph_matrix_t A = nullptr;
  MPH_CHECK_STATUS(mph_matrix_create_csr(&A, nrows, ncols, nnz, ro, ci, v, vt));
  return one_shot(A, x, y, init, stream);
}

int mph_coo_multiply(int64_t nrows, int64_t ncols, int64_t nnz,
                     const int32_t* ri, const int32_t* ci, const void* v,
                     int vt, const void* x, void* y, int init,
                     mph_stream_t stream) {
  mph_matrix_t A = nullptr;
  MPH_CHECK_STATUS(mph_matrix_create_coo(&A, nrows, ncols, nnz, ri, ci, v, vt));
  return one_shot(A, x, y, init, stream);
}

int mph_dia_multiply(int64_t nrows, int64_t ncols, int64_t nnz, int64_t ndiag,
                     int64_t pitch, const int32_t* off, const void* v, int vt,
                     const void* x, void* y, int init, mph_stream_t stream) {
  mph_matrix_t A = nullptr;
  MPH_CHECK_STATUS(
      mph_matrix_create_dia(&A, nrows, ncols, nnz, ndiag, pitch, off, v, vt));
  return one_shot(A, x, y, init, stream);
}

int mph_ell_multiply(int64_t nrows, int64_t ncols, int64_t nnz, int64_t K,
                     int64_t width, int64_t pitch, const int32_t* ci,
                     const void* v, int vt, const void* x, void* y, int init,
                     mph_stream_t stream) {
  mph_matrix_t A = nullptr;
  MPH_CHECK_STATUS(
      mph_matrix_create_ell(&A, nrows, ncols, nnz, K, width, pitch, ci, v, vt));
  return one_shot(A, x, y, init, stream);
}

int mph_exceeds_tolerance(int64_t num_rows, int64_t num_entries,
                          int64_t num_columns) {
  const float max_fill = 10.0f;
  const float threshold = 100e6f;
  const float size = (float)num_columns * (float)num_rows;
  const float fill_ratio = size / std::max(1.0f, (float)num_entries);
  return (fill_ratio > max_fill && size > threshold) ? 1 : 0;
}

}  // extern "C"
