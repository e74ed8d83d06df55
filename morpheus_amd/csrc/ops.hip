// ops.hip -- matrix operations on the GPU: update_diagonal and get_diagonal
// of a matrix handle of any format (the Jacobi preconditioner's pieces).
//
// update_diagonal replaces core/src/Morpheus_MatrixOperations.hpp:54-58 with
// the Serial semantics of core/src/impl/{Csr,Coo,Dia,Ell,Hyb,Hdc}/Serial/
// Morpheus_MatrixOperations_Impl.hpp:
//  * CSR: the first entry of row i in column i;  COO: every diagonal entry;
//  * DIA: the main diagonal's slots, a stored zero staying zero;
//  * ELL: the first slot of row i holding column i;
//  * HYB = ELL then COO, HDC = DIA then CSR.
// (The reference's HIP ELL kernel differs from its Serial loop: it writes
// every matching slot and keeps zeros; the Serial behaviour is the one kept,
// as for multiply and convert.)  The handle's values are written in place and
// its analysis is released, since the x-tiled plan holds copies of values.
//
// get_diagonal has no implementation in the reference (it throws
// NotImplementedException, core/src/impl/Csr/Serial/
// Morpheus_MatrixOperations_Impl.hpp:63-72); here it returns the diagonal
// the product sees: the sum of every stored (i, i) entry (in storage order
// within a row; COO entries are added atomically).
#include <algorithm>

#include "mph_internal.hpp"

namespace mph {

namespace {

constexpr int kT = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, kT)); }

template <typename V>
__device__ __forceinline__ V from_vec(typename vtraits<V>::vec d) {
  return (V)d;
}
// FP32 -> BF16, round to nearest even on the upper 16 bits (NaN kept quiet)
template <>
__device__ __forceinline__ bf16 from_vec<bf16>(float d) {
  uint32_t u = __float_as_uint(d);
  if ((u & 0x7fffffffu) > 0x7f800000u) return bf16{(uint16_t)((u >> 16) | 0x40u)};
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16{(uint16_t)(u >> 16)};
}

template <typename V>
__device__ __forceinline__ bool is_zero(V v) {
  return v == V{};
}
template <>
__device__ __forceinline__ bool is_zero<bf16>(bf16 v) {
  return (v.bits & 0x7fffu) == 0;
}

template <typename V>
using VecT = typename vtraits<V>::vec;

// ------------------------------------------------------------- update
template <typename V>
__global__ void csr_update_kernel(const int32_t* __restrict__ ro, const int32_t* __restrict__ ci,
                                  V* __restrict__ v, int64_t nrows, const VecT<V>* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++)
    if (ci[jj] == i) {
      v[jj] = from_vec<V>(d[i]);
      break;
    }
}

template <typename V>
__global__ void coo_update_kernel(const int32_t* __restrict__ ri, const int32_t* __restrict__ ci,
                                  V* __restrict__ v, int64_t nnz, int64_t n,
                                  const VecT<V>* __restrict__ d) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz && ri[k] == ci[k] && ci[k] >= 0 && ci[k] < n) v[k] = from_vec<V>(d[ci[k]]);
}

template <typename V>
__global__ void dia_update_kernel(const int32_t* __restrict__ off, V* __restrict__ v,
                                  int64_t ndiag, int64_t pitch, int64_t nrows, int64_t ncols,
                                  const VecT<V>* __restrict__ d) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows || row >= ncols) return;
  for (int64_t n = 0; n < ndiag; n++)
    if (off[n] == 0) {
      V& s = v[row + pitch * n];
      s = is_zero(s) ? V{} : from_vec<V>(d[row]);
    }
}

template <typename V>
__global__ void ell_update_kernel(const int32_t* __restrict__ ci, V* __restrict__ v, int64_t nrows,
                                  int64_t width, int64_t pitch, const VecT<V>* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  for (int64_t n = 0; n < width; n++)
    if (ci[i + pitch * n] == i) {
      v[i + pitch * n] = from_vec<V>(d[i]);
      break;
    }
}

// ------------------------------------------------------------- get
struct DiagArgs {
  int fmt;
  int64_t ndiag, dia_pitch, ell_width, ell_pitch;
  const int32_t *dia_off, *ell_ci, *csr_ro, *csr_ci;
  const void *dia_v, *ell_v, *csr_v;
};

// one thread per row, the row's diagonal entries added in storage order
template <typename V>
__global__ void get_diag_kernel(DiagArgs a, VecT<V>* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  using Acc = typename vtraits<V>::acc;
  Acc s = 0;
  if (a.fmt == MPH_DIA || a.fmt == MPH_HDC) {
    const V* dv = static_cast<const V*>(a.dia_v);
    for (int64_t k = 0; k < a.ndiag; k++)
      if (a.dia_off[k] == 0) s += vtraits<V>::widen(dv[i + a.dia_pitch * k]);
  }
  if (a.fmt == MPH_ELL || a.fmt == MPH_HYB) {
    const V* ev = static_cast<const V*>(a.ell_v);
    for (int64_t k = 0; k < a.ell_width; k++)
      if (a.ell_ci[i + a.ell_pitch * k] == i) s += vtraits<V>::widen(ev[i + a.ell_pitch * k]);
  }
  if (a.fmt == MPH_CSR || a.fmt == MPH_HDC) {
    const V* cv = static_cast<const V*>(a.csr_v);
    for (int32_t jj = a.csr_ro[i]; jj < a.csr_ro[i + 1]; jj++)
      if (a.csr_ci[jj] == i) s += vtraits<V>::widen(cv[jj]);
  }
  d[i] = (VecT<V>)s;
}

// COO entries (also HYB's tail), one thread per entry: a row with several
// stored diagonal entries gets them added in no fixed order
template <typename V>
__global__ void coo_get_diag_kernel(const int32_t* __restrict__ ri, const int32_t* __restrict__ ci,
                                    const V* __restrict__ v, int64_t nnz, int64_t n,
                                    VecT<V>* __restrict__ d) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz && ri[k] == ci[k] && ci[k] >= 0 && ci[k] < n)
    atomicAdd(&d[ci[k]], (VecT<V>)vtraits<V>::widen(v[k]));
}

template <typename V>
int update_t(mph_matrix_impl* A, const void* diag, hipStream_t s) {
  const VecT<V>* d = static_cast<const VecT<V>*>(diag);
  const int64_t n = A->nrows;
  auto csr = [&]() {
    if (n > 0)
      hipLaunchKernelGGL((csr_update_kernel<V>), dim3(grid_for(n)), dim3(kT), 0, s, A->csr_ro,
                         A->csr_ci, static_cast<V*>(const_cast<void*>(A->csr_v)), n, d);
  };
  auto coo = [&]() {
    if (A->coo_nnz > 0)
      hipLaunchKernelGGL((coo_update_kernel<V>), dim3(grid_for(A->coo_nnz)), dim3(kT), 0, s,
                         A->coo_ri, A->coo_ci, static_cast<V*>(const_cast<void*>(A->coo_v)),
                         A->coo_nnz, std::min(A->nrows, A->ncols), d);
  };
  auto dia = [&]() {
    if (n > 0 && A->ndiag > 0)
      hipLaunchKernelGGL((dia_update_kernel<V>), dim3(grid_for(n)), dim3(kT), 0, s, A->dia_off,
                         static_cast<V*>(const_cast<void*>(A->dia_v)), A->ndiag, A->dia_pitch, n,
                         A->ncols, d);
  };
  auto ell = [&]() {
    if (n > 0 && A->ell_width > 0)
      hipLaunchKernelGGL((ell_update_kernel<V>), dim3(grid_for(n)), dim3(kT), 0, s, A->ell_ci,
                         static_cast<V*>(const_cast<void*>(A->ell_v)), n, A->ell_width,
                         A->ell_pitch, d);
  };
  switch (A->format) {
    case MPH_CSR: csr(); break;
    case MPH_COO: coo(); break;
    case MPH_DIA: dia(); break;
    case MPH_ELL: ell(); break;
    case MPH_HYB: ell(); coo(); break;
    case MPH_HDC: dia(); csr(); break;
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V>
int get_t(mph_matrix_impl* A, void* diag, hipStream_t s) {
  VecT<V>* d = static_cast<VecT<V>*>(diag);
  const int64_t n = std::min(A->nrows, A->ncols);
  if (n == 0) return MPH_OK;
  const DiagArgs a{A->format, A->ndiag, A->dia_pitch, A->ell_width, A->ell_pitch, A->dia_off,
                   A->ell_ci, A->csr_ro, A->csr_ci, A->dia_v, A->ell_v, A->csr_v};
  hipLaunchKernelGGL((get_diag_kernel<V>), dim3(grid_for(n)), dim3(kT), 0, s, a, d, n);
  if ((A->format == MPH_COO || A->format == MPH_HYB) && A->coo_nnz > 0)
    hipLaunchKernelGGL((coo_get_diag_kernel<V>), dim3(grid_for(A->coo_nnz)), dim3(kT), 0, s,
                       A->coo_ri, A->coo_ci, static_cast<const V*>(A->coo_v), A->coo_nnz, n, d);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace

}  // namespace mph

using namespace mph;

extern "C" {

int mph_matrix_update_diagonal(mph_matrix_t A, const void* diagonal, mph_stream_t stream) {
  if (!A) return invalid("null matrix handle");
  if (std::min(A->nrows, A->ncols) > 0 && !diagonal) return invalid("null diagonal");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (A->format == MPH_CSR || A->format == MPH_HDC)
    if (A->nrows > 0) MPH_CHECK_STATUS(check_row_offsets(A->csr_ro, A->nrows, A->csr_nnz, s, "CSR"));
  int st = MPH_OK;
  switch (A->vt) {
    case MPH_F64: st = update_t<double>(A, diagonal, s); break;
    case MPH_F32: st = update_t<float>(A, diagonal, s); break;
    case MPH_BF16: st = update_t<bf16>(A, diagonal, s); break;
    default: return invalid("unknown value type");
  }
  // the x-tiled plan keeps copies of values: analyse again before the next product
  release_plans(A);
  return st;
}

int mph_matrix_get_diagonal(mph_matrix_t A, void* diagonal, mph_stream_t stream) {
  if (!A) return invalid("null matrix handle");
  const int64_t n = std::min(A->nrows, A->ncols);
  if (n > 0 && !diagonal) return invalid("null diagonal");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (A->format == MPH_CSR || A->format == MPH_HDC)
    if (A->nrows > 0) MPH_CHECK_STATUS(check_row_offsets(A->csr_ro, A->nrows, A->csr_nnz, s, "CSR"));
  switch (A->vt) {
    case MPH_F64: return get_t<double>(A, diagonal, s);
    case MPH_F32: return get_t<float>(A, diagonal, s);
    case MPH_BF16: return get_t<bf16>(A, diagonal, s);
  }
  return invalid("unknown value type");
}

}  // extern "C"
