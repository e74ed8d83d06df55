// analytics.hip -- matrix analytics on the GPU: non-zeros per row and per
// diagonal for the six formats, and the reductions the format selection
// reads off them.
//
// Replaces core/src/Morpheus_MatrixAnalytics.hpp:112-261 with the per-format
// counts of core/src/impl/{Coo,Csr,Dia,Ell,Hyb,Hdc}/Serial/
// Morpheus_NonZeros_Per_Row_Impl.hpp and core/src/impl/{Coo,Csr,Dia,Ell}/
// Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp (the Serial and HIP entry
// points of the per-diagonal count forward to those), and the reductions of
// core/src/impl/DenseVector/Kokkos/Morpheus_VectorAnalytics_Impl.hpp:50-127.
//
// What each format counts is the reference's, quirks included:
//  * per row: CSR / COO every stored entry, DIA every in-range slot (stored
//    zeros too), ELL every slot of the stored width whose index is not -1;
//  * per diagonal (entry col - row + nrows - 1): CSR / COO every stored
//    entry, DIA the non-zero values of each stored diagonal over all pitch
//    rows, ELL the slots with a valid index and a non-zero value among a
//    row's FIRST c slots, c being the row's count of such slots;
//  * HYB = ELL then COO, HDC = DIA then CSR.
// Counting uses wave-aggregated atomics: the lanes of a wave that hit the
// same counter add once, so a stencil's few diagonals do not serialise.
#include <algorithm>
#include <vector>

#include "mph_internal.hpp"

namespace mph {

namespace {

constexpr int kT = 256;

inline unsigned grid_for(int64_t n, int per = kT) {
  return (unsigned)std::max<int64_t>(1, ceil_div(n, per));
}

template <typename V>
__device__ __forceinline__ bool value_nonzero(V v) {
  return v != (V)0;
}
template <>
__device__ __forceinline__ bool value_nonzero<bf16>(bf16 v) {
  return (v.bits & 0x7fffu) != 0;
}

// Adds 1 to counter[key] for every lane with `valid`, one atomic per
// distinct key of the wave.  Every lane of the wave must call it.
__device__ __forceinline__ void wave_count(int32_t* __restrict__ counter, int64_t key,
                                           bool valid) {
  const int lane = threadIdx.x % kWave;
  uint64_t pending = __ballot(valid);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int64_t k = __shfl(key, leader, kWave);
    const uint64_t same = __ballot(valid && key == k) & pending;
    if (lane == leader) atomicAdd(&counter[k], (int32_t)__popcll(same));
    pending &= ~same;
  }
}

// ------------------------------------------------------------- per row
__global__ void csr_rows_kernel(const int32_t* __restrict__ ro, int64_t nrows,
                                int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrows) out[i] += ro[i + 1] - ro[i];
}

__global__ void coo_rows_kernel(const int32_t* __restrict__ ri, int64_t nnz, int64_t nrows,
                                int32_t* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = n < nnz ? (int64_t)ri[n] : -1;
  wave_count(out, r, r >= 0 && r < nrows);
}

__global__ void dia_rows_kernel(const int32_t* __restrict__ off, int64_t ndiag, int64_t nrows,
                                int64_t ncols, int32_t* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  int32_t c = 0;
  for (int64_t d = 0; d < ndiag; d++) {
    const int64_t col = row + off[d];
    c += (col >= 0 && col < ncols) ? 1 : 0;
  }
  out[row] += c;
}

__global__ void ell_rows_kernel(const int32_t* __restrict__ ci, int64_t nrows, int64_t width,
                                int64_t pitch, int32_t* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  int32_t c = 0;
  for (int64_t n = 0; n < width; n++) c += ci[row + pitch * n] != -1 ? 1 : 0;
  out[row] += c;
}

// ------------------------------------------------------------- per diagonal
__global__ void coo_diag_kernel(const int32_t* __restrict__ ri, const int32_t* __restrict__ ci,
                                int64_t nnz, int64_t nrows, int64_t ncols,
                                int32_t* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t r = -1, c = -1;
  if (n < nnz) {
    r = ri[n];
    c = ci[n];
  }
  wave_count(out, c - r + nrows - 1, r >= 0 && r < nrows && c >= 0 && c < ncols);
}

// one wave per row
__global__ void csr_diag_kernel(const int32_t* __restrict__ ro, const int32_t* __restrict__ ci,
                                int64_t nrows, int64_t ncols, int32_t* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * (kT / kWave) + threadIdx.x / kWave;
  if (row >= nrows) return;  // whole wave
  const int lane = threadIdx.x % kWave;
  const int64_t b = ro[row], e = ro[row + 1];
  for (int64_t j0 = b; j0 < e; j0 += kWave) {
    const int64_t j = j0 + lane;
    const int64_t c = j < e ? (int64_t)ci[j] : -1;
    wave_count(out, c - row + nrows - 1, c >= 0 && c < ncols);
  }
}

// one block per stored diagonal
template <typename V>
__global__ void dia_diag_kernel(const int32_t* __restrict__ off, const V* __restrict__ val,
                                int64_t pitch, int64_t nrows, int64_t ncols,
                                int32_t* __restrict__ out) {
  __shared__ int32_t s_cnt[kT / kWave];
  const int64_t d = blockIdx.x;
  int32_t c = 0;
  for (int64_t j = threadIdx.x; j < pitch; j += kT) c += value_nonzero(val[j + pitch * d]) ? 1 : 0;
  for (int m = kWave / 2; m > 0; m >>= 1) c += __shfl_xor(c, m, kWave);
  if (threadIdx.x % kWave == 0) s_cnt[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < kT / kWave; w++) t += s_cnt[w];
    const int64_t idx = (int64_t)off[d] + nrows - 1;
    if (t > 0 && idx >= 0 && idx < nrows + ncols - 1) atomicAdd(&out[idx], t);
  }
}

// one thread per row
template <typename V>
__global__ void ell_diag_kernel(const int32_t* __restrict__ ci, const V* __restrict__ val,
                                int64_t nrows, int64_t ncols, int64_t width, int64_t pitch,
                                int32_t* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t cnt = 0;
  if (row < nrows)
    for (int64_t n = 0; n < width; n++)
      cnt += (ci[row + pitch * n] != -1 && value_nonzero(val[row + pitch * n])) ? 1 : 0;
  // the reference's second pass: the row's first cnt slots (the wave loops
  // to its largest cnt, every lane taking part in each count)
  int64_t most = cnt;
  for (int m = kWave / 2; m > 0; m >>= 1) {
    const int64_t o = __shfl_xor(most, m, kWave);
    most = o > most ? o : most;
  }
  for (int64_t n = 0; n < most; n++) {
    int64_t c = -1;
    bool ok = false;
    if (row < nrows && n < cnt) {
      c = ci[row + pitch * n];
      ok = c >= 0 && c < ncols && value_nonzero(val[row + pitch * n]);
    }
    wave_count(out, c - row + nrows - 1, ok);
  }
}

// ------------------------------------------------------------- reductions
struct I32Part {
  int64_t mx, mn, count;
  double sq;
};

__device__ __forceinline__ I32Part merge_part(const I32Part& a, const I32Part& b) {
  return I32Part{a.mx > b.mx ? a.mx : b.mx, a.mn < b.mn ? a.mn : b.mn, a.count + b.count,
                 a.sq + b.sq};
}

// fixed-order block reduction (LDS tree), so the double sum is reproducible
__device__ I32Part block_reduce(I32Part p, I32Part* s) {
  s[threadIdx.x] = p;
  __syncthreads();
  for (int m = kT / 2; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) s[threadIdx.x] = merge_part(s[threadIdx.x], s[threadIdx.x + m]);
    __syncthreads();
  }
  return s[0];
}

constexpr I32Part kEmpty{INT64_MIN, INT64_MAX, 0, 0.0};

// block b reduces the contiguous chunk [b * chunk, (b + 1) * chunk)
__global__ void i32_stats_kernel(const int32_t* __restrict__ v, int64_t n, int64_t chunk,
                                 double mean, int32_t threshold, I32Part* __restrict__ part) {
  __shared__ I32Part s[kT];
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  I32Part p = kEmpty;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
    const int64_t x = v[i];
    p.mx = x > p.mx ? x : p.mx;
    p.mn = x < p.mn ? x : p.mn;
    p.sq += ((double)x - mean) * ((double)x - mean);
    p.count += (x > threshold && x != 0) ? 1 : 0;
  }
  p = block_reduce(p, s);
  if (threadIdx.x == 0) part[blockIdx.x] = p;
}

__global__ void i32_stats_final_kernel(const I32Part* __restrict__ part, int64_t nparts,
                                       I32Part* __restrict__ out) {
  __shared__ I32Part s[kT];
  I32Part p = kEmpty;
  for (int64_t i = threadIdx.x; i < nparts; i += kT) p = merge_part(p, part[i]);
  p = block_reduce(p, s);
  if (threadIdx.x == 0) *out = p;
}

int check_offsets_in_range(const mph_matrix_impl* A, hipStream_t s) {
  if (A->ndiag == 0) return MPH_OK;
  std::vector<int32_t> off((size_t)A->ndiag);
  MPH_HIP_CHECK(hipMemcpyAsync(off.data(), A->dia_off, sizeof(int32_t) * off.size(),
                               hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  for (int32_t o : off)
    if (o < -(A->nrows - 1) || o > A->ncols - 1) return invalid("DIA offset out of range");
  return MPH_OK;
}

int rows_part(const mph_matrix_impl* A, int fmt, int32_t* out, hipStream_t s) {
  const int64_t n = A->nrows;
  switch (fmt) {
    case MPH_CSR:
      if (n > 0) {
        MPH_CHECK_STATUS(check_row_offsets(A->csr_ro, n, A->csr_nnz, s, "CSR"));
        hipLaunchKernelGGL(csr_rows_kernel, dim3(grid_for(n)), dim3(kT), 0, s, A->csr_ro, n, out);
      }
      break;
    case MPH_COO:
      if (A->coo_nnz > 0)
        hipLaunchKernelGGL(coo_rows_kernel, dim3(grid_for(A->coo_nnz)), dim3(kT), 0, s,
                           A->coo_ri, A->coo_nnz, n, out);
      break;
    case MPH_DIA:
      if (n > 0 && A->ndiag > 0)
        hipLaunchKernelGGL(dia_rows_kernel, dim3(grid_for(n)), dim3(kT), 0, s, A->dia_off,
                           A->ndiag, n, A->ncols, out);
      break;
    case MPH_ELL:
      if (n > 0 && A->ell_width > 0)
        hipLaunchKernelGGL(ell_rows_kernel, dim3(grid_for(n)), dim3(kT), 0, s, A->ell_ci, n,
                           A->ell_width, A->ell_pitch, out);
      break;
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V>
int diag_part(const mph_matrix_impl* A, int fmt, int32_t* out, hipStream_t s) {
  const int64_t n = A->nrows, m = A->ncols;
  switch (fmt) {
    case MPH_CSR:
      if (n > 0) {
        MPH_CHECK_STATUS(check_row_offsets(A->csr_ro, n, A->csr_nnz, s, "CSR"));
        hipLaunchKernelGGL(csr_diag_kernel, dim3(grid_for(n, kT / kWave)), dim3(kT), 0, s,
                           A->csr_ro, A->csr_ci, n, m, out);
      }
      break;
    case MPH_COO:
      if (A->coo_nnz > 0)
        hipLaunchKernelGGL(coo_diag_kernel, dim3(grid_for(A->coo_nnz)), dim3(kT), 0, s,
                           A->coo_ri, A->coo_ci, A->coo_nnz, n, m, out);
      break;
    case MPH_DIA:
      if (A->ndiag > 0 && A->dia_pitch > 0)
        hipLaunchKernelGGL((dia_diag_kernel<V>), dim3((unsigned)A->ndiag), dim3(kT), 0, s,
                           A->dia_off, static_cast<const V*>(A->dia_v), A->dia_pitch, n, m, out);
      break;
    case MPH_ELL:
      if (n > 0 && A->ell_width > 0)
        hipLaunchKernelGGL((ell_diag_kernel<V>), dim3(grid_for(n)), dim3(kT), 0, s, A->ell_ci,
                           static_cast<const V*>(A->ell_v), n, m, A->ell_width, A->ell_pitch,
                           out);
      break;
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

// the parts of a format: HYB = ELL then COO, HDC = DIA then CSR
void parts_of(int fmt, int* first, int* second) {
  *second = -1;
  *first = fmt;
  if (fmt == MPH_HYB) {
    *first = MPH_ELL;
    *second = MPH_COO;
  } else if (fmt == MPH_HDC) {
    *first = MPH_DIA;
    *second = MPH_CSR;
  }
}

}  // namespace

}  // namespace mph

using namespace mph;

extern "C" {

int mph_matrix_count_nnz_per_row(mph_matrix_t A, int32_t* nnz_per_row, int32_t init,
                                 mph_stream_t stream) {
  if (!A) return invalid("null matrix handle");
  if (A->nrows > 0 && !nnz_per_row) return invalid("null nnz_per_row");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (init) MPH_CHECK_STATUS(fill_zero(nnz_per_row, sizeof(int32_t) * A->nrows, s));
  int f1, f2;
  parts_of(A->format, &f1, &f2);
  if (f1 == MPH_COO || f2 == MPH_COO)
    MPH_CHECK_STATUS(check_index_range(A->coo_ri, A->coo_nnz, A->nrows, false, s, "COO rows"));
  MPH_CHECK_STATUS(rows_part(A, f1, nnz_per_row, s));
  if (f2 >= 0) MPH_CHECK_STATUS(rows_part(A, f2, nnz_per_row, s));
  return MPH_OK;
}

int mph_matrix_count_nnz_per_diagonal(mph_matrix_t A, int32_t* nnz_per_diagonal, int32_t init,
                                      mph_stream_t stream) {
  if (!A) return invalid("null matrix handle");
  const int64_t nd = A->nrows + A->ncols - 1;
  if (nd > 0 && !nnz_per_diagonal) return invalid("null nnz_per_diagonal");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (init && nd > 0) MPH_CHECK_STATUS(fill_zero(nnz_per_diagonal, sizeof(int32_t) * nd, s));
  if (A->nrows == 0 || A->ncols == 0) return MPH_OK;
  int f1, f2;
  parts_of(A->format, &f1, &f2);
  // the counters are indexed by col - row + nrows - 1: indices are checked
  for (int f : {f1, f2}) {
    if (f == MPH_DIA) MPH_CHECK_STATUS(check_offsets_in_range(A, s));
    if (f == MPH_COO) {
      MPH_CHECK_STATUS(check_index_range(A->coo_ri, A->coo_nnz, A->nrows, false, s, "COO rows"));
      MPH_CHECK_STATUS(check_index_range(A->coo_ci, A->coo_nnz, A->ncols, false, s, "COO columns"));
    }
    if (f == MPH_CSR)
      MPH_CHECK_STATUS(check_index_range(A->csr_ci, A->csr_nnz, A->ncols, false, s, "CSR columns"));
    if (f == MPH_ELL)
      MPH_CHECK_STATUS(check_index_range(A->ell_ci, A->ell_pitch * A->ell_width, A->ncols, true, s,
                                         "ELL columns"));
  }
  for (int f : {f1, f2}) {
    if (f < 0) continue;
    int st = MPH_OK;
    switch (A->vt) {
      case MPH_F64: st = diag_part<double>(A, f, nnz_per_diagonal, s); break;
      case MPH_F32: st = diag_part<float>(A, f, nnz_per_diagonal, s); break;
      case MPH_BF16: st = diag_part<bf16>(A, f, nnz_per_diagonal, s); break;
      default: return invalid("unknown value type");
    }
    if (st != MPH_OK) return st;
  }
  return MPH_OK;
}

int mph_i32_stats(const int32_t* v, int64_t n, double mean, int32_t threshold,
                  mph_i32_stats_t* out, mph_stream_t stream) {
  if (!out) return invalid("null output");
  if (n < 0) return invalid("negative size");
  if (n > 0 && !v) return invalid("null vector");
  if (n == 0) {
    *out = mph_i32_stats_t{0, 0, 0, 0.0};
    return MPH_OK;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int64_t chunk = std::max<int64_t>(kT * 16, ceil_div(n, (int64_t)1024));
  const int64_t nparts = ceil_div(n, chunk);
  I32Part* d = nullptr;
  MPH_HIP_CHECK(hipMalloc(&d, sizeof(I32Part) * (nparts + 1)));
  hipLaunchKernelGGL(i32_stats_kernel, dim3((unsigned)nparts), dim3(kT), 0, s, v, n, chunk, mean,
                     threshold, d);
  hipLaunchKernelGGL(i32_stats_final_kernel, dim3(1), dim3(kT), 0, s, d, nparts, d + nparts);
  I32Part h{};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d + nparts, sizeof(I32Part), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(d);
  MPH_HIP_CHECK(e);
  *out = mph_i32_stats_t{h.mx, h.mn, h.count, h.sq};
  return MPH_OK;
}

}  // extern "C"
