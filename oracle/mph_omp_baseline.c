/*
 * mph_omp_baseline.c -- restatement of the reference OpenMP multiply, timed
 * as the CPU baseline beside the GPU kernels.  TEST/BENCH INFRASTRUCTURE
 * ONLY (loaded by the benchmark's cpu_baseline leg and by tests).
 *
 * The reference OpenMP backend cannot be compiled here (it needs Kokkos
 * 3.5.00, which is not vendored), so these loops restate it over raw arrays.
 */
#include <omp.h>

#include "mph_oracle.h"

int orc_omp_max_threads(void) { return omp_get_max_threads(); }

/* core/src/impl/Csr/OpenMP/Morpheus_Multiply_Impl.hpp:50-57 (static schedule) */
void orc_omp_csr_spmv(int64_t nrows, const int32_t* ro, const int32_t* ci,
                      const double* v, const double* x, double* y, int init) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nrows; i++) {
    double sum = init ? 0.0 : y[i];
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) sum += v[jj] * x[ci[jj]];
    y[i] = sum;
  }
}

/* core/src/impl/Coo/OpenMP/Morpheus_Multiply_Impl.hpp:54-106: nnz split in
 * equal parts, atomics on the rows shared with the neighbouring threads. */
void orc_omp_coo_spmv(int64_t nrows, int64_t nnz, const int32_t* ri,
                      const int32_t* ci, const double* v, const double* x,
                      double* y, int init) {
  if (init) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; i++) y[i] = 0.0;
  }
  const int64_t sentinel_row = nrows + 1;
#pragma omp parallel
  {
    const int64_t nt = omp_get_num_threads();
    const int64_t work = (nnz + nt - 1) / nt;
    const int64_t tid = omp_get_thread_num();
    const int64_t begin = work * tid;
    const int64_t end = begin + work < nnz ? begin + work : nnz;
    int64_t n = begin;
    if (begin < end) {
      const int64_t first = begin > 0 ? ri[begin - 1] : sentinel_row;
      const int64_t last = end < nnz ? ri[end] : sentinel_row;
      if (first != sentinel_row) {
        double partial = 0.0;
        for (; n < end && ri[n] == first; n++) partial += v[n] * x[ci[n]];
#pragma omp atomic
        y[first] += partial;
      }
      for (; n < end && ri[n] != last; n++) y[ri[n]] += v[n] * x[ci[n]];
      if (last != sentinel_row) {
        double partial = 0.0;
        for (; n < end; n++) partial += v[n] * x[ci[n]];
#pragma omp atomic
        y[last] += partial;
      }
    }
  }
}

/* core/src/impl/Dia/OpenMP/Morpheus_Multiply_Impl.hpp:51-64 */
void orc_omp_dia_spmv(int64_t nrows, int64_t ncols, int64_t ndiag,
                      int64_t pitch, const int32_t* offsets,
                      const double* values, const double* x, double* y,
                      int init) {
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < nrows; row++) {
    double sum = init ? 0.0 : y[row];
    for (int64_t n = 0; n < ndiag; n++) {
      const int64_t col = row + offsets[n];
      if (col >= 0 && col < ncols) sum += values[row + pitch * n] * x[col];
    }
    y[row] = sum;
  }
}

/* core/src/impl/Ell/OpenMP/Morpheus_Multiply_Impl.hpp:53-68 */
void orc_omp_ell_spmv(int64_t nrows, int64_t width, int64_t pitch,
                      const int32_t* ci, const double* v, const double* x,
                      double* y, int init) {
  if (init) {
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < nrows; n++) y[n] = 0.0;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t n = 0; n < width; n++) {
      const int32_t j = ci[i + pitch * n];
      if (j != -1) y[i] += v[i + pitch * n] * x[j];
    }
}

/* parallel first-touch copies so pages land on the threads that use them */
void orc_omp_copy_f64(int64_t n, const double* src, double* dst) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) dst[i] = src[i];
}

void orc_omp_copy_i32(int64_t n, const int32_t* src, int32_t* dst) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) dst[i] = src[i];
}
