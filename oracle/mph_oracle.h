/*
 * mph_oracle.h -- CPU oracle for the Morpheus SpMV / Convert hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Serial backend (the reference's own "textbook" loops) plus a restatement of
 * its OpenMP multiply, used as the timed CPU baseline.  Only tests/, the
 * smoke check and the benchmark's cpu_baseline leg may load it, and only to
 * check results or time the CPU baseline.  The product path (morpheus_amd)
 * never calls into it.
 *
 * Storage conventions (identical to what the HIP kernels consume):
 *   - indices are int32, sizes int64, values double;
 *   - DIA values are column-major with leading dimension `pitch`
 *     (pitch = pad(nrows, 32) for containers built by the converters), i.e.
 *     values(row, d) = values[row + pitch * d];
 *   - ELL column_indices / values are column-major with leading dimension
 *     `pitch` (= nrows) and `width` stored columns (= pad(K, 32)), i.e.
 *     slot (row, n) = [row + pitch * n]; padding slots hold index -1, value 0.
 *
 * Return codes match the product C-ABI (include/morpheus_amd.h):
 *   0 OK, 2 INVALID_INPUT, 3 FORMAT_CONVERSION (fill tolerance exceeded).
 *
 * Every function below is pinned by the tests in tests/test_oracle_golden.py
 * and the other *_cpu.py tests (the reference's golden 10x10 fixtures and
 * MatrixMarket files, closed forms, brute force or scipy cross-checks).
 */
#ifndef MPH_ORACLE_H
#define MPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_INVALID_INPUT 2
#define ORC_FORMAT_CONVERSION 3

/* ----------------------------------------------------------------- multiply */
void orc_csr_spmv(int64_t nrows, const int32_t* ro, const int32_t* ci,
                  const double* v, const double* x, double* y, int init);
void orc_coo_spmv(int64_t nrows, int64_t nnz, const int32_t* ri,
                  const int32_t* ci, const double* v, const double* x,
                  double* y, int init);
void orc_dia_spmv(int64_t nrows, int64_t ncols, int64_t ndiag, int64_t pitch,
                  const int32_t* offsets, const double* values,
                  const double* x, double* y, int init);
void orc_ell_spmv(int64_t nrows, int64_t width, int64_t pitch,
                  const int32_t* ci, const double* v, const double* x,
                  double* y, int init);

/* Per-row sum of |a_ij * x_j| (the scale of the FP error bound) and an
 * FP64 row sum with Neumaier compensation (the "exact" reference for the
 * reduced-precision sweeps).  CSR layout. */
void orc_csr_absrow(int64_t nrows, const int32_t* ro, const int32_t* ci,
                    const double* v, const double* x, double* out);
void orc_csr_spmv_compensated(int64_t nrows, const int32_t* ro,
                              const int32_t* ci, const double* v,
                              const double* x, double* y);
/* FP32 row sums in four accumulation schemes (0 plain serial, 1 64 lanes +
 * tree, 2 TwoSum on rounded products, 3 Dot2): the witnesses that the FP32
 * accuracy check separates plain from compensated accumulation. */
void orc_csr_spmv_f32(int64_t nrows, const int32_t* ro, const int32_t* ci,
                      const float* v, const float* x, float* y, int mode);

/* ----------------------------------------------------------------- helpers */
int orc_exceeds_tolerance(int64_t num_rows, int64_t num_entries,
                          int64_t num_columns);
int64_t orc_pad_size(int64_t len, int64_t alignment);
int orc_coo_is_sorted(int64_t nnz, const int32_t* ri, const int32_t* ci);
/* stable sort of the first n entries by (row, col); in place. */
int orc_coo_sort_by_row_and_column(int64_t n, int32_t* ri, int32_t* ci,
                                   double* v);
void orc_csr_update_diagonal(int64_t nrows, const int32_t* ro, const int32_t* ci, double* v,
                             const double* d);
void orc_coo_update_diagonal(int64_t nnz, const int32_t* ri, const int32_t* ci, double* v,
                             const double* d);
void orc_dia_update_diagonal(int64_t nrows, int64_t ncols, int64_t ndiag, int64_t pitch,
                             const int32_t* off, double* v, const double* d);
void orc_ell_update_diagonal(int64_t nrows, int64_t width, int64_t pitch, const int32_t* ci,
                             double* v, const double* d);
double orc_dot(int64_t n, const double* x, const double* y);
void orc_waxpby(int64_t n, double alpha, const double* x, double beta, const double* y,
                double* w);
double orc_reduce(int64_t n, const double* in);
double orc_cg_csr(int64_t n, const int32_t* ro, const int32_t* ci, const double* v,
                  const double* b, double* x, int64_t iters);
void orc_count_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t nnz,
                                const int32_t* ri, const int32_t* ci,
                                int32_t* out /* nrows+ncols-1, accumulated */);
void orc_csr_nnz_per_row(int64_t nrows, const int32_t* ro, int32_t* out, int init);
void orc_coo_nnz_per_row(int64_t nrows, int64_t nnz, const int32_t* ri, int32_t* out,
                         int init);
void orc_dia_nnz_per_row(int64_t nrows, int64_t ncols, int64_t ndiag, const int32_t* off,
                         int32_t* out, int init);
void orc_ell_nnz_per_row(int64_t nrows, int64_t width, int64_t pitch, const int32_t* ci,
                         int32_t* out, int init);
void orc_csr_nnz_per_diagonal(int64_t nrows, int64_t ncols, const int32_t* ro,
                              const int32_t* ci, int32_t* out, int init);
void orc_dia_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t ndiag, int64_t pitch,
                              const int32_t* off, const double* v, int32_t* out, int init);
void orc_ell_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t width, int64_t pitch,
                              const int32_t* ci, const double* v, int32_t* out, int init);
void orc_i32_stats(int64_t n, const int32_t* v, double mean, int32_t threshold,
                   int64_t* mx, int64_t* mn, double* sqdev, int64_t* count);

/* ----------------------------------------------------------------- convert */
void orc_csr_to_coo(int64_t nrows, int64_t nnz, const int32_t* ro,
                    const int32_t* ci, const double* v, int32_t* ri_out,
                    int32_t* ci_out, double* v_out);
void orc_coo_to_csr(int64_t nrows, int64_t nnz, const int32_t* ri,
                    const int32_t* ci, const double* v, int32_t* ro_out,
                    int32_t* ci_out, double* v_out);

/* COO -> DIA.  First call orc_coo_dia_ndiag to size the output. */
int64_t orc_coo_dia_ndiag(int64_t nnz, const int32_t* ri, const int32_t* ci);
int orc_coo_to_dia(int64_t nrows, int64_t ncols, int64_t nnz,
                   const int32_t* ri, const int32_t* ci, const double* v,
                   int64_t ndiag, int64_t pitch, int32_t* offsets_out,
                   double* values_out /* pitch*ndiag, zero-filled by caller */);
/* DIA -> COO.  Output arrays have nnz slots and must be zero-filled by the
 * caller (the reference's freshly resized containers are zero-initialised). */
int orc_dia_to_coo(int64_t nrows, int64_t ncols, int64_t nnz, int64_t ndiag,
                   int64_t pitch, const int32_t* offsets, const double* values,
                   int32_t* ri_out, int32_t* ci_out, double* v_out);

/* COO -> ELL.  orc_coo_ell_entries_per_row gives K; width = pad(K, 32). */
int64_t orc_coo_ell_entries_per_row(int64_t nrows, int64_t nnz,
                                    const int32_t* ri);
int orc_coo_to_ell(int64_t nrows, int64_t nnz, const int32_t* ri,
                   const int32_t* ci, const double* v, int64_t K,
                   int64_t width, int64_t pitch, int32_t* ci_out,
                   double* v_out);
int orc_ell_to_coo(int64_t nrows, int64_t nnz, int64_t width, int64_t pitch,
                   const int32_t* ci, const double* v, int32_t* ri_out,
                   int32_t* ci_out, double* v_out);

/* COO -> HYB.  Sizes first: K = max(1, nnz/nrows), ell_nnz, coo_nnz. */
int orc_coo_hyb_sizes(int64_t nrows, int64_t nnz, const int32_t* ri,
                      int64_t* K, int64_t* ell_nnz, int64_t* coo_nnz);
int orc_coo_to_hyb(int64_t nrows, int64_t nnz, const int32_t* ri,
                   const int32_t* ci, const double* v, int64_t K,
                   int64_t width, int64_t pitch, int32_t* ell_ci,
                   double* ell_v, int32_t* coo_ri, int32_t* coo_ci,
                   double* coo_v);
int orc_hyb_to_coo(int64_t nrows, int64_t nnz, int64_t width, int64_t pitch,
                   const int32_t* ell_ci, const double* ell_v,
                   int64_t coo_nnz, const int32_t* coo_ri,
                   const int32_t* coo_ci, const double* coo_v,
                   int32_t* ri_out, int32_t* ci_out, double* v_out);

/* COO -> HDC.  Sizes first: ndiag, dia_nnz, csr_nnz. */
int orc_coo_hdc_sizes(int64_t nrows, int64_t ncols, int64_t nnz,
                      const int32_t* ri, const int32_t* ci, int64_t* ndiag,
                      int64_t* dia_nnz, int64_t* csr_nnz);
int orc_coo_to_hdc(int64_t nrows, int64_t ncols, int64_t nnz,
                   const int32_t* ri, const int32_t* ci, const double* v,
                   int64_t ndiag, int64_t pitch, int32_t* dia_offsets,
                   double* dia_values /* pitch*ndiag, zero-filled */,
                   int32_t* csr_ro /* nrows+1 */, int32_t* csr_ci,
                   double* csr_v);
int orc_hdc_to_coo(int64_t nrows, int64_t ncols, int64_t nnz, int64_t ndiag,
                   int64_t pitch, const int32_t* dia_offsets,
                   const double* dia_values, int64_t csr_nnz,
                   const int32_t* csr_ro, const int32_t* csr_ci,
                   const double* csr_v, int32_t* ri_out, int32_t* ci_out,
                   double* v_out);

/* ----------------------------------------------------- OpenMP CPU baseline */
int orc_omp_max_threads(void);
void orc_omp_csr_spmv(int64_t nrows, const int32_t* ro, const int32_t* ci,
                      const double* v, const double* x, double* y, int init);
void orc_omp_coo_spmv(int64_t nrows, int64_t nnz, const int32_t* ri,
                      const int32_t* ci, const double* v, const double* x,
                      double* y, int init);
void orc_omp_dia_spmv(int64_t nrows, int64_t ncols, int64_t ndiag,
                      int64_t pitch, const int32_t* offsets,
                      const double* values, const double* x, double* y,
                      int init);
void orc_omp_ell_spmv(int64_t nrows, int64_t width, int64_t pitch,
                      const int32_t* ci, const double* v, const double* x,
                      double* y, int init);
/* parallel first-touch fill of a double / int32 buffer */
void orc_omp_copy_f64(int64_t n, const double* src, double* dst);
void orc_omp_copy_i32(int64_t n, const int32_t* src, int32_t* dst);

#ifdef __cplusplus
}
#endif
#endif
