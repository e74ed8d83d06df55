/*
 * mph_oracle.c -- CPU oracle (Serial restatement) for the Morpheus SpMV and
 * Convert hot path.  TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke
 * check and the benchmark's cpu_baseline leg, to check results or time the
 * CPU baseline; never by the product path.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the morpheus repository root).  Build with -ffp-contract=off so that the
 * Serial loops keep the reference's separate multiply and add roundings.
 */
#include "mph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ========================================================================
 * Multiply (Serial backend)
 * ===================================================================== */

/* core/src/impl/Csr/Serial/Morpheus_Multiply_Impl.hpp:51-57 */
void orc_csr_spmv(int64_t nrows, const int32_t* ro, const int32_t* ci,
                  const double* v, const double* x, double* y, int init) {
  for (int64_t i = 0; i < nrows; i++) {
    double sum = init ? 0.0 : y[i];
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) sum += v[jj] * x[ci[jj]];
    y[i] = sum;
  }
}

/* core/src/impl/Coo/Serial/Morpheus_Multiply_Impl.hpp:51-59 */
void orc_coo_spmv(int64_t nrows, int64_t nnz, const int32_t* ri,
                  const int32_t* ci, const double* v, const double* x,
                  double* y, int init) {
  if (init)
    for (int64_t n = 0; n < nrows; n++) y[n] = 0.0;
  for (int64_t n = 0; n < nnz; n++) y[ri[n]] += v[n] * x[ci[n]];
}

/* core/src/impl/Dia/Serial/Morpheus_Multiply_Impl.hpp:49-63
 * values(row, n) is column-major with leading dimension pitch. */
void orc_dia_spmv(int64_t nrows, int64_t ncols, int64_t ndiag, int64_t pitch,
                  const int32_t* offsets, const double* values,
                  const double* x, double* y, int init) {
  for (int64_t row = 0; row < nrows; row++) {
    double sum = init ? 0.0 : y[row];
    for (int64_t n = 0; n < ndiag; n++) {
      const int64_t col = row + offsets[n];
      if (col >= 0 && col < ncols) sum += values[row + pitch * n] * x[col];
    }
    y[row] = sum;
  }
}

/* core/src/impl/Ell/Serial/Morpheus_Multiply_Impl.hpp:49-66: every stored
 * column (the 32-padded width) is visited; slots holding -1 are skipped. */
void orc_ell_spmv(int64_t nrows, int64_t width, int64_t pitch,
                  const int32_t* ci, const double* v, const double* x,
                  double* y, int init) {
  if (init)
    for (int64_t n = 0; n < nrows; n++) y[n] = 0.0;
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t n = 0; n < width; n++) {
      const int32_t j = ci[i + pitch * n];
      if (j != -1) y[i] += v[i + pitch * n] * x[j];
    }
}

/* Scale of the FP error bound |y - y_ref| <= 4 nnz_row u sum_j |a_ij x_j|. */
void orc_csr_absrow(int64_t nrows, const int32_t* ro, const int32_t* ci,
                    const double* v, const double* x, double* out) {
  for (int64_t i = 0; i < nrows; i++) {
    double s = 0.0;
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) s += fabs(v[jj] * x[ci[jj]]);
    out[i] = s;
  }
}

/* FP64 row sums with Neumaier's compensated summation (Neumaier 1974);
 * used as the high-accuracy reference of the FP32/BF16 precision sweep. */
void orc_csr_spmv_compensated(int64_t nrows, const int32_t* ro,
                              const int32_t* ci, const double* v,
                              const double* x, double* y) {
  for (int64_t i = 0; i < nrows; i++) {
    double s = 0.0, c = 0.0;
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) {
      const double p = v[jj] * x[ci[jj]];
      const double e = fma(v[jj], x[ci[jj]], -p); /* exact: a*x - p */
      const double t = s + p;
      if (fabs(s) >= fabs(p))
        c += (s - t) + p;
      else
        c += (p - t) + s;
      s = t;
      c += e;
    }
    y[i] = s + c;
  }
}

/* FP32 row sums over FP32 values and x, in four accumulation schemes.  They
 * do not restate a reference function (the reference accumulates in the
 * value type, core/src/impl/Csr/Serial/Morpheus_Multiply_Impl.hpp:51-57);
 * they are the witnesses that show the FP32 accuracy check of the tests
 * tells a plain FP32 accumulator from a compensated one:
 *   mode 0  serial sum of rounded products (the Serial loop in float);
 *   mode 1  64 lane-strided partial sums, then a pairwise tree (the shape of
 *           a plain GPU row reduction);
 *   mode 2  TwoSum (Knuth) on the rounded products: the additions are
 *           compensated, the products are not;
 *   mode 3  Dot2 (Ogita, Rump, Oishi, "Accurate sum and dot product", SIAM J.
 *           Sci. Comput. 26(6), 2005, Algorithm 5.3): TwoSum on the additions
 *           and the FMA form of TwoProduct (e = fma(a, x, -p)) on the
 *           products, both error terms added into one FP32 correction.
 * Built with -ffp-contract=off, so every float operation rounds once. */
static void two_sum_f(float a, float b, float* s, float* e) {
  const float t = a + b;
  const float bb = t - a;
  *e = (a - (t - bb)) + (b - bb);
  *s = t;
}

void orc_csr_spmv_f32(int64_t nrows, const int32_t* ro, const int32_t* ci,
                      const float* v, const float* x, float* y, int mode) {
  for (int64_t i = 0; i < nrows; i++) {
    const int32_t b = ro[i], e = ro[i + 1];
    if (mode == 1) {
      float lane[64];
      for (int l = 0; l < 64; l++) lane[l] = 0.0f;
      for (int32_t jj = b; jj < e; jj++) lane[(jj - b) % 64] += v[jj] * x[ci[jj]];
      for (int w = 32; w > 0; w >>= 1)
        for (int l = 0; l < w; l++) lane[l] += lane[l + w];
      y[i] = lane[0];
      continue;
    }
    float s = 0.0f, c = 0.0f;
    for (int32_t jj = b; jj < e; jj++) {
      const float a = v[jj], xv = x[ci[jj]];
      const float p = a * xv;
      if (mode == 0) {
        s += p;
      } else {
        float t, es;
        two_sum_f(s, p, &t, &es);
        s = t;
        c += (mode == 3) ? es + fmaf(a, xv, -p) : es;
      }
    }
    y[i] = (mode == 0) ? s : s + c;
  }
}

/* ========================================================================
 * Helpers
 * ===================================================================== */

/* core/src/impl/Morpheus_Utils.hpp:91-105 -- the decision is made in float. */
int orc_exceeds_tolerance(int64_t num_rows, int64_t num_entries,
                          int64_t num_columns) {
  const float max_fill = 10.0f;
  const float threshold = 100e6f;
  const float size = (float)num_columns * (float)num_rows;
  const float ne = (float)num_entries;
  const float fill_ratio = size / (ne > 1.0f ? ne : 1.0f);
  return (fill_ratio > max_fill && size > threshold) ? 1 : 0;
}

/* core/src/impl/Dia/Morpheus_Utils_Impl.hpp:33-36 */
int64_t orc_pad_size(int64_t len, int64_t alignment) {
  return alignment * ((len + alignment - 1) / alignment);
}

/* core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:168-189 */
int orc_coo_is_sorted(int64_t nnz, const int32_t* ri, const int32_t* ci) {
  for (int64_t i = 0; i + 1 < nnz; i++)
    if (ri[i] > ri[i + 1] || (ri[i] == ri[i + 1] && ci[i] > ci[i + 1]))
      return 0;
  return 1;
}

/* core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:63-109: stable counting
 * sort of keys, carrying vals; min is forced to 0, bins = max + 2. */
static int counting_sort_by_key(int64_t n, int32_t* keys, int32_t* vals,
                                int32_t maxk) {
  if (maxk < 0) return ORC_INVALID_INPUT;
  const int64_t size = (int64_t)maxk;
  int64_t* counts = (int64_t*)calloc((size_t)(size + 2), sizeof(int64_t));
  int32_t* tk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int32_t* tv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  memcpy(tk, keys, sizeof(int32_t) * (size_t)n);
  memcpy(tv, vals, sizeof(int32_t) * (size_t)n);
  for (int64_t i = 0; i < n; i++) counts[keys[i] + 1]++;
  for (int64_t i = 0; i < size; i++) counts[i + 1] += counts[i];
  for (int64_t i = 0; i < n; i++) {
    keys[counts[tk[i]]] = tk[i];
    vals[counts[tk[i]]++] = tv[i];
  }
  free(counts);
  free(tk);
  free(tv);
  return ORC_OK;
}

static int32_t max_elem(int64_t n, const int32_t* a) {
  int32_t m = INT32_MIN; /* std::numeric_limits<int>::min(), Sort_Impl.hpp:54 */
  for (int64_t i = 0; i < n; i++)
    if (a[i] > m) m = a[i];
  return m;
}

/* core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:111-166: sort by column
 * (stable), gather rows through the permutation, sort by row (stable),
 * gather columns and values.  Net effect: stable lexicographic (row, col). */
int orc_coo_sort_by_row_and_column(int64_t n, int32_t* ri, int32_t* ci,
                                   double* v) {
  if (n <= 0) return ORC_OK;
  for (int64_t i = 0; i < n; i++)
    if (ri[i] < 0 || ci[i] < 0) return ORC_INVALID_INPUT;
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* temp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  double* tv = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  const int32_t maxr = max_elem(n, ri);
  const int32_t maxc = max_elem(n, ci);
  int rc = ORC_OK;

  memcpy(temp, ci, sizeof(int32_t) * (size_t)n);
  rc = counting_sort_by_key(n, temp, perm, maxc);
  if (rc == ORC_OK) {
    memcpy(temp, ri, sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) ri[i] = temp[perm[i]];
    rc = counting_sort_by_key(n, ri, perm, maxr);
  }
  if (rc == ORC_OK) {
    memcpy(temp, ci, sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) ci[i] = temp[perm[i]];
    memcpy(tv, v, sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; i++) v[i] = tv[perm[i]];
  }
  free(perm);
  free(temp);
  free(tv);
  return rc;
}

/* core/src/impl/Coo/Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp:62-73
 * (the Serial entry point forwards there,
 * core/src/impl/Coo/Serial/Morpheus_NonZeros_Per_Diagonal_Impl.hpp:40-51). */
void orc_count_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t nnz,
                                const int32_t* ri, const int32_t* ci,
                                int32_t* out) {
  (void)ncols;
  for (int64_t i = 0; i < nnz; i++) out[(int64_t)ci[i] - ri[i] + nrows - 1]++;
}

/* ========================================================================
 * Matrix analytics (Serial backend)
 * ===================================================================== */

/* core/src/impl/Csr/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp:53-61 */
void orc_csr_nnz_per_row(int64_t nrows, const int32_t* ro, int32_t* out, int init) {
  for (int64_t i = 0; i < nrows; i++) {
    int32_t nz = init ? 0 : out[i];
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) nz++;
    out[i] = nz;
  }
}

/* core/src/impl/Coo/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp:51-60 */
void orc_coo_nnz_per_row(int64_t nrows, int64_t nnz, const int32_t* ri, int32_t* out,
                         int init) {
  if (init)
    for (int64_t n = 0; n < nrows; n++) out[n] = 0;
  for (int64_t n = 0; n < nnz; n++) out[ri[n]]++;
}

/* core/src/impl/Dia/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp:53-66: every
 * in-range slot counts, stored zeros included */
void orc_dia_nnz_per_row(int64_t nrows, int64_t ncols, int64_t ndiag, const int32_t* off,
                         int32_t* out, int init) {
  for (int64_t row = 0; row < nrows; row++) {
    int32_t nz = init ? 0 : out[row];
    for (int64_t n = 0; n < ndiag; n++) {
      const int64_t col = row + off[n];
      if (col >= 0 && col < ncols) nz++;
    }
    out[row] = nz;
  }
}

/* core/src/impl/Ell/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp:54-69: every
 * slot of the stored width whose index is not the padding -1 */
void orc_ell_nnz_per_row(int64_t nrows, int64_t width, int64_t pitch, const int32_t* ci,
                         int32_t* out, int init) {
  if (init)
    for (int64_t n = 0; n < nrows; n++) out[n] = 0;
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t n = 0; n < width; n++)
      if (ci[i + pitch * n] != -1) out[i]++;
}

/* core/src/impl/Csr/Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp:59-76
 * (the Serial entry point forwards there): every stored entry */
void orc_csr_nnz_per_diagonal(int64_t nrows, int64_t ncols, const int32_t* ro,
                              const int32_t* ci, int32_t* out, int init) {
  if (init)
    for (int64_t n = 0; n < nrows + ncols - 1; n++) out[n] = 0;
  for (int64_t i = 0; i < nrows; i++)
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) out[(int64_t)ci[jj] - i + nrows - 1]++;
}

/* core/src/impl/Dia/Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp:61-76:
 * the non-zero values of each stored diagonal, over all pitch rows */
void orc_dia_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t ndiag, int64_t pitch,
                              const int32_t* off, const double* v, int32_t* out, int init) {
  if (init)
    for (int64_t n = 0; n < nrows + ncols - 1; n++) out[n] = 0;
  for (int64_t i = 0; i < ndiag; i++)
    for (int64_t j = 0; j < pitch; j++)
      if (v[j + pitch * i] != 0) out[(int64_t)off[i] + nrows - 1]++;
}

/* core/src/impl/Ell/Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp:63-122: a
 * row's count c is its slots with a valid index and a non-zero value; the
 * second pass then looks at the row's FIRST c slots only (so a zero value
 * ahead of a non-zero one hides the latter) */
void orc_ell_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t width, int64_t pitch,
                              const int32_t* ci, const double* v, int32_t* out, int init) {
  if (init)
    for (int64_t n = 0; n < nrows + ncols - 1; n++) out[n] = 0;
  for (int64_t i = 0; i < nrows; i++) {
    int64_t c = 0;
    for (int64_t n = 0; n < width; n++)
      if (ci[i + pitch * n] != -1 && v[i + pitch * n] != 0) c++;
    for (int64_t n = 0; n < c; n++) {
      const int32_t col = ci[i + pitch * n];
      if (col != -1 && v[i + pitch * n] != 0) out[(int64_t)col - i + nrows - 1]++;
    }
  }
}

/* core/src/impl/DenseVector/Kokkos/Morpheus_VectorAnalytics_Impl.hpp:50-127:
 * max, min, the sum of squared deviations from `mean` (std is
 * sqrt(sum / n)), and the count of entries > threshold and != 0 */
void orc_i32_stats(int64_t n, const int32_t* v, double mean, int32_t threshold,
                   int64_t* mx, int64_t* mn, double* sqdev, int64_t* count) {
  int64_t a = n > 0 ? v[0] : 0, b = n > 0 ? v[0] : 0, c = 0;
  double s = 0;
  for (int64_t i = 0; i < n; i++) {
    if (v[i] > a) a = v[i];
    if (v[i] < b) b = v[i];
    s += (v[i] - mean) * (v[i] - mean);
    if (v[i] > threshold && v[i] != 0) c++;
  }
  *mx = a;
  *mn = b;
  *sqdev = s;
  *count = c;
}

/* ========================================================================
 * Dense vector algorithms (Serial backend)
 * ===================================================================== */

/* core/src/impl/DenseVector/Serial/Morpheus_Dot_Impl.hpp:48-60 */
double orc_dot(int64_t n, const double* x, const double* y) {
  double r = 0;
  if (x == y)
    for (int64_t i = 0; i < n; i++) r += x[i] * x[i];
  else
    for (int64_t i = 0; i < n; i++) r += x[i] * y[i];
  return r;
}

/* core/src/impl/DenseVector/Serial/Morpheus_WAXPBY_Impl.hpp:49-64 (built
 * with -ffp-contract=off: every product and sum rounded on its own) */
void orc_waxpby(int64_t n, double alpha, const double* x, double beta, const double* y,
                double* w) {
  if (alpha == 1.0)
    for (int64_t i = 0; i < n; i++) w[i] = x[i] + beta * y[i];
  else if (beta == 1.0)
    for (int64_t i = 0; i < n; i++) w[i] = alpha * x[i] + y[i];
  else
    for (int64_t i = 0; i < n; i++) w[i] = alpha * x[i] + beta * y[i];
}

/* core/src/impl/DenseVector/Serial/Morpheus_Reduction_Impl.hpp:44-60 (Kahan) */
double orc_reduce(int64_t n, const double* in) {
  if (n <= 0) return 0;
  double sum = in[0], c = 0;
  for (int64_t i = 1; i < n; i++) {
    const double y = in[i] - c;
    const double t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

/* Conjugate gradients (textbook, Hestenes-Stiefel), for checking the fused
 * iteration: x0 given, r = b - A x0, p = r; each iteration q = A p,
 * alpha = rr / p.q, x += alpha p, r -= alpha q, rr' = r.r, p = r + (rr'/rr) p.
 * A is CSR.  Returns the final r.r. */
double orc_cg_csr(int64_t n, const int32_t* ro, const int32_t* ci, const double* v,
                  const double* b, double* x, int64_t iters) {
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  orc_csr_spmv(n, ro, ci, v, x, q, 1);
  for (int64_t i = 0; i < n; i++) {
    r[i] = b[i] - q[i];
    p[i] = r[i];
  }
  double rr = orc_dot(n, r, r);
  for (int64_t k = 0; k < iters; k++) {
    orc_csr_spmv(n, ro, ci, v, p, q, 1);
    const double alpha = rr / orc_dot(n, p, q);
    for (int64_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    const double rr2 = orc_dot(n, r, r);
    const double beta = rr2 / rr;
    for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    rr = rr2;
  }
  free(r);
  free(p);
  free(q);
  return rr;
}

/* ========================================================================
 * Matrix operations (Serial backend)
 * ===================================================================== */

/* core/src/impl/Csr/Serial/Morpheus_MatrixOperations_Impl.hpp:51-60: the
 * first entry of row i in column i */
void orc_csr_update_diagonal(int64_t nrows, const int32_t* ro, const int32_t* ci, double* v,
                             const double* d) {
  for (int64_t i = 0; i < nrows; i++)
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++)
      if (ci[jj] == i) {
        v[jj] = d[i];
        break;
      }
}

/* core/src/impl/Coo/Serial/Morpheus_MatrixOperations_Impl.hpp:49-55: every
 * diagonal entry */
void orc_coo_update_diagonal(int64_t nnz, const int32_t* ri, const int32_t* ci, double* v,
                             const double* d) {
  for (int64_t n = 0; n < nnz; n++)
    if (ri[n] == ci[n]) v[n] = d[ci[n]];
}

/* core/src/impl/Dia/Serial/Morpheus_MatrixOperations_Impl.hpp:52-66: the main
 * diagonal's slots, a stored zero staying zero */
void orc_dia_update_diagonal(int64_t nrows, int64_t ncols, int64_t ndiag, int64_t pitch,
                             const int32_t* off, double* v, const double* d) {
  for (int64_t row = 0; row < nrows; row++)
    for (int64_t n = 0; n < ndiag; n++) {
      const int64_t col = row + off[n];
      if (col >= 0 && col < ncols && col == row)
        v[row + pitch * n] = (v[row + pitch * n] == 0) ? 0 : d[col];
    }
}

/* core/src/impl/Ell/Serial/Morpheus_MatrixOperations_Impl.hpp:53-66: the
 * first slot of row i holding column i */
void orc_ell_update_diagonal(int64_t nrows, int64_t width, int64_t pitch, const int32_t* ci,
                             double* v, const double* d) {
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t n = 0; n < width; n++)
      if (ci[i + pitch * n] == i) {
        v[i + pitch * n] = d[i];
        break;
      }
}

/* ========================================================================
 * Convert (Serial backend)
 * ===================================================================== */

/* core/src/impl/Csr/Serial/Morpheus_Convert_Impl.hpp:76-96 */
void orc_csr_to_coo(int64_t nrows, int64_t nnz, const int32_t* ro,
                    const int32_t* ci, const double* v, int32_t* ri_out,
                    int32_t* ci_out, double* v_out) {
  for (int64_t i = 0; i < nrows; i++)
    for (int32_t jj = ro[i]; jj < ro[i + 1]; jj++) ri_out[jj] = (int32_t)i;
  for (int64_t n = 0; n < nnz; n++) ci_out[n] = ci[n];
  for (int64_t n = 0; n < nnz; n++) v_out[n] = v[n];
}

/* core/src/impl/Csr/Serial/Morpheus_Convert_Impl.hpp:109-148: histogram,
 * exclusive cumsum, stable scatter in input order, shift back. */
void orc_coo_to_csr(int64_t nrows, int64_t nnz, const int32_t* ri,
                    const int32_t* ci, const double* v, int32_t* ro_out,
                    int32_t* ci_out, double* v_out) {
  for (int64_t i = 0; i <= nrows; i++) ro_out[i] = 0;
  for (int64_t n = 0; n < nnz; n++) ro_out[ri[n]]++;
  int32_t cumsum = 0;
  for (int64_t i = 0; i < nrows; i++) {
    const int32_t temp = ro_out[i];
    ro_out[i] = cumsum;
    cumsum += temp;
  }
  ro_out[nrows] = (int32_t)nnz;
  for (int64_t n = 0; n < nnz; n++) {
    const int32_t row = ri[n];
    const int32_t dest = ro_out[row];
    ci_out[dest] = ci[n];
    v_out[dest] = v[n];
    ro_out[row]++;
  }
  int32_t last = 0;
  for (int64_t i = 0; i <= nrows; i++) {
    const int32_t temp = ro_out[i];
    ro_out[i] = last;
    last = temp;
  }
}

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Size of the unique set of col-row (the std::set of
 * core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:141-147). */
int64_t orc_coo_dia_ndiag(int64_t nnz, const int32_t* ri, const int32_t* ci) {
  if (nnz <= 0) return 0;
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
  for (int64_t n = 0; n < nnz; n++) d[n] = ci[n] - ri[n];
  qsort(d, (size_t)nnz, sizeof(int32_t), cmp_i32);
  int64_t u = 1;
  for (int64_t n = 1; n < nnz; n++)
    if (d[n] != d[n - 1]) u++;
  free(d);
  return u;
}

/* first index i with offsets[i] == key in an ascending array, else -1 */
static int64_t find_first(const int32_t* offsets, int64_t n, int32_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (offsets[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && offsets[lo] == key) ? lo : -1;
}

/* core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:130-168.  The
 * reference finds each entry's diagonal with a linear search over the
 * ascending offsets (:161-167); a binary search for the first match returns
 * the same index.  Repeated (row, col) entries: the later one wins, as in the
 * reference's sequential scatter. */
int orc_coo_to_dia(int64_t nrows, int64_t ncols, int64_t nnz,
                   const int32_t* ri, const int32_t* ci, const double* v,
                   int64_t ndiag, int64_t pitch, int32_t* offsets_out,
                   double* values_out) {
  (void)ncols;
  if (nnz == 0) return ORC_OK; /* :133-136 resize to 0 diagonals */
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
  for (int64_t n = 0; n < nnz; n++) d[n] = ci[n] - ri[n];
  qsort(d, (size_t)nnz, sizeof(int32_t), cmp_i32);
  int64_t u = 0;
  for (int64_t n = 0; n < nnz; n++)
    if (n == 0 || d[n] != d[n - 1]) d[u++] = d[n];
  if (u != ndiag) {
    free(d);
    return ORC_INVALID_INPUT;
  }
  if (orc_exceeds_tolerance(nrows, nnz, ndiag)) {
    free(d);
    return ORC_FORMAT_CONVERSION; /* :149-152 */
  }
  for (int64_t i = 0; i < ndiag; i++) offsets_out[i] = d[i];
  free(d);
  for (int64_t n = 0; n < nnz; n++) {
    const int64_t i = find_first(offsets_out, ndiag, ci[n] - ri[n]);
    values_out[ri[n] + pitch * i] = v[n];
  }
  return ORC_OK;
}

/* emits the non-zero stored entries of a DIA block, diagonal by diagonal
 * (core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:94-113 and
 * core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:86-105). */
static int emit_dia(int64_t nrows, int64_t ncols, int64_t ndiag,
                    int64_t pitch, const int32_t* offsets,
                    const double* values, int64_t cap, int32_t* ri_out,
                    int32_t* ci_out, double* v_out, int64_t* nnzid) {
  for (int64_t i = 0; i < ndiag; i++) {
    const int64_t k = offsets[i];
    const int64_t i_start = k < 0 ? -k : 0;
    const int64_t j_start = k > 0 ? k : 0;
    int64_t N = nrows - i_start;
    if (ncols - j_start < N) N = ncols - j_start;
    for (int64_t n = 0; n < N; n++) {
      const double temp = values[(i_start + n) + pitch * i];
      if (temp != 0.0) { /* drops +0.0 and -0.0 alike */
        if (*nnzid >= cap) return ORC_INVALID_INPUT;
        ri_out[*nnzid] = (int32_t)(i_start + n);
        ci_out[*nnzid] = (int32_t)(j_start + n);
        v_out[*nnzid] = temp;
        (*nnzid)++;
      }
    }
  }
  return ORC_OK;
}

/* core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:86-117.  The output has
 * nnz slots; slots left unfilled (when stored zeros were dropped) keep the
 * zero-initialised (0, 0, 0.0) and take part in the final sort. */
int orc_dia_to_coo(int64_t nrows, int64_t ncols, int64_t nnz, int64_t ndiag,
                   int64_t pitch, const int32_t* offsets, const double* values,
                   int32_t* ri_out, int32_t* ci_out, double* v_out) {
  if (nnz == 0) return ORC_OK;
  int64_t nnzid = 0;
  int rc = emit_dia(nrows, ncols, ndiag, pitch, offsets, values, nnz, ri_out,
                    ci_out, v_out, &nnzid);
  if (rc != ORC_OK) return rc;
  if (!orc_coo_is_sorted(nnz, ri_out, ci_out))
    rc = orc_coo_sort_by_row_and_column(nnz, ri_out, ci_out, v_out);
  return rc;
}

/* core/src/impl/Ell/Serial/Morpheus_Convert_Impl.hpp:126-147 */
int64_t orc_coo_ell_entries_per_row(int64_t nrows, int64_t nnz,
                                    const int32_t* ri) {
  int64_t* cnt = (int64_t*)calloc((size_t)(nrows + 1), sizeof(int64_t));
  for (int64_t n = 0; n < nnz; n++) cnt[ri[n]]++;
  int64_t K = 0;
  for (int64_t i = 0; i < nrows; i++)
    if (cnt[i] > K) K = cnt[i];
  free(cnt);
  return K;
}

/* core/src/impl/Ell/Serial/Morpheus_Convert_Impl.hpp:149-172.  Entries are
 * placed by their position inside the current run of equal row indices
 * (the row_id / n walk of :160-172), which equals their position in the row
 * for row-grouped input. */
int orc_coo_to_ell(int64_t nrows, int64_t nnz, const int32_t* ri,
                   const int32_t* ci, const double* v, int64_t K,
                   int64_t width, int64_t pitch, int32_t* ci_out,
                   double* v_out) {
  if (orc_exceeds_tolerance(nrows, nnz, K)) return ORC_FORMAT_CONVERSION;
  for (int64_t s = 0; s < width * pitch; s++) {
    ci_out[s] = -1;
    v_out[s] = 0.0;
  }
  int32_t row_id = 0;
  int64_t n = 0;
  for (int64_t i = 0; i < nnz; i++) {
    const int32_t row = ri[i];
    if (row_id != ri[i]) {
      n = 0;
      row_id = ri[i];
    }
    if (n >= width) return ORC_INVALID_INPUT;
    ci_out[row + pitch * n] = ci[i];
    v_out[row + pitch * n] = v[i];
    n++;
  }
  return ORC_OK;
}

/* core/src/impl/Ell/Serial/Morpheus_Convert_Impl.hpp:82-108: explicit zeros
 * are kept, -1 slots skipped, then sorted if needed. */
int orc_ell_to_coo(int64_t nrows, int64_t nnz, int64_t width, int64_t pitch,
                   const int32_t* ci, const double* v, int32_t* ri_out,
                   int32_t* ci_out, double* v_out) {
  if (nnz == 0) return ORC_OK;
  int64_t num_entries = 0;
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t n = 0; n < width; n++) {
      const int32_t j = ci[i + pitch * n];
      if (j != -1) {
        if (num_entries >= nnz) return ORC_INVALID_INPUT;
        ri_out[num_entries] = (int32_t)i;
        ci_out[num_entries] = j;
        v_out[num_entries] = v[i + pitch * n];
        num_entries++;
      }
    }
  if (!orc_coo_is_sorted(nnz, ri_out, ci_out))
    return orc_coo_sort_by_row_and_column(nnz, ri_out, ci_out, v_out);
  return ORC_OK;
}

/* core/src/impl/Hyb/Serial/Morpheus_Convert_Impl.hpp:129-162 */
int orc_coo_hyb_sizes(int64_t nrows, int64_t nnz, const int32_t* ri,
                      int64_t* K, int64_t* ell_nnz, int64_t* coo_nnz) {
  if (nrows <= 0) return ORC_INVALID_INPUT;
  const int64_t k = (nnz / nrows) > 1 ? nnz / nrows : 1;
  int64_t* cnt = (int64_t*)calloc((size_t)(nrows + 1), sizeof(int64_t));
  for (int64_t n = 0; n < nnz; n++) cnt[ri[n]]++;
  int64_t ell = 0;
  for (int64_t i = 0; i < nrows; i++) ell += cnt[i] > k ? k : cnt[i];
  free(cnt);
  *K = k;
  *ell_nnz = ell;
  *coo_nnz = nnz - ell;
  if (orc_exceeds_tolerance(nrows, ell, k)) return ORC_FORMAT_CONVERSION;
  return ORC_OK;
}

/* core/src/impl/Hyb/Serial/Morpheus_Convert_Impl.hpp:166-195: the first K
 * entries of each run of equal rows go to ELL slot n, the rest to the COO
 * part in input order. */
int orc_coo_to_hyb(int64_t nrows, int64_t nnz, const int32_t* ri,
                   const int32_t* ci, const double* v, int64_t K,
                   int64_t width, int64_t pitch, int32_t* ell_ci,
                   double* ell_v, int32_t* coo_ri, int32_t* coo_ci,
                   double* coo_v) {
  int64_t k2, ell_nnz, coo_cap;
  int rc = orc_coo_hyb_sizes(nrows, nnz, ri, &k2, &ell_nnz, &coo_cap);
  if (rc != ORC_OK) return rc;
  if (k2 != K) return ORC_INVALID_INPUT;
  for (int64_t s = 0; s < width * pitch; s++) {
    ell_ci[s] = -1;
    ell_v[s] = 0.0;
  }
  int32_t row_id = 0;
  int64_t coo = 0, n = 0;
  for (int64_t i = 0; i < nnz; i++) {
    const int32_t row = ri[i];
    if (row_id != ri[i]) {
      n = 0;
      row_id = ri[i];
    }
    if (n < K) {
      ell_ci[row + pitch * n] = ci[i];
      ell_v[row + pitch * n] = v[i];
      n++;
    } else {
      if (coo >= coo_cap) return ORC_INVALID_INPUT;
      coo_ri[coo] = ri[i];
      coo_ci[coo] = ci[i];
      coo_v[coo] = v[i];
      coo++;
    }
  }
  return ORC_OK;
}

/* core/src/impl/Hyb/Serial/Morpheus_Convert_Impl.hpp:80-112: row by row,
 * the ELL slots then the COO entries of that row (a forward walk over the
 * row-sorted COO part), then sorted if needed. */
int orc_hyb_to_coo(int64_t nrows, int64_t nnz, int64_t width, int64_t pitch,
                   const int32_t* ell_ci, const double* ell_v,
                   int64_t coo_nnz, const int32_t* coo_ri,
                   const int32_t* coo_ci, const double* coo_v,
                   int32_t* ri_out, int32_t* ci_out, double* v_out) {
  if (nnz == 0) return ORC_OK;
  int64_t num = 0, prog = 0;
  for (int64_t i = 0; i < nrows; i++) {
    for (int64_t n = 0; n < width; n++) {
      const int32_t j = ell_ci[i + pitch * n];
      if (j != -1) {
        if (num >= nnz) return ORC_INVALID_INPUT;
        ri_out[num] = (int32_t)i;
        ci_out[num] = j;
        v_out[num] = ell_v[i + pitch * n];
        num++;
      }
    }
    while (prog < coo_nnz && (int64_t)coo_ri[prog] == i) {
      if (num >= nnz) return ORC_INVALID_INPUT;
      ri_out[num] = (int32_t)i;
      ci_out[num] = coo_ci[prog];
      v_out[num] = coo_v[prog];
      num++;
      prog++;
    }
  }
  if (!orc_coo_is_sorted(nnz, ri_out, ci_out))
    return orc_coo_sort_by_row_and_column(nnz, ri_out, ci_out, v_out);
  return ORC_OK;
}

/* core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:142-163.  The threshold
 * is (int)(nrows * 0.2) (HdcMatrix::true_diagonal_ratio,
 * core/src/Morpheus_HdcMatrix.hpp:148); with fewer than 5 rows it is 0 and
 * every diagonal, populated or not, is selected. */
static int32_t* hdc_counts(int64_t nrows, int64_t ncols, int64_t nnz,
                           const int32_t* ri, const int32_t* ci) {
  int32_t* c = (int32_t*)calloc((size_t)(nrows + ncols - 1 > 0
                                             ? nrows + ncols - 1
                                             : 1),
                                sizeof(int32_t));
  orc_count_nnz_per_diagonal(nrows, ncols, nnz, ri, ci, c);
  return c;
}

int orc_coo_hdc_sizes(int64_t nrows, int64_t ncols, int64_t nnz,
                      const int32_t* ri, const int32_t* ci, int64_t* ndiag,
                      int64_t* dia_nnz, int64_t* csr_nnz) {
  if (nrows <= 0 || ncols <= 0) return ORC_INVALID_INPUT;
  const int32_t min_diag_elem = (int32_t)((double)nrows * 0.2);
  int32_t* c = hdc_counts(nrows, ncols, nnz, ri, ci);
  int64_t nd = 0, dn = 0, cn = 0;
  for (int64_t i = 0; i < nrows + ncols - 1; i++) {
    if (c[i] >= min_diag_elem) {
      nd++;
      dn += c[i];
    } else if (c[i] != 0) {
      cn += c[i];
    }
  }
  free(c);
  *ndiag = nd;
  *dia_nnz = dn;
  *csr_nnz = cn;
  if (orc_exceeds_tolerance(nrows, dn, nd)) return ORC_FORMAT_CONVERSION;
  return ORC_OK;
}

/* core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:165-225.  The offsets
 * are gathered left of the bin midpoint, at it, and right of it, then
 * sorted (:169-195); the midpoint bin is recorded as offset 0, which is the
 * true offset of that bin only for square matrices.  Entries whose diagonal
 * is not selected go to the CSR part in input order.  If the CSR part would
 * overflow its csr_nnz slots (the reference's behaviour is undefined there)
 * INVALID_INPUT is returned. */
int orc_coo_to_hdc(int64_t nrows, int64_t ncols, int64_t nnz,
                   const int32_t* ri, const int32_t* ci, const double* v,
                   int64_t ndiag, int64_t pitch, int32_t* dia_offsets,
                   double* dia_values, int32_t* csr_ro, int32_t* csr_ci,
                   double* csr_v) {
  int64_t nd, dn, cn;
  int rc = orc_coo_hdc_sizes(nrows, ncols, nnz, ri, ci, &nd, &dn, &cn);
  if (rc != ORC_OK) return rc;
  if (nd != ndiag) return ORC_INVALID_INPUT;
  const int32_t min_diag_elem = (int32_t)((double)nrows * 0.2);
  int32_t* c = hdc_counts(nrows, ncols, nnz, ri, ci);
  const int64_t nbins = nrows + ncols - 1;
  for (int64_t i = 0; i <= nrows; i++) csr_ro[i] = 0;

  const int mid = (int)nbins / 2;
  int64_t ctr = 0;
  for (int i = mid - 1; i >= 0; i--)
    if (c[i] >= min_diag_elem) dia_offsets[ctr++] = (int32_t)(i - nrows + 1);
  if (c[mid] >= min_diag_elem) dia_offsets[ctr++] = 0;
  for (int64_t i = mid + 1; i < nbins; i++)
    if (c[i] >= min_diag_elem) dia_offsets[ctr++] = (int32_t)(i - nrows + 1);
  free(c);
  qsort(dia_offsets, (size_t)ndiag, sizeof(int32_t), cmp_i32);

  int64_t csr_entries = 0;
  for (int64_t n = 0; n < nnz; n++) {
    const int64_t i = find_first(dia_offsets, ndiag, ci[n] - ri[n]);
    if (i >= 0) {
      dia_values[ri[n] + pitch * i] = v[n];
    } else {
      if (csr_entries >= cn) return ORC_INVALID_INPUT;
      csr_ro[ri[n]]++;
      csr_ci[csr_entries] = ci[n];
      csr_v[csr_entries] = v[n];
      csr_entries++;
    }
  }
  int32_t cumsum = 0;
  for (int64_t i = 0; i < nrows; i++) {
    const int32_t temp = csr_ro[i];
    csr_ro[i] = cumsum;
    cumsum += temp;
  }
  csr_ro[nrows] = (int32_t)cn;
  return ORC_OK;
}

/* core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:78-125 */
int orc_hdc_to_coo(int64_t nrows, int64_t ncols, int64_t nnz, int64_t ndiag,
                   int64_t pitch, const int32_t* dia_offsets,
                   const double* dia_values, int64_t csr_nnz,
                   const int32_t* csr_ro, const int32_t* csr_ci,
                   const double* csr_v, int32_t* ri_out, int32_t* ci_out,
                   double* v_out) {
  if (nnz == 0) return ORC_OK;
  int64_t nnzid = 0;
  int rc = emit_dia(nrows, ncols, ndiag, pitch, dia_offsets, dia_values, nnz,
                    ri_out, ci_out, v_out, &nnzid);
  if (rc != ORC_OK) return rc;
  if (nnzid + csr_nnz > nnz) return ORC_INVALID_INPUT;
  for (int64_t i = 0; i < nrows; i++)
    for (int32_t jj = csr_ro[i]; jj < csr_ro[i + 1]; jj++)
      ri_out[jj + nnzid] = (int32_t)i;
  for (int64_t n = 0; n < csr_nnz; n++) ci_out[n + nnzid] = csr_ci[n];
  for (int64_t n = 0; n < csr_nnz; n++) v_out[n + nnzid] = csr_v[n];
  if (!orc_coo_is_sorted(nnz, ri_out, ci_out))
    return orc_coo_sort_by_row_and_column(nnz, ri_out, ci_out, v_out);
  return ORC_OK;
}
