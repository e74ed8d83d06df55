/*
 * morpheus_amd.h -- C-ABI of the MI355X-native Morpheus SpMV hot path.
 *
 * Every entry point takes plain device pointers, sizes and enums; no C++ or
 * torch types cross this boundary.  The library never frees caller memory.
 * All device work is enqueued on the caller's stream (NULL = the legacy
 * default stream, which is the ordering the reference HIP backend relies on).
 *
 * Storage conventions (the reference's HIP/device layout, LayoutLeft):
 *   indices int32; values FP64, FP32 or BF16 (mph_value_type);
 *   x / y vectors: double for FP64 matrices, float for FP32 and BF16 ones;
 *   COO: row_indices[nnz], column_indices[nnz], values[nnz];
 *   CSR: row_offsets[nrows+1], column_indices[nnz], values[nnz];
 *   DIA: diagonal_offsets[ndiag] and values column-major with leading
 *        dimension `pitch` (pad(nrows,32) by default): values(r,d) =
 *        values[r + pitch*d] = A(r, r + offset[d]);
 *   ELL: column_indices / values column-major, leading dimension `pitch`
 *        (= nrows), `width` stored columns (pad(entries_per_row,32));
 *        slot (r,n) = [r + pitch*n]; padding slots hold index -1.
 *   HYB: an ELL part plus a COO part; HDC: a DIA part plus a CSR part.
 *
 * Status codes mirror the reference exception taxonomy
 * (core/src/Morpheus_Exceptions.hpp:32-74); the C++ drop-in headers rethrow
 * them as Morpheus::NotImplementedException, InvalidInputException,
 * FormatConversionException and RuntimeException.
 */
#ifndef MORPHEUS_AMD_H
#define MORPHEUS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t-compatible opaque stream handle */
typedef struct ihipStream_t* mph_stream_t;

enum mph_status {
  MPH_OK = 0,
  MPH_ERR_NOT_IMPLEMENTED = 1,   /* Morpheus::NotImplementedException   */
  MPH_ERR_INVALID_INPUT = 2,     /* Morpheus::InvalidInputException     */
  MPH_ERR_FORMAT_CONVERSION = 3, /* Morpheus::FormatConversionException */
  MPH_ERR_HIP = 4,               /* HIP runtime failure                 */
  MPH_ERR_OUT_OF_MEMORY = 5,
  MPH_ERR_RUNTIME = 6            /* Morpheus::RuntimeException          */
};

/* MPH_I32: 32-bit integer values (Morpheus containers of int).  Accepted by
 * the conversions, the COO sort and mph_values_cast, which move values as
 * raw words (the zero test of DIA -> COO is on every bit); not by the
 * products, which compute in FP64 / FP32 (the drop-in multiplies int
 * containers exactly through FP64, see mph_values_cast). */
enum mph_value_type { MPH_F64 = 0, MPH_F32 = 1, MPH_BF16 = 2, MPH_I32 = 3 };

/* Same order as Morpheus::formats_e (core/src/Morpheus_FormatsRegistry.hpp:56-64) */
enum mph_format {
  MPH_COO = 0,
  MPH_CSR = 1,
  MPH_DIA = 2,
  MPH_ELL = 3,
  MPH_HYB = 4,
  MPH_HDC = 5,
  MPH_NFORMATS = 6
};

/* ------------------------------------------------------------ library */

/* Library version string. */
const char* mph_version(void);
/* Text for a status code. */
const char* mph_status_string(int status);
/* Message of the last failure on the calling thread ("" if none). */
const char* mph_last_error(void);
/* Device-side part of Morpheus::initialize (core/src/impl/Morpheus_Core.cpp:172-194):
 * selects `device` (-1 = keep current) and caches its properties once, where
 * the reference re-queries them on every launch
 * (core/src/impl/Morpheus_HIPUtils.hpp:77-90). */
int mph_initialize(int device);
/* Morpheus::finalize (core/src/impl/Morpheus_Core.cpp:215-222). */
int mph_finalize(void);

/* ------------------------------------------------------------ matrices */

/* A matrix handle describes caller-owned device arrays in one of the six
 * formats and owns the analysis the kernels need (row blocks, carry
 * scratch, tuned launch shape).  It is the C form of the reference's
 * unmanaged containers built over user pointers
 * (core/src/Morpheus_CsrMatrix.hpp:186-201). */
typedef struct mph_matrix_s* mph_matrix_t;

/* core/src/Morpheus_CooMatrix.hpp (row_indices, column_indices, values) */
int mph_matrix_create_coo(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, const int32_t* row_indices,
                          const int32_t* column_indices, const void* values,
                          int value_type);
/* core/src/Morpheus_CsrMatrix.hpp:534-535 (row_offsets, column_indices, values) */
int mph_matrix_create_csr(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, const int32_t* row_offsets,
                          const int32_t* column_indices, const void* values,
                          int value_type);
/* core/src/Morpheus_DiaMatrix.hpp:192-201 (diagonal_offsets, values pad(N,32) x ndiag) */
int mph_matrix_create_dia(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, int64_t ndiag, int64_t pitch,
                          const int32_t* diagonal_offsets, const void* values,
                          int value_type);
/* core/src/Morpheus_EllMatrix.hpp:189-202 (column_indices / values N x pad(K,32)) */
int mph_matrix_create_ell(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t nnz, int64_t entries_per_row, int64_t width,
                          int64_t pitch, const int32_t* column_indices,
                          const void* values, int value_type);
/* core/src/Morpheus_HybMatrix.hpp:101-477 (ell() + coo()) */
int mph_matrix_create_hyb(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t ell_nnz, int64_t entries_per_row,
                          int64_t width, int64_t pitch,
                          const int32_t* ell_column_indices,
                          const void* ell_values, int64_t coo_nnz,
                          const int32_t* coo_row_indices,
                          const int32_t* coo_column_indices,
                          const void* coo_values, int value_type);
/* core/src/Morpheus_HdcMatrix.hpp:107-468 (dia() + csr()) */
int mph_matrix_create_hdc(mph_matrix_t* out, int64_t nrows, int64_t ncols,
                          int64_t dia_nnz, int64_t ndiag, int64_t pitch,
                          const int32_t* dia_offsets, const void* dia_values,
                          int64_t csr_nnz, const int32_t* csr_row_offsets,
                          const int32_t* csr_column_indices,
                          const void* csr_values, int value_type);
/* Releases the analysis owned by the handle (never the caller's arrays). */
int mph_matrix_destroy(mph_matrix_t A);
/* Builds the analysis now (multiply builds it lazily on first use).  Must be
 * called again (or the handle recreated) after the caller changes the
 * sparsity structure in place -- and, unless the handle's policy is
 * MPH_POLICY_NO_VALUE_COPIES, after the caller writes values in place: by
 * default the analysis may keep copies of values (the x-tiled long rows'
 * tile-major copy, a HYB's merged CSR copy), which products then read. */
int mph_matrix_analyse(mph_matrix_t A, mph_stream_t stream);

/* Analysis policy (no reference counterpart: reference containers hold no
 * derived state, core/src/Morpheus_CsrMatrix.hpp:148-163).
 * MPH_POLICY_NO_VALUE_COPIES: the analysis keeps no copy of any value: the
 * x-tiled pass copies the long rows' columns only and reads their values in
 * place, and a large HYB keeps no merged copy (it runs its ELL part as a
 * pass of its own, each row stopping at its last filled slot when the slots
 * are mostly padding, then adds its COO tail); every product reads the
 * caller's value arrays, so values written in place are always seen.  The
 * Morpheus drop-in containers set it.  Changing the policy releases the
 * analysis. */
#define MPH_POLICY_NO_VALUE_COPIES 1
int mph_matrix_set_policy(mph_matrix_t A, int flags);
int mph_matrix_format(mph_matrix_t A);
/* Bytes of the arrays one product y = A x (k = 1; init as given) streams
 * under A's analysed plan, each counted once: what the plan reads where it
 * differs from the format's minimal representation (a HYB over its merged
 * CSR copy, row-sorted COO through derived row offsets, the x-tiled copy of
 * long rows and its partials, the split passes of HYB / HDC).  No reference
 * counterpart: for roofline reporting.  Analyses A if needed. */
int mph_matrix_plan_bytes(mph_matrix_t A, int init, int64_t* bytes);

/* Launch-shape knobs the autotuner searches (0 = library default).
 *   rows_chunk   : nnz budget of a CSR row block (short rows) / long-row chunk
 *   lanes        : fixed lanes per row for CSR row blocks (1..64; 0 = per
 *                  block, chosen by the analysis from the longest rows of
 *                  each wave, for SpMV and SpMM alike)
 *   coo_items    : entries per lane of the COO segmented reduction (4, 8, 16);
 *                  0 = a row-sorted COO matrix runs the CSR row-block plan
 *                  over row offsets derived from its row indices (rows_chunk
 *                  and lanes apply), an unsorted one the atomic pass
 *   rows_per_thread : DIA / ELL rows per thread (1 or 2, strided by 256 rows);
 *                  0 = one row in FP64, and in FP32 / BF16 SpMV two adjacent
 *                  rows whose values (and ELL column indices) are loaded
 *                  as pairs, when the pitch is even; with 0, DIA SpMV (and
 *                  HDC's DIA pass) stages the x entries of each cluster of
 *                  nearby diagonals (offsets within 512, two or more
 *                  diagonals; at most 64 diagonals and 8 such clusters) in
 *                  LDS once per workgroup; the row-pair kernels do so when
 *                  the clusters average 8 diagonals or more.  The
 *                  environment variable MPH_DIA_STAGE, read at analysis,
 *                  sets 0 = no staging, 1 = the one-row kernels only.
 *                  The results do not depend on it.  DIA SpMM with 4 or
 *                  more lanes per row (FP64 from 8 columns per pass, FP32 /
 *                  BF16 from 16) takes 4 consecutive rows per lane group and
 *                  loads, per run of diagonals with consecutive offsets, the
 *                  X rows the run covers once and the run's values through
 *                  LDS (MPH_DIA_SPMM_RUNS=0 / 1 forces that off / on; the
 *                  results are the same bits).
 *   xtile        : x-tiled long rows of CSR / COO / HYB / HDC SpMV (1 = off,
 *                  2 = on; 0 = on when x exceeds 2 MiB).  Rows longer than
 *                  512 entries with sorted columns are summed per 32 KiB tile
 *                  of x staged in LDS, so their x reads are LDS reads instead
 *                  of scattered cache-line fetches (the tile is halved, down
 *                  to an eighth, when most of those rows are too dense for
 *                  it; SpMM runs the pass over groups of 4 / 2 columns of
 *                  the x tile for k <= 4, or k <= 16 when the rows average
 *                  32+ items per tile, or, without the value copy or with
 *                  the environment variable MPH_XTILE_K=0, once per column
 *                  for k <= 2 / 4; where both take the x-tiled rows this
 *                  way, each column comes out the same bit for bit).  On
 *                  (x-tiling) or from
 *                  65,536 rows, HYB / HDC run their ELL / DIA part as its own
 *                  pass and the COO / CSR part over the rows that have one;
 *                  a HYB whose ELL slots are less than a third filled, or
 *                  in which more than an eighth of the rows have COO
 *                  entries, runs instead as CSR over a merged copy made by
 *                  the analysis (per row its ELL entries, then its COO
 *                  entries); an FP64 HDC whose CSR rows hold at most 8
 *                  entries each (none x-tiled) runs one DIA pass that adds
 *                  each row's CSR entries after its diagonals
 *                  (MPH_HDC_TAIL=0 keeps the two passes); an HDC without
 *                  diagonals runs as its CSR part. */
typedef struct mph_tuning {
  int32_t rows_chunk;
  int32_t lanes;
  int32_t coo_items;
  int32_t rows_per_thread;
  int32_t xtile;
} mph_tuning;
int mph_matrix_set_tuning(mph_matrix_t A, const mph_tuning* t);
int mph_matrix_get_tuning(mph_matrix_t A, mph_tuning* t);

/* ------------------------------------------------------------ multiply */

/* Morpheus::multiply<ExecSpace>(A, x, y, init)
 * (core/src/Morpheus_Multiply.hpp:57-61,75-78; HIP backend
 * core/src/impl/{Coo,Csr,Dia,Ell,Hyb,Hdc}/HIP/Morpheus_Multiply_Impl.hpp):
 * y = A*x when init != 0, y += A*x otherwise.  Asynchronous on `stream`;
 * allocates nothing once the handle is analysed (hipGraph-capturable). */
int mph_multiply(mph_matrix_t A, const void* x, void* y, int init,
                 mph_stream_t stream);

/* k right-hand sides (SpMM): X is ncols x k and Y is nrows x k, row-major
 * (element (i, j) at [i*k + j]).  No reference counterpart: the reference
 * multiply takes one DenseVector (core/src/Morpheus_Multiply.hpp:57-78).
 * Asynchronous on `stream`.  On a handle with x-tiled long rows, the first
 * product with k >= 2 allocates the partials of the k-wide x-tiled pass
 * (kept by the handle until its analysis is released); later products
 * allocate nothing, so a product to be captured into a hipGraph follows one
 * eager call with the same k. */
int mph_multiply_multi(mph_matrix_t A, int32_t k, const void* X, void* Y,
                       int init, mph_stream_t stream);

/* One-shot per-format multiply over raw arrays, the direct counterparts of
 * Impl::multiply for each container
 * (core/src/impl/Csr/HIP/Morpheus_Multiply_Impl.hpp:50-63 etc.).  They build
 * a transient analysis (allocating) and release it; use a handle for
 * repeated products. */
int mph_csr_multiply(int64_t nrows, int64_t ncols, int64_t nnz,
                     const int32_t* row_offsets, const int32_t* column_indices,
                     const void* values, int value_type, const void* x,
                     void* y, int init, mph_stream_t stream);
int mph_coo_multiply(int64_t nrows, int64_t ncols, int64_t nnz,
                     const int32_t* row_indices, const int32_t* column_indices,
                     const void* values, int value_type, const void* x,
                     void* y, int init, mph_stream_t stream);

/* mph_coo_multiply without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Runs the
 * deterministic wavefront-segmented COO kernel (8 entries per lane) with its
 * per-chunk carries in a fixed device array of 8 MiB per GPU (held until the
 * process ends), so it calls no hipMalloc / hipFree and never synchronises.
 * Limits: nnz <= 268,435,456 (larger sizes return MPH_ERR_INVALID_INPUT with
 * the limit in mph_last_error()); rows must be sorted and every index in
 * range -- the preconditions of the reference HIP COO multiply
 * (core/src/impl/Coo/HIP/Morpheus_Multiply_Impl.hpp:51-179), not checked
 * here.  Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other threads) race and must be serialised by the
 * caller.  Same arguments and result as mph_coo_multiply (init = 1: y = A x,
 * rows without entries 0; init = 0: y += A x). */
int mph_coo_multiply_noalloc(int64_t nrows, int64_t ncols, int64_t nnz,
                             const int32_t* row_indices, const int32_t* column_indices,
                             const void* values, int value_type, const void* x, void* y,
                             int init, mph_stream_t stream);
int mph_dia_multiply(int64_t nrows, int64_t ncols, int64_t nnz, int64_t ndiag,
                     int64_t pitch, const int32_t* diagonal_offsets,
                     const void* values, int value_type, const void* x,
                     void* y, int init, mph_stream_t stream);
int mph_ell_multiply(int64_t nrows, int64_t ncols, int64_t nnz,
                     int64_t entries_per_row, int64_t width, int64_t pitch,
                     const int32_t* column_indices, const void* values,
                     int value_type, const void* x, void* y, int init,
                     mph_stream_t stream);

/* ------------------------------------------------------------ convert */
/* On-GPU format conversion, bit-exact with the reference Serial converters
 * (core/src/impl/<Fmt>/Serial/Morpheus_Convert_Impl.hpp).  The reference
 * HIP converters throw NotImplementedException
 * (core/src/impl/Csr/HIP/Morpheus_Convert_Impl.hpp:39-76); these succeed.
 * Size queries synchronise `stream`; fills are asynchronous but use
 * stream-ordered temporary allocations.  Output arrays are caller-owned and
 * sized from the size queries. */

/* core/src/impl/Csr/Serial/Morpheus_Convert_Impl.hpp:66-97 */
int mph_convert_csr_to_coo(int64_t nrows, int64_t nnz,
                           const int32_t* row_offsets, const int32_t* ci,
                           const void* v, int value_type, int32_t* ri_out,
                           int32_t* ci_out, void* v_out, mph_stream_t stream);
/* core/src/impl/Csr/Serial/Morpheus_Convert_Impl.hpp:99-151 */
int mph_convert_coo_to_csr(int64_t nrows, int64_t nnz, const int32_t* ri,
                           const int32_t* ci, const void* v, int value_type,
                           int32_t* ro_out, int32_t* ci_out, void* v_out,
                           mph_stream_t stream);
/* core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:130-154 (sizes and
 * fill-tolerance check, core/src/impl/Morpheus_Utils.hpp:91-105) */
int mph_convert_coo_to_dia_size(int64_t nrows, int64_t ncols, int64_t nnz,
                                const int32_t* ri, const int32_t* ci,
                                int64_t* ndiag, mph_stream_t stream);
/* core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:120-169 */
int mph_convert_coo_to_dia(int64_t nrows, int64_t ncols, int64_t nnz,
                           const int32_t* ri, const int32_t* ci,
                           const void* v, int value_type, int64_t ndiag,
                           int64_t pitch, int32_t* offsets_out,
                           void* values_out, mph_stream_t stream);
/* core/src/impl/Dia/Serial/Morpheus_Convert_Impl.hpp:72-118 */
int mph_convert_dia_to_coo(int64_t nrows, int64_t ncols, int64_t nnz,
                           int64_t ndiag, int64_t pitch,
                           const int32_t* offsets, const void* values,
                           int value_type, int32_t* ri_out, int32_t* ci_out,
                           void* v_out, mph_stream_t stream);
/* core/src/impl/Ell/Serial/Morpheus_Convert_Impl.hpp:126-152 */
int mph_convert_coo_to_ell_size(int64_t nrows, int64_t nnz, const int32_t* ri,
                                int64_t* entries_per_row, mph_stream_t stream);
/* core/src/impl/Ell/Serial/Morpheus_Convert_Impl.hpp:111-173 */
int mph_convert_coo_to_ell(int64_t nrows, int64_t nnz, const int32_t* ri,
                           const int32_t* ci, const void* v, int value_type,
                           int64_t entries_per_row, int64_t width,
                           int64_t pitch, int32_t* ci_out, void* v_out,
                           mph_stream_t stream);
/* core/src/impl/Ell/Serial/Morpheus_Convert_Impl.hpp:68-109 */
int mph_convert_ell_to_coo(int64_t nrows, int64_t nnz, int64_t width,
                           int64_t pitch, const int32_t* ci, const void* v,
                           int value_type, int32_t* ri_out, int32_t* ci_out,
                           void* v_out, mph_stream_t stream);
/* core/src/impl/Hyb/Serial/Morpheus_Convert_Impl.hpp:129-162 */
int mph_convert_coo_to_hyb_size(int64_t nrows, int64_t nnz, const int32_t* ri,
                                int64_t* entries_per_row, int64_t* ell_nnz,
                                int64_t* coo_nnz, mph_stream_t stream);
/* core/src/impl/Hyb/Serial/Morpheus_Convert_Impl.hpp:115-196 */
int mph_convert_coo_to_hyb(int64_t nrows, int64_t nnz, const int32_t* ri,
                           const int32_t* ci, const void* v, int value_type,
                           int64_t entries_per_row, int64_t width,
                           int64_t pitch, int32_t* ell_ci, void* ell_v,
                           int32_t* coo_ri, int32_t* coo_ci, void* coo_v,
                           mph_stream_t stream);
/* core/src/impl/Hyb/Serial/Morpheus_Convert_Impl.hpp:62-113 */
int mph_convert_hyb_to_coo(int64_t nrows, int64_t nnz, int64_t width,
                           int64_t pitch, const int32_t* ell_ci,
                           const void* ell_v, int64_t coo_nnz,
                           const int32_t* coo_ri, const int32_t* coo_ci,
                           const void* coo_v, int value_type, int32_t* ri_out,
                           int32_t* ci_out, void* v_out, mph_stream_t stream);
/* core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:138-163 */
int mph_convert_coo_to_hdc_size(int64_t nrows, int64_t ncols, int64_t nnz,
                                const int32_t* ri, const int32_t* ci,
                                int64_t* ndiag, int64_t* dia_nnz,
                                int64_t* csr_nnz, mph_stream_t stream);
/* core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:128-226 */
int mph_convert_coo_to_hdc(int64_t nrows, int64_t ncols, int64_t nnz,
                           const int32_t* ri, const int32_t* ci,
                           const void* v, int value_type, int64_t ndiag,
                           int64_t pitch, int32_t* dia_offsets,
                           void* dia_values, int32_t* csr_ro, int32_t* csr_ci,
                           void* csr_v, mph_stream_t stream);
/* core/src/impl/Hdc/Serial/Morpheus_Convert_Impl.hpp:64-126 */
int mph_convert_hdc_to_coo(int64_t nrows, int64_t ncols, int64_t nnz,
                           int64_t ndiag, int64_t pitch,
                           const int32_t* dia_offsets, const void* dia_values,
                           int64_t csr_nnz, const int32_t* csr_ro,
                           const int32_t* csr_ci, const void* csr_v,
                           int value_type, int32_t* ri_out, int32_t* ci_out,
                           void* v_out, mph_stream_t stream);

/* core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:111-166: stable sort of
 * the COO entries by (row, column), in place. */
int mph_coo_sort_by_row_and_column(int64_t nnz, int32_t* ri, int32_t* ci,
                                   void* v, int value_type,
                                   mph_stream_t stream);
/* core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:168-189 (synchronises) */
int mph_coo_is_sorted(int64_t nnz, const int32_t* ri, const int32_t* ci,
                      int* sorted, mph_stream_t stream);
/* core/src/impl/Coo/Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp:38-74:
 * out has nrows+ncols-1 int32 counters; init != 0 zeroes them first. */
int mph_coo_count_nnz_per_diagonal(int64_t nrows, int64_t ncols, int64_t nnz,
                                   const int32_t* ri, const int32_t* ci,
                                   int32_t* out, int init,
                                   mph_stream_t stream);
/* core/src/impl/Morpheus_Utils.hpp:91-105 (host-side decision, in float) */
int mph_exceeds_tolerance(int64_t num_rows, int64_t num_entries,
                          int64_t num_columns);

/* ------------------------------------------------------------ matrix operations */
/* Morpheus::update_diagonal<HIP>(A, diagonal) (core/src/Morpheus_MatrixOperations.hpp
 * :54-58), with the Serial semantics of core/src/impl/<Fmt>/Serial/
 * Morpheus_MatrixOperations_Impl.hpp: CSR / ELL the first entry of row i in
 * column i, COO every diagonal entry, DIA the main diagonal's slots (a stored
 * zero stays zero), HYB = ELL + COO, HDC = DIA + CSR.  diagonal: min(nrows,
 * ncols) device values of the handle's vector type.  Writes the handle's
 * value arrays in place and releases its analysis. */
int mph_matrix_update_diagonal(mph_matrix_t A, const void* diagonal, mph_stream_t stream);
/* The diagonal the product sees, the sum of every stored (i, i) entry, into
 * min(nrows, ncols) device values.  The reference's get_diagonal
 * (core/src/Morpheus_MatrixOperations.hpp:69-73) throws
 * NotImplementedException. */
int mph_matrix_get_diagonal(mph_matrix_t A, void* diagonal, mph_stream_t stream);

/* ------------------------------------------------------------ vectors */
/* Dense-vector algorithms on device vectors of `value_type` (MPH_F64:
 * double; MPH_F32 / MPH_BF16: float), n entries.  Sums are formed in FP64
 * per block and added in a fixed order, so each call gives the same bits. */
/* Morpheus::dot<HIP>(n, x, y) (core/src/Morpheus_Dot.hpp:45-57;
 * core/src/impl/DenseVector/HIP/Morpheus_Dot_Impl.hpp).  Synchronises. */
int mph_dot(int64_t n, const void* x, const void* y, int value_type,
            double* result, mph_stream_t stream);
/* Morpheus::waxpby<HIP>(n, alpha, x, beta, y, w): w = alpha x + beta y, each
 * product and sum rounded on its own, as the Serial loops
 * (core/src/Morpheus_WAXPBY.hpp:48-65). */
int mph_waxpby(int64_t n, double alpha, const void* x, double beta, const void* y,
               void* w, int value_type, mph_stream_t stream);
/* Morpheus::reduce<HIP>(in, n): the sum of the entries, TwoSum-compensated
 * (core/src/Morpheus_Reduction.hpp:44-52).  Synchronises. */
int mph_reduce(int64_t n, const void* x, int value_type, double* result,
               mph_stream_t stream);

/* Conjugate gradients on a square matrix handle (no reference counterpart:
 * the consumer of the SpMV path).  Vectors are device vectors of the
 * handle's vector type.  mph_cg_start(b, x0) forms r = b - A x0 and p = r;
 * mph_cg_iterate enqueues `iters` iterations, each the handle's SpMV and
 * three fused vector launches, with every scalar kept on the device (no host
 * synchronisation, no memset: a hipGraph may capture it, and replays of a
 * captured iteration continue the iteration); mph_cg_residual returns r.r
 * (synchronises).  Calls on one handle must be ordered. */
typedef struct mph_cg_s* mph_cg_t;
int mph_cg_create(mph_cg_t* out, mph_matrix_t A);
int mph_cg_destroy(mph_cg_t cg);
int mph_cg_start(mph_cg_t cg, const void* b, void* x, mph_stream_t stream);
int mph_cg_iterate(mph_cg_t cg, void* x, int32_t iters, mph_stream_t stream);
int mph_cg_residual(mph_cg_t cg, double* rr, mph_stream_t stream);

/* ------------------------------------------------------------ analytics */
/* Morpheus::count_nnz_per_row<Exec>(A, nnz_per_row, init)
 * (core/src/Morpheus_MatrixAnalytics.hpp:112-116) for a handle of any
 * format, counting as core/src/impl/{Coo,Csr,Dia,Ell,Hyb,Hdc}/Serial/
 * Morpheus_NonZeros_Per_Row_Impl.hpp do: CSR / COO stored entries, DIA
 * in-range slots, ELL slots whose index is not -1, HYB = ELL + COO,
 * HDC = DIA + CSR.  nnz_per_row: nrows int32 device counters; init != 0
 * zeroes them first, otherwise the counts are added. */
int mph_matrix_count_nnz_per_row(mph_matrix_t A, int32_t* nnz_per_row,
                                 int32_t init, mph_stream_t stream);
/* Morpheus::count_nnz_per_diagonal<Exec>(A, nnz_per_diagonal, init)
 * (core/src/Morpheus_MatrixAnalytics.hpp:206-211) for a handle of any
 * format, counting as core/src/impl/{Coo,Csr,Dia,Ell}/Kokkos/
 * Morpheus_NonZeros_Per_Diagonal_Impl.hpp do (DIA: non-zero values over the
 * pitch rows; ELL: among a row's first c slots, c being its count of slots
 * with a valid index and a non-zero value).  nnz_per_diagonal: nrows +
 * ncols - 1 int32 device counters, entry col - row + nrows - 1.  The
 * indices are checked first (synchronises). */
int mph_matrix_count_nnz_per_diagonal(mph_matrix_t A, int32_t* nnz_per_diagonal,
                                      int32_t init, mph_stream_t stream);
/* Morpheus::max / min / std / count_nnz of an int32 device vector, in one
 * pass (core/src/impl/DenseVector/Kokkos/Morpheus_VectorAnalytics_Impl.hpp
 * :50-127): std = sqrt(sq_dev / n), sq_dev summed in a fixed order;
 * count = entries > threshold and != 0.  Synchronises. */
typedef struct {
  int64_t max, min, count;
  double sq_dev;
} mph_i32_stats_t;
int mph_i32_stats(const int32_t* v, int64_t n, double mean, int32_t threshold,
                  mph_i32_stats_t* out, mph_stream_t stream);

/* ------------------------------------------------------------ batched */
/* Many matrices of uneven size and structure, each with k right-hand sides,
 * in one launch over a list of (matrix, row block) tiles, one workgroup per
 * tile.  FP32 / BF16 rows of more than 1024 entries are summed with Dot2
 * compensation by a second kernel on a stream of the batch's own, forked
 * from and joined to the caller's stream (both kernels are captured by
 * mph_batch_run_graph).  No reference counterpart (the reference multiplies
 * one matrix by one vector per call).  Matrices are given as handles; their
 * formats may differ.  A COO member must have its entries sorted by row (its
 * row offsets are derived on the device and kept by the batch); an unsorted
 * one is refused.  X[i] is ncols_i x k and Y[i] nrows_i x k, row-major,
 * aligned to min(16, k * element size) bytes.  Runs of one batch must not
 * overlap (on several streams or threads): they share the batch's second
 * stream and its events.  Tile shapes, read at mph_batch_create from the
 * environment (the result does not depend on them): MPH_BATCH_ROWS_PER_GROUP
 * and MPH_BATCH_CSR_ROWS_PER_GROUP (4 by default, or 1: rows a lane group
 * of a DIA / ELL resp. CSR-shaped row tile takes), MPH_BATCH_STAGE (1 by
 * default: narrow DIA members stage X rows in LDS), MPH_BATCH_WAVE_ROWS (1
 * by default: rows per wave of a one-wave tile), MPH_BATCH_SORT_ROWS (1 by
 * default: the short rows of CSR-shaped members go to tiles sorted by row
 * length), MPH_BATCH_SUB_MAX (128 by default: rows of 33 up to this many
 * entries take four rows per wave; 0: one row per wave). */
typedef struct mph_batch_s* mph_batch_t;
int mph_batch_create(mph_batch_t* out, int32_t nmat, const mph_matrix_t* mats,
                     int32_t k, const void* const* X, void* const* Y,
                     mph_stream_t stream);
int mph_batch_destroy(mph_batch_t b);
/* y_i = A_i x_i (init) or y_i += A_i x_i for all i, in one launch. */
int mph_batch_run(mph_batch_t b, int init, mph_stream_t stream);
/* Captures mph_batch_run into a hipGraph once and replays it. */
int mph_batch_run_graph(mph_batch_t b, int init, mph_stream_t stream);
/* Number of tiles of the tile kernel, and of rows of the Dot2 kernel. */
int mph_batch_info(mph_batch_t b, int64_t* ntiles, int32_t* ndot2_rows);

/* ------------------------------------------------------------ casts */
/* Element casts for containers whose index / value types the kernels do not
 * take (core/src/Morpheus_ContainerTraits.hpp:59-84 admits any integral index
 * type and any arithmetic value type; the reference's tests use long long
 * indices and int values, core/tests/utils/Utils.hpp:50-60).
 * mph_index_narrow: int64 -> int32, MPH_ERR_INVALID_INPUT when an entry does
 * not fit (synchronises the stream).  mph_index_widen: int32 -> int64.
 * mph_values_cast: between MPH_F64, MPH_F32 and MPH_I32 (floating point to
 * integer rounds to nearest).  No reference counterpart. */
int mph_index_narrow(const int64_t* src, int64_t n, int32_t* dst, mph_stream_t stream);
int mph_index_widen(const int32_t* src, int64_t n, int64_t* dst, mph_stream_t stream);
int mph_values_cast(const void* src, int src_type, void* dst, int dst_type, int64_t n,
                    mph_stream_t stream);

/* ------------------------------------------------------------ probes */
/* Device rate probes for the benchmark's ceilings (no reference
 * counterpart).  mph_probe_read: one pass of 16-byte non-temporal loads over
 * nbytes of `data` (16-byte aligned), nothing written (`sink`, one double,
 * is written only if the data sum to a NaN).  mph_probe_gather: for items
 * j < n, columns[j] read in order and x[columns[j]] gathered, eight items per
 * thread; out[t] (ceil(n / 8) doubles) = thread t's sum. */
int mph_probe_read(const void* data, int64_t nbytes, void* sink, mph_stream_t stream);
int mph_probe_gather(const int32_t* columns, int64_t n, const double* x, double* out,
                     mph_stream_t stream);

/* ------------------------------------------------------------ autotuner */
/* Times each admissible launch shape of A on `stream` (A's own x/y shapes,
 * `reps` timed runs each) and keeps the fastest in the handle.  x and y are
 * scratch vectors of A's column/row count (k = 1) or k-wide. */
int mph_matrix_autotune(mph_matrix_t A, int32_t k, const void* x, void* y,
                        int32_t reps, mph_stream_t stream, float* best_ms);

/* Converts the CSR matrix to every admissible format on the GPU (formats
 * whose conversion throws FormatConversionException in the reference are
 * skipped), times `reps` products with k right-hand sides each, and returns
 * the fastest format in *best_format (an mph_format).  ms, if not NULL,
 * receives the median time per format (-1 for skipped formats).  The
 * counterpart of choosing a DynamicMatrix's active format
 * (core/src/Morpheus_DynamicMatrix.hpp:409-442) by measurement instead of by
 * the --morpheus-format flag (core/src/impl/Morpheus_Core.cpp:92-140). */
int mph_autotune_format(int64_t nrows, int64_t ncols, int64_t nnz,
                        const int32_t* row_offsets,
                        const int32_t* column_indices, const void* values,
                        int value_type, int32_t k, int32_t reps,
                        mph_stream_t stream, int* best_format, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* MORPHEUS_AMD_H */
