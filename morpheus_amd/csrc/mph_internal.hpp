// mph_internal.hpp -- matrix handles, analysis ("plans") and the launcher
// entry points shared between the kernel translation units and the C-ABI.
#pragma once

#include <cstdint>
#include <vector>

#include "mph_common.hpp"

namespace mph {

// A row block of the CSR-family kernels.  sub < 0: rows [row, next.row) are
// reduced from an LDS-staged slice of at most 2*chunk items.  sub >= 0: the
// sub-th chunk of `chunk` items of the single long row `row`.
struct RowBlock {
  int32_t row;
  int32_t sub;
};

constexpr int kRowBlockThreads = 256;
constexpr int kRowBlockMaxRows = 512;    // rows of one short block
constexpr int kDefaultChunkF64 = 1024;   // LDS items of a short block (x2)
constexpr int kDefaultChunkF32 = 2048;
constexpr int kCompensateLen = 1024;     // FP32 rows longer than this use TwoSum

// x-tiled long rows (SpMV): the columns are cut into tiles of 32 KiB of x.
// The analysis copies the long rows' items tile-major (16-bit column within
// the tile + value); workgroup (group, tile) stages its x tile in LDS and its
// waves reduce the group's rows over that tile, one sub-block per wave.
struct XTSub {
  int32_t i0;  // first long-row index of the sub-block (the next record ends it)
  int32_t e0;  // first item of the sub-block (item numbering of the run offsets)
  int32_t p0;  // first slot of the sub-block in the copy (slots = 8 * ceil(n / 8))
  int32_t pad;
};
constexpr int kXTileThreads = 512;
constexpr int kXTileBytes = 32 << 10;          // x tile in LDS
constexpr int kXTileCap = 512;                 // items of a (one-wave) sub-block
constexpr int kXTileRows = 64;                 // rows of a sub-block
constexpr int kXTileMaxGroupRows = 1024;       // rows of a group (run offsets in LDS)
// SpMM with x-tiled rows: the x-tiled rows take the x-tiled pass for k up to
// kXTileSpmmMaxK, and for k up to kXTileSpmmLongRunMaxK when the tiled rows
// average at least kXTileSpmmLongRun items per (row, tile) run -- one pass
// per column (a plan without the value copy), or the k-wide pass over
// groups of 4 / 2 columns (the Kw limits)
constexpr int kXTileSpmmMaxK = 2;
constexpr int kXTileSpmmLongRunMaxK = 4;
constexpr int kXTileSpmmKwMaxK = 4;
constexpr int kXTileSpmmKwLongRunMaxK = 16;
constexpr int64_t kXTileSpmmLongRun = 32;
constexpr int64_t kXTileMaxBounds = 1 << 26;  // entries of the per-row tile-bound table (rows x (tiles + 1))
constexpr int64_t kXTileMinXBytes = 2 << 20;  // auto: x larger than this
constexpr int64_t kXTileMinRowLen = 512;       // rows longer than this are candidates
// HYB / HDC with at least this many rows run the ELL / DIA part as its own
// coalesced pass and the tail / CSR part over the rows that have one.
constexpr int64_t kSplitMinRows = 1 << 16;
// split FP64 HDC: CSR rows of at most this many entries are added by the DIA
// pass's threads
constexpr int64_t kHdcTailMax = 8;
// HYB from kSplitMinRows rows (or with x-tiling on) whose ELL slots are
// filled below kHybMergeMaxFill, or with more than 1 / kHybMergeTailRows of
// its rows holding COO entries, runs over a merged CSR copy
constexpr double kHybMergeMaxFill = 1.0 / 3.0;
constexpr int64_t kHybMergeTailRows = 8;

// The x-tiled pass of one matrix's long rows (see xtile_kernel).
struct XTilePlan {
  int32_t tile = 0;              // x entries per tile
  int32_t ntiles = 0;
  int32_t groups = 0;
  int32_t gr = 0;                // long rows per group
  int32_t rows = 0;              // tiled long rows (dealt to groups by length rank)
  int64_t items = 0;             // items of the tiled rows
  int32_t* row = nullptr;        // [rows] row ids
  int32_t* lro = nullptr;        // [(groups*ntiles) * (gr+1)] item offsets of each row's run
  XTSub* sub = nullptr;          // sub-blocks, a sentinel after each (group, tile)
  int32_t* sub_start = nullptr;  // [groups*ntiles + 1] first sub-block of (group, tile)
  uint8_t* head = nullptr;       // [sub-blocks * 64] run of each lane's first item
  uint16_t* col = nullptr;       // tile-major copy: column - tile * tile width
  void* val = nullptr;           // tile-major copy: values
  void* partial = nullptr;       // [ntiles * rows] (pairs in FP32), tile-major
  // the k-wide pass's partials, [ntiles * rows * partial_k_cols] (pairs in
  // FP32), allocated by its first launch
  void* partial_k = nullptr;
  int32_t partial_k_cols = 0;
  // direct plan (no copy of any value): the position in the caller's arrays
  // of each run's first item (lro's shape), and the caller's value array the
  // pass reads instead of `val`
  int32_t* rpos = nullptr;
  const void* src = nullptr;
};

// Analysis of a CSR-shaped item set (CSR itself, the CSR part of HDC, the
// COO tail of HYB viewed through row offsets built here).
struct RowBlockPlan {
  int32_t chunk = 0;
  // rows longer than this leave the short blocks (chunks of `chunk` items,
  // or the x-tiled pass): `chunk` in FP64, at most kCompensateLen in FP32 /
  // BF16, so that every row summed with compensation is summed on a path
  // that carries Dot2 pairs
  int32_t long_len = 0;
  // FP32 / BF16 HYB / HDC whose ELL / DIA part holds at least
  // kCompensateLen / 2 slots per row: that part runs as its own pass (the
  // DIA / ELL kernels compensate parts longer than kCompensateLen) and the
  // tail / CSR part accumulates onto it; never fused into the row passes
  bool prefix_apart = false;
  int32_t nblocks = 0;
  RowBlock* blocks = nullptr;    // nblocks + 1 (sentinel)
  void* partial = nullptr;       // per-block accumulators (long-row chunks)
  int32_t* tail_ro = nullptr;    // HYB / COO: row offsets of the COO entries
  // the CSR-shaped arrays the row pass reads: CSR / HDC's CSR part, the COO
  // arrays over tail_ro, or the merged copy below
  const int32_t* ro = nullptr;
  const int32_t* ci = nullptr;
  const void* val = nullptr;
  // HYB whose ELL part is mostly padding: a CSR copy of the whole matrix
  // (per row its ELL slots in slot order, then its COO entries), owned here
  int32_t* m_ro = nullptr;
  int32_t* m_ci = nullptr;
  void* m_val = nullptr;
  int32_t nlong = 0;             // rows split into chunks
  // the short-block kernel's records, blocks / sblocks order, sentinel
  // included: {row, sub, first item, lanes per row of the block's reduction
  // (picked by the analysis; 0 for chunk and tiled records)}
  int32_t* krec = nullptr;
  int32_t* skrec = nullptr;
  // x-tiled long rows (SpMV); xt.rows == 0: long rows use position chunks
  XTilePlan xt;
  // split SpMV plan (x-tiled rows, and / or HYB / HDC in two passes):
  // records over the list rows, {row, -2} per tiled row, chunks of the rest
  RowBlock* sblocks = nullptr;
  int32_t nsblocks = 0;
  int32_t nchunked = 0;          // long rows left to position chunks
  bool split = false;            // HYB / HDC: ELL / DIA pass, then the list
  // split FP64 HDC whose CSR rows all hold at most kHdcTailMax entries (none
  // x-tiled or chunked): the DIA pass adds each row's CSR entries itself
  // (hdc_tail_multiply), so no list pass read-modify-writes scattered y
  bool hdc_tail = false;
  int32_t* smap = nullptr;       // split: list row -> matrix row
  int32_t* sro = nullptr;        // split: item offsets of the list rows (+1)
  // split HYB whose ELL slots are mostly padding: per row, 1 + its last slot
  // holding a column (0 for a row without any); the ELL pass then loads no
  // slot beyond it, so lines holding only padding are never fetched
  uint8_t* ell_cnt = nullptr;
  int64_t ell_visited = 0;       // the sum of ell_cnt
};

// DIA SpMV with x staged in LDS: the diagonals, in storage order, cut into
// clusters of consecutive diagonals whose offsets lie within kDiaClusterSpan
// of each other.  A workgroup of 256 rows stages, for each cluster of two or
// more diagonals, the x entries its rows read through that cluster
// (256 + span of them) once; a one-diagonal cluster reads x directly.
constexpr int kDiaMaxClusters = 8;
constexpr int kDiaClusterSpan = 512;
constexpr int kDiaStageMax = 2048;  // x entries staged per workgroup
constexpr int kDiaStageMaxDiags = 64;
constexpr int kDiaPairMinDiags = 8;  // row-pair kernels: mean diagonals per staged cluster
constexpr int32_t kDiaDirect = INT32_MIN;
struct DiaStage {
  int32_t n = 0;                    // staged clusters (0: the plain DIA kernel)
  int32_t omin[kDiaMaxClusters];    // smallest offset of the cluster
  int32_t base[kDiaMaxClusters];    // first LDS entry of its window
  int32_t len[kDiaMaxClusters];     // entries of its window (256 + span)
  // per diagonal: the LDS entry holding x[r0 + 0] for its cluster's window
  // (base - omin), or kDiaDirect for a diagonal read from x directly
  int32_t lb[kDiaStageMaxDiags];
  // the same for the row-pair kernel's workgroups of 512 rows (windows of
  // 512 + span entries, cluster c's at base + 256 c), when they fit in
  // 2 kDiaStageMax entries (pair != 0)
  int32_t pair = 0;
  int32_t lb2[kDiaStageMaxDiags];
};

// Analysis of a row-sorted COO matrix for the segmented reduction.
// A COO wave chunk: entries [s, e).
struct CooChunk {
  int32_t s, e;
};

struct CooPlan {
  int items = 8;          // entries per lane
  int64_t nchunks = 0;    // wave chunks of 64*items entries
  int32_t nblocks = 0;
  void* first_val = nullptr;   // nchunks accumulators
  void* last_val = nullptr;    // nchunks accumulators
  bool sorted = true;          // rows non-decreasing
  bool prezero = false;        // y cleared by a fill pass instead of gap writes
  bool rowplan = false;        // row-sorted, coo_items 0: the row-block plan
  // with x-tiled long rows, the segmented pass covers the other entries only,
  // through an explicit chunk table
  XTilePlan xt;
  CooChunk* chunks = nullptr;  // [nchunks], nullptr: chunk c = [c*64*items, ...)
};

struct mph_matrix_impl {
  int format = MPH_CSR;
  int vt = MPH_F64;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  // COO / HYB-COO
  const int32_t* coo_ri = nullptr;
  const int32_t* coo_ci = nullptr;
  const void* coo_v = nullptr;
  int64_t coo_nnz = 0;
  // CSR / HDC-CSR
  const int32_t* csr_ro = nullptr;
  const int32_t* csr_ci = nullptr;
  const void* csr_v = nullptr;
  int64_t csr_nnz = 0;
  // DIA / HDC-DIA
  int64_t ndiag = 0, dia_pitch = 0, dia_nnz = 0;
  const int32_t* dia_off = nullptr;
  const void* dia_v = nullptr;
  DiaStage dia_stage;  // from the analysis
  // ELL / HYB-ELL
  int64_t ell_K = 0, ell_width = 0, ell_pitch = 0, ell_nnz = 0;
  int64_t ell_loop = 0;  // slots visited per row (K, or width if needed)
  const int32_t* ell_ci = nullptr;
  const void* ell_v = nullptr;

  mph_tuning tuning{0, 0, 0, 0};
  int32_t policy = 0;  // MPH_POLICY_* flags
  bool analysed = false;
  RowBlockPlan rb;
  CooPlan coo;
  int device = 0;
};

// ---------------------------------------------------------------- launchers
// All return an mph_status.  `x`/`y` element type follows vtraits<V>::vec.
int rowblock_analyse(mph_matrix_impl* A, hipStream_t s);
int rowblock_multiply(mph_matrix_impl* A, int k, const void* x, void* y,
                      int init, hipStream_t s);
int coo_analyse(mph_matrix_impl* A, hipStream_t s);
int coo_multiply(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s);
// Clusters A's diagonals for the LDS-staged DIA SpMV (A->dia_stage).
int dia_analyse(mph_matrix_impl* A, hipStream_t s);
// HDC with hdc_tail set, k = 1: the staged DIA pass with each row's CSR
// entries added after its diagonals
int hdc_tail_multiply(mph_matrix_impl* A, const void* x, void* y, int init, hipStream_t s);
int dia_multiply(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s);
int ell_multiply(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s);
void release_plans(mph_matrix_impl* A);
int rows_sorted(const int32_t* ri, int64_t nnz, int64_t nrows, hipStream_t s,
                int* sorted);
int ell_padding_only(const int32_t* ci, int64_t pitch, int64_t nrows, int64_t K,
                     int64_t width, hipStream_t s, int* padding_only);

// device scan helpers (scan.hip)
// exclusive prefix sum of n int32 counts into out (n+1 entries, out[n] =
// total); `tmp` must hold scan_tmp_elems(n) int64 values.
int64_t scan_tmp_elems(int64_t n);
int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n,
                       int64_t* tmp, hipStream_t s);
int fill_zero(void* p, size_t bytes, hipStream_t s);

int device_cu_count();

// validate.hip
int check_index_range(const int32_t* idx, int64_t n, int64_t hi, bool allow_m1,
                      hipStream_t s, const char* what);
int check_row_offsets(const int32_t* ro, int64_t nrows, int64_t nnz,
                      hipStream_t s, const char* what);
int validate_rowblock_plan(const int32_t* d_ro, int64_t nrows, RowBlockPlan& p,
                           hipStream_t s, std::vector<int32_t>* krec = nullptr);
// lanes per row for a short block of nr rows of lengths len(0..nr-1)
uint8_t block_lanes(int nr, const int32_t* ro_begin);

// x-tiled long rows.  xtile_build plans the candidate rows `lrows` (host row
// offsets `ro`, device row offsets / columns / values) when the x-tile mode
// and x size call for it; rows with unsorted columns cancel the plan and rows
// with a run longer than a sub-block (kXTileCap items) in some tile are left
// out (the tile is halved, down to an eighth, while more than an eighth of
// the candidate items would be).
// tiled[r] != 0 marks the rows taken.  xtile_multiply runs the pass of a
// plan without DIA / ELL parts (COO) and writes y of the tiled rows (x and y
// read with element stride ld).  direct: the plan copies the columns only;
// the pass reads the values from d_val at each run's position.
int xtile_build(mph_matrix_impl* A, const std::vector<int32_t>& ro, const int32_t* d_ro,
                const int32_t* d_ci, const void* d_val, std::vector<int32_t> lrows,
                XTilePlan& xp, std::vector<char>& tiled, hipStream_t s, bool direct = false);
void xtile_release(XTilePlan& xp);
int xtile_multiply(mph_matrix_impl* A, const XTilePlan& xp, const void* x, void* y, int init,
                   int64_t ld, hipStream_t s);
// The same pass for kr = 2 or 4 columns at once (a plan with the value copy):
// the x tile's kr columns staged together, each item read once; column j
// equals xtile_multiply's result over column j bit for bit.
int xtile_multiply_k(mph_matrix_impl* A, XTilePlan& xp, int kr, const void* x, void* y, int init,
                     int64_t ld, hipStream_t s);
// the widest such pass (4, 2; 1: none) whose LDS tile the device allows
int xtile_k_cols(const mph_matrix_impl* A, const XTilePlan& xp);
// HYB / COO: row offsets of a row-sorted COO entry list (ro has nrows + 1 entries)
int coo_row_offsets(const int32_t* ri, int64_t nnz, int64_t nrows, int32_t* ro, hipStream_t s);
int default_chunk(int vt);
// the device address of the current device's mph_coo_multiply_noalloc
// scratch (spmv_coo.hip), looked up on first use per device
int noalloc_scratch(void** out);

// the x-tiling mode of the segmented COO pass (tuning.xtile: 0 auto, 1 off,
// 2 on); off when the policy forbids value copies (that pass has no direct
// plan; the row-block plans read the values in place instead)
inline int xtile_mode(const mph_matrix_impl* A) {
  return (A->policy & MPH_POLICY_NO_VALUE_COPIES) ? 1 : A->tuning.xtile;
}

}  // namespace mph

struct mph_matrix_s : mph::mph_matrix_impl {};
