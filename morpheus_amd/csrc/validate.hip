// validate.hip -- one-time checks run by the analysis, so that malformed
// input or a malformed analysis is reported as MPH_ERR_INVALID_INPUT /
// MPH_ERR_RUNTIME instead of reaching a kernel as an out-of-range access.
//
//  * index arrays: every row / column index inside its dimension (ELL: or
//    the -1 padding marker); CSR row offsets start at 0, never decrease and
//    end at nnz;
//  * row-block plans: copied back to the host and checked against the row
//    offsets (sentinel, row ranges, per-block item budget, long-row chunk
//    numbering) before any launch uses them.
#include <functional>
#include <vector>

#include "mph_internal.hpp"

namespace mph {

namespace {

__global__ void index_range_kernel(const int32_t* __restrict__ idx, int64_t n,
                                   int64_t hi, int allow_m1,
                                   int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t v = idx[i];
  const bool ok = (v >= 0 && v < hi) || (allow_m1 && v == -1);
  if (!ok) atomicOr(flag, 1);
}

__global__ void row_offsets_kernel(const int32_t* __restrict__ ro,
                                   int64_t nrows, int64_t nnz,
                                   int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nrows) return;
  bool ok = true;
  if (i == 0) ok = ro[0] == 0;
  if (i == nrows) ok = ok && (int64_t)ro[nrows] == nnz;
  if (i < nrows) ok = ok && ro[i] <= ro[i + 1];
  if (!ok) atomicOr(flag, 1);
}

int run_flag_check(hipStream_t s, int* bad,
                   const std::function<void(int32_t*)>& launch) {
  int32_t* flag = nullptr;
  MPH_HIP_CHECK(hipMalloc(&flag, sizeof(int32_t)));
  int st = fill_zero(flag, sizeof(int32_t), s);
  int32_t h = 0;
  if (st == MPH_OK) {
    launch(flag);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(std::string("index check: ") + hipGetErrorString(e));
      st = MPH_ERR_HIP;
    }
  }
  dfree(flag);
  *bad = h;
  return st;
}

}  // namespace

int check_index_range(const int32_t* idx, int64_t n, int64_t hi, bool allow_m1,
                      hipStream_t s, const char* what) {
  if (n <= 0) return MPH_OK;
  int bad = 0;
  MPH_CHECK_STATUS(run_flag_check(s, &bad, [&](int32_t* flag) {
    hipLaunchKernelGGL(index_range_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                       idx, n, hi, allow_m1 ? 1 : 0, flag);
  }));
  if (bad) return invalid(std::string(what) + ": index out of range");
  return MPH_OK;
}

int check_row_offsets(const int32_t* ro, int64_t nrows, int64_t nnz,
                      hipStream_t s, const char* what) {
  int bad = 0;
  MPH_CHECK_STATUS(run_flag_check(s, &bad, [&](int32_t* flag) {
    hipLaunchKernelGGL(row_offsets_kernel, dim3((unsigned)ceil_div(nrows + 1, 256)), dim3(256),
                       0, s, ro, nrows, nnz, flag);
  }));
  if (bad)
    return invalid(std::string(what) +
                   ": row offsets must start at 0, not decrease and end at nnz");
  return MPH_OK;
}

int validate_rowblock_plan(const int32_t* d_ro, int64_t nrows, RowBlockPlan& p,
                           hipStream_t s, std::vector<int32_t>* krec) {
  const int64_t nb = p.nblocks;
  std::vector<int32_t> ro(nrows + 1);
  std::vector<RowBlock> bl(nb + 1);
  MPH_HIP_CHECK(hipMemcpyAsync(ro.data(), d_ro, sizeof(int32_t) * (nrows + 1),
                               hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipMemcpyAsync(bl.data(), p.blocks, sizeof(RowBlock) * (nb + 1),
                               hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  auto fail = [](const std::string& m) {
    set_error("row-block analysis check failed: " + m);
    return MPH_ERR_RUNTIME;
  };
  if (bl[nb].row != nrows || bl[nb].sub != -1) return fail("sentinel");
  if (nb > 0 && bl[0].row != 0) return fail("first block");
  p.nlong = 0;
  for (int64_t b = 0; b < nb; b++) {
    const RowBlock k = bl[b];
    if (k.row < 0 || k.row >= nrows) return fail("row out of range");
    const int64_t len = (int64_t)ro[k.row + 1] - ro[k.row];
    if (k.sub < 0) {
      const int64_t r1 = bl[b + 1].row;
      if (r1 <= k.row || r1 > nrows) return fail("short block row range");
      if (r1 - k.row > kRowBlockMaxRows) return fail("short block row count");
      if ((int64_t)ro[r1] - ro[k.row] >= 2 * (int64_t)p.chunk) return fail("short block items");
      for (int64_t r = k.row; r < r1; r++)
        if ((int64_t)ro[r + 1] - ro[r] > p.long_len) return fail("long row in short block");
    } else {
      if (len <= p.long_len) return fail("chunked row is short");
      const int64_t nsub = (len + p.chunk - 1) / p.chunk;
      if (k.sub >= nsub || b - k.sub < 0) return fail("chunk index");
      const RowBlock f = bl[b - k.sub];
      if (f.row != k.row || f.sub != 0) return fail("chunk numbering");
      if (k.sub == 0) p.nlong++;
      if (k.sub == nsub - 1 && bl[b + 1].row != k.row + 1 && bl[b + 1].row <= k.row)
        return fail("chunk successor");
    }
  }
  // the short-block kernel's records: {row, sub, first item, lanes}
  if (krec) {
    krec->assign(4 * (nb + 1), 0);
    for (int64_t b = 0; b <= nb; b++) {
      (*krec)[4 * b] = bl[b].row;
      (*krec)[4 * b + 1] = bl[b].sub;
      (*krec)[4 * b + 2] = ro[bl[b].row];
      if (b < nb && bl[b].sub < 0)
        (*krec)[4 * b + 3] = block_lanes(bl[b + 1].row - bl[b].row, ro.data() + bl[b].row);
    }
  }
  return MPH_OK;
}

}  // namespace mph
