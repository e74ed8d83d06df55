// scan.hip -- device prefix sums and fills used by the analysis and the
// on-GPU converters.  Three passes: per-tile sums, a single-workgroup scan of
// the tile sums, and the per-tile exclusive scan with the tile offset added.
#include "mph_internal.hpp"

namespace mph {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Inclusive scan of one value per thread over a 256-thread block.
template <typename T>
__device__ T block_inclusive_scan(T v, T* s_wave, T& total) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  for (int d = 1; d < kWave; d <<= 1) {
    T o = shfl_up(v, d);
    if (lane >= d) v += o;
  }
  if (lane == kWave - 1) s_wave[wave] = v;
  __syncthreads();
  T base = 0;
  total = 0;
  for (int w = 0; w < kScanThreads / kWave; w++) {
    if (w < wave) base += s_wave[w];
    total += s_wave[w];
  }
  __syncthreads();
  return v + base;
}

__global__ __launch_bounds__(kScanThreads) void tile_sums_kernel(
    const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums) {
  __shared__ int64_t s_wave[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t v = 0;
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + threadIdx.x + (int64_t)k * kScanThreads;
    if (i < n) v += in[i];
  }
  int64_t total;
  block_inclusive_scan<int64_t>(v, s_wave, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(
    int64_t* __restrict__ sums, int64_t nt) {
  __shared__ int64_t s_wave[kScanThreads / kWave];
  int64_t carry = 0;
  for (int64_t base = 0; base < nt; base += kScanThreads) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < nt ? sums[i] : 0;
    int64_t total;
    const int64_t inc = block_inclusive_scan<int64_t>(v, s_wave, total);
    if (i < nt) sums[i] = carry + inc - v;
    carry += total;
  }
  if (threadIdx.x == 0) sums[nt] = carry;
}

__global__ __launch_bounds__(kScanThreads) void tile_scan_kernel(
    const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t n,
    const int64_t* __restrict__ sums, int64_t nt) {
  __shared__ int64_t s_wave[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  // each thread owns kScanItems consecutive elements
  const int64_t mine = base + (int64_t)threadIdx.x * kScanItems;
  int32_t vals[kScanItems];
  int64_t local = 0;
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = mine + k;
    vals[k] = i < n ? in[i] : 0;
    local += vals[k];
  }
  int64_t total;
  const int64_t inc = block_inclusive_scan<int64_t>(local, s_wave, total);
  int64_t run = sums[blockIdx.x] + inc - local;
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = mine + k;
    if (i < n) out[i] = (int32_t)run;
    run += vals[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = (int32_t)sums[nt];
}

__global__ void fill_zero_kernel(uint32_t* __restrict__ p, int64_t nwords) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += stride)
    p[i] = 0u;
}

__global__ void fill_zero_bytes_kernel(uint8_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

int64_t scan_tmp_elems(int64_t n) { return ceil_div(n, kScanTile) + 2; }

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n,
                       int64_t* tmp, hipStream_t s) {
  if (n < 0) return invalid("scan: negative length");
  const int64_t nt = ceil_div(n, kScanTile);
  if (nt == 0) {
    // out[0] = 0
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<uint32_t*>(out), (int64_t)1);
    MPH_HIP_CHECK(hipGetLastError());
    return MPH_OK;
  }
  hipLaunchKernelGGL(tile_sums_kernel, dim3((unsigned)nt), dim3(kScanThreads),
                     0, s, in, n, tmp);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, s, tmp,
                     nt);
  hipLaunchKernelGGL(tile_scan_kernel, dim3((unsigned)nt), dim3(kScanThreads),
                     0, s, in, out, n, tmp, nt);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

// Clears device memory with a kernel rather than hipMemsetAsync so that the
// call stays safe inside captured hipGraphs.
int fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return MPH_OK;
  const int64_t words = (int64_t)(bytes / 4);
  if (words > 0) {
    const int64_t blocks = std::min<int64_t>(ceil_div(words, 256), 65536);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       s, reinterpret_cast<uint32_t*>(p), words);
  }
  const int64_t rem = (int64_t)(bytes % 4);
  if (rem > 0) {
    hipLaunchKernelGGL(fill_zero_bytes_kernel, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<uint8_t*>(p) + words * 4, rem);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace mph
