// calibrate.hip -- device rate probes that give the SpMV kernels' ceilings
// on the device they run on (no counterpart in the reference):
//  * a read stream: one pass over a buffer with 16-byte non-temporal loads,
//    grid-stride, four loads in flight per thread, nothing written -- what a
//    product that reads its bytes once could reach at that size;
//  * a gather: per item a column index read in order and x[column] gathered,
//    eight items per thread, coalesced index loads -- the cost of the x
//    gathers of a row pass (e.g. a power-law matrix's short rows) without
//    the row bookkeeping.
#include "mph_internal.hpp"

namespace mph {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void read_probe_kernel(int64_t n16, const d2* __restrict__ p,
                                                         double* __restrict__ out) {
  const int64_t tid = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  d2 s = {0, 0};
  int64_t i = tid;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const d2 a = __builtin_nontemporal_load(p + i);
    const d2 b = __builtin_nontemporal_load(p + i + stride);
    const d2 c = __builtin_nontemporal_load(p + i + 2 * stride);
    const d2 d = __builtin_nontemporal_load(p + i + 3 * stride);
    s += a + b + c + d;
  }
  for (; i < n16; i += stride) s += __builtin_nontemporal_load(p + i);
  // one write per thread only if the sum is a NaN: the loads are kept and
  // nothing is written in practice
  if (s.x != s.x || s.y != s.y) out[0] = s.x + s.y;
}

__global__ __launch_bounds__(256) void gather_probe_kernel(int64_t n, const int32_t* __restrict__ ci,
                                                           const double* __restrict__ x,
                                                           double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x % kWave;
  const int64_t base = (t - lane) * 8;
  int32_t c[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int64_t j = base + lane + kWave * q;
    c[q] = ld_stream(ci + (j < n ? j : n - 1));
  }
  double s = 0;
#pragma unroll
  for (int q = 0; q < 8; q++) s += x[c[q]];
  if (t < (n + 7) / 8) out[t] = s;  // out holds ceil(n / 8) sums
}

}  // namespace
}  // namespace mph

using namespace mph;

extern "C" int mph_probe_read(const void* data, int64_t nbytes, void* sink, mph_stream_t stream) {
  if (nbytes < 0 || (nbytes > 0 && !data) || !sink) return invalid("null or negative argument");
  if (reinterpret_cast<uintptr_t>(data) % 16 != 0) return invalid("data must be 16-byte aligned");
  const int64_t n16 = nbytes / 16;
  if (n16 == 0) return MPH_OK;
  // four workgroups per CU, about 16 loads per thread at 2 GiB
  const int64_t blocks = std::min<int64_t>(ceil_div(n16, 256), (int64_t)device_cu_count() * 16);
  hipLaunchKernelGGL(read_probe_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n16, static_cast<const d2*>(data),
                     static_cast<double*>(sink));
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

extern "C" int mph_probe_gather(const int32_t* columns, int64_t n, const double* x, double* out,
                                mph_stream_t stream) {
  if (n < 0 || (n > 0 && (!columns || !x || !out))) return invalid("null or negative argument");
  if (n == 0) return MPH_OK;
  const int64_t threads = ceil_div(n, 8);
  hipLaunchKernelGGL(gather_probe_kernel, dim3((unsigned)ceil_div(threads, 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n, columns, x, out);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}
