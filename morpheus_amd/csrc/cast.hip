// cast.hip -- element casts between the containers' index / value types and
// the types the kernels take (the Morpheus drop-in's int64-indexed and
// int-valued containers; core/src/Morpheus_ContainerTraits.hpp:59-84 admits
// any integral index type and any arithmetic value type).
#include "mph_internal.hpp"

namespace mph {
namespace {

constexpr int kCT = 256;

// dst[i] = src[i] narrowed to 32 bits; *bad = 1 where it does not fit
__global__ void narrow_kernel(const int64_t* __restrict__ src, int64_t n, int32_t* __restrict__ dst,
                              int32_t* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t v = src[i];
  if (v < INT32_MIN || v > INT32_MAX) *bad = 1;
  dst[i] = (int32_t)v;
}

__global__ void clear_flag_kernel(int32_t* flag) { *flag = 0; }

__global__ void widen_kernel(const int32_t* __restrict__ src, int64_t n, int64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

template <typename S, typename D>
__device__ __forceinline__ D cast_one(S v) {
  if constexpr (std::is_integral<D>::value && !std::is_integral<S>::value)
    return (D)__builtin_rint(v);  // nearest integer (ties to even)
  else
    return (D)v;
}

template <typename S, typename D>
__global__ void cast_kernel(const S* __restrict__ src, int64_t n, D* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = cast_one<S, D>(src[i]);
}

template <typename S>
int cast_from(const S* src, void* dst, int dst_type, int64_t n, hipStream_t s) {
  const dim3 g((unsigned)ceil_div(n, kCT)), b(kCT);
  switch (dst_type) {
    case MPH_F64: hipLaunchKernelGGL((cast_kernel<S, double>), g, b, 0, s, src, n, static_cast<double*>(dst)); break;
    case MPH_F32: hipLaunchKernelGGL((cast_kernel<S, float>), g, b, 0, s, src, n, static_cast<float*>(dst)); break;
    case MPH_I32: hipLaunchKernelGGL((cast_kernel<S, int32_t>), g, b, 0, s, src, n, static_cast<int32_t*>(dst)); break;
    default: return invalid("mph_values_cast: destination type must be MPH_F64, MPH_F32 or MPH_I32");
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace
}  // namespace mph

using namespace mph;

extern "C" {

int mph_index_narrow(const int64_t* src, int64_t n, int32_t* dst, mph_stream_t stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (n < 0 || (n > 0 && (!src || !dst))) return invalid("mph_index_narrow: bad arguments");
  if (n == 0) return MPH_OK;
  int32_t* bad = nullptr;
  MPH_HIP_CHECK(hipMalloc(&bad, sizeof(int32_t)));
  int32_t flag = 0;
  hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, s, bad);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)ceil_div(n, kCT)), dim3(kCT), 0, s, src, n,
                       dst, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&flag, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(bad);
  MPH_HIP_CHECK(e);
  if (flag) return invalid("an index does not fit in 32 bits (the kernels take int32 indices)");
  return MPH_OK;
}

int mph_index_widen(const int32_t* src, int64_t n, int64_t* dst, mph_stream_t stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (n < 0 || (n > 0 && (!src || !dst))) return invalid("mph_index_widen: bad arguments");
  if (n == 0) return MPH_OK;
  hipLaunchKernelGGL(widen_kernel, dim3((unsigned)ceil_div(n, kCT)), dim3(kCT), 0, s, src, n, dst);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

int mph_values_cast(const void* src, int src_type, void* dst, int dst_type, int64_t n,
                    mph_stream_t stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (n < 0 || (n > 0 && (!src || !dst))) return invalid("mph_values_cast: bad arguments");
  if (n == 0) return MPH_OK;
  switch (src_type) {
    case MPH_F64: return cast_from(static_cast<const double*>(src), dst, dst_type, n, s);
    case MPH_F32: return cast_from(static_cast<const float*>(src), dst, dst_type, n, s);
    case MPH_I32: return cast_from(static_cast<const int32_t*>(src), dst, dst_type, n, s);
  }
  return invalid("mph_values_cast: source type must be MPH_F64, MPH_F32 or MPH_I32");
}

}  // extern "C"
