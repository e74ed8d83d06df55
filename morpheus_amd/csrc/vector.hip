// vector.hip -- the dense-vector algorithms around SpMV, and a conjugate
// gradient iteration built from them and the SpMV of a matrix handle.
//
// Replaces core/src/Morpheus_Dot.hpp, core/src/Morpheus_WAXPBY.hpp and
// core/src/Morpheus_Reduction.hpp (HIP:
// core/src/impl/DenseVector/HIP/Morpheus_{Dot,WAXPBY,Reduction}_Impl.hpp):
//  * dot: per-block partial sums over contiguous chunks, then one fixed-order
//    tree, in FP64 whatever the vector type, so every call gives the same
//    bits;
//  * waxpby: w = alpha x + beta y with each product and sum rounded on its
//    own (no FMA contraction), as the Serial loops compute it; alpha == 1 or
//    beta == 1 need no special case since 1 * v is exact;
//  * reduce: the same two-level tree with TwoSum-compensated partials (the
//    Serial reduce is a Kahan sum).
// The CG iteration has no reference counterpart: it is the consumer of the
// SpMV path.  Each iteration is the matrix handle's product q = A p and three
// launches that keep every scalar on the device: (p.q partials), (alpha =
// r.r / p.q; x += alpha p; r -= alpha q; r.r partials) and (beta = r.r' / r.r;
// p = r + beta p).  Every block sums the partials in the same fixed order, so
// the scalars are identical across blocks and runs; nothing synchronises
// with the host and nothing is cleared with a memset, so the loop can be
// captured in a hipGraph.
#include <algorithm>
#include <cmath>

#include "mph_internal.hpp"

namespace mph {

namespace {

constexpr int kT = 256;
constexpr int64_t kMaxParts = 1024;
constexpr int kPerThread = 8;

inline int64_t parts_for(int64_t n) {
  return std::max<int64_t>(1, std::min<int64_t>(kMaxParts, ceil_div(n, (int64_t)kT * kPerThread)));
}
inline int64_t chunk_for(int64_t n, int64_t parts) { return ceil_div(std::max<int64_t>(n, 1), parts); }

// fixed-order block sum: per-thread values into LDS, then a tree
__device__ double block_sum(double v, double* s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (int m = kT / 2; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) s[threadIdx.x] += s[threadIdx.x + m];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

// sum of parts[0..np) in a fixed order (each thread a strided run, then the
// tree), the same in every block that calls it
__device__ double sum_parts(const double* __restrict__ parts, int64_t np, double* s) {
  double t = 0;
  for (int64_t i = threadIdx.x; i < np; i += kT) t += parts[i];
  return block_sum(t, s);
}

// block b: the contiguous chunk [b * chunk, (b + 1) * chunk)
template <typename Vec>
__global__ void dot_parts_kernel(const Vec* __restrict__ x, const Vec* __restrict__ y, int64_t n,
                                 int64_t chunk, double* __restrict__ parts) {
  __shared__ double s[kT];
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  double t = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) t += (double)x[i] * (double)y[i];
  t = block_sum(t, s);
  if (threadIdx.x == 0) parts[blockIdx.x] = t;
}

template <typename Vec>
__global__ void reduce_parts_kernel(const Vec* __restrict__ x, int64_t n, int64_t chunk,
                                    double* __restrict__ parts) {
  __shared__ double s[kT];
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  double t = 0, c = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
    double u, e;
    two_sum(t, (double)x[i], u, e);
    t = u;
    c += e;
  }
  t = block_sum(t, s);
  c = block_sum(c, s);
  if (threadIdx.x == 0) parts[blockIdx.x] = t + c;
}

__global__ void sum_parts_kernel(const double* __restrict__ parts, int64_t np,
                                 double* __restrict__ out) {
  __shared__ double s[kT];
  const double t = sum_parts(parts, np, s);
  if (threadIdx.x == 0) *out = t;
}

template <typename Vec>
__global__ void waxpby_kernel(int64_t n, double alpha, const Vec* __restrict__ x, double beta,
                              const Vec* __restrict__ y, Vec* __restrict__ w) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const Vec a = (Vec)alpha, b = (Vec)beta;
    const Vec ax = a * x[i];
    const Vec by = b * y[i];
    w[i] = ax + by;
  }
}

// ------------------------------------------------------------------ CG
template <typename Vec>
__global__ void cg_update_kernel(Vec* __restrict__ x, Vec* __restrict__ r,
                                 const Vec* __restrict__ p, const Vec* __restrict__ q, int64_t n,
                                 int64_t chunk, const double* __restrict__ pq_parts, int64_t np,
                                 const double* __restrict__ state, double* __restrict__ rr_parts) {
  __shared__ double s[kT];
  const double pq = sum_parts(pq_parts, np, s);
  const Vec alpha = (Vec)(state[0] / pq);
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  double t = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
    x[i] += alpha * p[i];
    const Vec ri = r[i] - alpha * q[i];
    r[i] = ri;
    t += (double)ri * (double)ri;
  }
  t = block_sum(t, s);
  if (threadIdx.x == 0) rr_parts[blockIdx.x] = t;
}

template <typename Vec>
__global__ void cg_direction_kernel(Vec* __restrict__ p, const Vec* __restrict__ r, int64_t n,
                                    const double* __restrict__ rr_parts, int64_t np,
                                    const double* __restrict__ state) {
  __shared__ double s[kT];
  const double rr = sum_parts(rr_parts, np, s);
  const Vec beta = (Vec)(rr / state[0]);
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    p[i] = r[i] + beta * p[i];
}

// r = b - q and p = r, with the r.r partials
template <typename Vec>
__global__ void cg_start_kernel(const Vec* __restrict__ b, const Vec* __restrict__ q,
                                Vec* __restrict__ r, Vec* __restrict__ p, int64_t n, int64_t chunk,
                                double* __restrict__ rr_parts) {
  __shared__ double s[kT];
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  double t = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
    const Vec ri = b[i] - q[i];
    r[i] = ri;
    p[i] = ri;
    t += (double)ri * (double)ri;
  }
  t = block_sum(t, s);
  if (threadIdx.x == 0) rr_parts[blockIdx.x] = t;
}

template <typename Vec>
int dot_t(int64_t n, const void* x, const void* y, double* out, hipStream_t s) {
  const int64_t np = parts_for(n), chunk = chunk_for(n, np);
  double* d = nullptr;
  MPH_HIP_CHECK(hipMalloc(&d, sizeof(double) * (np + 1)));
  hipLaunchKernelGGL((dot_parts_kernel<Vec>), dim3((unsigned)np), dim3(kT), 0, s,
                     static_cast<const Vec*>(x), static_cast<const Vec*>(y), n, chunk, d);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(kT), 0, s, d, np, d + np);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d + np, sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(d);
  MPH_HIP_CHECK(e);
  return MPH_OK;
}

template <typename Vec>
int reduce_t(int64_t n, const void* x, double* out, hipStream_t s) {
  const int64_t np = parts_for(n), chunk = chunk_for(n, np);
  double* d = nullptr;
  MPH_HIP_CHECK(hipMalloc(&d, sizeof(double) * (np + 1)));
  hipLaunchKernelGGL((reduce_parts_kernel<Vec>), dim3((unsigned)np), dim3(kT), 0, s,
                     static_cast<const Vec*>(x), n, chunk, d);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(kT), 0, s, d, np, d + np);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d + np, sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(d);
  MPH_HIP_CHECK(e);
  return MPH_OK;
}

bool vector_is_double(int vt) { return vt == MPH_F64; }

int check_vt(int vt) {
  if (vt != MPH_F64 && vt != MPH_F32 && vt != MPH_BF16) return invalid("unknown value type");
  return MPH_OK;
}

}  // namespace

}  // namespace mph

// ------------------------------------------------------------------ CG state
struct mph_cg_s {
  mph_matrix_t A = nullptr;
  int64_t n = 0;
  int vt = MPH_F64;
  void *r = nullptr, *p = nullptr, *q = nullptr;
  double* parts = nullptr;  // [np] p.q partials | [np] r.r partials | [2] state
  int64_t np = 0, chunk = 0;
  bool started = false;
};

using namespace mph;

namespace {

template <typename Vec>
int cg_start_t(mph_cg_s* c, const void* b, void* x, hipStream_t s) {
  MPH_CHECK_STATUS(mph_multiply(c->A, x, c->q, 1, s));
  double* rr_parts = c->parts + c->np;
  hipLaunchKernelGGL((cg_start_kernel<Vec>), dim3((unsigned)c->np), dim3(kT), 0, s,
                     static_cast<const Vec*>(b), static_cast<const Vec*>(c->q),
                     static_cast<Vec*>(c->r), static_cast<Vec*>(c->p), c->n, c->chunk, rr_parts);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename Vec>
int cg_iterate_t(mph_cg_s* c, void* x, int32_t iters, hipStream_t s) {
  double* pq_parts = c->parts;
  double* rr_parts = c->parts + c->np;
  double* state = c->parts + 2 * c->np;
  const unsigned grid_p = (unsigned)std::min<int64_t>(ceil_div(c->n, kT), 8 * 1024);
  for (int32_t k = 0; k < iters; k++) {
    // state[0] = r.r of the current residual (its partials are rr_parts)
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(kT), 0, s, rr_parts, c->np, state);
    MPH_CHECK_STATUS(mph_multiply(c->A, c->p, c->q, 1, s));
    hipLaunchKernelGGL((dot_parts_kernel<Vec>), dim3((unsigned)c->np), dim3(kT), 0, s,
                       static_cast<const Vec*>(c->p), static_cast<const Vec*>(c->q), c->n,
                       c->chunk, pq_parts);
    hipLaunchKernelGGL((cg_update_kernel<Vec>), dim3((unsigned)c->np), dim3(kT), 0, s,
                       static_cast<Vec*>(x), static_cast<Vec*>(c->r),
                       static_cast<const Vec*>(c->p), static_cast<const Vec*>(c->q), c->n,
                       c->chunk, pq_parts, c->np, state, rr_parts);
    hipLaunchKernelGGL((cg_direction_kernel<Vec>), dim3(grid_p), dim3(kT), 0, s,
                       static_cast<Vec*>(c->p), static_cast<const Vec*>(c->r), c->n, rr_parts,
                       c->np, state);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace

extern "C" {

int mph_dot(int64_t n, const void* x, const void* y, int vt, double* result, mph_stream_t stream) {
  if (!result) return invalid("null result");
  MPH_CHECK_STATUS(check_vt(vt));
  if (n < 0) return invalid("negative size");
  if (n > 0 && (!x || !y)) return invalid("null vector");
  if (n == 0) {
    *result = 0;
    return MPH_OK;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  return vector_is_double(vt) ? dot_t<double>(n, x, y, result, s)
                              : dot_t<float>(n, x, y, result, s);
}

int mph_waxpby(int64_t n, double alpha, const void* x, double beta, const void* y, void* w, int vt,
               mph_stream_t stream) {
  MPH_CHECK_STATUS(check_vt(vt));
  if (n < 0) return invalid("negative size");
  if (n == 0) return MPH_OK;
  if (!x || !y || !w) return invalid("null vector");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const unsigned g = (unsigned)ceil_div(n, kT);
  if (vector_is_double(vt))
    hipLaunchKernelGGL((waxpby_kernel<double>), dim3(g), dim3(kT), 0, s, n, alpha,
                       static_cast<const double*>(x), beta, static_cast<const double*>(y),
                       static_cast<double*>(w));
  else
    hipLaunchKernelGGL((waxpby_kernel<float>), dim3(g), dim3(kT), 0, s, n, alpha,
                       static_cast<const float*>(x), beta, static_cast<const float*>(y),
                       static_cast<float*>(w));
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

int mph_reduce(int64_t n, const void* x, int vt, double* result, mph_stream_t stream) {
  if (!result) return invalid("null result");
  MPH_CHECK_STATUS(check_vt(vt));
  if (n < 0) return invalid("negative size");
  if (n > 0 && !x) return invalid("null vector");
  if (n == 0) {
    *result = 0;
    return MPH_OK;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  return vector_is_double(vt) ? reduce_t<double>(n, x, result, s)
                              : reduce_t<float>(n, x, result, s);
}

int mph_cg_create(mph_cg_t* out, mph_matrix_t A) {
  if (!out) return invalid("null output");
  if (!A) return invalid("null matrix handle");
  if (A->nrows != A->ncols) return invalid("CG needs a square matrix");
  mph_cg_s* c = new mph_cg_s();
  c->A = A;
  c->n = A->nrows;
  c->vt = A->vt;
  c->np = parts_for(c->n);
  c->chunk = chunk_for(c->n, c->np);
  const size_t vs = c->vt == MPH_F64 ? 8 : 4;
  const size_t vb = vs * (size_t)std::max<int64_t>(c->n, 1);
  if (hipMalloc(&c->r, vb) != hipSuccess || hipMalloc(&c->p, vb) != hipSuccess ||
      hipMalloc(&c->q, vb) != hipSuccess ||
      hipMalloc(&c->parts, sizeof(double) * (2 * c->np + 2)) != hipSuccess) {
    mph_cg_destroy(c);
    set_error("hipMalloc failed in mph_cg_create");
    return MPH_ERR_OUT_OF_MEMORY;
  }
  *out = c;
  return MPH_OK;
}

int mph_cg_destroy(mph_cg_t c) {
  if (!c) return MPH_OK;
  dfree(c->r);
  dfree(c->p);
  dfree(c->q);
  dfree(c->parts);
  delete c;
  return MPH_OK;
}

int mph_cg_start(mph_cg_t c, const void* b, void* x, mph_stream_t stream) {
  if (!c) return invalid("null CG handle");
  if (c->n > 0 && (!b || !x)) return invalid("null b or x");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MPH_CHECK_STATUS(mph_matrix_analyse(c->A, stream));
  if (c->n > 0)
    MPH_CHECK_STATUS(vector_is_double(c->vt) ? cg_start_t<double>(c, b, x, s)
                                             : cg_start_t<float>(c, b, x, s));
  c->started = true;
  return MPH_OK;
}

int mph_cg_iterate(mph_cg_t c, void* x, int32_t iters, mph_stream_t stream) {
  if (!c) return invalid("null CG handle");
  if (!c->started) return invalid("mph_cg_start must come first");
  if (iters < 0) return invalid("negative iteration count");
  if (c->n == 0 || iters == 0) return MPH_OK;
  if (!x) return invalid("null x");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  return vector_is_double(c->vt) ? cg_iterate_t<double>(c, x, iters, s)
                                 : cg_iterate_t<float>(c, x, iters, s);
}

int mph_cg_residual(mph_cg_t c, double* rr, mph_stream_t stream) {
  if (!c || !rr) return invalid("null argument");
  if (!c->started) return invalid("mph_cg_start must come first");
  if (c->n == 0) {
    *rr = 0;
    return MPH_OK;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double* slot = c->parts + 2 * c->np + 1;
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(kT), 0, s, c->parts + c->np, c->np, slot);
  MPH_HIP_CHECK(hipGetLastError());
  MPH_HIP_CHECK(hipMemcpyAsync(rr, slot, sizeof(double), hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  return MPH_OK;
}

}  // extern "C"
