// autotune.cpp -- timed-trial autotuning, the GPU counterpart of switching a
// DynamicMatrix's active format (core/src/Morpheus_DynamicMatrix.hpp:409-442,
// the --morpheus-format flag of core/src/impl/Morpheus_Core.cpp:92-140):
//  * mph_matrix_autotune: launch shape of one handle (lanes per row, row
//    block budget, COO entries per lane, DIA/ELL rows per thread);
//  * mph_autotune_format: which of the six formats runs a given CSR matrix
//    fastest with k right-hand sides, converting on the GPU and timing each
//    admissible candidate with HIP events.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "mph_internal.hpp"

namespace mph {
namespace {

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// median time of `reps` products (after one untimed product)
int time_product(mph_matrix_t A, int k, const void* x, void* y, int reps, hipStream_t s,
                 float* ms) {
  Events ev;
  MPH_HIP_CHECK(hipEventCreate(&ev.a));
  MPH_HIP_CHECK(hipEventCreate(&ev.b));
  MPH_CHECK_STATUS(mph_multiply_multi(A, k, x, y, 1, reinterpret_cast<mph_stream_t>(s)));
  std::vector<float> t;
  for (int r = 0; r < std::max(reps, 1); r++) {
    MPH_HIP_CHECK(hipEventRecord(ev.a, s));
    MPH_CHECK_STATUS(mph_multiply_multi(A, k, x, y, 1, reinterpret_cast<mph_stream_t>(s)));
    MPH_HIP_CHECK(hipEventRecord(ev.b, s));
    MPH_HIP_CHECK(hipEventSynchronize(ev.b));
    float v = 0;
    MPH_HIP_CHECK(hipEventElapsedTime(&v, ev.a, ev.b));
    t.push_back(v);
  }
  std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
  *ms = t[t.size() / 2];
  return MPH_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { dfree(p); }
  int alloc(size_t bytes) {
    MPH_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    return MPH_OK;
  }
};

__global__ void fill_pattern_kernel(float* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (float)((i * 2654435761u) % 1000u) * 1e-3f - 0.5f;
}
__global__ void fill_pattern_kernel_d(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (double)((i * 2654435761u) % 1000u) * 1e-3 - 0.5;
}

int fill_vector(void* p, int64_t n, int vt, hipStream_t s) {
  if (n <= 0) return MPH_OK;
  const unsigned g = (unsigned)ceil_div(n, 256);
  if (vt == MPH_F64)
    hipLaunchKernelGGL(fill_pattern_kernel_d, dim3(g), dim3(256), 0, s, (double*)p, n);
  else
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(g), dim3(256), 0, s, (float*)p, n);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace
}  // namespace mph

using namespace mph;

extern "C" {

int mph_matrix_autotune(mph_matrix_t A, int32_t k, const void* x, void* y, int32_t reps,
                        mph_stream_t stream, float* best_ms) {
  if (!A) return invalid("null matrix handle");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::vector<mph_tuning> cands;
  switch (A->format) {
    case MPH_COO:
      // the segmented reduction, and the row-block plan (row-sorted only)
      for (int it : {4, 8, 16}) cands.push_back(mph_tuning{0, 0, it, 0});
      for (int chunk : {0, 512})
        for (int lanes : {0, 4, 16}) cands.push_back(mph_tuning{chunk, lanes, 0, 0});
      break;
    case MPH_DIA:
    case MPH_ELL:
      for (int rpt : {0, 1, 2}) cands.push_back(mph_tuning{0, 0, 0, rpt});
      break;
    default: {
      const int maxc = A->vt == MPH_F64 ? 2048 : 4096;
      for (int chunk : {0, 512, maxc})
        for (int lanes : {0, 1, 4, 16, 64}) cands.push_back(mph_tuning{chunk, lanes, 0, 0});
    }
  }
  float best = FLT_MAX;
  mph_tuning bt = A->tuning;
  for (const mph_tuning& t : cands) {
    MPH_CHECK_STATUS(mph_matrix_set_tuning(A, &t));
    float ms = 0;
    const int st = time_product(A, k, x, y, reps, s, &ms);
    if (st != MPH_OK) return st;
    if (ms < best) {
      best = ms;
      bt = t;
    }
  }
  MPH_CHECK_STATUS(mph_matrix_set_tuning(A, &bt));
  MPH_CHECK_STATUS(mph_matrix_analyse(A, stream));
  if (best_ms) *best_ms = best;
  return MPH_OK;
}

int mph_autotune_format(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* ro,
                        const int32_t* ci, const void* v, int vt, int32_t k, int32_t reps,
                        mph_stream_t stream, int* best_format, float* ms_out) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!best_format) return invalid("null best_format");
  if (k < 1) return invalid("k must be >= 1");
  const size_t vs = value_size(vt), xs = vector_size(vt);
  float ms[MPH_NFORMATS];
  for (float& m : ms) m = FLT_MAX;
  DevBuf x, y, cri, cci, cv;
  MPH_CHECK_STATUS(x.alloc(xs * ncols * k));
  MPH_CHECK_STATUS(y.alloc(xs * nrows * k));
  MPH_CHECK_STATUS(fill_vector(x.p, ncols * k, vt, s));
  // the COO hub every conversion goes through
  MPH_CHECK_STATUS(cri.alloc(sizeof(int32_t) * nnz));
  MPH_CHECK_STATUS(cci.alloc(sizeof(int32_t) * nnz));
  MPH_CHECK_STATUS(cv.alloc(vs * nnz));
  MPH_CHECK_STATUS(mph_convert_csr_to_coo(nrows, nnz, ro, ci, v, vt, (int32_t*)cri.p,
                                          (int32_t*)cci.p, cv.p, stream));
  const int32_t* Ri = (const int32_t*)cri.p;
  const int32_t* Ci = (const int32_t*)cci.p;
  for (int f = 0; f < MPH_NFORMATS; f++) {
    DevBuf a0, a1, a2, a3, a4, a5;
    mph_matrix_t A = nullptr;
    int st = MPH_OK;
    const int64_t pitch = pad_size(nrows, 32);
    if (f == MPH_CSR) {
      st = mph_matrix_create_csr(&A, nrows, ncols, nnz, ro, ci, v, vt);
    } else if (f == MPH_COO) {
      st = mph_matrix_create_coo(&A, nrows, ncols, nnz, Ri, Ci, cv.p, vt);
    } else if (f == MPH_DIA) {
      int64_t nd = 0;
      st = mph_convert_coo_to_dia_size(nrows, ncols, nnz, Ri, Ci, &nd, stream);
      if (st == MPH_OK) st = a0.alloc(sizeof(int32_t) * nd);
      if (st == MPH_OK) st = a1.alloc(vs * pitch * nd);
      if (st == MPH_OK)
        st = mph_convert_coo_to_dia(nrows, ncols, nnz, Ri, Ci, cv.p, vt, nd, pitch,
                                    (int32_t*)a0.p, a1.p, stream);
      if (st == MPH_OK)
        st = mph_matrix_create_dia(&A, nrows, ncols, nnz, nd, pitch, (int32_t*)a0.p, a1.p, vt);
    } else if (f == MPH_ELL) {
      int64_t K = 0;
      st = mph_convert_coo_to_ell_size(nrows, nnz, Ri, &K, stream);
      const int64_t w = pad_size(K, 32);
      if (st == MPH_OK) st = a0.alloc(sizeof(int32_t) * w * nrows);
      if (st == MPH_OK) st = a1.alloc(vs * w * nrows);
      if (st == MPH_OK)
        st = mph_convert_coo_to_ell(nrows, nnz, Ri, Ci, cv.p, vt, K, w, nrows, (int32_t*)a0.p,
                                    a1.p, stream);
      if (st == MPH_OK)
        st = mph_matrix_create_ell(&A, nrows, ncols, nnz, K, w, nrows, (int32_t*)a0.p, a1.p, vt);
    } else if (f == MPH_HYB) {
      int64_t K = 0, en = 0, cn = 0;
      st = mph_convert_coo_to_hyb_size(nrows, nnz, Ri, &K, &en, &cn, stream);
      const int64_t w = pad_size(K, 32);
      if (st == MPH_OK) st = a0.alloc(sizeof(int32_t) * w * nrows);
      if (st == MPH_OK) st = a1.alloc(vs * w * nrows);
      if (st == MPH_OK) st = a2.alloc(sizeof(int32_t) * cn);
      if (st == MPH_OK) st = a3.alloc(sizeof(int32_t) * cn);
      if (st == MPH_OK) st = a4.alloc(vs * cn);
      if (st == MPH_OK)
        st = mph_convert_coo_to_hyb(nrows, nnz, Ri, Ci, cv.p, vt, K, w, nrows, (int32_t*)a0.p,
                                    a1.p, (int32_t*)a2.p, (int32_t*)a3.p, a4.p, stream);
      if (st == MPH_OK)
        st = mph_matrix_create_hyb(&A, nrows, ncols, en, K, w, nrows, (int32_t*)a0.p, a1.p, cn,
                                   (int32_t*)a2.p, (int32_t*)a3.p, a4.p, vt);
    } else {
      int64_t nd = 0, dn = 0, cn = 0;
      st = mph_convert_coo_to_hdc_size(nrows, ncols, nnz, Ri, Ci, &nd, &dn, &cn, stream);
      if (st == MPH_OK) st = a0.alloc(sizeof(int32_t) * nd);
      if (st == MPH_OK) st = a1.alloc(vs * pitch * nd);
      if (st == MPH_OK) st = a2.alloc(sizeof(int32_t) * (nrows + 1));
      if (st == MPH_OK) st = a3.alloc(sizeof(int32_t) * cn);
      if (st == MPH_OK) st = a4.alloc(vs * cn);
      if (st == MPH_OK)
        st = mph_convert_coo_to_hdc(nrows, ncols, nnz, Ri, Ci, cv.p, vt, nd, pitch,
                                    (int32_t*)a0.p, a1.p, (int32_t*)a2.p, (int32_t*)a3.p, a4.p,
                                    stream);
      if (st == MPH_OK)
        st = mph_matrix_create_hdc(&A, nrows, ncols, dn, nd, pitch, (int32_t*)a0.p, a1.p, cn,
                                   (int32_t*)a2.p, (int32_t*)a3.p, a4.p, vt);
    }
    if (st == MPH_ERR_FORMAT_CONVERSION || st == MPH_ERR_NOT_IMPLEMENTED) {
      mph_matrix_destroy(A);
      continue;  // inadmissible for this matrix, as in the reference
    }
    if (st == MPH_OK) st = time_product(A, k, x.p, y.p, reps, s, &ms[f]);
    if (st == MPH_ERR_NOT_IMPLEMENTED) st = MPH_OK;
    (void)hipStreamSynchronize(s);
    mph_matrix_destroy(A);
    if (st != MPH_OK) return st;
  }
  int best = MPH_CSR;
  for (int f = 0; f < MPH_NFORMATS; f++)
    if (ms[f] < ms[best]) best = f;
  *best_format = best;
  if (ms_out)
    for (int f = 0; f < MPH_NFORMATS; f++) ms_out[f] = ms[f] == FLT_MAX ? -1.0f : ms[f];
  return MPH_OK;
}

}  // extern "C"
