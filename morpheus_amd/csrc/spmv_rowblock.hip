// spmv_rowblock.hip -- adaptive row-block SpMV / SpMM for CSR-shaped data,
// and the fused single-pass HYB (ELL + COO tail) and HDC (DIA + CSR) forms.
//
// Replaces the reference HIP CSR path (core/src/impl/Csr/HIP/
// Morpheus_Multiply_Impl.hpp:50-162, kernels core/src/impl/Csr/Kernels/
// Morpheus_Multiply_Impl.hpp:36-168) and the two-pass HYB / HDC compositions
// (core/src/impl/Hyb/HIP/Morpheus_Multiply_Impl.hpp:42-53,
//  core/src/impl/Hdc/HIP/Morpheus_Multiply_Impl.hpp:42-53).
//
// Design (gfx950, wave64):
//  * An on-GPU analysis cuts the rows into blocks: a short block holds rows
//    whose items start inside one `chunk`-item window (so it never exceeds
//    2*chunk items) and at most kRowBlockMaxRows rows; a row longer than
//    `chunk` gets ceil(len/chunk) blocks of its own.
//  * Short block: 256 threads stream the block's values and column indices
//    once, coalesced and non-temporal, form the products a*x in LDS, then
//    reduce each row with 1..64 lanes chosen from the block's mean row length
//    (pow2 >= mean), and write y once (no separate zeroing pass).
//  * Long-row chunk: the whole workgroup reduces `chunk` items into a
//    per-block partial with a plain store; a second, small launch adds each
//    long row's partials in chunk order (and its DIA / ELL part), so FP64
//    results are deterministic run to run.  There is no cross-workgroup
//    synchronisation inside a kernel: on gfx950 every agent-scope
//    release / acquire writes back / invalidates the whole L2 of the issuing
//    XCD (the eight XCD L2s are not coherent with each other), which costs
//    far more than the extra launch.
//  * HYB / HDC: the ELL or DIA part of each row is computed by a
//    thread-per-row pass (coalesced, column-major) into LDS and added to the
//    row's CSR / COO-tail sum in the same launch, so y is written once.
//  * FP32 accumulation (FP32 / BF16 storage) uses TwoSum-compensated sums on
//    rows longer than kCompensateLen items.
#include <algorithm>

#include "mph_internal.hpp"

namespace mph {

enum { PRE_NONE = 0, PRE_DIA = 1, PRE_ELL = 2 };

template <typename V>
struct RBArgs {
  int64_t nrows, ncols;
  const int32_t* __restrict__ ro;
  const int32_t* __restrict__ ci;
  const V* __restrict__ val;
  const RowBlock* __restrict__ blocks;
  int32_t chunk;
  int32_t lanes_fixed;
  // DIA prefix
  int32_t ndiag;
  int64_t dpitch;
  const int32_t* __restrict__ doff;
  const V* __restrict__ dval;
  // ELL prefix
  int32_t eloop;
  int64_t epitch;
  const int32_t* __restrict__ eci;
  const V* __restrict__ evalp;
  // vectors: element (i, j) of X at x[i*ld + j]
  const typename vtraits<V>::vec* __restrict__ x;
  typename vtraits<V>::vec* __restrict__ y;
  int32_t ld;
  int init;
  typename vtraits<V>::acc* __restrict__ partial;
  int32_t nblocks;
};

// --------------------------------------------------------------- prefixes
template <typename V, int PRE, int KR>
__device__ __forceinline__ void row_prefix(const RBArgs<V>& a, int64_t r,
                                           typename vtraits<V>::acc* s) {
  using Acc = typename vtraits<V>::acc;
  if constexpr (PRE == PRE_DIA) {
    for (int d = 0; d < a.ndiag; d++) {
      const int64_t col = r + a.doff[d];
      if (col >= 0 && col < a.ncols) {
        const Acc v = vtraits<V>::widen(ld_stream(a.dval + r + a.dpitch * d));
#pragma unroll
        for (int q = 0; q < KR; q++) s[q] += v * (Acc)a.x[col * a.ld + q];
      }
    }
  } else if constexpr (PRE == PRE_ELL) {
    for (int n = 0; n < a.eloop; n++) {
      const int64_t slot = r + a.epitch * n;
      const int32_t c = ld_stream(a.eci + slot);
      if (c != -1) {
        const Acc v = vtraits<V>::widen(ld_stream(a.evalp + slot));
#pragma unroll
        for (int q = 0; q < KR; q++) s[q] += v * (Acc)a.x[(int64_t)c * a.ld + q];
      }
    }
  }
}

__device__ __forceinline__ int choose_lanes(int64_t items, int64_t rows,
                                            int fixed) {
  if (fixed > 0) return fixed;
  const int64_t mean = rows > 0 ? items / rows : 0;
  int L = 1;
  while (L < kWave && L < mean) L <<= 1;
  return L;
}

// block-wide sum over 256 threads; result valid in thread 0
template <typename A>
__device__ __forceinline__ A block_sum(A v, A* s_red) {
  for (int m = kWave >> 1; m > 0; m >>= 1) v += shfl_xor(v, m);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) s_red[wave] = v;
  __syncthreads();
  A t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kRowBlockThreads / kWave; w++) t += s_red[w];
  return t;
}

template <typename A>
__device__ __forceinline__ void block_csum(csum<A>& cs, A* s_red) {
  for (int m = kWave >> 1; m > 0; m >>= 1) {
    const A s2 = shfl_xor(cs.s, m), c2 = shfl_xor(cs.c, m);
    cs.combine(s2, c2);
  }
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) {
    s_red[2 * wave] = cs.s;
    s_red[2 * wave + 1] = cs.c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    csum<A> t{s_red[0], s_red[1]};
    for (int w = 1; w < kRowBlockThreads / kWave; w++)
      t.combine(s_red[2 * w], s_red[2 * w + 1]);
    cs = t;
  }
}

template <typename Vec, typename A>
__device__ __forceinline__ void store_y(Vec* y, int64_t i, A v, int init) {
  y[i] = init ? (Vec)v : (Vec)((A)y[i] + v);
}

// Products a_j * x[c_j] for items j = base + tid + k*256, k < kBatch: all
// index and value loads of the batch are issued before the dependent x
// gathers, so each thread keeps kBatch gathers in flight.
constexpr int kBatch = 8;

template <typename V, typename F>
__device__ __forceinline__ void batched_products(const int32_t* __restrict__ ci,
                                                 const V* __restrict__ val,
                                                 const typename vtraits<V>::vec* __restrict__ x,
                                                 int64_t begin, int64_t end, F&& sink) {
  using Acc = typename vtraits<V>::acc;
  for (int64_t base = begin + threadIdx.x; base < end;
       base += (int64_t)kRowBlockThreads * kBatch) {
    int32_t c[kBatch];
    Acc v[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int64_t i = base + (int64_t)k * kRowBlockThreads;
      c[k] = i < end ? ld_stream(ci + i) : 0;
      v[k] = i < end ? vtraits<V>::widen(ld_stream(val + i)) : (Acc)0;
    }
    Acc xv[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int64_t i = base + (int64_t)k * kRowBlockThreads;
      xv[k] = i < end ? (Acc)x[c[k]] : (Acc)0;
    }
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int64_t i = base + (int64_t)k * kRowBlockThreads;
      if (i < end) sink(i, v[k] * xv[k]);
    }
  }
}

// --------------------------------------------------------------- k = 1
// Long-row chunk: workgroup sum of one chunk of one row into partial[b].
template <typename V, bool COMP>
__device__ void long_chunk_spmv(const RBArgs<V>& a, uint32_t b, RowBlock blk,
                                typename vtraits<V>::acc* s_red) {
  using Acc = typename vtraits<V>::acc;
  const int64_t r = blk.row;
  const int64_t rs = a.ro[r], re = a.ro[r + 1];
  const int64_t s = rs + (int64_t)blk.sub * a.chunk;
  const int64_t e = min(re, s + (int64_t)a.chunk);
  if constexpr (COMP) {
    csum<Acc> cs{0, 0};
    batched_products<V>(a.ci, a.val, a.x, s, e, [&](int64_t, Acc p) { cs.add(p); });
    block_csum(cs, s_red);
    if (threadIdx.x == 0) {
      a.partial[2 * b] = cs.s;
      a.partial[2 * b + 1] = cs.c;
    }
  } else {
    Acc t = 0;
    batched_products<V>(a.ci, a.val, a.x, s, e, [&](int64_t, Acc p) { t += p; });
    t = block_sum(t, s_red);
    if (threadIdx.x == 0) a.partial[b] = t;
  }
}

// Second launch: one thread per long row (the block holding its chunk 0)
// adds the row's partials in chunk order plus its DIA / ELL part.
template <typename V, int PRE, bool COMP>
__global__ void rowblock_combine_kernel(RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nblocks) return;
  const RowBlock blk = a.blocks[b];
  if (blk.sub != 0) return;
  const int64_t r = blk.row;
  const int64_t nsub = ((int64_t)a.ro[r + 1] - a.ro[r] + a.chunk - 1) / a.chunk;
  Acc pre[1] = {0};
  row_prefix<V, PRE, 1>(a, r, pre);
  if constexpr (COMP) {
    csum<Acc> cs{pre[0], 0};
    for (int64_t q = 0; q < nsub; q++)
      cs.combine(a.partial[2 * (b + q)], a.partial[2 * (b + q) + 1]);
    store_y(a.y, r, cs.value(), a.init);
  } else {
    Acc t = 0;
    for (int64_t q = 0; q < nsub; q++) t += a.partial[b + q];
    store_y(a.y, r, t + pre[0], a.init);
  }
}

template <typename V, int PRE, bool COMP>
__global__ __launch_bounds__(kRowBlockThreads) void rowblock_spmv_kernel(
    RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ Acc s_red[2 * kRowBlockThreads / kWave];
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const RowBlock blk = a.blocks[b];
  if (blk.sub >= 0) {
    long_chunk_spmv<V, COMP>(a, b, blk, s_red);
    return;
  }
  Acc* s_prod = reinterpret_cast<Acc*>(smem);
  Acc* s_pre = s_prod + 2 * a.chunk;
  int32_t* s_ro = reinterpret_cast<int32_t*>(s_pre + (PRE ? kRowBlockMaxRows : 0));

  const int64_t row0 = blk.row;
  const int64_t row1 = a.blocks[b + 1].row;
  const int nr = (int)(row1 - row0);
  const int64_t p0 = a.ro[row0];
  const int n = (int)(a.ro[row1] - p0);

  for (int i = threadIdx.x; i <= nr; i += kRowBlockThreads)
    s_ro[i] = (int32_t)(a.ro[row0 + i] - p0);
  batched_products<V>(a.ci + p0, a.val + p0, a.x, 0, n,
                      [&](int64_t i, Acc p) { s_prod[i] = p; });
  if constexpr (PRE != PRE_NONE) {
    for (int i = threadIdx.x; i < nr; i += kRowBlockThreads) {
      Acc p[1] = {0};
      row_prefix<V, PRE, 1>(a, row0 + i, p);
      s_pre[i] = p[0];
    }
  }
  __syncthreads();

  const int L = choose_lanes(n, nr, a.lanes_fixed);
  const int groups = kRowBlockThreads / L;
  const int g = threadIdx.x / L, l = threadIdx.x % L;
  for (int r = g; r < nr; r += groups) {
    const int st = s_ro[r], en = s_ro[r + 1];
    Acc out;
    if (COMP && en - st > kCompensateLen) {
      csum<Acc> cs{0, 0};
      for (int j = st + l; j < en; j += L) cs.add(s_prod[j]);
      for (int m = L >> 1; m > 0; m >>= 1) {
        const Acc s2 = shfl_xor(cs.s, m), c2 = shfl_xor(cs.c, m);
        cs.combine(s2, c2);
      }
      if constexpr (PRE != PRE_NONE) cs.add(s_pre[r]);
      out = cs.value();
    } else {
      Acc t = 0;
      for (int j = st + l; j < en; j += L) t += s_prod[j];
      t = group_sum(t, L);
      if constexpr (PRE != PRE_NONE) t += s_pre[r];
      out = t;
    }
    if (l == 0) store_y(a.y, row0 + r, out, a.init);
  }
}

// --------------------------------------------------------------- k > 1
// SpMM form: KR right-hand sides, row-major X / Y with leading dimension ld.
// Each row is reduced by a group of lanes reading the row's items directly
// (coalesced within the row); every item feeds KR FMAs from one contiguous
// load of the x row.  The DIA / ELL prefix is split across the group's lanes.
template <typename V, int PRE, int KR>
__global__ __launch_bounds__(kRowBlockThreads) void rowblock_spmm_kernel(
    RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[KR * kRowBlockThreads / kWave];
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const RowBlock blk = a.blocks[b];
  if (blk.sub >= 0) {
    const int64_t r = blk.row;
    const int64_t rs = a.ro[r], re = a.ro[r + 1];
    const int64_t s = rs + (int64_t)blk.sub * a.chunk;
    const int64_t e = min(re, s + (int64_t)a.chunk);
    Acc t[KR];
#pragma unroll
    for (int q = 0; q < KR; q++) t[q] = 0;
    for (int64_t i = s + threadIdx.x; i < e; i += kRowBlockThreads) {
      const Acc v = vtraits<V>::widen(ld_stream(a.val + i));
      const int64_t c = ld_stream(a.ci + i);
#pragma unroll
      for (int q = 0; q < KR; q++) t[q] += v * (Acc)a.x[c * a.ld + q];
    }
#pragma unroll
    for (int q = 0; q < KR; q++) {
      for (int m = kWave >> 1; m > 0; m >>= 1) t[q] += shfl_xor(t[q], m);
    }
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < KR; q++) s_red[q * (kRowBlockThreads / kWave) + wave] = t[q];
    __syncthreads();
    if (threadIdx.x < KR) {
      Acc u = 0;
      for (int w = 0; w < kRowBlockThreads / kWave; w++)
        u += s_red[threadIdx.x * (kRowBlockThreads / kWave) + w];
      a.partial[(int64_t)b * KR + threadIdx.x] = u;
    }
    return;
  }
  const int64_t row0 = blk.row;
  const int64_t row1 = a.blocks[b + 1].row;
  const int nr = (int)(row1 - row0);
  const int n = (int)(a.ro[row1] - a.ro[row0]);
  int L = choose_lanes(n, nr, a.lanes_fixed);
  if constexpr (PRE != PRE_NONE) {
    // keep at least a few lanes per row for the prefix part
    if (a.lanes_fixed == 0 && L < 4) L = 4;
  }
  const int groups = kRowBlockThreads / L;
  const int g = threadIdx.x / L, l = threadIdx.x % L;
  for (int r = g; r < nr; r += groups) {
    const int64_t row = row0 + r;
    Acc t[KR];
#pragma unroll
    for (int q = 0; q < KR; q++) t[q] = 0;
    if constexpr (PRE == PRE_DIA) {
      for (int d = l; d < a.ndiag; d += L) {
        const int64_t col = row + a.doff[d];
        if (col >= 0 && col < a.ncols) {
          const Acc v = vtraits<V>::widen(ld_stream(a.dval + row + a.dpitch * d));
#pragma unroll
          for (int q = 0; q < KR; q++) t[q] += v * (Acc)a.x[col * a.ld + q];
        }
      }
    } else if constexpr (PRE == PRE_ELL) {
      for (int nn = l; nn < a.eloop; nn += L) {
        const int64_t slot = row + a.epitch * nn;
        const int32_t c = ld_stream(a.eci + slot);
        if (c != -1) {
          const Acc v = vtraits<V>::widen(ld_stream(a.evalp + slot));
#pragma unroll
          for (int q = 0; q < KR; q++) t[q] += v * (Acc)a.x[(int64_t)c * a.ld + q];
        }
      }
    }
    const int64_t st = a.ro[row], en = a.ro[row + 1];
    for (int64_t j = st + l; j < en; j += L) {
      const Acc v = vtraits<V>::widen(ld_stream(a.val + j));
      const int64_t c = ld_stream(a.ci + j);
#pragma unroll
      for (int q = 0; q < KR; q++) t[q] += v * (Acc)a.x[c * a.ld + q];
    }
#pragma unroll
    for (int q = 0; q < KR; q++) t[q] = group_sum(t[q], L);
    if (l == 0) {
#pragma unroll
      for (int q = 0; q < KR; q++) store_y(a.y, row * a.ld + q, t[q], a.init);
    }
  }
}

template <typename V, int PRE, int KR>
__global__ void rowblock_spmm_combine_kernel(RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nblocks) return;
  const RowBlock blk = a.blocks[b];
  if (blk.sub != 0) return;
  const int64_t r = blk.row;
  const int64_t nsub = ((int64_t)a.ro[r + 1] - a.ro[r] + a.chunk - 1) / a.chunk;
  Acc tot[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) tot[q] = 0;
  row_prefix<V, PRE, KR>(a, r, tot);
  for (int64_t j = 0; j < nsub; j++)
#pragma unroll
    for (int q = 0; q < KR; q++) tot[q] += a.partial[(b + j) * KR + q];
#pragma unroll
  for (int q = 0; q < KR; q++) store_y(a.y, r * a.ld + q, tot[q], a.init);
}

// --------------------------------------------------------------- analysis
__device__ __forceinline__ bool rb_is_start(const int32_t* ro, int64_t r,
                                            int32_t chunk, bool& is_long,
                                            int64_t& len) {
  len = (int64_t)ro[r + 1] - ro[r];
  is_long = len > chunk;
  if (r == 0 || is_long) return true;
  const int64_t plen = (int64_t)ro[r] - ro[r - 1];
  return (r % kRowBlockMaxRows) == 0 || (ro[r] / chunk) != (ro[r - 1] / chunk) ||
         plen > chunk;
}

__global__ void rb_count_kernel(const int32_t* __restrict__ ro, int64_t nrows,
                                int32_t chunk, int32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  bool lg;
  int64_t len;
  const bool st = rb_is_start(ro, r, chunk, lg, len);
  cnt[r] = st ? (lg ? (int32_t)((len + chunk - 1) / chunk) : 1) : 0;
}

__global__ void rb_fill_kernel(const int32_t* __restrict__ ro, int64_t nrows,
                               int32_t chunk, const int32_t* __restrict__ first,
                               RowBlock* __restrict__ blocks) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  // the end sentinel is written whether or not the last row starts a block
  if (r == nrows - 1) blocks[first[nrows]] = RowBlock{(int32_t)nrows, -1};
  const int32_t f = first[r], c = first[r + 1] - f;
  if (c == 0) return;
  bool lg;
  int64_t len;
  rb_is_start(ro, r, chunk, lg, len);
  for (int32_t q = 0; q < c; q++) blocks[f + q] = RowBlock{(int32_t)r, lg ? q : -1};
}

// HYB: row offsets of a row-sorted COO tail.  Entry i is the first of its
// row when its row differs from entry i-1; every row in (row[i-1], row[i]]
// starts at i.
__global__ void tail_offsets_kernel(const int32_t* __restrict__ ri, int64_t nnz,
                                    int64_t nrows, int32_t* __restrict__ ro) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nnz) return;
  const int64_t prev = i == 0 ? -1 : ri[i - 1];
  const int64_t cur = i == nnz ? nrows : ri[i];
  for (int64_t r = prev + 1; r <= cur; r++) ro[r] = (int32_t)i;
}

__global__ void rows_unsorted_kernel(const int32_t* __restrict__ ri, int64_t nnz,
                                     int64_t nrows, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int32_t r = ri[i];
  if (r < 0 || r >= nrows || (i > 0 && ri[i - 1] > r)) atomicOr(flag, 1);
}

// ELL: are all stored slots at or beyond entries_per_row padding (-1)?
__global__ void ell_tail_check_kernel(const int32_t* __restrict__ ci,
                                      int64_t pitch, int64_t nrows, int64_t K,
                                      int64_t width, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = nrows * (width - K);
  if (i >= total) return;
  const int64_t n = K + i / nrows, r = i % nrows;
  if (ci[r + pitch * n] != -1) atomicOr(flag, 1);
}

// --------------------------------------------------------------- host side
namespace {

int default_chunk(int vt) {
  return vt == MPH_F64 ? kDefaultChunkF64 : kDefaultChunkF32;
}

int read_flag(int32_t* dflag, hipStream_t s, int* out) {
  int32_t h = 0;
  MPH_HIP_CHECK(hipMemcpyAsync(&h, dflag, sizeof(int32_t),
                               hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  *out = h;
  return MPH_OK;
}

}  // namespace

int ell_padding_only(const int32_t* ci, int64_t pitch, int64_t nrows, int64_t K,
                     int64_t width, hipStream_t s, int* padding_only) {
  *padding_only = 1;
  if (width <= K || nrows == 0) return MPH_OK;
  int32_t* flag = nullptr;
  MPH_HIP_CHECK(hipMalloc(&flag, sizeof(int32_t)));
  int st = fill_zero(flag, sizeof(int32_t), s);
  if (st == MPH_OK) {
    const int64_t total = nrows * (width - K);
    hipLaunchKernelGGL(ell_tail_check_kernel, dim3((unsigned)ceil_div(total, 256)),
                       dim3(256), 0, s, ci, pitch, nrows, K, width, flag);
    int f = 0;
    st = read_flag(flag, s, &f);
    *padding_only = !f;
  }
  dfree(flag);
  return st;
}

int rows_sorted(const int32_t* ri, int64_t nnz, int64_t nrows, hipStream_t s,
                int* sorted) {
  *sorted = 1;
  if (nnz == 0) return MPH_OK;
  int32_t* flag = nullptr;
  MPH_HIP_CHECK(hipMalloc(&flag, sizeof(int32_t)));
  int st = fill_zero(flag, sizeof(int32_t), s);
  if (st == MPH_OK) {
    hipLaunchKernelGGL(rows_unsorted_kernel, dim3((unsigned)ceil_div(nnz, 256)),
                       dim3(256), 0, s, ri, nnz, nrows, flag);
    int f = 0;
    st = read_flag(flag, s, &f);
    *sorted = !f;
  }
  dfree(flag);
  return st;
}

// Builds the row blocks of A's CSR-shaped part (CSR, HDC's CSR part or the
// HYB tail viewed through offsets built here).
int rowblock_analyse(mph_matrix_impl* A, hipStream_t s) {
  RowBlockPlan& p = A->rb;
  const int64_t nrows = A->nrows;
  const int32_t* ro = A->csr_ro;
  if (A->format == MPH_HYB) {
    int sorted = 1;
    MPH_CHECK_STATUS(rows_sorted(A->coo_ri, A->coo_nnz, nrows, s, &sorted));
    if (!sorted) return MPH_ERR_NOT_IMPLEMENTED;  // caller falls back
    MPH_HIP_CHECK(hipMalloc(&p.tail_ro, sizeof(int32_t) * (nrows + 1)));
    hipLaunchKernelGGL(tail_offsets_kernel,
                       dim3((unsigned)ceil_div(A->coo_nnz + 1, 256)), dim3(256),
                       0, s, A->coo_ri, A->coo_nnz, nrows, p.tail_ro);
    MPH_HIP_CHECK(hipGetLastError());
    ro = p.tail_ro;
  }
  p.chunk = A->tuning.rows_chunk > 0 ? A->tuning.rows_chunk : default_chunk(A->vt);
  if (nrows == 0) {
    p.nblocks = 0;
    return MPH_OK;
  }
  int32_t* cnt = nullptr;
  int32_t* first = nullptr;
  int64_t* tmp = nullptr;
  MPH_HIP_CHECK(hipMalloc(&cnt, sizeof(int32_t) * nrows));
  MPH_HIP_CHECK(hipMalloc(&first, sizeof(int32_t) * (nrows + 1)));
  MPH_HIP_CHECK(hipMalloc(&tmp, sizeof(int64_t) * scan_tmp_elems(nrows)));
  const unsigned g = (unsigned)ceil_div(nrows, 256);
  hipLaunchKernelGGL(rb_count_kernel, dim3(g), dim3(256), 0, s, ro, nrows,
                     p.chunk, cnt);
  MPH_HIP_CHECK(hipGetLastError());
  int st = exclusive_scan_i32(cnt, first, nrows, tmp, s);
  int32_t nb = 0;
  if (st == MPH_OK) {
    MPH_HIP_CHECK(hipMemcpyAsync(&nb, first + nrows, sizeof(int32_t),
                                 hipMemcpyDeviceToHost, s));
    MPH_HIP_CHECK(hipStreamSynchronize(s));
    p.nblocks = nb;
    const int accs = (A->vt == MPH_F64 ? 8 : 4);
    // partials: up to 16 RHS per block (SpMM) or a compensated pair (SpMV)
    MPH_HIP_CHECK(hipMalloc(&p.blocks, sizeof(RowBlock) * (nb + 1)));
    MPH_HIP_CHECK(hipMalloc(&p.partial, (size_t)accs * 16 * (nb + 1)));
    {
      hipLaunchKernelGGL(rb_fill_kernel, dim3(g), dim3(256), 0, s, ro, nrows,
                         p.chunk, first, p.blocks);
      MPH_HIP_CHECK(hipGetLastError());
      MPH_HIP_CHECK(hipStreamSynchronize(s));
      st = validate_rowblock_plan(ro, nrows, p, s);
    }
  }
  dfree(cnt);
  dfree(first);
  dfree(tmp);
  return st;
}

namespace {

template <typename V>
RBArgs<V> make_args(mph_matrix_impl* A, const void* x, void* y, int ld,
                    int init) {
  RBArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  if (A->format == MPH_HYB) {
    a.ro = A->rb.tail_ro;
    a.ci = A->coo_ci;
    a.val = static_cast<const V*>(A->coo_v);
  } else {
    a.ro = A->csr_ro;
    a.ci = A->csr_ci;
    a.val = static_cast<const V*>(A->csr_v);
  }
  a.blocks = A->rb.blocks;
  a.chunk = A->rb.chunk;
  a.lanes_fixed = A->tuning.lanes;
  a.ndiag = (int32_t)A->ndiag;
  a.dpitch = A->dia_pitch;
  a.doff = A->dia_off;
  a.dval = static_cast<const V*>(A->dia_v);
  a.eloop = (int32_t)A->ell_loop;
  a.epitch = A->ell_pitch;
  a.eci = A->ell_ci;
  a.evalp = static_cast<const V*>(A->ell_v);
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.ld = ld;
  a.init = init;
  a.partial = static_cast<typename vtraits<V>::acc*>(A->rb.partial);
  a.nblocks = A->rb.nblocks;
  return a;
}

template <typename V, int PRE>
int launch_spmv(mph_matrix_impl* A, const void* x, void* y, int init,
                hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  RBArgs<V> a = make_args<V>(A, x, y, 1, init);
  const size_t lds = sizeof(Acc) * (2 * (size_t)a.chunk + (PRE ? kRowBlockMaxRows : 0)) +
                     sizeof(int32_t) * (kRowBlockMaxRows + 1);
  constexpr bool comp = std::is_same<Acc, float>::value;
  hipLaunchKernelGGL((rowblock_spmv_kernel<V, PRE, comp>),
                     dim3((unsigned)A->rb.nblocks), dim3(kRowBlockThreads), lds,
                     s, a);
  if (A->rb.nlong > 0)
    hipLaunchKernelGGL((rowblock_combine_kernel<V, PRE, comp>),
                       dim3((unsigned)ceil_div(A->rb.nblocks, 256)), dim3(256), 0, s, a);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, int PRE, int KR>
int launch_spmm_k(mph_matrix_impl* A, const void* x, void* y, int ld, int init,
                  hipStream_t s) {
  RBArgs<V> a = make_args<V>(A, x, y, ld, init);
  hipLaunchKernelGGL((rowblock_spmm_kernel<V, PRE, KR>),
                     dim3((unsigned)A->rb.nblocks), dim3(kRowBlockThreads), 0, s,
                     a);
  if (A->rb.nlong > 0)
    hipLaunchKernelGGL((rowblock_spmm_combine_kernel<V, PRE, KR>),
                       dim3((unsigned)ceil_div(A->rb.nblocks, 256)), dim3(256), 0, s, a);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, int PRE>
int launch_spmm(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                hipStream_t s) {
  using Vec = typename vtraits<V>::vec;
  // split k into column groups of 16, 8, 4, 2, 1 sharing the leading dim k
  int done = 0;
  while (done < k) {
    const int rem = k - done;
    const Vec* xs = static_cast<const Vec*>(x) + done;
    Vec* ys = static_cast<Vec*>(y) + done;
    int st;
    if (rem >= 16) {
      st = launch_spmm_k<V, PRE, 16>(A, xs, ys, k, init, s);
      done += 16;
    } else if (rem >= 8) {
      st = launch_spmm_k<V, PRE, 8>(A, xs, ys, k, init, s);
      done += 8;
    } else if (rem >= 4) {
      st = launch_spmm_k<V, PRE, 4>(A, xs, ys, k, init, s);
      done += 4;
    } else if (rem >= 2) {
      st = launch_spmm_k<V, PRE, 2>(A, xs, ys, k, init, s);
      done += 2;
    } else {
      st = launch_spmm_k<V, PRE, 1>(A, xs, ys, k, init, s);
      done += 1;
    }
    if (st != MPH_OK) return st;
  }
  return MPH_OK;
}

template <typename V>
int dispatch_pre(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s) {
  int pre = A->format == MPH_HDC ? PRE_DIA : A->format == MPH_HYB ? PRE_ELL : PRE_NONE;
  // an HDC without diagonals / a HYB without ELL slots is its CSR-shaped part
  if ((pre == PRE_DIA && A->ndiag == 0) || (pre == PRE_ELL && A->ell_loop == 0)) pre = PRE_NONE;
  if (k == 1) {
    if (pre == PRE_DIA) return launch_spmv<V, PRE_DIA>(A, x, y, init, s);
    if (pre == PRE_ELL) return launch_spmv<V, PRE_ELL>(A, x, y, init, s);
    return launch_spmv<V, PRE_NONE>(A, x, y, init, s);
  }
  if (pre == PRE_DIA) return launch_spmm<V, PRE_DIA>(A, k, x, y, init, s);
  if (pre == PRE_ELL) return launch_spmm<V, PRE_ELL>(A, k, x, y, init, s);
  return launch_spmm<V, PRE_NONE>(A, k, x, y, init, s);
}

}  // namespace

int rowblock_multiply(mph_matrix_impl* A, int k, const void* x, void* y,
                      int init, hipStream_t s) {
  if (A->rb.nblocks == 0) return MPH_OK;
  switch (A->vt) {
    case MPH_F64: return dispatch_pre<double>(A, k, x, y, init, s);
    case MPH_F32: return dispatch_pre<float>(A, k, x, y, init, s);
    case MPH_BF16: return dispatch_pre<bf16>(A, k, x, y, init, s);
  }
  return invalid("unknown value type");
}

}  // namespace mph
