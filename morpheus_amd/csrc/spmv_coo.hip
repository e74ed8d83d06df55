// spmv_coo.hip -- COO SpMV as a wavefront-segmented reduction.
//
// Replaces core/src/impl/Coo/HIP/Morpheus_Multiply_Impl.hpp:51-179 (flat
// kernel core/src/impl/Coo/Kernels/Morpheus_Multiply_Impl.hpp:121-208 and
// the carry reduction core/src/impl/DenseVector/Kernels/
// Morpheus_Segmented_Reduction_Impl.hpp:40-162), which needs a y-zeroing
// launch, a per-call temporary allocation, a one-block carry pass and a
// one-thread tail pass.
//
// Design (row-sorted COO, which the reference HIP kernel also requires):
//  * each wavefront owns a chunk of 64*T consecutive entries; every lane
//    loads T consecutive entries with 16-byte vector loads and reduces them
//    serially, writing rows that start and end inside the lane directly;
//  * one segmented inclusive scan across the 64 lanes (6 shuffle steps per
//    64*T entries) joins the lane boundary segments;
//  * rows complete inside the chunk are written once with '=' (init) or
//    '+=' (accumulate); rows cut by chunk boundaries leave a partial in
//    per-chunk scratch;
//  * a second, small launch (one thread per chunk) adds the partials of
//    every cut row in chunk order -- deterministic, with no cross-workgroup
//    synchronisation inside the main kernel (agent-scope fences write back
//    and invalidate a whole XCD L2 on gfx950);
//  * with init, rows without entries are zero-filled by the lane that sees
//    the gap, so no separate y pass exists (a fill pass is used instead when
//    the analysis finds a gap longer than 256 rows);
//  * FP32 accumulation carries TwoSum-compensated (sum, error) pairs.
// Unsorted COO (valid for the reference Serial backend) falls back to an
// atomic scatter, which is correct but not bitwise reproducible.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "mph_internal.hpp"

namespace mph {

template <typename A, bool C>
struct Sum;

template <typename A>
struct Sum<A, false> {
  A s;
  __device__ static Sum zero() { return Sum{0}; }
  __device__ static Sum of(A v) { return Sum{v}; }
  __device__ void add_prod(A a, A x) { s += a * x; }
  __device__ void merge(const Sum& o) { s += o.s; }
  __device__ A value() const { return s; }
  __device__ Sum up(int d) const { return Sum{shfl_up(s, d)}; }
  __device__ Sum down(int d) const { return Sum{shfl_down(s, d)}; }
  __device__ void store(A* p, int64_t i) const { p[i] = s; }
  __device__ static Sum load(const A* p, int64_t i) { return Sum{p[i]}; }
};

// Dot2 pair (TwoSum on the additions, the FMA TwoProduct on the products;
// Ogita, Rump, Oishi 2005)
template <typename A>
struct Sum<A, true> {
  A s, c;
  __device__ static Sum zero() { return Sum{0, 0}; }
  __device__ static Sum of(A v) { return Sum{v, 0}; }
  __device__ void add_prod(A a, A x) {
#pragma clang fp contract(off)
    A p, ep, t, e;
    two_prod(a, x, p, ep);
    two_sum(s, p, t, e);
    s = t;
    c += e + ep;
  }
  __device__ void merge(const Sum& o) {
#pragma clang fp contract(off)
    A t, e;
    two_sum(s, o.s, t, e);
    s = t;
    c += e + o.c;
  }
  __device__ A value() const {
#pragma clang fp contract(off)
    return s + c;
  }
  __device__ Sum up(int d) const { return Sum{shfl_up(s, d), shfl_up(c, d)}; }
  __device__ Sum down(int d) const { return Sum{shfl_down(s, d), shfl_down(c, d)}; }
  __device__ void store(A* p, int64_t i) const {
    p[2 * i] = s;
    p[2 * i + 1] = c;
  }
  __device__ static Sum load(const A* p, int64_t i) { return Sum{p[2 * i], p[2 * i + 1]}; }
};

template <typename V>
struct CooArgs {
  int64_t nrows, nnz, nchunks;
  const int32_t* __restrict__ ri;
  const int32_t* __restrict__ ci;
  const V* __restrict__ val;
  const typename vtraits<V>::vec* __restrict__ x;
  typename vtraits<V>::vec* __restrict__ y;
  int init;
  int gapfill;
  int vec_ok;  // row / column index arrays are 16-byte aligned
  typename vtraits<V>::acc* __restrict__ first_val;
  typename vtraits<V>::acc* __restrict__ last_val;
  const CooChunk* __restrict__ chunks;  // nullptr: chunk c = [c*64*T, (c+1)*64*T)
  int64_t ld;                           // element stride of x and y (one column of X / Y)
};

template <typename V, int T>
__device__ __forceinline__ void chunk_range(const CooArgs<V>& a, int64_t c, int64_t& cs,
                                            int64_t& ce) {
  if (a.chunks) {
    const CooChunk k = a.chunks[c];
    cs = k.s;
    ce = k.e;
  } else {
    cs = c * (int64_t)(kWave * T);
    ce = min(a.nnz, cs + (int64_t)(kWave * T));
  }
}

// y[r] = v, or y[r] + v (the old y joining the sum before its final rounding)
template <typename Vec, typename S>
__device__ __forceinline__ void put_y(Vec* y, int64_t r, S v, int init) {
  if (!init) v.merge(S::of((decltype(v.s))y[r]));
  y[r] = (Vec)v.value();
}

__device__ __forceinline__ void zero_rows(double* y, int64_t a, int64_t b, int64_t ld) {
  for (int64_t r = a; r < b; r++) y[r * ld] = 0.0;
}
__device__ __forceinline__ void zero_rows(float* y, int64_t a, int64_t b, int64_t ld) {
  for (int64_t r = a; r < b; r++) y[r * ld] = 0.0f;
}

typedef int v4i __attribute__((ext_vector_type(4)));

// load T consecutive int32 starting at p (16-byte aligned when full)
template <int T>
__device__ __forceinline__ void load_idx(const int32_t* p, int n, int32_t* out,
                                         bool vec_ok) {
  if (vec_ok && n == T && (T % 4) == 0) {
#pragma unroll
    for (int k = 0; k < T / 4; k++) {
      const v4i v = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p) + k);
      out[4 * k] = v.x;
      out[4 * k + 1] = v.y;
      out[4 * k + 2] = v.z;
      out[4 * k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < T; k++) out[k] = k < n ? ld_stream(p + k) : 0;
  }
}

template <typename V, int T>
__device__ __forceinline__ void load_val(const V* p, int n,
                                         typename vtraits<V>::acc* out) {
#pragma unroll
  for (int k = 0; k < T; k++) out[k] = k < n ? vtraits<V>::widen(ld_stream(p + k)) : 0;
}

template <typename V, int T, bool COMP>
__device__ void coo_chunk(const CooArgs<V>& a, int64_t c, int lane) {
  using A = typename vtraits<V>::acc;
  using S = Sum<A, COMP>;
  int64_t cs, ce;
  chunk_range<V, T>(a, c, cs, ce);
  const int64_t i0 = cs + (int64_t)lane * T;
  const int n = (int)max<int64_t>(0, min<int64_t>(T, ce - i0));
  const bool active = n > 0;
  // lanes past the end of the chunk are inactive; the last active lane
  const int last_lane = (int)((ce - cs - 1) / T);

  const int32_t first_row = a.ri[cs];
  const int32_t last_row = a.ri[ce - 1];
  const bool starts_before = cs > 0 && a.ri[cs - 1] == first_row;
  const bool continues_after = ce < a.nnz && a.ri[ce] == last_row;

  auto write_seg = [&](int32_t row, const S& v) {
    if (row == first_row && starts_before) {
      v.store(a.first_val, c);
    } else if (row == last_row && continues_after) {
      v.store(a.last_val, c);
    } else {
      put_y(a.y, row * a.ld, v, a.init);
    }
  };

  int32_t rows[T], cols[T];
  A vals[T];
  load_idx<T>(a.ri + i0, n, rows, a.vec_ok);
  load_idx<T>(a.ci + i0, n, cols, a.vec_ok);
  load_val<V, T>(a.val + i0, n, vals);

  // all gathers of the lane are issued before the serial pass consumes them
  A xk[T];
#pragma unroll
  for (int k = 0; k < T; k++) xk[k] = k < n ? (A)a.x[(int64_t)cols[k] * a.ld] : (A)0;

  // serial pass: head segment, complete middle segments, tail segment
  int32_t h = rows[0], cur = rows[0];
  S head = S::zero(), run = S::zero();
  bool single = true;
  const bool fill = a.init && a.gapfill;
#pragma unroll
  for (int k = 0; k < T; k++) {
    if (k < n) {
      if (rows[k] != cur) {
        if (single) {
          head = run;
          single = false;
        } else {
          write_seg(cur, run);  // a middle segment: complete
        }
        if (fill) zero_rows(a.y, (int64_t)cur + 1, rows[k], a.ld);
        cur = rows[k];
        run = S::zero();
      }
      run.add_prod(vals[k], xk[k]);
    }
  }
  const int32_t t = cur;

  // segmented inclusive scan of the lane tails across the wavefront
  const int32_t t_prev = shfl_up(t, 1);
  const bool cont = active && lane > 0 && single && h == t_prev;
  S acc = active ? run : S::zero();
  int flag = cont ? 0 : 1;
  for (int d = 1; d < kWave; d <<= 1) {
    const S o = acc.up(d);
    const int of = shfl_up(flag, d);
    if (lane >= d && !flag) {
      acc.merge(o);
    }
    if (lane >= d) flag |= of;
  }
  // carry into this lane's head from the previous lanes
  const S prev_acc = acc.up(1);
  const int32_t h_next = shfl_down(h, 1);
  if (!active) return;
  if (!single) {
    S hv = head;
    if (lane > 0 && h == t_prev) hv.merge(prev_acc);
    write_seg(h, hv);
  }
  // the tail segment ends at this lane's boundary
  if (lane == last_lane) {
    write_seg(t, acc);
    if (fill && ce == a.nnz) zero_rows(a.y, (int64_t)t + 1, a.nrows, a.ld);
  } else if (h_next != t) {
    write_seg(t, acc);
    if (fill) zero_rows(a.y, (int64_t)t + 1, h_next, a.ld);
  }
  // gap before the chunk's first row
  if (lane == 0 && fill) {
    const int64_t prev_row = cs == 0 ? -1 : a.ri[cs - 1];
    zero_rows(a.y, prev_row + 1, first_row, a.ld);
  }
}

// Second launch: the chunk where a cut row starts adds the partials of the
// chunks that continue it, in chunk order.
template <typename V, int T, bool COMP>
__global__ void coo_fixup_kernel(CooArgs<V> a) {
  using A = typename vtraits<V>::acc;
  using S = Sum<A, COMP>;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nchunks) return;
  int64_t cs, ce;
  chunk_range<V, T>(a, c, cs, ce);
  const int32_t fr = a.ri[cs], lr = a.ri[ce - 1];
  const bool sb = cs > 0 && a.ri[cs - 1] == fr;
  const bool ca = ce < a.nnz && a.ri[ce] == lr;
  if (!ca || (sb && fr == lr)) return;
  S tot = S::load(a.last_val, c);
  for (int64_t cc = c + 1; cc < a.nchunks; cc++) {
    tot.merge(S::load(a.first_val, cc));
    int64_t s2, e2;
    chunk_range<V, T>(a, cc, s2, e2);
    const bool whole = a.ri[s2] == a.ri[e2 - 1];
    const bool more = e2 < a.nnz && a.ri[e2] == lr;
    if (!(whole && more)) break;
  }
  put_y(a.y, lr * a.ld, tot, a.init);
}

template <typename V, int T, bool COMP>
__global__ __launch_bounds__(256) void coo_segmented_kernel(CooArgs<V> a) {
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t c = (int64_t)b * (256 / kWave) + wave;
  if (c < a.nchunks) coo_chunk<V, T, COMP>(a, c, lane);
}

// Unsorted COO: y (cleared beforehand when init) += a*x with atomics.
template <typename V>
__global__ void coo_atomic_kernel(CooArgs<V> a) {
  using A = typename vtraits<V>::acc;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nnz) return;
  const A p = vtraits<V>::widen(a.val[i]) * (A)a.x[(int64_t)a.ci[i] * a.ld];
  atomicAdd(&a.y[(int64_t)a.ri[i] * a.ld], p);
}

template <typename V>
__global__ void coo_zero_kernel(typename vtraits<V>::vec* y, int64_t n, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i * ld] = 0;
}

// max row gap (for the gap-fill decision)
__global__ void coo_gap_kernel(const int32_t* __restrict__ ri, int64_t nnz,
                               int64_t nrows, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nnz) return;
  const int64_t prev = i == 0 ? -1 : ri[i - 1];
  const int64_t cur = i == nnz ? nrows : ri[i];
  const int64_t gap = cur - prev - 1;
  if (gap > 0) atomicMax(out, (int32_t)min<int64_t>(gap, 1 << 30));
}

namespace {
constexpr int kGapFillMax = 256;
}

// Row-sorted COO with long rows and a large x: rows longer than
// kXTileMinRowLen (contiguous, column-sorted runs of entries) take the x-tiled
// pass, which needs only their
// entry ranges; the segmented pass keeps every other entry through a table of
// chunks that skip the tiled rows, and y is cleared first (init) since the
// segmented pass no longer sees every gap between rows.
int coo_xtile_analyse(mph_matrix_impl* A, hipStream_t s) {
  CooPlan& p = A->coo;
  const int64_t nrows = A->nrows, nnz = A->nnz;
  const size_t vec_bytes = A->vt == MPH_F64 ? 8 : 4;
  const int mode = xtile_mode(A);
  const bool on = mode == 2 || (mode == 0 && (int64_t)A->ncols * (int64_t)vec_bytes >= kXTileMinXBytes);
  if (!on || nnz == 0 || nrows == 0 || A->ncols == 0) return MPH_OK;
  int32_t* d_ro = nullptr;
  MPH_HIP_CHECK(hipMalloc(&d_ro, sizeof(int32_t) * (nrows + 1)));
  std::vector<int32_t> ro(nrows + 1);
  int st = coo_row_offsets(A->coo_ri, nnz, nrows, d_ro, s);
  if (st == MPH_OK &&
      (hipMemcpyAsync(ro.data(), d_ro, sizeof(int32_t) * (nrows + 1), hipMemcpyDeviceToHost, s) !=
           hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)) {
    set_error("COO row offsets copy failed");
    st = MPH_ERR_HIP;
  }
  std::vector<int32_t> lrows;
  std::vector<char> tiled;
  if (st == MPH_OK) {
    if (ro[0] != 0 || ro[nrows] != nnz) {
      set_error("COO analysis check failed: row offsets");
      st = MPH_ERR_RUNTIME;
    }
    for (int64_t r = 0; r < nrows && st == MPH_OK; r++) {
      if (ro[r + 1] < ro[r]) {
        set_error("COO analysis check failed: row offsets not monotone");
        st = MPH_ERR_RUNTIME;
      } else if ((int64_t)ro[r + 1] - ro[r] > kXTileMinRowLen) {
        lrows.push_back((int32_t)r);
      }
    }
  }
  if (st == MPH_OK && !lrows.empty())
    st = xtile_build(A, ro, d_ro, A->coo_ci, A->coo_v, lrows, p.xt, tiled, s);
  dfree(d_ro);
  if (st != MPH_OK || p.xt.rows == 0) return st;

  // chunks of the entries outside the tiled rows
  const int64_t CH = (int64_t)kWave * p.items;
  std::vector<CooChunk> ch;
  auto cut = [&](int64_t a, int64_t b) {
    for (int64_t c = a; c < b; c += CH) ch.push_back(CooChunk{(int32_t)c, (int32_t)std::min(c + CH, b)});
  };
  int64_t start = 0;
  for (int64_t r = 0; r < nrows; r++) {
    if (tiled[r]) {
      cut(start, ro[r]);
      start = ro[r + 1];
    }
  }
  cut(start, nnz);
  p.nchunks = (int64_t)ch.size();
  p.nblocks = (int32_t)ceil_div(p.nchunks, (int64_t)(256 / kWave));
  p.prezero = true;
  if (!ch.empty()) {
    MPH_HIP_CHECK(hipMalloc(&p.chunks, sizeof(CooChunk) * ch.size()));
    MPH_HIP_CHECK(hipMemcpyAsync(p.chunks, ch.data(), sizeof(CooChunk) * ch.size(),
                                 hipMemcpyHostToDevice, s));
    MPH_HIP_CHECK(hipStreamSynchronize(s));
  }
  return MPH_OK;
}

int coo_analyse(mph_matrix_impl* A, hipStream_t s) {
  CooPlan& p = A->coo;
  p.items = A->tuning.coo_items > 0 ? A->tuning.coo_items : 8;
  if (p.items != 4 && p.items != 8 && p.items != 16)
    return invalid("coo_items must be 4, 8 or 16");
  int sorted = 1;
  MPH_CHECK_STATUS(rows_sorted(A->coo_ri, A->nnz, A->nrows, s, &sorted));
  p.sorted = sorted;
  // row-sorted and no segmented shape asked for: the row-block plan over
  // row offsets derived from the row indices (no copy of the entries)
  if (sorted && A->tuning.coo_items == 0) {
    p.rowplan = true;
    return rowblock_analyse(A, s);
  }
  p.nchunks = ceil_div(A->nnz, (int64_t)kWave * p.items);
  p.nblocks = (int32_t)ceil_div(p.nchunks, 256 / kWave);
  p.prezero = false;
  if (A->nnz > 0 && sorted) {
    int32_t* g = nullptr;
    MPH_HIP_CHECK(hipMalloc(&g, sizeof(int32_t)));
    int st = fill_zero(g, sizeof(int32_t), s);
    if (st == MPH_OK) {
      hipLaunchKernelGGL(coo_gap_kernel, dim3((unsigned)ceil_div(A->nnz + 1, 256)),
                         dim3(256), 0, s, A->coo_ri, A->nnz, A->nrows, g);
      int32_t h = 0;
      MPH_HIP_CHECK(hipMemcpyAsync(&h, g, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      MPH_HIP_CHECK(hipStreamSynchronize(s));
      p.prezero = h > kGapFillMax;
    }
    dfree(g);
    if (st != MPH_OK) return st;
  }
  if (sorted) MPH_CHECK_STATUS(coo_xtile_analyse(A, s));
  const size_t acc = A->vt == MPH_F64 ? 8 : 4;
  const size_t pair = A->vt == MPH_F64 ? 1 : 2;  // compensated pairs in FP32
  const size_t n = (size_t)std::max<int64_t>(p.nchunks, 1);
  MPH_HIP_CHECK(hipMalloc(&p.first_val, acc * pair * n));
  MPH_HIP_CHECK(hipMalloc(&p.last_val, acc * pair * n));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  return MPH_OK;
}

namespace {

template <typename V>
CooArgs<V> coo_args(mph_matrix_impl* A, const void* x, void* y, int init, int64_t ld) {
  CooArgs<V> a{};
  a.nrows = A->nrows;
  a.nnz = A->nnz;
  a.nchunks = A->coo.nchunks;
  a.ri = A->coo_ri;
  a.ci = A->coo_ci;
  a.val = static_cast<const V*>(A->coo_v);
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.init = init;
  a.gapfill = !A->coo.prezero;
  a.vec_ok = (reinterpret_cast<uintptr_t>(A->coo_ri) % 16 == 0) &&
             (reinterpret_cast<uintptr_t>(A->coo_ci) % 16 == 0);
  a.first_val = static_cast<typename vtraits<V>::acc*>(A->coo.first_val);
  a.last_val = static_cast<typename vtraits<V>::acc*>(A->coo.last_val);
  a.chunks = A->coo.chunks;
  a.ld = ld;
  return a;
}

template <typename V>
int coo_run(mph_matrix_impl* A, const void* x, void* y, int init, int64_t ld,
            hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  constexpr bool comp = std::is_same<Acc, float>::value;
  CooArgs<V> a = coo_args<V>(A, x, y, init, ld);
  const CooPlan& p = A->coo;
  if (init && (p.prezero || !p.sorted || A->nnz == 0)) {
    if (A->nrows > 0)
      hipLaunchKernelGGL(coo_zero_kernel<V>, dim3((unsigned)ceil_div(A->nrows, 256)),
                         dim3(256), 0, s, a.y, A->nrows, ld);
    MPH_HIP_CHECK(hipGetLastError());
  }
  if (A->nnz == 0) return MPH_OK;
  if (!p.sorted) {
    hipLaunchKernelGGL(coo_atomic_kernel<V>, dim3((unsigned)ceil_div(A->nnz, 256)),
                       dim3(256), 0, s, a);
    MPH_HIP_CHECK(hipGetLastError());
    return MPH_OK;
  }
  if (p.xt.rows > 0) MPH_CHECK_STATUS(xtile_multiply(A, p.xt, x, y, init, ld, s));
  if (p.nchunks == 0) return MPH_OK;
  const dim3 grid((unsigned)p.nblocks), block(256);
  const dim3 fgrid((unsigned)ceil_div(p.nchunks, 256));
  switch (p.items) {
    case 4:
      hipLaunchKernelGGL((coo_segmented_kernel<V, 4, comp>), grid, block, 0, s, a);
      hipLaunchKernelGGL((coo_fixup_kernel<V, 4, comp>), fgrid, block, 0, s, a);
      break;
    case 16:
      hipLaunchKernelGGL((coo_segmented_kernel<V, 16, comp>), grid, block, 0, s, a);
      hipLaunchKernelGGL((coo_fixup_kernel<V, 16, comp>), fgrid, block, 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((coo_segmented_kernel<V, 8, comp>), grid, block, 0, s, a);
      hipLaunchKernelGGL((coo_fixup_kernel<V, 8, comp>), fgrid, block, 0, s, a);
      break;
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace

// ---------------------------------------------------------- no allocation
// The scratch of mph_coo_multiply_noalloc: the per-chunk carries (first and
// last row partials, pairs in FP32) of the segmented kernel, in one
// file-scope device array per GPU.  16 bytes per chunk of 64 * 8 entries in
// every value type, so kNoAllocScratchBytes bounds the entries to
// kNoAllocScratchBytes / 16 * 512 (268,435,456 for 8 MiB).
constexpr size_t kNoAllocScratchBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocItems = 8;
constexpr int kMaxDevices = 64;
// device address of `scratch` on each GPU (looked up once per device: each
// GPU holds its own copy of the array at its own address)
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Looks up (once per device, with that device current) and returns the
// address of this device's copy of `scratch`.  Called by mph_initialize and,
// for a device first used without it, by the entry point's first call there.
int noalloc_scratch(void** out) {
  int dev = 0;
  MPH_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return runtime_error("device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    MPH_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return runtime_error("no device address for the no-allocation scratch");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return MPH_OK;
}

template <typename V>
int coo_noalloc_run(int64_t nrows, int64_t nnz, const int32_t* ri, const int32_t* ci,
                    const void* v, const void* x, void* y, int init, void* scratch_ptr,
                    hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  constexpr bool comp = std::is_same<Acc, float>::value;
  constexpr int T = kNoAllocItems;
  CooArgs<V> a{};
  a.nrows = nrows;
  a.nnz = nnz;
  a.nchunks = ceil_div(nnz, (int64_t)kWave * T);
  a.ri = ri;
  a.ci = ci;
  a.val = static_cast<const V*>(v);
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.init = init;
  a.gapfill = 0;  // y is cleared by a fill pass instead
  a.vec_ok = (reinterpret_cast<uintptr_t>(ri) % 16 == 0) && (reinterpret_cast<uintptr_t>(ci) % 16 == 0);
  // scratch: first partials of every chunk, then the last partials
  const size_t half = (size_t)a.nchunks * (comp ? 2 : 1);
  a.first_val = static_cast<Acc*>(scratch_ptr);
  a.last_val = a.first_val + half;
  a.chunks = nullptr;
  a.ld = 1;
  if (init && nrows > 0)
    hipLaunchKernelGGL(coo_zero_kernel<V>, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, s,
                       a.y, nrows, (int64_t)1);
  if (nnz > 0) {
    const dim3 block(256);
    hipLaunchKernelGGL((coo_segmented_kernel<V, T, comp>),
                       dim3((unsigned)ceil_div(a.nchunks, 256 / kWave)), block, 0, s, a);
    hipLaunchKernelGGL((coo_fixup_kernel<V, T, comp>), dim3((unsigned)ceil_div(a.nchunks, 256)),
                       block, 0, s, a);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

// k right-hand sides (row-major X / Y): one segmented pass per column, each
// reading x and y with element stride k.
int coo_multiply(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s) {
  if (A->coo.rowplan) return A->nrows == 0 ? MPH_OK : rowblock_multiply(A, k, x, y, init, s);
  const size_t vec = A->vt == MPH_F64 ? 8 : 4;
  for (int j = 0; j < k; j++) {
    const void* xj = static_cast<const char*>(x) + vec * j;
    void* yj = static_cast<char*>(y) + vec * j;
    int st;
    switch (A->vt) {
      case MPH_F64: st = coo_run<double>(A, xj, yj, init, k, s); break;
      case MPH_F32: st = coo_run<float>(A, xj, yj, init, k, s); break;
      case MPH_BF16: st = coo_run<bf16>(A, xj, yj, init, k, s); break;
      default: return invalid("unknown value type");
    }
    if (st != MPH_OK) return st;
  }
  return MPH_OK;
}

}  // namespace mph

extern "C" int mph_coo_multiply_noalloc(int64_t nrows, int64_t ncols, int64_t nnz,
                                        const int32_t* ri, const int32_t* ci, const void* v,
                                        int vt, const void* x, void* y, int init,
                                        mph_stream_t stream) {
  using namespace mph;
  if (nrows < 0 || ncols < 0 || nnz < 0) return invalid("negative size");
  if (vt != MPH_F64 && vt != MPH_F32 && vt != MPH_BF16) return invalid("unknown value type");
  const int64_t max_nnz = (int64_t)(kNoAllocScratchBytes / 16) * kWave * kNoAllocItems;
  if (nnz > max_nnz)
    return invalid("mph_coo_multiply_noalloc: " + std::to_string(nnz) +
                   " entries exceed its scratch limit of " + std::to_string(max_nnz));
  if (nnz > INT32_MAX || nrows > INT32_MAX) return invalid("sizes beyond int32 indices");
  if ((nnz > 0 && (!ri || !ci || !v)) || (nnz > 0 && !x) || (nrows > 0 && !y))
    return invalid("null array");
  void* sp = nullptr;
  MPH_CHECK_STATUS(noalloc_scratch(&sp));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (vt) {
    case MPH_F64: return coo_noalloc_run<double>(nrows, nnz, ri, ci, v, x, y, init ? 1 : 0, sp, s);
    case MPH_F32: return coo_noalloc_run<float>(nrows, nnz, ri, ci, v, x, y, init ? 1 : 0, sp, s);
    default: return coo_noalloc_run<bf16>(nrows, nnz, ri, ci, v, x, y, init ? 1 : 0, sp, s);
  }
}
