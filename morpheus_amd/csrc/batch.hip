// batch.hip -- batched SpMM: many matrices of uneven size and structure,
// each with k right-hand sides, in ONE persistent launch that walks a work
// queue of (matrix, row block) tiles; replayable through a hipGraph.
//
// There is no counterpart in the reference (its multiply takes one matrix and
// one vector, core/src/Morpheus_Multiply.hpp:57-78); this is the GPU form of
// running Morpheus::multiply over a set of DynamicMatrix objects.
//
// Design (gfx950):
//  * tiles: a short CSR row block (rows whose items start in one window of
//    the matrix's row-block analysis, optionally with a fused DIA / ELL part
//    for HDC / HYB), a whole long row (one workgroup, no partials), or a
//    256-row slab of a DIA or ELL matrix;
//  * tiles, sorted largest first on the host, are dealt into eight
//    sub-queues (one per XCD); a workgroup takes `grab` tiles per counter
//    update from its home sub-queue, then from the others, skipping drained
//    ones with a relaxed load -- the queue balances the measured time of
//    mixed batches (tiles of different formats cost differently per entry);
//  * the alternative static schedule (MPH_BATCH_SCHEDULE=1) packs the tiles
//    into one list per workgroup by longest-processing-time bin packing on an
//    estimated cost and needs no atomics: it is faster for uniform batches,
//    slower for the mixed benchmark batch;
//  * the counters are cleared by a small kernel before the main kernel, so
//    both forms stay capturable (no memset nodes);
//  * X / Y are row-major with k columns: every matrix entry feeds k FMAs from
//    one contiguous x row (16-byte vector loads when k*sizeof(x) allows).
#include <algorithm>
#include <cstdlib>
#include <queue>
#include <vector>

#include "mph_internal.hpp"

namespace mph {

enum { TILE_SHORT = 0, TILE_WHOLEROW = 1, TILE_DIA = 2, TILE_ELL = 3 };
enum { BPRE_NONE = 0, BPRE_DIA = 1, BPRE_ELL = 2 };

struct BatchMat {
  int64_t nrows, ncols;
  const int32_t* ro;
  const int32_t* ci;
  const void* val;
  int32_t pre;
  int32_t ndiag;
  int64_t dpitch;
  const int32_t* doff;
  const void* dval;
  int32_t eloop;
  int64_t epitch;
  const int32_t* eci;
  const void* evalp;
  const void* x;
  void* y;
};

struct BatchTile {
  int32_t mat;
  int32_t kind;
  int32_t r0;
  int32_t r1;
};

constexpr int kBT = 256;
constexpr int kSlab = 256;
constexpr int kQueues = 8;  // one sub-queue per XCD
constexpr int kGrab = 4;    // tiles taken per counter update

template <typename V, int KR>
struct XRow {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  // acc[q] += v * x[c*KR + q]
  __device__ __forceinline__ static void fma(Acc* acc, Acc v, const Vec* __restrict__ x,
                                             int64_t c) {
    const Vec* p = x + c * KR;
    if constexpr ((KR * sizeof(Vec)) % 16 == 0) {
      typedef Vec v16 __attribute__((ext_vector_type(16 / sizeof(Vec))));
      constexpr int W = 16 / sizeof(Vec);
#pragma unroll
      for (int q = 0; q < KR; q += W) {
        const v16 xv = *reinterpret_cast<const v16*>(p + q);
#pragma unroll
        for (int w = 0; w < W; w++) acc[q + w] += v * (Acc)xv[w];
      }
    } else {
#pragma unroll
      for (int q = 0; q < KR; q++) acc[q] += v * (Acc)p[q];
    }
  }
};

template <typename V, int KR>
__device__ __forceinline__ void store_row(const BatchMat& m, int64_t r,
                                          const typename vtraits<V>::acc* acc, int init) {
  using Vec = typename vtraits<V>::vec;
  using Acc = typename vtraits<V>::acc;
  Vec* y = static_cast<Vec*>(m.y) + r * KR;
#pragma unroll
  for (int q = 0; q < KR; q++) y[q] = init ? (Vec)acc[q] : (Vec)((Acc)y[q] + acc[q]);
}

// Up to kU independent (value, column) pairs are loaded before their x rows
// are gathered and consumed, so a thread keeps several memory round trips in
// flight instead of one per diagonal / slot / item.
// kU shrinks as k grows so that k accumulators plus kU x rows in flight stay
// within a register budget that keeps 8 waves per SIMD
template <int KR>
struct Unroll {
  static constexpr int value = KR >= 8 ? 2 : (KR >= 4 ? 4 : 8);
};

// (values arrive in their stored type; BF16 is widened here, after its
// loads have all been issued)
template <typename V, int KR>
__device__ __forceinline__ void consume(typename vtraits<V>::acc* acc,
                                        V (&v)[Unroll<KR>::value], const int64_t* c,
                                        const typename vtraits<V>::vec* __restrict__ x) {
  constexpr int kU = Unroll<KR>::value;
  hold_raw(v);
#pragma unroll
  for (int u = 0; u < kU; u++)
    if (c[u] >= 0) XRow<V, KR>::fma(acc, vtraits<V>::widen(v[u]), x, c[u]);
}

// DIA part of row r over diagonals d = d0, d0 + step, ...
template <typename V, int KR>
__device__ __forceinline__ void dia_row(const BatchMat& m, int64_t r, int d0, int step,
                                        typename vtraits<V>::acc* acc) {
  constexpr int kU = Unroll<KR>::value;
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  const V* dv = static_cast<const V*>(m.dval);
  const Vec* x = static_cast<const Vec*>(m.x);
  for (int d = d0; d < m.ndiag; d += kU * step) {
    V v[kU];
    int64_t c[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int dd = d + u * step;
      c[u] = -1;
      v[u] = V{};
      if (dd < m.ndiag) {
        const int64_t col = r + m.doff[dd];
        if (col >= 0 && col < m.ncols) {
          c[u] = col;
          v[u] = ld_stream(dv + r + m.dpitch * dd);
        }
      }
    }
    consume<V, KR>(acc, v, c, x);
  }
}

// ELL part of row r over slots n = n0, n0 + step, ...
template <typename V, int KR>
__device__ __forceinline__ void ell_row(const BatchMat& m, int64_t r, int n0, int step,
                                        typename vtraits<V>::acc* acc) {
  constexpr int kU = Unroll<KR>::value;
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  const V* ev = static_cast<const V*>(m.evalp);
  const Vec* x = static_cast<const Vec*>(m.x);
  for (int n = n0; n < m.eloop; n += kU * step) {
    int64_t c[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int nn = n + u * step;
      c[u] = nn < m.eloop ? (int64_t)ld_stream(m.eci + r + m.epitch * nn) : -1;
    }
    V v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int nn = n + u * step;
      v[u] = c[u] >= 0 ? ld_stream(ev + r + m.epitch * nn) : V{};
    }
    consume<V, KR>(acc, v, c, x);
  }
}

template <typename V, int KR>
__device__ void row_items(const BatchMat& m, int64_t r, int lane, int L,
                          typename vtraits<V>::acc* acc) {
  constexpr int kU = Unroll<KR>::value;
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  const V* val = static_cast<const V*>(m.val);
  const Vec* x = static_cast<const Vec*>(m.x);
  if (m.pre == BPRE_DIA) dia_row<V, KR>(m, r, lane, L, acc);
  else if (m.pre == BPRE_ELL) ell_row<V, KR>(m, r, lane, L, acc);
  if (!m.ro) return;  // a DIA / ELL row taken as a whole-row tile
  const int64_t st = m.ro[r], en = m.ro[r + 1];
  for (int64_t j = st + lane; j < en; j += (int64_t)kU * L) {
    int64_t c[kU];
    V v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int64_t jj = j + (int64_t)u * L;
      c[u] = jj < en ? (int64_t)ld_stream(m.ci + jj) : -1;
      v[u] = jj < en ? ld_stream(val + jj) : V{};
    }
    consume<V, KR>(acc, v, c, x);
  }
}

// A row of the tile longer than kDefer * L items is not left to its group of
// L lanes (it would serialise the tile): it is queued in LDS and reduced by
// the whole workgroup after the short rows.
constexpr int kDefer = 8;
constexpr int kMaxDeferred = 64;

template <typename V, int KR>
__device__ void block_row(const BatchMat& m, int64_t r, int init, typename vtraits<V>::acc* s_red);

template <typename V, int KR>
__device__ void tile_short(const BatchMat& m, const BatchTile& t, int init,
                           typename vtraits<V>::acc* s_red, int32_t* s_def, int32_t* s_ndef) {
  using Acc = typename vtraits<V>::acc;
  const int nr = t.r1 - t.r0;
  const int64_t items = (int64_t)m.ro[t.r1] - m.ro[t.r0];
  int L = 1;
  const int64_t mean = nr > 0 ? items / nr : 0;
  while (L < kWave && L < mean) L <<= 1;
  if (m.pre != BPRE_NONE && L < 4) L = 4;
  const int groups = kBT / L;
  const int g = threadIdx.x / L, l = threadIdx.x % L;
  if (threadIdx.x == 0) *s_ndef = 0;
  __syncthreads();
  for (int rr = g; rr < nr; rr += groups) {
    const int64_t r = t.r0 + rr;
    const int64_t len = (int64_t)m.ro[r + 1] - m.ro[r];
    if (len > (int64_t)kDefer * L && L < kWave) {
      // the group's lane 0 queues the row; the list overflows into the
      // group path, which stays correct
      bool queued = false;
      if (l == 0) {
        const int slot = atomicAdd(s_ndef, 1);
        if (slot < kMaxDeferred) {
          s_def[slot] = (int32_t)r;
          queued = true;
        }
      }
      queued = __shfl(queued, (threadIdx.x % kWave) - l, kWave);
      if (queued) continue;
    }
    Acc acc[KR];
#pragma unroll
    for (int q = 0; q < KR; q++) acc[q] = 0;
    row_items<V, KR>(m, r, l, L, acc);
#pragma unroll
    for (int q = 0; q < KR; q++) acc[q] = group_sum(acc[q], L);
    if (l == 0) store_row<V, KR>(m, r, acc, init);
  }
  __syncthreads();
  const int nd = min(*s_ndef, kMaxDeferred);
  for (int i = 0; i < nd; i++) {
    block_row<V, KR>(m, s_def[i], init, s_red);
    __syncthreads();
  }
}

// One row reduced by the whole workgroup as Dot2 pairs per column (FP32 /
// BF16 rows of more than kCompensateLen items: every such row is a
// whole-row tile): lane-strided pairs, then a fixed tree over lanes and
// waves; the old y joins through TwoSum when accumulating.
template <typename V, int KR>
__device__ void block_row_comp(const BatchMat& m, int64_t r, int init, float* s_red) {
  using Vec = typename vtraits<V>::vec;
  csum<float> cs[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) cs[q] = csum<float>{0, 0};
  const Vec* x = static_cast<const Vec*>(m.x);
  if (m.pre == BPRE_DIA) {
    const V* dv = static_cast<const V*>(m.dval);
    for (int d = threadIdx.x; d < m.ndiag; d += kBT) {
      const int64_t col = r + m.doff[d];
      if (col >= 0 && col < m.ncols) {
        const float v = vtraits<V>::widen(ld_stream(dv + r + m.dpitch * d));
#pragma unroll
        for (int q = 0; q < KR; q++) cs[q].add_prod(v, (float)x[col * KR + q]);
      }
    }
  } else if (m.pre == BPRE_ELL) {
    const V* ev = static_cast<const V*>(m.evalp);
    for (int n = threadIdx.x; n < m.eloop; n += kBT) {
      const int32_t c = ld_stream(m.eci + r + m.epitch * n);
      if (c >= 0) {
        const float v = vtraits<V>::widen(ld_stream(ev + r + m.epitch * n));
#pragma unroll
        for (int q = 0; q < KR; q++) cs[q].add_prod(v, (float)x[(int64_t)c * KR + q]);
      }
    }
  }
  if (m.ro) {
    const V* val = static_cast<const V*>(m.val);
    for (int64_t j = m.ro[r] + threadIdx.x; j < m.ro[r + 1]; j += kBT) {
      const float v = vtraits<V>::widen(ld_stream(val + j));
      const int64_t c = ld_stream(m.ci + j);
#pragma unroll
      for (int q = 0; q < KR; q++) cs[q].add_prod(v, (float)x[c * KR + q]);
    }
  }
  constexpr int kW = kBT / kWave;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < KR; q++) {
    for (int sh = kWave >> 1; sh > 0; sh >>= 1) {
      const float s2 = shfl_xor(cs[q].s, sh), c2 = shfl_xor(cs[q].c, sh);
      cs[q].combine(s2, c2);
    }
    if (lane == 0) {
      s_red[2 * (q * kW + wave)] = cs[q].s;
      s_red[2 * (q * kW + wave) + 1] = cs[q].c;
    }
  }
  __syncthreads();
  if (threadIdx.x < KR) {
    const int q = threadIdx.x;
    csum<float> t{s_red[2 * q * kW], s_red[2 * q * kW + 1]};
    for (int w = 1; w < kW; w++) t.combine(s_red[2 * (q * kW + w)], s_red[2 * (q * kW + w) + 1]);
    Vec* y = static_cast<Vec*>(m.y) + r * KR + q;
    if (!init) t.combine((float)*y, 0.0f);
    *y = (Vec)t.value();
  }
}

// one row reduced by the whole workgroup
template <typename V, int KR>
__device__ void block_row(const BatchMat& m, int64_t r, int init,
                          typename vtraits<V>::acc* s_red) {
  using Acc = typename vtraits<V>::acc;
  Acc acc[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) acc[q] = 0;
  row_items<V, KR>(m, r, threadIdx.x, kBT, acc);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < KR; q++) {
    for (int s = kWave >> 1; s > 0; s >>= 1) acc[q] += shfl_xor(acc[q], s);
    if (lane == 0) s_red[q * (kBT / kWave) + wave] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc tot[KR];
#pragma unroll
    for (int q = 0; q < KR; q++) {
      tot[q] = 0;
      for (int w = 0; w < kBT / kWave; w++) tot[q] += s_red[q * (kBT / kWave) + w];
    }
    store_row<V, KR>(m, r, tot, init);
  }
}

template <typename V, int KR>
__device__ void tile_wholerow(const BatchMat& m, const BatchTile& t, int init,
                              typename vtraits<V>::acc* s_red) {
  block_row<V, KR>(m, t.r0, init, s_red);
}

template <typename V, int KR>
__device__ void tile_dia(const BatchMat& m, const BatchTile& t, int init) {
  using Acc = typename vtraits<V>::acc;
  const int64_t r = t.r0 + threadIdx.x;
  if (r >= t.r1) return;
  Acc acc[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) acc[q] = 0;
  dia_row<V, KR>(m, r, 0, 1, acc);
  store_row<V, KR>(m, r, acc, init);
}

template <typename V, int KR>
__device__ void tile_ell(const BatchMat& m, const BatchTile& t, int init) {
  using Acc = typename vtraits<V>::acc;
  const int64_t r = t.r0 + threadIdx.x;
  if (r >= t.r1) return;
  Acc acc[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) acc[q] = 0;
  ell_row<V, KR>(m, r, 0, 1, acc);
  store_row<V, KR>(m, r, acc, init);
}

__global__ void batch_reset_kernel(int32_t* counters) {
  if (threadIdx.x < kQueues) counters[threadIdx.x] = 0;
}

template <typename V, int KR>
__device__ __forceinline__ void run_tile(const BatchMat* __restrict__ mats,
                                         const BatchTile& t, int init,
                                         typename vtraits<V>::acc* s_red) {
  __shared__ int32_t s_def[kMaxDeferred];
  __shared__ int32_t s_ndef;
  const BatchMat m = mats[t.mat];
  switch (t.kind) {
    case TILE_SHORT: tile_short<V, KR>(m, t, init, s_red, s_def, &s_ndef); break;
    case TILE_WHOLEROW: tile_wholerow<V, KR>(m, t, init, s_red); break;
    case TILE_DIA: tile_dia<V, KR>(m, t, init); break;
    default: tile_ell<V, KR>(m, t, init); break;
  }
}

// Static schedule: workgroup w walks its own slice woff[w] .. woff[w+1] of
// the tile list, filled on the host by longest-processing-time bin packing
// (each tile, largest first, goes to the workgroup with the least estimated
// work so far).  No counters, no atomics.
template <typename V, int KR>
__global__ __launch_bounds__(kBT) void batch_spmm_static_kernel(
    const BatchMat* __restrict__ mats, const BatchTile* __restrict__ tiles,
    const int32_t* __restrict__ woff, int init) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[KR * kBT / kWave];
  const int32_t b = woff[blockIdx.x], e = woff[blockIdx.x + 1];
  for (int32_t i = b; i < e; i++) {
    run_tile<V, KR>(mats, tiles[i], init, s_red);
    __syncthreads();
  }
}

// qoff[q] .. qoff[q+1]: the tiles of sub-queue q.  A workgroup drains its
// home sub-queue, then the others; a relaxed load skips a drained sub-queue
// without a read-modify-write on its counter.
template <typename V, int KR>
__global__ __launch_bounds__(kBT) void batch_spmm_kernel(const BatchMat* __restrict__ mats,
                                                         const BatchTile* __restrict__ tiles,
                                                         const int32_t* __restrict__ qoff,
                                                         int32_t* counters, int32_t grab,
                                                         int init) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[KR * kBT / kWave];
  __shared__ int32_t s_next;
  const int home = blockIdx.x % kQueues;
  for (int qi = 0; qi < kQueues; qi++) {
    const int q = (home + qi) % kQueues;
    const int32_t qb = qoff[q], qn = qoff[q + 1] - qoff[q];
    while (true) {
      if (threadIdx.x == 0) {
        int32_t v = __hip_atomic_load(&counters[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_next = v >= qn ? qn : atomicAdd(&counters[q], grab);
      }
      __syncthreads();
      const int32_t first = s_next;
      __syncthreads();
      if (first >= qn) break;  // this sub-queue is drained: try the next one
      const int32_t last = min(first + grab, qn);
      for (int32_t i = first; i < last; i++) {
        run_tile<V, KR>(mats, tiles[qb + i], init, s_red);
        __syncthreads();
      }
    }
  }
}

// FP32 / BF16 rows of more than kCompensateLen items (whole-row tiles kept
// out of the queue): one Dot2 row per workgroup turn, in a launch of its own
// so that the pairs do not raise the register count of the queue kernel.
template <typename V, int KR>
__global__ __launch_bounds__(kBT) void batch_comp_rows_kernel(const BatchMat* __restrict__ mats,
                                                              const BatchTile* __restrict__ tiles,
                                                              int32_t n, int init) {
  __shared__ float s_red[2 * KR * kBT / kWave];
  for (int32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const BatchTile t = tiles[i];
    block_row_comp<V, KR>(mats[t.mat], t.r0, init, s_red);
    __syncthreads();
  }
}

}  // namespace mph

// ------------------------------------------------------------------ host
struct mph_batch_s {
  int vt = MPH_F32;
  int k = 1;
  int32_t ntiles = 0;
  int32_t nwg = 0;
  mph::BatchMat* d_mats = nullptr;
  mph::BatchTile* d_tiles = nullptr;      // queue order (sub-queues)
  mph::BatchTile* d_tiles_lpt = nullptr;  // per-workgroup lists (static)
  int32_t* d_woff = nullptr;              // nwg + 1 list offsets
  int32_t* d_qoff = nullptr;
  int sched = 0;                          // 0 queues (default), 1 static lists
  int32_t grab = 4;                       // tiles per counter update
  int32_t* d_counter = nullptr;
  // FP32 / BF16 rows of more than kCompensateLen items (Dot2 kernel)
  mph::BatchTile* d_comp = nullptr;
  int32_t ncomp = 0;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<void*> owned;  // row offsets derived for COO members
};

namespace mph {
namespace {

// estimated fixed cost of one tile, in matrix entries
constexpr int64_t kTileOverhead = 512;

// the ELL / DIA slots a row of a HYB / HDC member adds to its tail / CSR part
int64_t row_prefix_len(const mph_matrix_impl* A) {
  return A->format == MPH_HDC ? A->ndiag : A->format == MPH_HYB ? A->ell_loop : 0;
}

int64_t tile_cost(const BatchTile& t, const std::vector<int32_t>* ro, const mph_matrix_impl* A) {
  if (t.kind == TILE_DIA) return (int64_t)(t.r1 - t.r0) * std::max<int64_t>(A->ndiag, 1);
  if (t.kind == TILE_ELL) return (int64_t)(t.r1 - t.r0) * std::max<int64_t>(A->ell_loop, 1);
  int64_t c = (int64_t)(*ro)[t.r1] - (*ro)[t.r0] + (t.r1 - t.r0);
  if (A->format == MPH_HDC) c += (int64_t)(t.r1 - t.r0) * A->ndiag;
  if (A->format == MPH_HYB) c += (int64_t)(t.r1 - t.r0) * A->ell_loop;
  return c;
}

template <typename V, int KR>
int launch_batch(mph_batch_s* b, int init, hipStream_t s) {
  if constexpr (!std::is_same<V, double>::value)
    if (b->ncomp > 0)
      hipLaunchKernelGGL((batch_comp_rows_kernel<V, KR>),
                         dim3((unsigned)std::min<int32_t>(b->ncomp, 4 * device_cu_count())),
                         dim3(kBT), 0, s, b->d_mats, b->d_comp, b->ncomp, init);
  if (b->ntiles == 0) {
    MPH_HIP_CHECK(hipGetLastError());
    return MPH_OK;
  }
  if (b->sched == 1) {
    hipLaunchKernelGGL((batch_spmm_static_kernel<V, KR>), dim3((unsigned)b->nwg), dim3(kBT), 0,
                       s, b->d_mats, b->d_tiles_lpt, b->d_woff, init);
  } else {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, b->d_counter);
    hipLaunchKernelGGL((batch_spmm_kernel<V, KR>), dim3((unsigned)b->nwg), dim3(kBT), 0, s,
                       b->d_mats, b->d_tiles, b->d_qoff, b->d_counter, b->grab, init);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V>
int launch_k(mph_batch_s* b, int init, hipStream_t s) {
  switch (b->k) {
    case 1: return launch_batch<V, 1>(b, init, s);
    case 2: return launch_batch<V, 2>(b, init, s);
    case 4: return launch_batch<V, 4>(b, init, s);
    case 8: return launch_batch<V, 8>(b, init, s);
    case 16: return launch_batch<V, 16>(b, init, s);
  }
  return invalid("batched k must be 1, 2, 4, 8 or 16");
}

int batch_launch(mph_batch_s* b, int init, hipStream_t s) {
  if (b->ntiles == 0 && b->ncomp == 0) return MPH_OK;
  switch (b->vt) {
    case MPH_F64: return launch_k<double>(b, init, s);
    case MPH_F32: return launch_k<float>(b, init, s);
    case MPH_BF16: return launch_k<bf16>(b, init, s);
  }
  return invalid("unknown value type");
}

}  // namespace
}  // namespace mph

using namespace mph;

extern "C" {

int mph_batch_create(mph_batch_t* out, int32_t nmat, const mph_matrix_t* mats, int32_t k,
                     const void* const* X, void* const* Y, mph_stream_t stream) {
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!out || nmat < 0 || (nmat > 0 && (!mats || !X || !Y))) return invalid("null batch argument");
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16)
    return invalid("batched k must be 1, 2, 4, 8 or 16");
  std::vector<BatchMat> hm(nmat);
  std::vector<BatchTile> tiles;
  std::vector<int64_t> cost;
  std::vector<int64_t> tlen;  // items of a whole-row tile's row (else 0)
  int vt = nmat > 0 ? mats[0]->vt : MPH_F32;
  // COO members are batched as the CSR view of their row-sorted entries: row
  // offsets derived on the device (owned by the batch) and the row blocks of
  // that view
  std::vector<void*> owned;
  struct OwnedGuard {
    std::vector<void*>* v;
    bool keep = false;
    ~OwnedGuard() {
      if (!keep)
        for (void* p : *v) dfree(p);
    }
  } guard{&owned};
  for (int32_t i = 0; i < nmat; i++) {
    mph_matrix_impl* A = mats[i];
    if (!A) return invalid("null matrix in batch");
    if (A->vt != vt) return invalid("all batched matrices must share one value type");
    if (A->format == MPH_COO) {
      int sorted = 1;
      MPH_CHECK_STATUS(check_index_range(A->coo_ri, A->coo_nnz, A->nrows, false, s, "COO rows"));
      MPH_CHECK_STATUS(check_index_range(A->coo_ci, A->coo_nnz, A->ncols, false, s, "COO columns"));
      MPH_CHECK_STATUS(rows_sorted(A->coo_ri, A->coo_nnz, A->nrows, s, &sorted));
      if (!sorted) return invalid("a COO matrix with unsorted rows cannot be batched");
      int32_t* d_ro = nullptr;
      MPH_HIP_CHECK(hipMalloc(&d_ro, sizeof(int32_t) * (A->nrows + 1)));
      owned.push_back(d_ro);
      MPH_CHECK_STATUS(coo_row_offsets(A->coo_ri, A->coo_nnz, A->nrows, d_ro, s));
      mph_matrix_s view;
      view.format = MPH_CSR;
      view.vt = A->vt;
      view.nrows = A->nrows;
      view.ncols = A->ncols;
      view.nnz = view.csr_nnz = A->coo_nnz;
      view.csr_ro = d_ro;
      view.csr_ci = A->coo_ci;
      view.csr_v = A->coo_v;
      view.tuning = A->tuning;
      view.tuning.xtile = 1;  // the batch reads row blocks only
      BatchMat m{};
      m.nrows = A->nrows;
      m.ncols = A->ncols;
      m.x = X[i];
      m.y = Y[i];
      if (A->nrows > 0 && (!X[i] || !Y[i])) return invalid("null X / Y in batch");
      if ((k * vector_size(vt)) % 16 == 0 && ((reinterpret_cast<uintptr_t>(X[i]) % 16) != 0 ||
                                              (reinterpret_cast<uintptr_t>(Y[i]) % 16) != 0))
        return invalid("batched X / Y rows must be 16-byte aligned");
      m.ro = d_ro;
      m.ci = A->coo_ci;
      m.val = A->coo_v;
      m.pre = BPRE_NONE;
      int st = rowblock_analyse(&view, s);
      if (st == MPH_OK && A->nrows > 0) {
        std::vector<int32_t> ro(A->nrows + 1);
        std::vector<RowBlock> bl(view.rb.nblocks + 1);
        hipError_t e = hipMemcpyAsync(ro.data(), d_ro, sizeof(int32_t) * ro.size(),
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(bl.data(), view.rb.blocks, sizeof(RowBlock) * bl.size(),
                             hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
          release_plans(&view);
          MPH_HIP_CHECK(e);
        }
        for (int32_t q = 0; q < view.rb.nblocks; q++) {
          if (bl[q].sub > 0) continue;  // long rows: one whole-row tile
          BatchTile t{i, bl[q].sub < 0 ? TILE_SHORT : TILE_WHOLEROW, bl[q].row,
                      bl[q].sub < 0 ? bl[q + 1].row : bl[q].row + 1};
          tiles.push_back(t);
          cost.push_back(tile_cost(t, &ro, A));
          tlen.push_back(t.kind == TILE_WHOLEROW ? (int64_t)ro[t.r0 + 1] - ro[t.r0] + row_prefix_len(A) : 0);
        }
      }
      release_plans(&view);
      if (st != MPH_OK) return st;
      hm[i] = m;
      continue;
    }
    if (!A->analysed) MPH_CHECK_STATUS(mph_matrix_analyse(mats[i], stream));
    if (A->format == MPH_HYB && A->rb.tail_ro == nullptr && A->coo_nnz > 0)
      return invalid("HYB with an unsorted COO part cannot be batched");
    BatchMat m{};
    m.nrows = A->nrows;
    m.ncols = A->ncols;
    m.x = X[i];
    m.y = Y[i];
    if (A->nrows > 0 && (!X[i] || !Y[i])) return invalid("null X / Y in batch");
    const size_t vsz = vector_size(vt);
    if ((k * vsz) % 16 == 0 && ((reinterpret_cast<uintptr_t>(X[i]) % 16) != 0 ||
                                (reinterpret_cast<uintptr_t>(Y[i]) % 16) != 0))
      return invalid("batched X / Y rows must be 16-byte aligned");
    std::vector<int32_t> ro;
    // an HDC without CSR entries is its DIA part, a HYB without COO entries
    // its ELL part: they are cut into slabs like DIA / ELL matrices
    const bool as_dia = A->format == MPH_DIA || (A->format == MPH_HDC && A->csr_nnz == 0);
    const bool as_ell = A->format == MPH_ELL || (A->format == MPH_HYB && A->coo_nnz == 0);
    const bool rowfam = !as_dia && !as_ell;
    if (rowfam) {
      m.ro = A->format == MPH_HYB ? A->rb.tail_ro : A->csr_ro;
      m.ci = A->format == MPH_HYB ? A->coo_ci : A->csr_ci;
      m.val = A->format == MPH_HYB ? A->coo_v : A->csr_v;
      m.pre = A->format == MPH_HDC ? BPRE_DIA : A->format == MPH_HYB ? BPRE_ELL : BPRE_NONE;
      ro.resize(A->nrows + 1);
      if (A->nrows > 0) {
        MPH_HIP_CHECK(hipMemcpyAsync(ro.data(), m.ro, sizeof(int32_t) * (A->nrows + 1),
                                     hipMemcpyDeviceToHost, s));
        std::vector<RowBlock> bl(A->rb.nblocks + 1);
        MPH_HIP_CHECK(hipMemcpyAsync(bl.data(), A->rb.blocks, sizeof(RowBlock) * bl.size(),
                                     hipMemcpyDeviceToHost, s));
        MPH_HIP_CHECK(hipStreamSynchronize(s));
        if (A->rb.prefix_apart) {
          // FP32 / BF16 HYB / HDC with an ELL / DIA part of kCompensateLen / 2
          // or more slots per row: every row a whole-row tile, so that the
          // row's two parts are summed as one Dot2 sum
          for (int64_t r = 0; r < A->nrows; r++) {
            BatchTile t{i, TILE_WHOLEROW, (int32_t)r, (int32_t)r + 1};
            tiles.push_back(t);
            cost.push_back(tile_cost(t, &ro, A) + kTileOverhead);
            tlen.push_back((int64_t)ro[t.r0 + 1] - ro[t.r0] + row_prefix_len(A));
          }
          bl.assign(1, RowBlock{0, -1});
        }
        for (int32_t q = 0; q + 1 < (int32_t)bl.size(); q++) {
          if (bl[q].sub > 0) continue;  // long rows: one whole-row tile
          BatchTile t{i, bl[q].sub < 0 ? TILE_SHORT : TILE_WHOLEROW, bl[q].row,
                      bl[q].sub < 0 ? bl[q + 1].row : bl[q].row + 1};
          tiles.push_back(t);
          cost.push_back(tile_cost(t, &ro, A));
          tlen.push_back(t.kind == TILE_WHOLEROW ? (int64_t)ro[t.r0 + 1] - ro[t.r0] + row_prefix_len(A) : 0);
        }
      }
    } else if (vt != MPH_F64 && (as_dia ? A->ndiag : A->ell_loop) > kCompensateLen) {
      // FP32 / BF16 rows of more than kCompensateLen diagonals / slots: one
      // whole-row tile each (Dot2 pairs), with no CSR part
      m.ro = nullptr;
      m.pre = as_dia ? BPRE_DIA : BPRE_ELL;
      for (int64_t r = 0; r < A->nrows; r++) {
        BatchTile t{i, TILE_WHOLEROW, (int32_t)r, (int32_t)r + 1};
        tiles.push_back(t);
        cost.push_back((as_dia ? A->ndiag : A->ell_loop) + kTileOverhead);
        tlen.push_back(as_dia ? A->ndiag : A->ell_loop);
      }
    } else {
      for (int64_t r = 0; r < A->nrows; r += kSlab) {
        BatchTile t{i, as_dia ? TILE_DIA : TILE_ELL, (int32_t)r,
                    (int32_t)std::min<int64_t>(A->nrows, r + kSlab)};
        tiles.push_back(t);
        cost.push_back(tile_cost(t, nullptr, A));
        tlen.push_back(0);
      }
    }
    m.ndiag = (int32_t)A->ndiag;
    m.dpitch = A->dia_pitch;
    m.doff = A->dia_off;
    m.dval = A->dia_v;
    m.eloop = (int32_t)A->ell_loop;
    m.epitch = A->ell_pitch;
    m.eci = A->ell_ci;
    m.evalp = A->ell_v;
    hm[i] = m;
  }
  // FP32 / BF16 whole-row tiles of more than kCompensateLen items leave the
  // queue for the Dot2 kernel
  std::vector<BatchTile> comp;
  if (vt != MPH_F64) {
    std::vector<BatchTile> keep;
    std::vector<int64_t> keep_cost;
    for (size_t q = 0; q < tiles.size(); q++) {
      const BatchTile& t = tiles[q];
      if (t.kind == TILE_WHOLEROW && tlen[q] > kCompensateLen) {
        comp.push_back(t);
      } else {
        keep.push_back(t);
        keep_cost.push_back(cost[q]);
      }
    }
    tiles.swap(keep);
    cost.swap(keep_cost);
  }
  // largest tiles first: the dynamic queue then drains evenly
  std::vector<size_t> order(tiles.size());
  for (size_t q = 0; q < order.size(); q++) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return cost[a] > cost[b]; });
  // deal the sorted tiles round-robin into the sub-queues
  std::vector<BatchTile> sorted;
  sorted.reserve(tiles.size());
  std::vector<int32_t> qoff(kQueues + 1, 0);
  for (int q = 0; q < kQueues; q++) {
    qoff[q] = (int32_t)sorted.size();
    for (size_t j = q; j < order.size(); j += kQueues) sorted.push_back(tiles[order[j]]);
  }
  qoff[kQueues] = (int32_t)sorted.size();

  mph_batch_s* b = new mph_batch_s();
  b->owned = owned;
  guard.keep = true;
  b->vt = vt;
  b->k = k;
  b->ntiles = (int32_t)sorted.size();
  b->nwg = (int32_t)std::max<int64_t>(1, std::min<int64_t>((int64_t)device_cu_count() * 8,
                                                           (int64_t)b->ntiles));
  hipError_t e = hipMalloc(&b->d_mats, sizeof(BatchMat) * std::max<int32_t>(nmat, 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_tiles, sizeof(BatchTile) * std::max<int32_t>(b->ntiles, 1));
  // longest-processing-time packing into per-workgroup lists
  std::vector<std::vector<int32_t>> lists(b->nwg);
  {
    std::priority_queue<std::pair<int64_t, int32_t>, std::vector<std::pair<int64_t, int32_t>>,
                        std::greater<std::pair<int64_t, int32_t>>>
        heap;
    for (int32_t w = 0; w < b->nwg; w++) heap.push({0, w});
    for (size_t q = 0; q < order.size(); q++) {
      auto top = heap.top();
      heap.pop();
      lists[top.second].push_back((int32_t)order[q]);
      heap.push({top.first + cost[order[q]] + kTileOverhead, top.second});
    }
  }
  std::vector<BatchTile> packed;
  packed.reserve(order.size());
  std::vector<int32_t> woff(b->nwg + 1, 0);
  for (int32_t w = 0; w < b->nwg; w++) {
    woff[w] = (int32_t)packed.size();
    for (int32_t id : lists[w]) packed.push_back(tiles[id]);
  }
  woff[b->nwg] = (int32_t)packed.size();
  if (e == hipSuccess) e = hipMalloc(&b->d_tiles_lpt, sizeof(BatchTile) * std::max<int32_t>(b->ntiles, 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_woff, sizeof(int32_t) * (b->nwg + 1));
  if (e == hipSuccess && b->ntiles > 0)
    e = hipMemcpyAsync(b->d_tiles_lpt, packed.data(), sizeof(BatchTile) * b->ntiles,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_woff, woff.data(), sizeof(int32_t) * (b->nwg + 1),
                       hipMemcpyHostToDevice, s);
  if (const char* env = std::getenv("MPH_BATCH_SCHEDULE")) b->sched = std::atoi(env) == 1 ? 1 : 0;
  // about four grabs per workgroup per sub-queue
  b->grab = (int32_t)std::max<int64_t>(
      1, std::min<int64_t>(16, (int64_t)b->ntiles / std::max<int64_t>(1, (int64_t)b->nwg * 4)));
  if (e == hipSuccess) e = hipMalloc(&b->d_counter, sizeof(int32_t) * kQueues);
  if (e == hipSuccess) e = hipMalloc(&b->d_qoff, sizeof(int32_t) * (kQueues + 1));
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_qoff, qoff.data(), sizeof(int32_t) * (kQueues + 1),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && nmat > 0)
    e = hipMemcpyAsync(b->d_mats, hm.data(), sizeof(BatchMat) * nmat, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && b->ntiles > 0)
    e = hipMemcpyAsync(b->d_tiles, sorted.data(), sizeof(BatchTile) * b->ntiles,
                       hipMemcpyHostToDevice, s);
  b->ncomp = (int32_t)comp.size();
  if (e == hipSuccess && b->ncomp > 0) e = hipMalloc(&b->d_comp, sizeof(BatchTile) * b->ncomp);
  if (e == hipSuccess && b->ncomp > 0)
    e = hipMemcpyAsync(b->d_comp, comp.data(), sizeof(BatchTile) * b->ncomp,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error(std::string("batch create: ") + hipGetErrorString(e));
    mph_batch_destroy(b);
    return MPH_ERR_HIP;
  }
  *out = b;
  return MPH_OK;
}

int mph_batch_destroy(mph_batch_t b) {
  if (!b) return MPH_OK;
  for (int i = 0; i < 2; i++) {
    if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
    if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
  }
  dfree(b->d_mats);
  dfree(b->d_tiles);
  dfree(b->d_tiles_lpt);
  dfree(b->d_woff);
  dfree(b->d_qoff);
  dfree(b->d_counter);
  dfree(b->d_comp);
  for (void* p : b->owned) dfree(p);
  delete b;
  return MPH_OK;
}

int mph_batch_run(mph_batch_t b, int init, mph_stream_t stream) {
  if (!b) return invalid("null batch");
  return batch_launch(b, init ? 1 : 0, reinterpret_cast<hipStream_t>(stream));
}

int mph_batch_run_graph(mph_batch_t b, int init, mph_stream_t stream) {
  if (!b) return invalid("null batch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int g = init ? 1 : 0;
  if (!b->exec[g]) {
    hipStream_t cap = nullptr;
    MPH_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    MPH_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    const int st = batch_launch(b, g, cap);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(cap, &graph);
    (void)hipStreamDestroy(cap);
    if (st != MPH_OK) return st;
    MPH_HIP_CHECK(e);
    b->graph[g] = graph;
    MPH_HIP_CHECK(hipGraphInstantiate(&b->exec[g], graph, nullptr, nullptr, 0));
  }
  MPH_HIP_CHECK(hipGraphLaunch(b->exec[g], s));
  return MPH_OK;
}

int mph_batch_info(mph_batch_t b, int64_t* ntiles, int32_t* nwg) {
  if (!b) return invalid("null batch");
  if (ntiles) *ntiles = b->ntiles;
  if (nwg) *nwg = b->nwg;
  return MPH_OK;
}

}  // extern "C"
