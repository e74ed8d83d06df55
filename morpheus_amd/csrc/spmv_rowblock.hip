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
#include <cstdlib>

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
  // split plan: block rows are list rows, list row i is matrix row rmap[i]
  // (nullptr: the identity); used with PRE_NONE only
  const int32_t* __restrict__ rmap;
  // {row, sub, first item, lanes} per record of `blocks`, sentinel
  // included: built by the analysis for every list the short-block kernel
  // runs over
  const int32_t* __restrict__ krec;
};

template <typename V>
__device__ __forceinline__ int64_t out_row(const RBArgs<V>& a, int64_t i) {
  return a.rmap ? (int64_t)a.rmap[i] : i;
}

// --------------------------------------------------------------- prefixes
template <typename V, int PRE, int KR>
__device__ __forceinline__ void row_prefix(const RBArgs<V>& a, int64_t r,
                                           typename vtraits<V>::acc* s) {
  using Acc = typename vtraits<V>::acc;
  // loads of 8 diagonals / slots are issued before their x reads; the
  // products are added in diagonal / slot order
  constexpr int kP = 8;
  if constexpr (PRE == PRE_DIA) {
    for (int d0 = 0; d0 < a.ndiag; d0 += kP) {
      int64_t col[kP];
      Acc v[kP];
#pragma unroll
      for (int q = 0; q < kP; q++) {
        const int d = d0 + q;
        col[q] = d < a.ndiag ? r + a.doff[d] : -1;
        const bool in = d < a.ndiag && col[q] >= 0 && col[q] < a.ncols;
        v[q] = in ? vtraits<V>::widen(ld_stream(a.dval + r + a.dpitch * d)) : (Acc)0;
        if (!in) col[q] = -1;
      }
#pragma unroll
      for (int q = 0; q < kP; q++) {
        if (col[q] >= 0) {
#pragma unroll
          for (int k = 0; k < KR; k++) s[k] += v[q] * (Acc)a.x[col[q] * a.ld + k];
        }
      }
    }
  } else if constexpr (PRE == PRE_ELL) {
    for (int n0 = 0; n0 < a.eloop; n0 += kP) {
      int32_t c[kP];
      Acc v[kP];
#pragma unroll
      for (int q = 0; q < kP; q++) {
        const int64_t slot = r + a.epitch * (n0 + q);
        c[q] = n0 + q < a.eloop ? ld_stream(a.eci + slot) : -1;
      }
#pragma unroll
      for (int q = 0; q < kP; q++) {
        const int64_t slot = r + a.epitch * (n0 + q);
        v[q] = c[q] != -1 ? vtraits<V>::widen(ld_stream(a.evalp + slot)) : (Acc)0;
      }
#pragma unroll
      for (int q = 0; q < kP; q++) {
        if (c[q] != -1) {
#pragma unroll
          for (int k = 0; k < KR; k++) s[k] += v[q] * (Acc)a.x[(int64_t)c[q] * a.ld + k];
        }
      }
    }
  }
}

// The same prefix as a Dot2 sum per column (FP32 / BF16 long rows).
template <typename V, int PRE, int KR>
__device__ __forceinline__ void row_prefix_comp(const RBArgs<V>& a, int64_t r,
                                                csum<typename vtraits<V>::acc>* cs) {
  using Acc = typename vtraits<V>::acc;
  if constexpr (PRE == PRE_DIA) {
    for (int d = 0; d < a.ndiag; d++) {
      const int64_t col = r + a.doff[d];
      if (col >= 0 && col < a.ncols) {
        const Acc v = vtraits<V>::widen(ld_stream(a.dval + r + a.dpitch * d));
#pragma unroll
        for (int k = 0; k < KR; k++) cs[k].add_prod(v, (Acc)a.x[col * a.ld + k]);
      }
    }
  } else if constexpr (PRE == PRE_ELL) {
    for (int n = 0; n < a.eloop; n++) {
      const int64_t slot = r + a.epitch * n;
      const int32_t c = ld_stream(a.eci + slot);
      if (c != -1) {
        const Acc v = vtraits<V>::widen(ld_stream(a.evalp + slot));
#pragma unroll
        for (int k = 0; k < KR; k++) cs[k].add_prod(v, (Acc)a.x[(int64_t)c * a.ld + k]);
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

// A compensated row sum stored into y; when accumulating, the old y joins
// the pair through TwoSum before the one final rounding.
template <typename Vec, typename A>
__device__ __forceinline__ void store_y_cs(Vec* y, int64_t i, csum<A> cs, int init) {
  if (!init) cs.combine((A)y[i], (A)0);
  y[i] = (Vec)cs.value();
}

// Items j = base + tid + k*256, k < kBatch, passed to sink(j, a_j, x[c_j])
// (the value widened to the accumulation type): all index and value loads
// of the batch are issued before the dependent x gathers, so each thread
// keeps kBatch gathers in flight.  The sink forms the product (plainly, or
// as a Dot2 step).
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
    V v[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int64_t i = base + (int64_t)k * kRowBlockThreads;
      c[k] = i < end ? ld_stream(ci + i) : 0;
      v[k] = i < end ? ld_stream(val + i) : V{};
    }
    Acc xv[kBatch];
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int64_t i = base + (int64_t)k * kRowBlockThreads;
      xv[k] = i < end ? (Acc)x[c[k]] : (Acc)0;
    }
    hold_raw(v);  // BF16: widened below the gathers
#pragma unroll
    for (int k = 0; k < kBatch; k++) {
      const int64_t i = base + (int64_t)k * kRowBlockThreads;
      if (i < end) sink(i, vtraits<V>::widen(v[k]), xv[k]);
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
    batched_products<V>(a.ci, a.val, a.x, s, e,
                        [&](int64_t, Acc av, Acc xv) { cs.add_prod(av, xv); });
    block_csum(cs, s_red);
    if (threadIdx.x == 0) {
      a.partial[2 * b] = cs.s;
      a.partial[2 * b + 1] = cs.c;
    }
  } else {
    Acc t = 0;
    batched_products<V>(a.ci, a.val, a.x, s, e, [&](int64_t, Acc av, Acc xv) { t += av * xv; });
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
  if constexpr (COMP) {
    csum<Acc> cs{0, 0};
    row_prefix_comp<V, PRE, 1>(a, r, &cs);
    for (int64_t q = 0; q < nsub; q++)
      cs.combine(a.partial[2 * (b + q)], a.partial[2 * (b + q) + 1]);
    store_y_cs(a.y, out_row(a, r), cs, a.init);
  } else {
    Acc pre[1] = {0};
    row_prefix<V, PRE, 1>(a, r, pre);
    Acc t = 0;
    for (int64_t q = 0; q < nsub; q++) t += a.partial[b + q];
    store_y(a.y, out_row(a, r), t + pre[0], a.init);
  }
}

template <typename V, int PRE, bool COMP>
__global__ __launch_bounds__(kRowBlockThreads) void rowblock_spmv_kernel(
    RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ Acc s_red[2 * kRowBlockThreads / kWave];
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  // the block's record and the next one (the sentinel bounds the last
  // block) in one scalar load: rows, first items and the lane count, so
  // that the column / value loads need not wait for the row offsets
  int32_t rec[8];
  __builtin_memcpy(rec, a.krec + 4 * (int64_t)b, sizeof(rec));
  const RowBlock blk{rec[0], rec[1]};
  const int32_t next_row = rec[4];
  const int lanes_b = rec[3];
  const int32_t item[2] = {rec[2], rec[6]};
  // an empty asm statement that takes the fields as operands, so that the
  // compiler loads them here, together, instead of sinking each load to its
  // first use behind the branches below (a round trip each)
  asm volatile("" : : "s"(next_row), "s"(lanes_b), "s"(item[0]), "s"(item[1]));
  if (blk.sub < -1) return;  // a long row handled by the x-tiled pass
  if (blk.sub >= 0) {
    long_chunk_spmv<V, COMP>(a, b, blk, s_red);
    return;
  }
  Acc* s_prod = reinterpret_cast<Acc*>(smem);
  Acc* s_pre = s_prod + 2 * a.chunk;
  int32_t* s_ro = reinterpret_cast<int32_t*>(s_pre + (PRE ? kRowBlockMaxRows : 0));

  const int64_t row0 = blk.row;
  const int64_t row1 = next_row;
  const int nr = (int)(row1 - row0);
  const int64_t p0 = item[0];
  const int n = item[1] - item[0];

  // the block's row offsets (at most kRowBlockMaxRows + 1) are loaded into
  // registers first and stored to LDS after the products, so that offsets,
  // columns and values are in flight together; every lane loads (indices
  // past the block clamped to its last offset), since a load under a branch
  // is waited on where the branch joins
  constexpr int kRoLoads = (kRowBlockMaxRows + kRowBlockThreads) / kRowBlockThreads;
  int32_t rv[kRoLoads];
#pragma unroll
  for (int k = 0; k < kRoLoads; k++) {
    const int i = threadIdx.x + k * kRowBlockThreads;
    rv[k] = a.ro[row0 + min(i, nr)];
  }
  batched_products<V>(a.ci + p0, a.val + p0, a.x, 0, n,
                      [&](int64_t i, Acc av, Acc xv) { s_prod[i] = av * xv; });
#pragma unroll
  for (int k = 0; k < kRoLoads; k++) {
    const int i = threadIdx.x + k * kRowBlockThreads;
    if (i <= nr) s_ro[i] = (int32_t)(rv[k] - p0);
  }
  if constexpr (PRE != PRE_NONE) {
    for (int i = threadIdx.x; i < nr; i += kRowBlockThreads) {
      Acc p[1] = {0};
      row_prefix<V, PRE, 1>(a, row0 + i, p);
      s_pre[i] = p[0];
    }
  }
  __syncthreads();

  const int L = a.lanes_fixed > 0 ? a.lanes_fixed : lanes_b;
  const int groups = kRowBlockThreads / L;
  const int g = threadIdx.x / L, l = threadIdx.x % L;
  // (a short block holds rows of at most long_len <= kCompensateLen items in
  // FP32 / BF16: longer rows take the chunk or x-tiled paths, which carry
  // Dot2 pairs, so the sums here are plain)
  for (int r = g; r < nr; r += groups) {
    const int st = s_ro[r], en = s_ro[r + 1];
    Acc t = 0;
    for (int j = st + l; j < en; j += L) t += s_prod[j];
    t = group_sum(t, L);
    if constexpr (PRE != PRE_NONE) t += s_pre[r];
    if (l == 0) store_y(a.y, out_row(a, row0 + r), t, a.init);
  }
}

// --------------------------------------------------------------- x-tiled
// Long rows with sorted columns, summed per tile of x (64 KiB).  The analysis
// copies their items tile-major: for (group, tile), the items of the group's
// rows whose columns fall in the tile, row after row, as (16-bit column
// within the tile, value), each sub-block (at most kXTileCap items of at
// most kXTileRows whole runs) stored transposed: its item j = 8a + b sits in
// slot p0 + b*S + a (S = ceil(n/8)), so lane a holds items 8a..8a+7 while
// every load instruction reads 64 consecutive slots.  Workgroup (group, tile)
// copies x[tile*T, +T) into LDS once; its waves then take the sub-blocks in
// turn, form the products with x read from LDS and reduce the runs with a
// segmented sum (lane-serial, then a segmented scan across the wave) into the
// (tile, row) partials.  Scattered x
// reads thus cost LDS bandwidth rather than one cache-line fill each.  Tiles
// are dealt to the XCDs by tile index (workgroup id % 8 is its XCD), so an x
// tile is fetched from memory once per XCD.
// LDS of xtile_kernel: x tile, a group's run offsets and its sub-block
// records (a group has at most kXTileMaxGroupRows rows, so
// at most as many sub-blocks, plus a sentinel).
// (A direct plan keeps its sub-block records in memory and the runs'
// positions in LDS instead.)
inline size_t xtile_lds_bytes(size_t vec, size_t /*acc*/, int32_t tile, bool direct = false) {
  return vec * (size_t)tile + sizeof(int32_t) * (kXTileMaxGroupRows + 2) +
         (direct ? sizeof(int32_t) : sizeof(XTSub)) * (kXTileMaxGroupRows + 1);
}

struct XTArgs {
  int32_t tile, ntiles, groups, gr, rows;
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ lro;
  const XTSub* __restrict__ sub;
  const int32_t* __restrict__ sub_start;
  const uint8_t* __restrict__ head;
  const uint16_t* __restrict__ col;
  const void* __restrict__ val;  // the copy's values, or (direct) the caller's
  void* __restrict__ partial;
  const int32_t* __restrict__ rpos;  // direct plans: position of each run's first item
};

// Items of one sub-block as loaded by a wave: lane l holds the kBatch
// consecutive items l*kBatch .. l*kBatch + kBatch - 1.
inline XTArgs xt_args(const XTilePlan& p) {
  return XTArgs{p.tile, p.ntiles, p.groups, p.gr, p.rows, p.row, p.lro,
                p.sub, p.sub_start, p.head, p.col, p.rpos ? p.src : p.val, p.partial, p.rpos};
}

// A sub-block record is the same for every lane of the wave that reads it:
// kept in scalar registers.
__device__ __forceinline__ XTSub uniform_sub(const XTSub& v) {
  return XTSub{__builtin_amdgcn_readfirstlane(v.i0), __builtin_amdgcn_readfirstlane(v.e0),
               __builtin_amdgcn_readfirstlane(v.p0), 0};
}

// Loaded values as held until their products: FP64 / FP32 as accumulated,
// BF16 as its raw bits (widened at the product; widened at the load, the
// compiler waited for each value before issuing the next load).
template <typename V>
struct XTHeld {
  using T = typename vtraits<V>::acc;
  __device__ __forceinline__ static T hold(V v) { return vtraits<V>::widen(v); }
  __device__ __forceinline__ static typename vtraits<V>::acc use(T t) { return t; }
};
template <>
struct XTHeld<bf16> {
  using T = uint32_t;
  __device__ __forceinline__ static T hold(bf16 v) { return v.bits; }
  __device__ __forceinline__ static float use(T t) { return __uint_as_float(t << 16); }
};

template <typename V>
struct XTLoad {
  uint32_t c[kBatch];
  typename XTHeld<V>::T v[kBatch];
  uint32_t h;  // run of the lane's first item
};

template <typename V>
__device__ __forceinline__ void xtile_load(const XTArgs& t, const V* cval, const XTSub& s0,
                                           int32_t n, int32_t sidx, int lane, XTLoad<V>& d) {
  const int32_t jb = lane * kBatch;
  const int32_t S = (n + kBatch - 1) / kBatch;
  d.h = jb < n ? (uint32_t)t.head[(int64_t)sidx * kWave + lane] : 0u;
#pragma unroll
  for (int q = 0; q < kBatch; q++) {
    const int32_t j = jb + q;
    const int64_t slot = (int64_t)s0.p0 + (int64_t)q * S + lane;
    d.c[q] = j < n ? (uint32_t)ld_stream(t.col + slot) : 0u;
    d.v[q] = j < n ? XTHeld<V>::hold(ld_stream(cval + slot)) : (typename XTHeld<V>::T)0;
  }
}

// A running sum of products a * x: plain (FP64) or Dot2-compensated (FP32).
template <typename A, bool C>
struct XSum {
  A s, c;
  __device__ static XSum zero() { return XSum{0, 0}; }
  __device__ void add(A p) { s += p; }  // plain form only
  __device__ void add_prod(A a, A x) {
    if constexpr (C) {
#pragma clang fp contract(off)
      A p, ep, t, e;
      two_prod(a, x, p, ep);
      two_sum(s, p, t, e);
      s = t;
      c += e + ep;
    } else {
      s += a * x;
    }
  }
  __device__ void merge(const XSum& o) {
    if constexpr (C) {
#pragma clang fp contract(off)
      A t, e;
      two_sum(s, o.s, t, e);
      s = t;
      c += e + o.c;
    } else {
      s += o.s;
    }
  }
  // the running sum of another lane, moved by a DPP control
  template <int CTRL, int ROWS>
  __device__ XSum mv() const {
    return XSum{dpp_val<CTRL, ROWS>(s), C ? dpp_val<CTRL, ROWS>(c) : (A)0};
  }
};

// One step of the wave's segmented inclusive scan: a lane whose segment did
// not start at or after the source lane adds the source's running sum.
template <int CTRL, int ROWS, typename S>
__device__ __forceinline__ void xscan_step(S& acc, int& flag, bool valid) {
  const S o = acc.template mv<CTRL, ROWS>();
  const int of = dpp_i32<CTRL, ROWS>(flag);
  if (valid && !flag) acc.merge(o);
  if (valid) flag |= of;
}

template <typename V, bool COMP>
__global__ __launch_bounds__(kXTileThreads) void xtile_kernel(RBArgs<V> a, XTArgs t) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  using S = XSum<Acc, COMP>;
  constexpr int kWaves = kXTileThreads / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: x tile | run offsets of the group | sub-block records
  Vec* xs = reinterpret_cast<Vec*>(smem);
  int32_t* s_lro = reinterpret_cast<int32_t*>(xs + t.tile);
  XTSub* s_sub = reinterpret_cast<XTSub*>(s_lro + kXTileMaxGroupRows + 2);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint32_t id = blockIdx.x;
  const uint32_t k = id / kNumXcd;
  const int32_t g = (int32_t)(k % (uint32_t)t.groups);
  const int32_t tile = (int32_t)((k / (uint32_t)t.groups) * kNumXcd + id % kNumXcd);
  if (tile >= t.ntiles) return;  // whole workgroup: no barrier reached
  const int64_t c0 = (int64_t)tile * t.tile;
  const int32_t nx = (int32_t)min<int64_t>(t.tile, a.ncols - c0);
  const int64_t blk = (int64_t)g * t.ntiles + tile;
  const int32_t* lro = t.lro + blk * (t.gr + 1);
  const int32_t gstart = g * t.gr;
  const int32_t sb0 = t.sub_start[blk];
  const int32_t nsub = t.sub_start[blk + 1] - 1 - sb0;  // sub-blocks (a sentinel follows)
  const V* cval = static_cast<const V*>(t.val);
  Acc* part = static_cast<Acc*>(t.partial);

  // prologue: the first sub-block's items, the x tile, the run offsets and
  // the sub-block records, all requested before one barrier
  XTSub s0{0, 0}, s1{0, 0};
  XTLoad<V> cur;
  if (wave < nsub) {
    s0 = uniform_sub(t.sub[sb0 + wave]);
    s1 = uniform_sub(t.sub[sb0 + wave + 1]);
    xtile_load<V>(t, cval, s0, s1.e0 - s0.e0, sb0 + wave, lane, cur);
  }
  {
    constexpr int kPer = 8;  // x loads in flight per thread
    for (int32_t j0 = threadIdx.x; j0 < nx; j0 += kXTileThreads * kPer) {
      Vec xv[kPer];
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int32_t j = j0 + q * kXTileThreads;
        xv[q] = j < nx ? a.x[(c0 + j) * a.ld] : (Vec)0;
      }
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int32_t j = j0 + q * kXTileThreads;
        if (j < nx) xs[j] = xv[q];
      }
    }
  }
  for (int32_t j = threadIdx.x; j <= t.gr; j += kXTileThreads) s_lro[j] = lro[j];
  for (int32_t j = threadIdx.x; j <= nsub; j += kXTileThreads) s_sub[j] = t.sub[sb0 + j];
  __syncthreads();

  // each wave takes every kWaves-th sub-block and reduces its runs with a
  // segmented sum: lane-serial inside a lane, a segmented scan across lanes
  for (int32_t q = wave; q < nsub; q += kWaves) {
    const int32_t nr = s1.i0 - s0.i0, n = s1.e0 - s0.e0;
    const int32_t i0 = s0.i0, e0 = s0.e0;
    const int32_t* ro = s_lro + (i0 - gstart);  // ro[r] - e0: first item of run r
    // FP64: the products; FP32 / BF16: value and x kept apart for Dot2
    Acc pa[kBatch], px[kBatch];
#pragma unroll
    for (int b = 0; b < kBatch; b++) {
      // the copy holds columns inside the tile; the clamp keeps a corrupted
      // entry inside the LDS tile
      const Acc xv = (Acc)xs[min(cur.c[b], (uint32_t)(nx - 1))];
      if constexpr (COMP) {
        pa[b] = XTHeld<V>::use(cur.v[b]);
        px[b] = xv;
      } else {
        pa[b] = XTHeld<V>::use(cur.v[b]) * xv;
        px[b] = (Acc)1;
      }
    }
    const int32_t hd = (int32_t)min(cur.h, (uint32_t)(nr - 1));
    auto put = [&](int32_t r, const S& v) {
      const int64_t o = (int64_t)tile * t.rows + i0 + r;
      if constexpr (COMP) {
        part[2 * o] = v.s;
        part[2 * o + 1] = v.c;
      } else {
        part[o] = v.s;
      }
    };
    const int32_t jb = lane * kBatch;
    const int cnt = (int)max(0, min(kBatch, n - jb));
    const bool active = cnt > 0;
    // run of the lane's first item (the last r with ro[r] <= jb), from the
    // analysis
    const int32_t r = active ? hd : 0;
    const int32_t h = r;
    int32_t cur_r = r;
    S head = S::zero(), run = S::zero();
    bool single = true;
    // first item of the next run, relative to the lane's first item (the
    // loop compares it with the constant b)
    const int32_t base = e0 + jb;
    int32_t next_rel = active ? ro[cur_r + 1] - base : 0;
#pragma unroll
    for (int b = 0; b < kBatch; b++) {
      if (b < cnt) {
        if (b >= next_rel) {
          // a run ends here: the lane's head run, or a run complete in the lane
          if (single) {
            head = run;
            single = false;
          } else {
            put(cur_r, run);
          }
          do {
            cur_r++;
            next_rel = ro[cur_r + 1] - base;
          } while (b >= next_rel);
          run = S::zero();
        }
        if constexpr (COMP)
          run.add_prod(pa[b], px[b]);
        else
          run.add(pa[b]);
      }
    }
    const int32_t tl = cur_r;
    // the next sub-block's items are requested once this one's products are
    // folded into the lane runs (fewer registers live across the scan)
    if (q + kWaves < nsub) {
      s0 = uniform_sub(s_sub[q + kWaves]);
      s1 = uniform_sub(s_sub[q + kWaves + 1]);
      xtile_load<V>(t, cval, s0, s1.e0 - s0.e0, sb0 + q + kWaves, lane, cur);
    }
    // segmented inclusive scan of the lane tails across the wave
    // (Hillis-Steele within rows of 16 lanes, then across rows; all 64
    // lanes take part)
    const int32_t t_prev = dpp_i32<kDppWaveShr1>(tl);
    const bool cont = active && lane > 0 && single && h == t_prev;
    S acc = active ? run : S::zero();
    int flag = cont ? 0 : 1;
    const int rl = lane & 15;
    xscan_step<kDppRowShr1, 0xf>(acc, flag, rl >= 1);
    xscan_step<kDppRowShr2, 0xf>(acc, flag, rl >= 2);
    xscan_step<kDppRowShr4, 0xf>(acc, flag, rl >= 4);
    xscan_step<kDppRowShr8, 0xf>(acc, flag, rl >= 8);
    xscan_step<kDppRowBcast15, 0xa>(acc, flag, (lane & 16) != 0);
    xscan_step<kDppRowBcast31, 0xc>(acc, flag, lane >= 32);
    const S prev_acc = acc.template mv<kDppWaveShr1, 0xf>();
    const int32_t h_next = dpp_i32<kDppWaveShl1>(h);
    const int last_lane = (n - 1) / kBatch;
    if (active) {
      if (!single) {
        S hv = head;
        if (lane > 0 && h == t_prev) hv.merge(prev_acc);
        put(h, hv);
      }
      if (lane == last_lane || h_next != tl) put(tl, acc);
    }
    // runs without items in this tile
    if (lane < nr && ro[lane] == ro[lane + 1]) put(lane, S::zero());
  }
}

// The x-tiled pass for KR right-hand sides at once (X, Y row-major with
// leading dimension a.ld, the group's first column at a.x / a.y): the
// workgroup stages the KR columns of its x tile in LDS (entry (c, j) at
// c * KR + j) and each item's value meets its KR entries.  Every column is
// reduced exactly as xtile_kernel reduces its one column -- the same runs,
// products, lane-serial sums and scan -- so column j equals the
// single-column pass over X[:, j] bit for bit, while the copy's items are
// read once for the KR columns.  Partials per (tile, row, column) in partk.
template <int CTRL, int ROWS, typename S, int KR>
__device__ __forceinline__ void xscan_step_k(S (&acc)[KR], int& flag, bool valid) {
  S o[KR];
#pragma unroll
  for (int j = 0; j < KR; j++) o[j] = acc[j].template mv<CTRL, ROWS>();
  const int of = dpp_i32<CTRL, ROWS>(flag);
  if (valid && !flag) {
#pragma unroll
    for (int j = 0; j < KR; j++) acc[j].merge(o[j]);
  }
  if (valid) flag |= of;
}

// threads of a k-wide workgroup: its LDS leaves one workgroup per CU, so it
// takes 16 waves (4 per SIMD; the kernel fits their 128 registers)
constexpr int kXTileKThreads = 1024;

inline size_t xtile_k_lds_bytes(size_t vec, int32_t tile, int kr) {
  return vec * (size_t)tile * (size_t)kr + sizeof(int32_t) * (kXTileMaxGroupRows + 2) +
         sizeof(XTSub) * (kXTileMaxGroupRows + 1);
}

template <typename V, bool COMP, int KR, int NT>
__global__ __launch_bounds__(NT) void xtile_kernel_k(RBArgs<V> a, XTArgs t, void* partk) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  using S = XSum<Acc, COMP>;
  constexpr int kWaves = NT / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: x tile (KR columns) | run offsets of the group | sub-block records
  Vec* xs = reinterpret_cast<Vec*>(smem);
  int32_t* s_lro = reinterpret_cast<int32_t*>(xs + (size_t)t.tile * KR);
  XTSub* s_sub = reinterpret_cast<XTSub*>(s_lro + kXTileMaxGroupRows + 2);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint32_t id = blockIdx.x;
  const uint32_t k = id / kNumXcd;
  const int32_t g = (int32_t)(k % (uint32_t)t.groups);
  const int32_t tile = (int32_t)((k / (uint32_t)t.groups) * kNumXcd + id % kNumXcd);
  if (tile >= t.ntiles) return;  // whole workgroup: no barrier reached
  const int64_t c0 = (int64_t)tile * t.tile;
  const int32_t nx = (int32_t)min<int64_t>(t.tile, a.ncols - c0);
  const int64_t blk = (int64_t)g * t.ntiles + tile;
  const int32_t* lro = t.lro + blk * (t.gr + 1);
  const int32_t gstart = g * t.gr;
  const int32_t sb0 = t.sub_start[blk];
  const int32_t nsub = t.sub_start[blk + 1] - 1 - sb0;
  const V* cval = static_cast<const V*>(t.val);
  Acc* part = static_cast<Acc*>(partk);

  XTSub s0{0, 0}, s1{0, 0};
  XTLoad<V> cur;
  if (wave < nsub) {
    s0 = uniform_sub(t.sub[sb0 + wave]);
    s1 = uniform_sub(t.sub[sb0 + wave + 1]);
    xtile_load<V>(t, cval, s0, s1.e0 - s0.e0, sb0 + wave, lane, cur);
  }
  {
    constexpr int kPer = 8;
    const int32_t ne = nx * KR;
    for (int32_t j0 = threadIdx.x; j0 < ne; j0 += NT * kPer) {
      Vec xv[kPer];
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int32_t e = j0 + q * NT;
        xv[q] = e < ne ? a.x[(c0 + e / KR) * a.ld + e % KR] : (Vec)0;
      }
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int32_t e = j0 + q * NT;
        if (e < ne) xs[e] = xv[q];
      }
    }
  }
  for (int32_t j = threadIdx.x; j <= t.gr; j += NT) s_lro[j] = lro[j];
  for (int32_t j = threadIdx.x; j <= nsub; j += NT) s_sub[j] = t.sub[sb0 + j];
  __syncthreads();

  for (int32_t q = wave; q < nsub; q += kWaves) {
    const int32_t nr = s1.i0 - s0.i0, n = s1.e0 - s0.e0;
    const int32_t i0 = s0.i0, e0 = s0.e0;
    const int32_t* ro = s_lro + (i0 - gstart);
    // FP64: the products; FP32 / BF16: value and x kept apart for Dot2
    Acc pa[kBatch], px[kBatch][KR];
#pragma unroll
    for (int b = 0; b < kBatch; b++) {
      const uint32_t cc = min(cur.c[b], (uint32_t)(nx - 1));
      const Acc av = XTHeld<V>::use(cur.v[b]);
#pragma unroll
      for (int j = 0; j < KR; j++) {
        const Acc xv = (Acc)xs[cc * KR + j];
        if constexpr (COMP) {
          px[b][j] = xv;
        } else {
          px[b][j] = av * xv;
        }
      }
      pa[b] = av;
    }
    const int32_t hd = (int32_t)min(cur.h, (uint32_t)(nr - 1));
    auto put = [&](int32_t r, const S (&v)[KR]) {
      const int64_t o = ((int64_t)tile * t.rows + i0 + r) * KR;
#pragma unroll
      for (int j = 0; j < KR; j++) {
        if constexpr (COMP) {
          part[2 * (o + j)] = v[j].s;
          part[2 * (o + j) + 1] = v[j].c;
        } else {
          part[o + j] = v[j].s;
        }
      }
    };
    const int32_t jb = lane * kBatch;
    const int cnt = (int)max(0, min(kBatch, n - jb));
    const bool active = cnt > 0;
    const int32_t r = active ? hd : 0;
    const int32_t h = r;
    int32_t cur_r = r;
    S head[KR], run[KR];
#pragma unroll
    for (int j = 0; j < KR; j++) {
      head[j] = S::zero();
      run[j] = S::zero();
    }
    bool single = true;
    const int32_t base = e0 + jb;
    int32_t next_rel = active ? ro[cur_r + 1] - base : 0;
#pragma unroll
    for (int b = 0; b < kBatch; b++) {
      if (b < cnt) {
        if (b >= next_rel) {
          if (single) {
#pragma unroll
            for (int j = 0; j < KR; j++) head[j] = run[j];
            single = false;
          } else {
            put(cur_r, run);
          }
          do {
            cur_r++;
            next_rel = ro[cur_r + 1] - base;
          } while (b >= next_rel);
#pragma unroll
          for (int j = 0; j < KR; j++) run[j] = S::zero();
        }
#pragma unroll
        for (int j = 0; j < KR; j++) {
          if constexpr (COMP)
            run[j].add_prod(pa[b], px[b][j]);
          else
            run[j].add(px[b][j]);
        }
      }
    }
    const int32_t tl = cur_r;
    if (q + kWaves < nsub) {
      s0 = uniform_sub(s_sub[q + kWaves]);
      s1 = uniform_sub(s_sub[q + kWaves + 1]);
      xtile_load<V>(t, cval, s0, s1.e0 - s0.e0, sb0 + q + kWaves, lane, cur);
    }
    const int32_t t_prev = dpp_i32<kDppWaveShr1>(tl);
    const bool cont = active && lane > 0 && single && h == t_prev;
    S acc[KR];
#pragma unroll
    for (int j = 0; j < KR; j++) acc[j] = active ? run[j] : S::zero();
    int flag = cont ? 0 : 1;
    const int rl = lane & 15;
    xscan_step_k<kDppRowShr1, 0xf>(acc, flag, rl >= 1);
    xscan_step_k<kDppRowShr2, 0xf>(acc, flag, rl >= 2);
    xscan_step_k<kDppRowShr4, 0xf>(acc, flag, rl >= 4);
    xscan_step_k<kDppRowShr8, 0xf>(acc, flag, rl >= 8);
    xscan_step_k<kDppRowBcast15, 0xa>(acc, flag, (lane & 16) != 0);
    xscan_step_k<kDppRowBcast31, 0xc>(acc, flag, lane >= 32);
    S prev_acc[KR];
#pragma unroll
    for (int j = 0; j < KR; j++) prev_acc[j] = acc[j].template mv<kDppWaveShr1, 0xf>();
    const int32_t h_next = dpp_i32<kDppWaveShl1>(h);
    const int last_lane = (n - 1) / kBatch;
    if (active) {
      if (!single) {
        S hv[KR];
#pragma unroll
        for (int j = 0; j < KR; j++) {
          hv[j] = head[j];
          if (lane > 0 && h == t_prev) hv[j].merge(prev_acc[j]);
        }
        put(h, hv);
      }
      if (lane == last_lane || h_next != tl) put(tl, acc);
    }
    if (lane < nr && ro[lane] == ro[lane + 1]) {
      S z[KR];
#pragma unroll
      for (int j = 0; j < KR; j++) z[j] = S::zero();
      put(lane, z);
    }
  }
}

// The x-tiled pass of a direct plan (no copy of any value): the columns come
// from the column copy (stored in item order), the values from the caller's
// array.  Run r of a sub-block holds its block items ro[r] .. ro[r + 1] - 1,
// the first at position rp[r] of the caller's arrays (a run is a contiguous
// piece of its row).  Lane l takes items l, l + 64, ..., l + 448 of the
// sub-block, so that each load instruction reads 64 consecutive items --
// consecutive positions inside a run -- instead of items eight apart; an
// item's run comes from a binary search over the run offsets in LDS.  The
// runs are then reduced chunk by chunk (items 64 c .. 64 c + 63): a
// segmented scan across the wave, a run's total written by the lane holding
// its last item, and a run that continues into the next chunk carried to
// that chunk's first lane.  LDS: x tile | run offsets | run positions (the
// sub-block records are read through scalar loads).
template <typename V>
struct XDLoad {
  uint32_t c[kBatch];
  int32_t r[kBatch];
  typename XTHeld<V>::T v[kBatch];
};

// chunk c (items 64 c + lane) of a direct plan's sub-block: column, run
// and value
template <typename V>
__device__ __forceinline__ void xtile_direct_load1(const XTArgs& t, const V* val,
                                                   const int32_t* ro, const int32_t* rp,
                                                   const XTSub& s0, int32_t nr, int32_t n,
                                                   int lane, int c, XDLoad<V>& d) {
  const int32_t j = c * kWave + lane;
  const int32_t E = s0.e0 + j;
  // the last run r < nr with ro[r] <= E (nr <= 64 runs)
  int32_t r = 0;
#pragma unroll
  for (int step = kWave / 2; step > 0; step >>= 1)
    if (r + step < nr && ro[r + step] <= E) r += step;
  d.r[c] = r;
  const int32_t pos = j < n ? rp[r] + (E - ro[r]) : 0;
  d.c[c] = (uint32_t)ld_stream(t.col + (int64_t)s0.p0 + (j < n ? j : 0));
  d.v[c] = XTHeld<V>::hold(ld_stream(val + pos));
}

template <typename S>
__device__ __forceinline__ S readlane_sum(const S& v, int l) {
  auto rl = [&](auto x) {
    using T = decltype(x);
    if constexpr (sizeof(T) == 8) {
      const uint64_t u = __builtin_bit_cast(uint64_t, x);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
      return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
      return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
    }
  };
  return S{rl(v.s), rl(v.c)};
}

template <typename V, bool COMP>
__global__ __launch_bounds__(kXTileThreads) void xtile_direct_kernel(RBArgs<V> a, XTArgs t) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  using S = XSum<Acc, COMP>;
  constexpr int kWaves = kXTileThreads / kWave;
  extern __shared__ __align__(16) unsigned char smem[];
  Vec* xs = reinterpret_cast<Vec*>(smem);
  int32_t* s_lro = reinterpret_cast<int32_t*>(xs + t.tile);
  int32_t* s_rpos = s_lro + kXTileMaxGroupRows + 2;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint32_t id = blockIdx.x;
  const uint32_t k = id / kNumXcd;
  const int32_t g = (int32_t)(k % (uint32_t)t.groups);
  const int32_t tile = (int32_t)((k / (uint32_t)t.groups) * kNumXcd + id % kNumXcd);
  if (tile >= t.ntiles) return;  // whole workgroup: no barrier reached
  const int64_t c0 = (int64_t)tile * t.tile;
  const int32_t nx = (int32_t)min<int64_t>(t.tile, a.ncols - c0);
  const int64_t blk = (int64_t)g * t.ntiles + tile;
  const int32_t* lro = t.lro + blk * (t.gr + 1);
  const int32_t* rpo = t.rpos + blk * (t.gr + 1);
  const int32_t gstart = g * t.gr;
  const int32_t sb0 = t.sub_start[blk];
  const int32_t nsub = t.sub_start[blk + 1] - 1 - sb0;  // sub-blocks (a sentinel follows)
  const V* val = static_cast<const V*>(t.val);
  Acc* part = static_cast<Acc*>(t.partial);
  {
    constexpr int kPer = 8;  // x loads in flight per thread
    for (int32_t j0 = threadIdx.x; j0 < nx; j0 += kXTileThreads * kPer) {
      Vec xv[kPer];
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int32_t j = j0 + q * kXTileThreads;
        xv[q] = j < nx ? a.x[(c0 + j) * a.ld] : (Vec)0;
      }
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int32_t j = j0 + q * kXTileThreads;
        if (j < nx) xs[j] = xv[q];
      }
    }
  }
  for (int32_t j = threadIdx.x; j <= t.gr; j += kXTileThreads) {
    s_lro[j] = lro[j];
    s_rpos[j] = rpo[j];
  }
  __syncthreads();

  // (a wave-uniform sub-block index: the records come through scalar loads)
  auto rec = [&](int32_t q) { return uniform_sub(t.sub[__builtin_amdgcn_readfirstlane(sb0 + q)]); };
  XTSub s0{0, 0}, s1{0, 0};
  XDLoad<V> cur;
  if (wave < nsub) {
    s0 = rec(wave);
    s1 = rec(wave + 1);
#pragma unroll
    for (int c = 0; c < kBatch; c++)
      xtile_direct_load1<V>(t, val, s_lro + (s0.i0 - gstart), s_rpos + (s0.i0 - gstart), s0,
                            s1.i0 - s0.i0, s1.e0 - s0.e0, lane, c, cur);
  }
  for (int32_t q = wave; q < nsub; q += kWaves) {
    const int32_t nr = s1.i0 - s0.i0, n = s1.e0 - s0.e0;
    const int32_t i0 = s0.i0;
    const int32_t* ro = s_lro + (i0 - gstart);
    auto put = [&](int32_t r, const S& v) {
      const int64_t o = (int64_t)tile * t.rows + i0 + r;
      if constexpr (COMP) {
        part[2 * o] = v.s;
        part[2 * o + 1] = v.c;
      } else {
        part[o] = v.s;
      }
    };
    S carry = S::zero();
    int32_t carry_run = -1;
#pragma unroll
    for (int c = 0; c < kBatch; c++) {
      if (c * kWave >= n) break;  // uniform
      const int32_t j = c * kWave + lane;
      const bool valid = j < n;
      const int32_t r = cur.r[c];
      // (the clamp keeps a corrupted column inside the LDS tile)
      const Acc xv = (Acc)xs[min(cur.c[c], (uint32_t)(nx - 1))];
      S acc = S::zero();
      if (valid) {
        if constexpr (COMP)
          acc.add_prod(XTHeld<V>::use(cur.v[c]), xv);
        else
          acc.add(XTHeld<V>::use(cur.v[c]) * xv);
      }
      if (lane == 0 && valid && r == carry_run) {
        S tmp = carry;
        tmp.merge(acc);
        acc = tmp;
      }
      const int32_t r_prev = dpp_i32<kDppWaveShr1>(r);
      int flag = (lane == 0 || !valid || r_prev != r) ? 1 : 0;
      const int rl = lane & 15;
      xscan_step<kDppRowShr1, 0xf>(acc, flag, rl >= 1);
      xscan_step<kDppRowShr2, 0xf>(acc, flag, rl >= 2);
      xscan_step<kDppRowShr4, 0xf>(acc, flag, rl >= 4);
      xscan_step<kDppRowShr8, 0xf>(acc, flag, rl >= 8);
      xscan_step<kDppRowBcast15, 0xa>(acc, flag, (lane & 16) != 0);
      xscan_step<kDppRowBcast31, 0xc>(acc, flag, lane >= 32);
      // the run of the next chunk's first item (uniform; -1 past the items)
      const int32_t next_run = (c + 1 < kBatch && (c + 1) * kWave < n)
                                   ? __builtin_amdgcn_readlane(cur.r[c + 1 < kBatch ? c + 1 : c], 0)
                                   : -1;
      const int32_t r_next = dpp_i32<kDppWaveShl1>(r);
      const bool last = valid && (lane == kWave - 1 ? next_run != r : (j + 1 >= n || r_next != r));
      if (last) put(r, acc);
      carry = readlane_sum(acc, kWave - 1);
      carry_run = (c * kWave + kWave - 1 < n && next_run == __builtin_amdgcn_readlane(r, kWave - 1))
                      ? next_run
                      : -1;
    }
    // runs without items in this tile
    if (lane < nr && ro[lane] == ro[lane + 1]) put(lane, S::zero());
    // the next sub-block's items, requested once this one is reduced (its
    // loads and this one's products are never live together; requested per
    // chunk between the reductions, they made the pass slower: 76 -> 85 us
    // on Z15)
    if (q + kWaves < nsub) {
      s0 = rec(q + kWaves);
      s1 = rec(q + kWaves + 1);
#pragma unroll
      for (int c = 0; c < kBatch; c++)
        xtile_direct_load1<V>(t, val, s_lro + (s0.i0 - gstart), s_rpos + (s0.i0 - gstart), s0,
                              s1.i0 - s0.i0, s1.e0 - s0.e0, lane, c, cur);
    }
  }
}

// Analysis: the tile-major copy.  Thread (tile q, long row i) copies the
// row's items [bnd(q, i), bnd(q + 1, i)) to its run of the (group, tile)
// block, placing item j of a sub-block at slot p0 + (j % 8) * S + j / 8.
template <typename V>
__global__ void xtile_copy_kernel(const int32_t* __restrict__ bnd, int32_t nl, int32_t ntiles,
                                  int32_t gr, int64_t tile, const int32_t* __restrict__ lro,
                                  const int32_t* __restrict__ run_sub,
                                  const XTSub* __restrict__ sub, const int32_t* __restrict__ ci,
                                  const V* __restrict__ val, uint16_t* __restrict__ col_out,
                                  V* __restrict__ val_out, int direct) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nl * ntiles) return;
  const int64_t q = t / nl, i = t % nl;
  const int64_t g = i / gr;
  const int64_t w = (g * ntiles + q) * (gr + 1) + (i - g * gr);
  const int32_t b = bnd[q * nl + i], e = bnd[(q + 1) * nl + i];
  const XTSub s0 = sub[run_sub[w]], s1 = sub[run_sub[w] + 1];
  const int32_t n = s1.e0 - s0.e0;
  const int32_t S = (n + 7) / 8;
  const int32_t first = lro[w] - s0.e0;  // the run's first item within the sub-block
  for (int32_t j = b; j < e; j++) {
    const int32_t L = first + (j - b);
    // a direct plan keeps the columns in item order (its pass gives lane l
    // the items l, l + 64, ...), the copy plan transposed
    const int64_t slot = (int64_t)s0.p0 + (direct ? (int64_t)L : (int64_t)(L % 8) * S + L / 8);
    col_out[slot] = (uint16_t)(ci[j] - q * tile);
    if (val_out) val_out[slot] = val[j];
  }
}

// Combine: a 256-thread block adds the tile partials of 32 long rows; thread
// (range, row) adds its row's partials over one of 8 contiguous tile ranges
// (loads coalesced across the 32 rows), then the range sums are added in
// range order with the row's DIA / ELL part -- a fixed order, so results are
// reproducible.
constexpr int kXCombRows = 32;
constexpr int kXCombRanges = 256 / kXCombRows;

template <typename V, int PRE, bool COMP>
__global__ __launch_bounds__(256) void xtile_combine_kernel(RBArgs<V> a, XTArgs t) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_sum[kXCombRanges][kXCombRows];
  __shared__ Acc s_err[kXCombRanges][kXCombRows];
  const int ri = threadIdx.x % kXCombRows, rg = threadIdx.x / kXCombRows;
  const int64_t i = (int64_t)blockIdx.x * kXCombRows + ri;
  const Acc* part = static_cast<const Acc*>(t.partial);
  const int32_t per = (t.ntiles + kXCombRanges - 1) / kXCombRanges;
  const int32_t q0 = rg * per, q1 = min(t.ntiles, q0 + per);
  csum<Acc> cs{0, 0};
  Acc s = 0;
  if (i < t.rows) {
    for (int32_t qb = q0; qb < q1; qb += kBatch) {
      Acc ps[kBatch], pc[kBatch];
#pragma unroll
      for (int b = 0; b < kBatch; b++) {
        const int64_t o = (int64_t)(qb + b) * t.rows + i;
        if constexpr (COMP) {
          ps[b] = qb + b < q1 ? part[2 * o] : (Acc)0;
          pc[b] = qb + b < q1 ? part[2 * o + 1] : (Acc)0;
        } else {
          ps[b] = qb + b < q1 ? part[o] : (Acc)0;
          pc[b] = 0;
        }
      }
#pragma unroll
      for (int b = 0; b < kBatch; b++) {
        if (qb + b < q1) {
          if constexpr (COMP)
            cs.combine(ps[b], pc[b]);
          else
            s += ps[b];
        }
      }
    }
  }
  s_sum[rg][ri] = COMP ? cs.s : s;
  s_err[rg][ri] = COMP ? cs.c : (Acc)0;
  __syncthreads();
  if (rg != 0 || i >= t.rows) return;
  const int64_t r = t.row[i];
  if constexpr (COMP) {
    csum<Acc> tot{0, 0};
    row_prefix_comp<V, PRE, 1>(a, r, &tot);
    for (int q = 0; q < kXCombRanges; q++) tot.combine(s_sum[q][ri], s_err[q][ri]);
    store_y_cs(a.y, r * a.ld, tot, a.init);
  } else {
    Acc pre[1] = {0};
    row_prefix<V, PRE, 1>(a, r, pre);
    Acc tot = 0;
    for (int q = 0; q < kXCombRanges; q++) tot += s_sum[q][ri];
    store_y(a.y, r * a.ld, tot + pre[0], a.init);
  }
}

// The combine of the k-wide pass: xtile_combine_kernel (no DIA / ELL part)
// for each of the KR columns, in the same order, into y[r * ld + j].
template <typename V, bool COMP, int KR>
__global__ __launch_bounds__(256) void xtile_combine_k_kernel(RBArgs<V> a, XTArgs t,
                                                              const void* partk) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_sum[KR][kXCombRanges][kXCombRows];
  __shared__ Acc s_err[KR][kXCombRanges][kXCombRows];
  const int ri = threadIdx.x % kXCombRows, rg = threadIdx.x / kXCombRows;
  const int64_t i = (int64_t)blockIdx.x * kXCombRows + ri;
  const Acc* part = static_cast<const Acc*>(partk);
  const int32_t per = (t.ntiles + kXCombRanges - 1) / kXCombRanges;
  const int32_t q0 = rg * per, q1 = min(t.ntiles, q0 + per);
  csum<Acc> cs[KR];
  Acc s[KR];
#pragma unroll
  for (int j = 0; j < KR; j++) {
    cs[j] = csum<Acc>{0, 0};
    s[j] = 0;
  }
  if (i < t.rows) {
    for (int32_t qb = q0; qb < q1; qb += kBatch) {
      Acc ps[kBatch][KR], pc[kBatch][KR];
#pragma unroll
      for (int b = 0; b < kBatch; b++) {
        const int64_t o = ((int64_t)(qb + b) * t.rows + i) * KR;
#pragma unroll
        for (int j = 0; j < KR; j++) {
          if constexpr (COMP) {
            ps[b][j] = qb + b < q1 ? part[2 * (o + j)] : (Acc)0;
            pc[b][j] = qb + b < q1 ? part[2 * (o + j) + 1] : (Acc)0;
          } else {
            ps[b][j] = qb + b < q1 ? part[o + j] : (Acc)0;
            pc[b][j] = 0;
          }
        }
      }
#pragma unroll
      for (int b = 0; b < kBatch; b++) {
        if (qb + b < q1) {
#pragma unroll
          for (int j = 0; j < KR; j++) {
            if constexpr (COMP)
              cs[j].combine(ps[b][j], pc[b][j]);
            else
              s[j] += ps[b][j];
          }
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < KR; j++) {
    s_sum[j][rg][ri] = COMP ? cs[j].s : s[j];
    s_err[j][rg][ri] = COMP ? cs[j].c : (Acc)0;
  }
  __syncthreads();
  if (rg != 0 || i >= t.rows) return;
  const int64_t r = t.row[i];
#pragma unroll
  for (int j = 0; j < KR; j++) {
    if constexpr (COMP) {
      csum<Acc> tot{0, 0};
      for (int q = 0; q < kXCombRanges; q++) tot.combine(s_sum[j][q][ri], s_err[j][q][ri]);
      store_y_cs(a.y, r * a.ld + j, tot, a.init);
    } else {
      Acc tot = 0;
      for (int q = 0; q < kXCombRanges; q++) tot += s_sum[j][q][ri];
      store_y(a.y, r * a.ld + j, tot + (Acc)0, a.init);
    }
  }
}

// Analysis: unsorted[i] flags a long row whose columns are not sorted (one
// workgroup per row, its threads striding over the row's adjacent pairs).
__global__ __launch_bounds__(256) void xtile_sorted_kernel(const int32_t* __restrict__ ro,
                                                           const int32_t* __restrict__ ci,
                                                           const int32_t* __restrict__ rows,
                                                           int32_t* __restrict__ unsorted) {
  __shared__ int32_t s_bad;
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();
  const int32_t r = rows[blockIdx.x];
  const int32_t b = ro[r], e = ro[r + 1];
  int32_t bad = 0;
  for (int32_t j = b + 1 + (int32_t)threadIdx.x; j < e; j += (int32_t)blockDim.x)
    bad |= ci[j - 1] > ci[j] ? 1 : 0;
  if (bad) s_bad = 1;
  __syncthreads();
  if (threadIdx.x == 0) unsorted[blockIdx.x] = s_bad;
}

// Analysis: per long row i and tile q = 0..ntiles, the first item whose
// column is >= q * tile (lower bound in the sorted columns), stored
// tile-major at out[q * nl + i].
__global__ void xtile_bounds_kernel(const int32_t* __restrict__ ro,
                                    const int32_t* __restrict__ ci,
                                    const int32_t* __restrict__ rows, int32_t nl,
                                    int32_t ntiles, int64_t tile, int32_t* __restrict__ out,
                                    int32_t* __restrict__ unsorted) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nl * (ntiles + 1)) return;
  const int64_t q = t / nl;
  const int64_t i = t % nl;
  const int32_t r = rows[i];
  const int32_t b = ro[r], e = ro[r + 1];
  if (q == 0) {
    out[t] = b;
    return;
  }
  if (q == ntiles) {
    out[t] = e;
    return;
  }
  const int64_t key = q * tile;
  int32_t lo = b, hi = e;
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)ci[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  out[t] = lo;
}

// --------------------------------------------------------------- k > 1
// SpMM form: KR right-hand sides, row-major X / Y with leading dimension ld.
// Each row is reduced by a group of lanes reading the row's items directly
// (coalesced within the row); every item feeds KR FMAs from one contiguous
// load of the x row.  The DIA / ELL prefix is split across the group's lanes.
// A long-row chunk of the SpMM in FP32 / BF16: Dot2 pairs per column, the
// block's pair per column stored at partial[2 * (b * KR + q)].
template <typename V, int KR>
__device__ void spmm_chunk_comp(const RBArgs<V>& a, uint32_t b, int64_t s, int64_t e,
                                typename vtraits<V>::acc* s_red) {
  using Acc = typename vtraits<V>::acc;
  csum<Acc> t[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) t[q] = csum<Acc>{0, 0};
  for (int64_t i = s + threadIdx.x; i < e; i += kRowBlockThreads) {
    const Acc v = vtraits<V>::widen(ld_stream(a.val + i));
    const int64_t c = ld_stream(a.ci + i);
#pragma unroll
    for (int q = 0; q < KR; q++) t[q].add_prod(v, (Acc)a.x[c * a.ld + q]);
  }
#pragma unroll
  for (int q = 0; q < KR; q++)
    for (int m = kWave >> 1; m > 0; m >>= 1) {
      const Acc s2 = shfl_xor(t[q].s, m), c2 = shfl_xor(t[q].c, m);
      t[q].combine(s2, c2);
    }
  constexpr int kW = kRowBlockThreads / kWave;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < KR; q++) {
      s_red[2 * (q * kW + wave)] = t[q].s;
      s_red[2 * (q * kW + wave) + 1] = t[q].c;
    }
  __syncthreads();
  if (threadIdx.x < KR) {
    const int q = threadIdx.x;
    csum<Acc> u{s_red[2 * q * kW], s_red[2 * q * kW + 1]};
    for (int w = 1; w < kW; w++) u.combine(s_red[2 * (q * kW + w)], s_red[2 * (q * kW + w) + 1]);
    a.partial[2 * ((int64_t)b * KR + q)] = u.s;
    a.partial[2 * ((int64_t)b * KR + q) + 1] = u.c;
  }
}

// FP32 / BF16: the long-row chunks of the SpMM in a launch of their own, so
// that the Dot2 pairs do not raise the register count of the short-row
// kernel (every record but the chunks returns at once).
template <typename V, int KR>
__global__ __launch_bounds__(kRowBlockThreads) void rowblock_spmm_chunk_kernel(RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[2 * KR * kRowBlockThreads / kWave];
  const uint32_t b = blockIdx.x;
  const RowBlock blk = a.blocks[b];
  if (blk.sub < 0) return;
  const int64_t r = blk.row;
  const int64_t rs = a.ro[r], re = a.ro[r + 1];
  const int64_t s = rs + (int64_t)blk.sub * a.chunk;
  spmm_chunk_comp<V, KR>(a, b, s, min(re, s + (int64_t)a.chunk), s_red);
}

template <typename V, int PRE, int KR>
__global__ __launch_bounds__(kRowBlockThreads) void rowblock_spmm_kernel(
    RBArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  constexpr bool COMP = std::is_same<Acc, float>::value;
  __shared__ Acc s_red[KR * kRowBlockThreads / kWave];
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const RowBlock blk = a.blocks[b];
  if (blk.sub < -1) return;  // a long row handled by the x-tiled pass
  if (blk.sub >= 0) {
    if constexpr (COMP) return;  // rowblock_spmm_chunk_kernel
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
  // lanes per row: the analysis' choice for the block (from its longest rows
  // per wave, as the SpMV pass uses it), else from the block's mean length
  // (which gives one lane to a 500-entry row among one-entry rows)
  int L = a.lanes_fixed > 0 ? a.lanes_fixed
          : a.krec         ? a.krec[4 * (int64_t)b + 3]
                           : choose_lanes(n, nr, 0);
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
  if constexpr (std::is_same<Acc, float>::value) {
    csum<Acc> cs[KR];
#pragma unroll
    for (int q = 0; q < KR; q++) cs[q] = csum<Acc>{0, 0};
    row_prefix_comp<V, PRE, KR>(a, r, cs);
    for (int64_t j = 0; j < nsub; j++)
#pragma unroll
      for (int q = 0; q < KR; q++)
        cs[q].combine(a.partial[2 * ((b + j) * KR + q)], a.partial[2 * ((b + j) * KR + q) + 1]);
#pragma unroll
    for (int q = 0; q < KR; q++) store_y_cs(a.y, r * a.ld + q, cs[q], a.init);
    return;
  }
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
// A row longer than long_len (<= chunk) is long: ceil(len / chunk) chunk
// blocks of its own.  Short blocks start at multiples of kRowBlockMaxRows
// rows, where the item window of `chunk` items changes, and after a long row.
__device__ __forceinline__ bool rb_is_start(const int32_t* ro, int64_t r,
                                            int32_t chunk, int32_t long_len,
                                            bool& is_long, int64_t& len) {
  len = (int64_t)ro[r + 1] - ro[r];
  is_long = len > long_len;
  if (r == 0 || is_long) return true;
  const int64_t plen = (int64_t)ro[r] - ro[r - 1];
  return (r % kRowBlockMaxRows) == 0 || (ro[r] / chunk) != (ro[r - 1] / chunk) ||
         plen > long_len;
}

__global__ void rb_count_kernel(const int32_t* __restrict__ ro, int64_t nrows,
                                int32_t chunk, int32_t long_len,
                                int32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  bool lg;
  int64_t len;
  const bool st = rb_is_start(ro, r, chunk, long_len, lg, len);
  cnt[r] = st ? (lg ? (int32_t)((len + chunk - 1) / chunk) : 1) : 0;
}

__global__ void rb_fill_kernel(const int32_t* __restrict__ ro, int64_t nrows,
                               int32_t chunk, int32_t long_len,
                               const int32_t* __restrict__ first,
                               RowBlock* __restrict__ blocks) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  // the end sentinel is written whether or not the last row starts a block
  if (r == nrows - 1) blocks[first[nrows]] = RowBlock{(int32_t)nrows, -1};
  const int32_t f = first[r], c = first[r + 1] - f;
  if (c == 0) return;
  bool lg;
  int64_t len;
  rb_is_start(ro, r, chunk, long_len, lg, len);
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

// HYB merged copy, counts: per row its ELL slots that hold a column (in any
// position, as the ELL product tests every slot) plus its COO entries.
__global__ void hyb_merge_count_kernel(const int32_t* __restrict__ eci, int64_t pitch,
                                       int64_t loop, int64_t nrows,
                                       const int32_t* __restrict__ tro,
                                       int32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int32_t c = tro[r + 1] - tro[r];
  for (int64_t n = 0; n < loop; n++) c += eci[r + pitch * n] >= 0 ? 1 : 0;
  cnt[r] = c;
}

// HYB merged copy, fill: row r's ELL entries in slot order, then its COO
// entries in storage order (the order of the Serial HYB product, an ELL pass
// then a COO pass), into [mro[r], mro[r + 1]).  The ELL part: one thread per
// row; the COO part: one thread per entry (entry j of row r lands at
// mro[r + 1] - tro[r + 1] + j, the row's range ending with its COO entries).
template <typename T>
__global__ void hyb_merge_fill_kernel(const int32_t* __restrict__ eci, const T* __restrict__ ev,
                                      int64_t pitch, int64_t loop, int64_t nrows,
                                      const int32_t* __restrict__ mro, int32_t* __restrict__ mci,
                                      T* __restrict__ mv) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int64_t o = mro[r];
  for (int64_t n = 0; n < loop; n++) {
    const int32_t c = eci[r + pitch * n];
    if (c >= 0) {
      mci[o] = c;
      mv[o] = ev[r + pitch * n];
      o++;
    }
  }
}

template <typename T>
__global__ void hyb_merge_tail_kernel(const int32_t* __restrict__ cri, int64_t nnz,
                                      const int32_t* __restrict__ tro,
                                      const int32_t* __restrict__ cci, const T* __restrict__ cv,
                                      const int32_t* __restrict__ mro, int32_t* __restrict__ mci,
                                      T* __restrict__ mv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  const int32_t r = cri[j];
  const int64_t o = (int64_t)mro[r + 1] - tro[r + 1] + j;
  mci[o] = cci[j];
  mv[o] = cv[j];
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
// Lanes per row for the reduction of one short block (rows of lengths
// ro[r + 1] - ro[r], r < nr): the power of two 1..64 with the least modelled
// time of the block's slowest wave.  Groups of L lanes take rows g,
// g + 256 / L, ...; each row iteration of a wave costs the longest trip
// count ceil(len / L) among its groups' rows (an LDS read and an add per
// trip), 1.5 per shuffle step (log2 L) and 3 of row overhead (offsets,
// store).  Regular short rows get few lanes (S7, 7 per row: 1), blocks with
// a few long rows among short ones more, so that one row does not hold a
// whole wave.
uint8_t block_lanes(int nr, const int32_t* ro) {
  constexpr int kWaves = kRowBlockThreads / kWave;
  int best_l = 1;
  double best = 0;
  for (int L = 1, lg = 0; L <= kWave; L <<= 1, lg++) {
    const int groups = kRowBlockThreads / L, per_wave = kWave / L;
    const int iters = (nr + groups - 1) / groups;
    double worst = 0;
    for (int w = 0; w < kWaves; w++) {
      double c = 0;
      for (int k = 0; k < iters; k++) {
        int mx = -1;
        for (int g = w * per_wave; g < (w + 1) * per_wave; g++) {
          const int r = k * groups + g;
          if (r < nr) mx = std::max(mx, (ro[r + 1] - ro[r] + L - 1) / L);
        }
        if (mx >= 0) c += mx + 1.5 * lg + 3;
      }
      worst = std::max(worst, c);
    }
    if (L == 1 || worst < best) {
      best = worst;
      best_l = L;
    }
  }
  return (uint8_t)best_l;
}

int default_chunk(int vt) {
  return vt == MPH_F64 ? kDefaultChunkF64 : kDefaultChunkF32;
}

namespace {

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

namespace {

template <typename V>
void launch_xtile_copy(const int32_t* bnd, int32_t nl, int32_t ntiles, int32_t gr, int64_t tile,
                       const int32_t* lro, const int32_t* run_sub, const XTSub* sub,
                       const int32_t* ci, const void* val, uint16_t* col_out, void* val_out,
                       hipStream_t s) {
  const int64_t n = (int64_t)nl * ntiles;
  hipLaunchKernelGGL((xtile_copy_kernel<V>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                     bnd, nl, ntiles, gr, tile, lro, run_sub, sub, ci,
                     static_cast<const V*>(val), col_out, static_cast<V*>(val_out),
                     val_out ? 0 : 1);
}

}  // namespace

// The x-tiled plan of the long rows (see xtile_kernel), built on the host
// from the per-row tile bounds (computed on the device and checked here),
// then the tile-major copy is made on the device.
int xtile_build(mph_matrix_impl* A, const std::vector<int32_t>& ro, const int32_t* d_ro,
                const int32_t* d_ci, const void* d_val, std::vector<int32_t> lrows,
                XTilePlan& xp, std::vector<char>& tiled, hipStream_t s, bool direct) {
  const int64_t nrows = A->nrows;
  const size_t vec_bytes = A->vt == MPH_F64 ? 8 : 4;
  tiled.assign(nrows, 0);
  if (lrows.empty() || nrows == 0 || A->ncols == 0) return MPH_OK;
  // longest first, so neighbouring rows of a sub-block have similar runs
  std::stable_sort(lrows.begin(), lrows.end(), [&](int32_t u, int32_t v) {
    return ro[u + 1] - ro[u] > ro[v + 1] - ro[v];
  });
  // the per-row tile-bound table and the (tile, row) partials grow with
  // rows x tiles: past kXTileMaxBounds entries only the longest rows are
  // planned, the others keep the position chunks
  {
    const int64_t nt0 = ceil_div((int64_t)A->ncols, (int64_t)(kXTileBytes / vec_bytes));
    const int64_t cap = std::max<int64_t>(1, kXTileMaxBounds / (nt0 + 1));
    if ((int64_t)lrows.size() > cap) lrows.resize((size_t)cap);
  }
  const int32_t nl = (int32_t)lrows.size();
  const int32_t tile_max = (int32_t)(kXTileBytes / vec_bytes);
  int32_t tile = tile_max, ntiles = 0;

  // per-row tile bounds on the device for a tile of tl columns, checked here
  std::vector<int32_t> bnd, uns(nl);
  auto bounds = [&](int32_t tl) -> int {
    tile = tl;
    ntiles = (int32_t)ceil_div((int64_t)A->ncols, (int64_t)tile);
    bnd.assign((size_t)nl * (ntiles + 1), 0);
    int32_t *d_rows = nullptr, *d_bnd = nullptr, *d_uns = nullptr;
    int st = MPH_OK;
    if (hipMalloc(&d_rows, sizeof(int32_t) * nl) != hipSuccess ||
        hipMalloc(&d_bnd, sizeof(int32_t) * bnd.size()) != hipSuccess ||
        hipMalloc(&d_uns, sizeof(int32_t) * nl) != hipSuccess) {
      set_error("hipMalloc failed in the x-tile analysis");
      st = MPH_ERR_OUT_OF_MEMORY;
    }
    if (st == MPH_OK &&
        hipMemcpyAsync(d_rows, lrows.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, s) !=
            hipSuccess)
      st = MPH_ERR_HIP;
    if (st == MPH_OK) {
      hipLaunchKernelGGL(xtile_sorted_kernel, dim3((unsigned)nl), dim3(256), 0, s, d_ro, d_ci,
                         d_rows, d_uns);
      hipLaunchKernelGGL(xtile_bounds_kernel, dim3((unsigned)ceil_div((int64_t)bnd.size(), 256)),
                         dim3(256), 0, s, d_ro, d_ci, d_rows, nl, ntiles, (int64_t)tile, d_bnd,
                         d_uns);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(bnd.data(), d_bnd, sizeof(int32_t) * bnd.size(), hipMemcpyDeviceToHost,
                         s) != hipSuccess ||
          hipMemcpyAsync(uns.data(), d_uns, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("x-tile bounds kernel failed");
        st = MPH_ERR_HIP;
      }
    }
    dfree(d_rows);
    dfree(d_bnd);
    dfree(d_uns);
    if (st != MPH_OK) return st;
    for (int32_t i = 0; i < nl; i++) {
      const int32_t r = lrows[i];
      if (bnd[i] != ro[r] || bnd[(size_t)ntiles * nl + i] != ro[r + 1]) {
        set_error("x-tile analysis check failed: row bounds");
        return MPH_ERR_RUNTIME;
      }
      for (int32_t q = 0; q < ntiles; q++)
        if (bnd[(size_t)q * nl + i] > bnd[(size_t)(q + 1) * nl + i]) {
          set_error("x-tile analysis check failed: bounds not monotone");
          return MPH_ERR_RUNTIME;
        }
    }
    return MPH_OK;
  };
  auto B = [&](int32_t q, int32_t i) { return bnd[(size_t)q * nl + i]; };
  // a row takes the x-tiled pass when every one of its runs fits a
  // sub-block (a longer run, over an eighth of a tile's columns, leaves its
  // row to the position chunks)
  auto fits = [&](int32_t i) {
    int32_t mx = 0;
    for (int32_t q = 0; q < ntiles; q++) mx = std::max(mx, B(q + 1, i) - B(q, i));
    return mx <= kXTileCap;
  };

  MPH_CHECK_STATUS(bounds(tile_max));
  for (int32_t i = 0; i < nl; i++)
    if (uns[i]) return MPH_OK;  // unsorted columns: keep the position chunks
  // Rows dense enough that a run exceeds a sub-block: a tile of half the
  // columns halves their runs.  The tile is halved (down to an eighth of
  // the LDS tile) while more than an eighth of the candidate items would be
  // left to the position chunks; FP32 / BF16 tiles hold twice the columns
  // of FP64 ones, so a matrix whose longest rows fit FP64 tiles may not fit
  // FP32 ones.
  for (;;) {
    int64_t all = 0, left = 0;
    for (int32_t i = 0; i < nl; i++) {
      const int64_t len = (int64_t)ro[lrows[i] + 1] - ro[lrows[i]];
      all += len;
      if (!fits(i)) left += len;
    }
    if (8 * left <= all || tile <= tile_max / 8) break;
    if ((int64_t)nl * (2 * (int64_t)ntiles + 1) > kXTileMaxBounds) break;  // host table size
    MPH_CHECK_STATUS(bounds(tile / 2));
  }
  std::vector<int32_t> trow;     // indices into lrows
  for (int32_t i = 0; i < nl; i++) {
    if (fits(i)) {
      trow.push_back(i);
      tiled[lrows[i]] = 1;
    }
  }
  const int32_t nt = (int32_t)trow.size();
  if (nt == 0) return MPH_OK;

  // groups: about eight (group, tile) workgroups per CU (a group re-reads the
  // x tiles, which its XCD's L2 holds); rows are dealt to the groups by length
  // rank (group g gets ranks g, g + G, ...), so every group has the same mix
  const int64_t want = 8 * (int64_t)device_cu_count();
  int32_t ngroups = (int32_t)std::min<int64_t>(nt, std::max<int64_t>(1, ceil_div(want, (int64_t)ntiles)));
  ngroups = (int32_t)std::max<int64_t>(ngroups, ceil_div((int64_t)nt, (int64_t)kXTileMaxGroupRows));
  const int32_t gr = (int32_t)ceil_div((int64_t)nt, (int64_t)ngroups);
  ngroups = (int32_t)ceil_div((int64_t)nt, (int64_t)gr);
  const int64_t nblk = (int64_t)ngroups * ntiles;
  {
    std::vector<int32_t> dealt;
    dealt.reserve(nt);
    for (int32_t g = 0; g < ngroups; g++)
      for (int64_t rank = g; rank < nt; rank += ngroups) dealt.push_back(trow[rank]);
    trow.swap(dealt);
  }

  // run offsets, sub-blocks and the tiled rows' bounds (compacted)
  std::vector<int32_t> lro((size_t)nblk * (gr + 1));
  std::vector<int32_t> run_sub((size_t)nblk * (gr + 1), 0);
  // direct plans: the position of each run's first item (0 past the group)
  std::vector<int32_t> rpos(direct ? lro.size() : 0, 0);
  std::vector<XTSub> subs;
  std::vector<int32_t> sub_start(nblk + 1);
  std::vector<int32_t> tb((size_t)nt * (ntiles + 1));
  for (int32_t j = 0; j < nt; j++)
    for (int32_t q = 0; q <= ntiles; q++) tb[(size_t)q * nt + j] = B(q, trow[j]);
  int64_t total = 0;  // items
  int64_t slots = 0;  // copy slots (sub-blocks padded to multiples of 8)
  for (int32_t g = 0; g < ngroups; g++) {
    const int32_t j0 = g * gr, j1 = std::min(nt, j0 + gr);
    for (int32_t q = 0; q < ntiles; q++) {
      const int64_t k = (int64_t)g * ntiles + q;
      int32_t* L = &lro[(size_t)k * (gr + 1)];
      int32_t* RS = &run_sub[(size_t)k * (gr + 1)];
      sub_start[k] = (int32_t)subs.size();
      int32_t cur_rows = 0;
      int64_t cur_items = 0;
      for (int32_t j = j0; j < j1; j++) {
        const int32_t len = B(q + 1, trow[j]) - B(q, trow[j]);
        if (cur_rows == 0 || cur_rows == kXTileRows || cur_items + len > kXTileCap) {
          if (cur_rows > 0) slots += 8 * ceil_div(cur_items, (int64_t)8);
          subs.push_back(XTSub{j, (int32_t)total, (int32_t)slots, 0});
          cur_rows = 0;
          cur_items = 0;
        }
        L[j - j0] = (int32_t)total;
        RS[j - j0] = (int32_t)subs.size() - 1;
        if (direct) rpos[(size_t)k * (gr + 1) + (j - j0)] = B(q, trow[j]);
        total += len;
        cur_rows++;
        cur_items += len;
      }
      if (cur_rows > 0) slots += 8 * ceil_div(cur_items, (int64_t)8);
      for (int32_t j = j1; j <= j0 + gr; j++) L[j - j0] = (int32_t)total;
      subs.push_back(XTSub{j1, (int32_t)total, (int32_t)slots, 0});  // sentinel
      if (slots > INT32_MAX) {
        set_error("x-tile copy exceeds the int32 item range");
        return MPH_ERR_INVALID_INPUT;
      }
    }
  }
  sub_start[nblk] = (int32_t)subs.size();
  {
    int64_t expect = 0;
    for (int32_t j = 0; j < nt; j++) expect += ro[lrows[trow[j]] + 1] - ro[lrows[trow[j]]];
    if (expect != total) {
      set_error("x-tile analysis check failed: copy size");
      return MPH_ERR_RUNTIME;
    }
  }
  // per sub-block, the run of each lane's first item 8 * lane: the last run
  // r with lro[r] - e0 <= 8 * lane (0 for lanes past the items)
  std::vector<uint8_t> head(subs.size() * kWave, 0);
  for (int32_t g = 0; g < ngroups; g++) {
    const int32_t j0 = g * gr;
    for (int32_t q = 0; q < ntiles; q++) {
      const int64_t k = (int64_t)g * ntiles + q;
      const int32_t* L = &lro[(size_t)k * (gr + 1)];
      for (int32_t si = sub_start[k]; si + 1 < sub_start[k + 1]; si++) {
        const XTSub a = subs[si], b = subs[si + 1];
        const int32_t nr = b.i0 - a.i0, n = b.e0 - a.e0;
        if (nr <= 0 || nr > kXTileRows || n > kXTileCap || a.i0 < j0 || b.i0 > j0 + gr) {
          set_error("x-tile analysis check failed: sub-block shape");
          return MPH_ERR_RUNTIME;
        }
        int32_t r = 0;
        for (int32_t l = 0; l < kWave && l * kBatch < n; l++) {
          const int32_t jb = l * kBatch;
          while (r + 1 < nr && L[a.i0 - j0 + r + 1] - a.e0 <= jb) r++;
          head[(size_t)si * kWave + l] = (uint8_t)r;
        }
      }
    }
  }
  std::vector<int32_t> trow_id(nt);
  for (int32_t j = 0; j < nt; j++) trow_id[j] = lrows[trow[j]];

  // device side
  const size_t vsz = A->vt == MPH_F64 ? 8 : A->vt == MPH_F32 ? 4 : 2;
  const size_t acc = A->vt == MPH_F64 ? 8 : 4, pair = A->vt == MPH_F64 ? 1 : 2;
  int32_t *d_row = nullptr, *d_lro = nullptr, *d_ss = nullptr, *d_tb = nullptr, *d_rs = nullptr;
  int32_t* d_rpos = nullptr;
  XTSub* d_sub = nullptr;
  uint8_t* d_head = nullptr;
  uint16_t* d_col = nullptr;
  void *d_cv = nullptr, *d_part = nullptr;
  auto release = [&]() {
    dfree(d_row);
    dfree(d_lro);
    dfree(d_ss);
    dfree(d_tb);
    dfree(d_rs);
    dfree(d_sub);
    dfree(d_head);
    dfree(d_col);
    dfree(d_cv);
    dfree(d_part);
    dfree(d_rpos);
  };
  if (hipMalloc(&d_row, sizeof(int32_t) * nt) != hipSuccess ||
      hipMalloc(&d_lro, sizeof(int32_t) * lro.size()) != hipSuccess ||
      hipMalloc(&d_ss, sizeof(int32_t) * sub_start.size()) != hipSuccess ||
      hipMalloc(&d_tb, sizeof(int32_t) * tb.size()) != hipSuccess ||
      hipMalloc(&d_rs, sizeof(int32_t) * run_sub.size()) != hipSuccess ||
      hipMalloc(&d_sub, sizeof(XTSub) * subs.size()) != hipSuccess ||
      hipMalloc(&d_head, head.size()) != hipSuccess ||
      hipMalloc(&d_col, sizeof(uint16_t) * std::max<int64_t>(slots, 1)) != hipSuccess ||
      (!direct && hipMalloc(&d_cv, vsz * std::max<int64_t>(slots, 1)) != hipSuccess) ||
      (direct && hipMalloc(&d_rpos, sizeof(int32_t) * rpos.size()) != hipSuccess) ||
      hipMalloc(&d_part, acc * pair * (size_t)ntiles * nt) != hipSuccess) {
    release();
    set_error("hipMalloc failed in the x-tile analysis");
    return MPH_ERR_OUT_OF_MEMORY;
  }
  bool ok = hipMemcpyAsync(d_row, trow_id.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(d_lro, lro.data(), sizeof(int32_t) * lro.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(d_ss, sub_start.data(), sizeof(int32_t) * sub_start.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(d_tb, tb.data(), sizeof(int32_t) * tb.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(d_rs, run_sub.data(), sizeof(int32_t) * run_sub.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(d_sub, subs.data(), sizeof(XTSub) * subs.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(d_head, head.data(), head.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
            (!direct || hipMemcpyAsync(d_rpos, rpos.data(), sizeof(int32_t) * rpos.size(),
                                       hipMemcpyHostToDevice, s) == hipSuccess);
  if (ok) {
    if (A->vt == MPH_F64)
      launch_xtile_copy<double>(d_tb, nt, ntiles, gr, tile, d_lro, d_rs, d_sub, d_ci, d_val, d_col,
                                     d_cv, s);
    else if (A->vt == MPH_F32)
      launch_xtile_copy<float>(d_tb, nt, ntiles, gr, tile, d_lro, d_rs, d_sub, d_ci, d_val, d_col,
                                     d_cv, s);
    else
      launch_xtile_copy<uint16_t>(d_tb, nt, ntiles, gr, tile, d_lro, d_rs, d_sub, d_ci, d_val, d_col,
                                     d_cv, s);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  }
  if (!ok) {
    release();
    set_error("x-tile copy failed");
    return MPH_ERR_HIP;
  }
  dfree(d_tb);
  dfree(d_rs);
  // the x tile plus the sub-block products exceed the default 64 KiB of LDS;
  // the attribute is per kernel, so it is set for the largest tile (a plan
  // with a halved tile launches with less)
  {
    hipError_t e = hipSuccess;
    const size_t xlds = xtile_lds_bytes(vec_bytes, acc, tile_max, direct);
    auto attr = [&](const void* f) {
      return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)xlds);
    };
    if (A->vt == MPH_F64)
      e = direct ? attr(reinterpret_cast<const void*>(&xtile_direct_kernel<double, false>))
                 : attr(reinterpret_cast<const void*>(&xtile_kernel<double, false>));
    else if (A->vt == MPH_F32)
      e = direct ? attr(reinterpret_cast<const void*>(&xtile_direct_kernel<float, true>))
                 : attr(reinterpret_cast<const void*>(&xtile_kernel<float, true>));
    else
      e = direct ? attr(reinterpret_cast<const void*>(&xtile_direct_kernel<bf16, true>))
                 : attr(reinterpret_cast<const void*>(&xtile_kernel<bf16, true>));
    if (e != hipSuccess) {
      release();
      set_error(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
      return MPH_ERR_HIP;
    }
  }
  xp.row = d_row;
  xp.lro = d_lro;
  xp.sub = d_sub;
  xp.sub_start = d_ss;
  xp.head = d_head;
  xp.col = d_col;
  xp.val = d_cv;
  xp.partial = d_part;
  xp.rpos = d_rpos;
  xp.src = direct ? d_val : nullptr;
  xp.tile = tile;
  xp.ntiles = ntiles;
  xp.groups = ngroups;
  xp.gr = gr;
  xp.rows = nt;
  xp.items = total;
  return MPH_OK;
}

void xtile_release(XTilePlan& xp) {
  dfree(xp.row);
  dfree(xp.lro);
  dfree(xp.sub);
  dfree(xp.sub_start);
  dfree(xp.head);
  dfree(xp.col);
  dfree(xp.val);
  dfree(xp.partial);
  dfree(xp.partial_k);
  dfree(xp.rpos);
  xp = XTilePlan{};
}

namespace {

// The split SpMV plan of CSR / HYB / HDC, built on the host around the
// x-tiled rows:
// * x-tiling on: rows longer than kXTileMinRowLen go to the x-tiled pass;
// * HYB / HDC with x-tiling on or at least kSplitMinRows rows: the ELL / DIA
//   part of every row is computed first by its own coalesced pass, and the
//   rows whose tail / CSR part is non-empty form a compacted row list (list
//   row i is matrix row map[i]; its items start at offset lro[i], and
//   lro[i + 1] ends them, since the rows left out hold no items).
// Over the list (all rows for CSR), a tiled row becomes a {row, -2} record
// the kernel skips, a row longer than the chunk that the pass does not take
// keeps its position chunks, and the other rows form short blocks with the
// usual limits; the records are checked before upload.
// the longest row of offsets ro (a workgroup maximum, then one atomic)
__global__ __launch_bounds__(256) void max_row_len_kernel(const int32_t* __restrict__ ro,
                                                          int64_t nrows, int32_t* __restrict__ out) {
  __shared__ int32_t s_max;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  int32_t m = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (int64_t)gridDim.x * blockDim.x)
    m = max(m, ro[r + 1] - ro[r]);
  atomicMax(&s_max, m);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, s_max);
}

int max_row_len(const int32_t* d_ro, int64_t nrows, hipStream_t s, int32_t* out) {
  int32_t* d = nullptr;
  MPH_HIP_CHECK(hipMalloc(&d, sizeof(int32_t)));
  const int32_t zero = 0;
  hipError_t e = hipMemcpyAsync(d, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(max_row_len_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(nrows, 256), 2048)),
                       dim3(256), 0, s, d_ro, nrows, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  MPH_HIP_CHECK(e);
  return MPH_OK;
}

// per row, 1 + its last ELL slot holding a column (0 without any)
__global__ void ell_count_kernel(const int32_t* __restrict__ eci, int64_t pitch, int64_t loop,
                                 int64_t nrows, uint8_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int32_t c = 0;
  for (int64_t n = 0; n < loop; n++)
    if (eci[r + pitch * n] >= 0) c = (int32_t)n + 1;
  cnt[r] = (uint8_t)c;
}

// A split HYB's ELL pass visits each row's slots up to its last filled one
// when that at least halves the slots visited (p.ell_cnt); otherwise every
// row visits ell_loop slots.
int ell_counts(mph_matrix_impl* A, hipStream_t s) {
  RowBlockPlan& p = A->rb;
  const int64_t nrows = A->nrows, loop = A->ell_loop;
  if (loop <= 0 || loop > 255 || nrows == 0) return MPH_OK;
  uint8_t* d = nullptr;
  MPH_HIP_CHECK(hipMalloc(&d, (size_t)nrows));
  hipLaunchKernelGGL(ell_count_kernel, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, s,
                     A->ell_ci, A->ell_pitch, loop, nrows, d);
  std::vector<uint8_t> h(nrows);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, (size_t)nrows, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    dfree(d);
    set_error("ELL slot count analysis failed");
    return MPH_ERR_HIP;
  }
  int64_t visited = 0;
  for (int64_t r = 0; r < nrows; r++) visited += h[r];
  if (2 * visited > nrows * loop) {
    dfree(d);
    return MPH_OK;
  }
  p.ell_cnt = d;
  p.ell_visited = visited;
  return MPH_OK;
}

int split_analyse(mph_matrix_impl* A, const int32_t* d_ro, hipStream_t s) {
  RowBlockPlan& p = A->rb;
  const int64_t nrows = A->nrows;
  const size_t vec_bytes = A->vt == MPH_F64 ? 8 : 4;
  // (with MPH_POLICY_NO_VALUE_COPIES the x-tiled pass reads the values in
  // place: a direct plan)
  const int mode = A->tuning.xtile;
  const bool direct = (A->policy & MPH_POLICY_NO_VALUE_COPIES) != 0;
  const bool xt_on = mode == 2 || (mode == 0 && (int64_t)A->ncols * (int64_t)vec_bytes >= kXTileMinXBytes);
  // (an HDC without diagonals / a HYB without ELL slots is its CSR-shaped
  // part: no prefix pass)
  const bool prefixed = (A->format == MPH_HYB && !p.m_ro && A->ell_loop > 0) ||
                        (A->format == MPH_HDC && A->ndiag > 0);
  const bool split = prefixed && (xt_on || nrows >= kSplitMinRows || p.prefix_apart);
  if ((!xt_on && !split) || nrows == 0 || A->ncols == 0) return MPH_OK;
  if (xt_on && !split) {
    // nothing to tile: the plain row-block plan stands (no host pass over
    // the row offsets)
    int32_t longest = 0;
    MPH_CHECK_STATUS(max_row_len(d_ro, nrows, s, &longest));
    if (longest <= kXTileMinRowLen) return MPH_OK;
  }
  const int32_t* d_ci = p.ci;
  const void* d_val = p.val;

  std::vector<int32_t> ro(nrows + 1);
  MPH_HIP_CHECK(hipMemcpyAsync(ro.data(), d_ro, sizeof(int32_t) * (nrows + 1),
                               hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<char> tiled(nrows, 0);
  if (xt_on) {
    std::vector<int32_t> lrows;
    for (int64_t r = 0; r < nrows; r++)
      if ((int64_t)ro[r + 1] - ro[r] > kXTileMinRowLen) lrows.push_back((int32_t)r);
    if (!lrows.empty())
      MPH_CHECK_STATUS(xtile_build(A, ro, d_ro, d_ci, d_val, lrows, p.xt, tiled, s, direct));
  }
  if (p.xt.rows == 0 && !split) return MPH_OK;

  // the list rows: every row (CSR), or the rows with tail / CSR items
  std::vector<int32_t> map, lro;
  if (split) {
    for (int64_t r = 0; r < nrows; r++)
      if (ro[r + 1] > ro[r]) {
        map.push_back((int32_t)r);
        lro.push_back(ro[r]);
      }
    lro.push_back(ro[nrows]);
  }
  const std::vector<int32_t>& L = split ? lro : ro;
  const int64_t m = split ? (int64_t)map.size() : nrows;
  auto is_tiled = [&](int64_t i) { return tiled[split ? map[i] : i] != 0; };
  auto len = [&](int64_t i) { return (int64_t)L[i + 1] - L[i]; };

  const int64_t chunk = p.chunk, long_len = p.long_len;
  std::vector<RowBlock> sb;
  int32_t nchunked = 0;
  for (int64_t i = 0; i < m;) {
    if (is_tiled(i)) {
      sb.push_back(RowBlock{(int32_t)i, -2});
      i++;
    } else if (len(i) > long_len) {
      for (int64_t q = 0; q * chunk < len(i); q++) sb.push_back(RowBlock{(int32_t)i, (int32_t)q});
      nchunked++;
      i++;
    } else {
      const int64_t i0 = i;
      int64_t items = 0;
      while (i < m && !is_tiled(i) && len(i) <= long_len && i - i0 < kRowBlockMaxRows &&
             items + len(i) < 2 * chunk) {
        items += len(i);
        i++;
      }
      sb.push_back(RowBlock{(int32_t)i0, -1});
    }
  }
  sb.push_back(RowBlock{(int32_t)m, -1});
  // check: every list row once, block limits as the kernels assume, the
  // map strictly increasing over rows with items and lro its offsets
  {
    bool ok = true;
    for (int64_t i = 0; ok && split && i < m; i++)
      ok = map[i] >= 0 && map[i] < nrows && (i == 0 || map[i] > map[i - 1]) &&
           L[i] == ro[map[i]] && L[i + 1] == ro[map[i] + 1];
    int64_t next = 0;
    for (size_t b = 0; ok && b + 1 < sb.size(); b++) {
      const RowBlock k = sb[b];
      ok = k.row == next || (k.sub > 0 && k.row == next - 1);
      if (k.sub == -1) {
        const int64_t i1 = sb[b + 1].row;
        ok = ok && i1 > k.row && i1 - k.row <= kRowBlockMaxRows && (int64_t)L[i1] - L[k.row] < 2 * chunk;
        for (int64_t i = k.row; ok && i < i1; i++) ok = !is_tiled(i) && len(i) <= long_len;
        next = i1;
      } else if (k.sub == -2) {
        ok = ok && is_tiled(k.row);
        next = k.row + 1;
      } else {
        ok = ok && !is_tiled(k.row) && len(k.row) > long_len && (int64_t)k.sub * chunk < len(k.row) &&
             (k.sub == 0 || (sb[b - 1].row == k.row && sb[b - 1].sub == k.sub - 1));
        next = k.row + 1;
      }
    }
    if (!ok || next != m) {
      xtile_release(p.xt);
      set_error("split SpMV analysis check failed: short-block list");
      return MPH_ERR_RUNTIME;
    }
  }
  std::vector<int32_t> kr(4 * sb.size());
  for (size_t b = 0; b < sb.size(); b++) {
    kr[4 * b] = sb[b].row;
    kr[4 * b + 1] = sb[b].sub;
    kr[4 * b + 2] = L[sb[b].row];
    kr[4 * b + 3] = b + 1 < sb.size() && sb[b].sub == -1
                        ? block_lanes(sb[b + 1].row - sb[b].row, L.data() + sb[b].row)
                        : 0;
  }
  RowBlock* d_sb = nullptr;
  int32_t* d_kr = nullptr;
  void* d_part = nullptr;
  int32_t* d_map = nullptr;
  int32_t* d_lro = nullptr;
  const int accs = (A->vt == MPH_F64 ? 8 : 4);
  const bool grow = (int64_t)sb.size() > (int64_t)p.nblocks + 1;
  bool fail = hipMalloc(&d_sb, sizeof(RowBlock) * sb.size()) != hipSuccess ||
              hipMalloc(&d_kr, sizeof(int32_t) * kr.size()) != hipSuccess ||
              (grow && hipMalloc(&d_part, (size_t)accs * 32 * sb.size()) != hipSuccess);
  if (!fail && split)
    fail = hipMalloc(&d_map, sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)) != hipSuccess ||
           hipMalloc(&d_lro, sizeof(int32_t) * (size_t)(m + 1)) != hipSuccess;
  if (fail) {
    dfree(d_sb);
    dfree(d_kr);
    dfree(d_part);
    dfree(d_map);
    dfree(d_lro);
    xtile_release(p.xt);
    set_error("hipMalloc failed in the split SpMV analysis");
    return MPH_ERR_OUT_OF_MEMORY;
  }
  fail = hipMemcpyAsync(d_sb, sb.data(), sizeof(RowBlock) * sb.size(), hipMemcpyHostToDevice, s) !=
             hipSuccess ||
         hipMemcpyAsync(d_kr, kr.data(), sizeof(int32_t) * kr.size(), hipMemcpyHostToDevice, s) !=
             hipSuccess;
  if (!fail && split && m > 0)
    fail = hipMemcpyAsync(d_map, map.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s) !=
           hipSuccess;
  if (!fail && split)
    fail = hipMemcpyAsync(d_lro, lro.data(), sizeof(int32_t) * (m + 1), hipMemcpyHostToDevice, s) !=
           hipSuccess;
  if (fail || hipStreamSynchronize(s) != hipSuccess) {
    dfree(d_sb);
    dfree(d_kr);
    dfree(d_part);
    dfree(d_map);
    dfree(d_lro);
    xtile_release(p.xt);
    set_error("split SpMV analysis: upload failed");
    return MPH_ERR_HIP;
  }
  if (grow) {
    // the chunk partials are indexed by the record position in either list
    dfree(p.partial);
    p.partial = d_part;
  }
  p.sblocks = d_sb;
  p.skrec = d_kr;
  p.nsblocks = (int32_t)sb.size() - 1;
  p.nchunked = nchunked;
  p.split = split;
  {
    // FP64 HDC whose CSR rows are all short, none tiled or chunked
    // (MPH_HDC_TAIL=0 keeps the list pass)
    const char* env = std::getenv("MPH_HDC_TAIL");
    bool ok = split && A->format == MPH_HDC && A->vt == MPH_F64 && nchunked == 0 &&
              p.xt.rows == 0 && !(env && std::atoi(env) == 0);
    for (int64_t i = 0; ok && i < m; i++) ok = len(i) <= kHdcTailMax;
    p.hdc_tail = ok;
  }
  p.smap = d_map;
  p.sro = d_lro;
  if (split && A->format == MPH_HYB) MPH_CHECK_STATUS(ell_counts(A, s));
  return MPH_OK;
}

}  // namespace

namespace {

// A HYB with a row-sorted COO part, from kSplitMinRows rows or with x-tiling
// on: when its ELL slots are filled below kHybMergeMaxFill (an ELL pass would
// read mostly padding), or when more than 1 / kHybMergeTailRows of its rows
// have COO entries (the tail pass would be a second pass over most of the
// row structure), the product runs over a merged CSR copy (p.m_*).  Counts
// first; the copy is made only when one of the two holds.
int hyb_merge(mph_matrix_impl* A, hipStream_t s) {
  RowBlockPlan& p = A->rb;
  const int64_t nrows = A->nrows, loop = A->ell_loop;
  const size_t vec_bytes = A->vt == MPH_F64 ? 8 : 4;
  const int mode = xtile_mode(A);
  const bool xt_on = mode == 2 || (mode == 0 && (int64_t)A->ncols * (int64_t)vec_bytes >= kXTileMinXBytes);
  if (A->coo_nnz == 0 || nrows == 0 || loop == 0) return MPH_OK;
  if (A->policy & MPH_POLICY_NO_VALUE_COPIES) return MPH_OK;  // the ELL / tail passes instead
  if (!xt_on && nrows < kSplitMinRows) return MPH_OK;
  if (nrows * loop + A->coo_nnz >= (int64_t)INT32_MAX) return MPH_OK;  // int32 offsets
  int32_t* cnt = nullptr;
  int32_t* mro = nullptr;
  int64_t* tmp = nullptr;
  auto fail = [&](int st) {
    dfree(cnt);
    dfree(mro);
    dfree(tmp);
    return st;
  };
  if (hipMalloc(&cnt, sizeof(int32_t) * nrows) != hipSuccess ||
      hipMalloc(&mro, sizeof(int32_t) * (nrows + 1)) != hipSuccess ||
      hipMalloc(&tmp, sizeof(int64_t) * scan_tmp_elems(nrows)) != hipSuccess) {
    set_error("hipMalloc failed in the HYB merge analysis");
    return fail(MPH_ERR_OUT_OF_MEMORY);
  }
  const unsigned g = (unsigned)ceil_div(nrows, 256);
  hipLaunchKernelGGL(hyb_merge_count_kernel, dim3(g), dim3(256), 0, s, A->ell_ci, A->ell_pitch,
                     loop, nrows, p.tail_ro, cnt);
  if (hipGetLastError() != hipSuccess) {
    set_error("HYB merge count launch failed");
    return fail(MPH_ERR_HIP);
  }
  int st = exclusive_scan_i32(cnt, mro, nrows, tmp, s);
  if (st != MPH_OK) return fail(st);
  int32_t total = 0;
  if (hipMemcpyAsync(&total, mro + nrows, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("HYB merge analysis: copy failed");
    return fail(MPH_ERR_HIP);
  }
  const int64_t ell_valid = (int64_t)total - A->coo_nnz;
  if (ell_valid < 0 || ell_valid > nrows * loop) {
    set_error("HYB merge analysis check failed: counts");
    return fail(MPH_ERR_RUNTIME);
  }
  const bool sparse_ell = (double)ell_valid < kHybMergeMaxFill * (double)nrows * (double)loop;
  if (!sparse_ell) {
    std::vector<int32_t> tro(nrows + 1);
    if (hipMemcpyAsync(tro.data(), p.tail_ro, sizeof(int32_t) * (nrows + 1),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("HYB merge analysis: copy failed");
      return fail(MPH_ERR_HIP);
    }
    int64_t tail_rows = 0;
    for (int64_t r = 0; r < nrows; r++) tail_rows += tro[r + 1] > tro[r] ? 1 : 0;
    if (tail_rows * kHybMergeTailRows <= nrows) return fail(MPH_OK);
  }
  const size_t vs = value_size(A->vt);
  if (hipMalloc(&p.m_ci, sizeof(int32_t) * std::max<int64_t>(total, 1)) != hipSuccess ||
      hipMalloc(&p.m_val, vs * std::max<int64_t>(total, 1)) != hipSuccess) {
    dfree(p.m_ci);
    dfree(p.m_val);
    p.m_ci = nullptr;
    p.m_val = nullptr;
    set_error("hipMalloc failed for the HYB merged copy");
    return fail(MPH_ERR_OUT_OF_MEMORY);
  }
  auto fill = [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(hyb_merge_fill_kernel<T>, dim3(g), dim3(256), 0, s, A->ell_ci,
                       static_cast<const T*>(A->ell_v), A->ell_pitch, loop, nrows, mro, p.m_ci,
                       static_cast<T*>(p.m_val));
    hipLaunchKernelGGL(hyb_merge_tail_kernel<T>, dim3((unsigned)ceil_div(A->coo_nnz, 256)),
                       dim3(256), 0, s, A->coo_ri, A->coo_nnz, p.tail_ro, A->coo_ci,
                       static_cast<const T*>(A->coo_v), mro, p.m_ci, static_cast<T*>(p.m_val));
  };
  switch (vs) {
    case 8: fill(uint64_t{}); break;
    case 4: fill(uint32_t{}); break;
    default: fill(uint16_t{}); break;
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    set_error("HYB merge fill failed");
    return fail(MPH_ERR_HIP);
  }
  dfree(cnt);
  dfree(tmp);
  p.m_ro = mro;
  return MPH_OK;
}

}  // namespace

// Builds the row blocks of A's CSR-shaped part: CSR, HDC's CSR part, the
// HYB tail or a row-sorted COO matrix viewed through offsets built here, or
// a HYB's merged copy.
int rowblock_analyse(mph_matrix_impl* A, hipStream_t s) {
  RowBlockPlan& p = A->rb;
  const int64_t nrows = A->nrows;
  const int32_t* ro = A->csr_ro;
  p.ro = A->csr_ro;
  p.ci = A->csr_ci;
  p.val = A->csr_v;
  if (A->format == MPH_HYB || A->format == MPH_COO) {
    // (a COO matrix comes here only once its rows are known to be sorted)
    if (A->format == MPH_HYB) {
      int sorted = 1;
      MPH_CHECK_STATUS(rows_sorted(A->coo_ri, A->coo_nnz, nrows, s, &sorted));
      if (!sorted) return MPH_ERR_NOT_IMPLEMENTED;  // caller falls back
    }
    MPH_HIP_CHECK(hipMalloc(&p.tail_ro, sizeof(int32_t) * (nrows + 1)));
    hipLaunchKernelGGL(tail_offsets_kernel,
                       dim3((unsigned)ceil_div(A->coo_nnz + 1, 256)), dim3(256),
                       0, s, A->coo_ri, A->coo_nnz, nrows, p.tail_ro);
    MPH_HIP_CHECK(hipGetLastError());
    ro = p.ro = p.tail_ro;
    p.ci = A->coo_ci;
    p.val = A->coo_v;
    if (A->format == MPH_HYB) {
      MPH_CHECK_STATUS(hyb_merge(A, s));
      if (p.m_ro) {
        ro = p.ro = p.m_ro;
        p.ci = p.m_ci;
        p.val = p.m_val;
      }
    }
  }
  p.chunk = A->tuning.rows_chunk > 0 ? A->tuning.rows_chunk : default_chunk(A->vt);
  {
    // the ELL / DIA slots a row adds in a fused pass
    const int64_t prefix = (A->format == MPH_HYB && !p.m_ro) ? A->ell_loop
                           : A->format == MPH_HDC            ? A->ndiag
                                                             : 0;
    p.prefix_apart = A->vt != MPH_F64 && prefix >= kCompensateLen / 2;
    p.long_len = A->vt == MPH_F64 ? p.chunk
                                  : std::min<int64_t>(p.chunk, p.prefix_apart ? kCompensateLen
                                                                              : kCompensateLen - prefix);
  }
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
                     p.chunk, p.long_len, cnt);
  MPH_HIP_CHECK(hipGetLastError());
  int st = exclusive_scan_i32(cnt, first, nrows, tmp, s);
  int32_t nb = 0;
  if (st == MPH_OK) {
    MPH_HIP_CHECK(hipMemcpyAsync(&nb, first + nrows, sizeof(int32_t),
                                 hipMemcpyDeviceToHost, s));
    MPH_HIP_CHECK(hipStreamSynchronize(s));
    p.nblocks = nb;
    const int accs = (A->vt == MPH_F64 ? 8 : 4);
    // partials: up to 16 RHS per block (SpMM), as Dot2 pairs in FP32 / BF16
    MPH_HIP_CHECK(hipMalloc(&p.blocks, sizeof(RowBlock) * (nb + 1)));
    MPH_HIP_CHECK(hipMalloc(&p.partial, (size_t)accs * 32 * (nb + 1)));
    {
      hipLaunchKernelGGL(rb_fill_kernel, dim3(g), dim3(256), 0, s, ro, nrows,
                         p.chunk, p.long_len, first, p.blocks);
      MPH_HIP_CHECK(hipGetLastError());
      MPH_HIP_CHECK(hipStreamSynchronize(s));
      std::vector<int32_t> krec;
      st = validate_rowblock_plan(ro, nrows, p, s, &krec);
      if (st == MPH_OK && nb > 0) {
        if (hipMalloc(&p.krec, sizeof(int32_t) * krec.size()) != hipSuccess ||
            hipMemcpyAsync(p.krec, krec.data(), sizeof(int32_t) * krec.size(),
                           hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
          set_error("row-block analysis: kernel record upload failed");
          st = MPH_ERR_HIP;
        }
      }
    }
  }
  dfree(cnt);
  dfree(first);
  dfree(tmp);
  if (st == MPH_OK) st = split_analyse(A, ro, s);
  return st;
}

namespace {

template <typename V>
RBArgs<V> make_args(mph_matrix_impl* A, const void* x, void* y, int ld,
                    int init) {
  RBArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.ro = A->rb.ro;
  a.ci = A->rb.ci;
  a.val = static_cast<const V*>(A->rb.val);
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
  a.krec = A->rb.krec;
  return a;
}

// The split plan: short blocks over the list rows (tiled rows skipped,
// untiled long rows as chunks), the x-tiled pass, and the combines.
template <typename V, int PRE>
int launch_split_spmv(mph_matrix_impl* A, const void* x, void* y, int init, hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  constexpr bool comp = std::is_same<Acc, float>::value;
  RBArgs<V> a = make_args<V>(A, x, y, 1, init);
  const RowBlockPlan& p = A->rb;
  const size_t lds = sizeof(Acc) * (2 * (size_t)a.chunk + (PRE ? kRowBlockMaxRows : 0)) +
                     sizeof(int32_t) * (kRowBlockMaxRows + 1);
  RBArgs<V> sa = a;
  sa.blocks = p.sblocks;
  sa.nblocks = p.nsblocks;
  sa.krec = p.skrec;
  if (p.split) {
    sa.ro = p.sro;
    sa.rmap = p.smap;
  }
  if (p.nsblocks > 0) {
    if (!sa.krec) return runtime_error("split SpMV plan without kernel records");
    hipLaunchKernelGGL((rowblock_spmv_kernel<V, PRE, comp>), dim3((unsigned)p.nsblocks),
                       dim3(kRowBlockThreads), lds, s, sa);
  }
  if (p.nchunked > 0)
    hipLaunchKernelGGL((rowblock_combine_kernel<V, PRE, comp>),
                       dim3((unsigned)ceil_div(p.nsblocks, 256)), dim3(256), 0, s, sa);
  if (p.xt.rows > 0) {
    XTArgs t = xt_args(p.xt);
    const int64_t tiles8 = ceil_div((int64_t)p.xt.ntiles, (int64_t)kNumXcd) * kNumXcd;
    const bool direct = p.xt.rpos != nullptr;
    const size_t xlds = xtile_lds_bytes(sizeof(typename vtraits<V>::vec), sizeof(Acc), p.xt.tile, direct);
    if (direct)
      hipLaunchKernelGGL((xtile_direct_kernel<V, comp>), dim3((unsigned)(tiles8 * p.xt.groups)),
                         dim3(kXTileThreads), xlds, s, a, t);
    else
      hipLaunchKernelGGL((xtile_kernel<V, comp>), dim3((unsigned)(tiles8 * p.xt.groups)),
                         dim3(kXTileThreads), xlds, s, a, t);
    hipLaunchKernelGGL((xtile_combine_kernel<V, PRE, comp>),
                       dim3((unsigned)ceil_div(p.xt.rows, kXCombRows)), dim3(256), 0, s, a, t);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, int PRE>
int launch_spmv(mph_matrix_impl* A, const void* x, void* y, int init,
                hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  RBArgs<V> a = make_args<V>(A, x, y, 1, init);
  const size_t lds = sizeof(Acc) * (2 * (size_t)a.chunk + (PRE ? kRowBlockMaxRows : 0)) +
                     sizeof(int32_t) * (kRowBlockMaxRows + 1);
  constexpr bool comp = std::is_same<Acc, float>::value;
  if (A->rb.sblocks) {
    if (!A->rb.split) return launch_split_spmv<V, PRE_NONE>(A, x, y, init, s);
    if (A->rb.hdc_tail && A->dia_stage.n > 0 && A->tuning.rows_per_thread == 0)
      return hdc_tail_multiply(A, x, y, init, s);
    // HYB / HDC: the ELL / DIA part of every row in one coalesced pass that
    // writes y (zeros when that part is empty; nothing to do then when
    // accumulating), then the tail / CSR part of the rows that have one is
    // added (short blocks and the x-tiled pass accumulate)
    if (PRE != PRE_NONE || init)
      MPH_CHECK_STATUS(A->format == MPH_HYB ? ell_multiply(A, 1, x, y, init, s)
                                            : dia_multiply(A, 1, x, y, init, s));
    return launch_split_spmv<V, PRE_NONE>(A, x, y, 0, s);
  }
  if (!a.krec) return runtime_error("row-block plan without kernel records");
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
  if constexpr (!std::is_same<V, double>::value)
    if (A->rb.nlong > 0)
      hipLaunchKernelGGL((rowblock_spmm_chunk_kernel<V, KR>), dim3((unsigned)A->rb.nblocks),
                         dim3(kRowBlockThreads), 0, s, a);
  if (A->rb.nlong > 0)
    hipLaunchKernelGGL((rowblock_spmm_combine_kernel<V, PRE, KR>),
                       dim3((unsigned)ceil_div(A->rb.nblocks, 256)), dim3(256), 0, s, a);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

// The split plan's short blocks and untiled long-row chunks (tiled rows
// skipped) with KR right-hand sides.
template <typename V, int KR>
int launch_split_spmm_k(mph_matrix_impl* A, const void* x, void* y, int ld, int init,
                        hipStream_t s) {
  RBArgs<V> a = make_args<V>(A, x, y, ld, init);
  const RowBlockPlan& p = A->rb;
  a.blocks = p.sblocks;
  a.nblocks = p.nsblocks;
  a.krec = p.skrec;
  if (p.nsblocks > 0)
    hipLaunchKernelGGL((rowblock_spmm_kernel<V, PRE_NONE, KR>), dim3((unsigned)p.nsblocks),
                       dim3(kRowBlockThreads), 0, s, a);
  if constexpr (!std::is_same<V, double>::value)
    if (p.nchunked > 0)
      hipLaunchKernelGGL((rowblock_spmm_chunk_kernel<V, KR>), dim3((unsigned)p.nsblocks),
                         dim3(kRowBlockThreads), 0, s, a);
  if (p.nchunked > 0)
    hipLaunchKernelGGL((rowblock_spmm_combine_kernel<V, PRE_NONE, KR>),
                       dim3((unsigned)ceil_div(p.nsblocks, 256)), dim3(256), 0, s, a);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, int PRE>
int launch_spmm(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                hipStream_t s) {
  using Vec = typename vtraits<V>::vec;
  // A CSR-shaped matrix with x-tiled rows and few right-hand sides: its
  // x-tiled rows take the x-tiled pass (x and y with stride k) -- the k-wide
  // pass over groups of 4 / 2 columns, which reads the tile-major copy once
  // per group, or (a plan without the value copy) one pass per column --
  // instead of gathering a random k-wide row of X per item, which costs a
  // 64-128 B line from beyond L2 once X (k times x) outgrows the caches;
  // the other rows take the k-wide kernel over the split plan.  Measured
  // against those gathers: on a 2M-row power-law(1.2) matrix (runs of about
  // 43 items) the k-wide pass runs k = 4 / 8 / 16 1.65 / 1.6 / 1.08x faster,
  // one pass per column k = 2 / 4 2.1 / 1.4x faster and k = 8 1.2x slower;
  // on a 1M-row power-law(1.5) matrix (runs of about 13) the k-wide pass
  // k = 4 1.17x faster and k = 8 1.17x slower, one pass per column k = 2
  // 1.1x faster and k = 4 1.1x slower.
  const XTilePlan& xp = A->rb.xt;
  const bool long_runs =
      xp.rows > 0 && xp.items >= kXTileSpmmLongRun * (int64_t)xp.rows * xp.ntiles;
  const char* kw_env = std::getenv("MPH_XTILE_K");
  const int kw = xp.rows == 0 || xp.rpos || (kw_env && std::atoi(kw_env) == 0)
                     ? 1
                     : xtile_k_cols(A, xp);
  const int max_k = kw > 1 ? (long_runs ? kXTileSpmmKwLongRunMaxK : kXTileSpmmKwMaxK)
                           : (long_runs ? kXTileSpmmLongRunMaxK : kXTileSpmmMaxK);
  const bool xt_cols =
      PRE == PRE_NONE && A->rb.sblocks && !A->rb.split && xp.rows > 0 && k <= max_k;
  // split k into column groups of 16, 8, 4, 2, 1 sharing the leading dim k
  int done = 0;
  while (done < k) {
    const int rem = k - done;
    const Vec* xs = static_cast<const Vec*>(x) + done;
    Vec* ys = static_cast<Vec*>(y) + done;
    int st;
    if (rem >= 16) {
      st = xt_cols ? launch_split_spmm_k<V, 16>(A, xs, ys, k, init, s)
                   : launch_spmm_k<V, PRE, 16>(A, xs, ys, k, init, s);
      done += 16;
    } else if (rem >= 8) {
      st = xt_cols ? launch_split_spmm_k<V, 8>(A, xs, ys, k, init, s)
                   : launch_spmm_k<V, PRE, 8>(A, xs, ys, k, init, s);
      done += 8;
    } else if (rem >= 4) {
      st = xt_cols ? launch_split_spmm_k<V, 4>(A, xs, ys, k, init, s)
                   : launch_spmm_k<V, PRE, 4>(A, xs, ys, k, init, s);
      done += 4;
    } else if (rem >= 2) {
      st = xt_cols ? launch_split_spmm_k<V, 2>(A, xs, ys, k, init, s)
                   : launch_spmm_k<V, PRE, 2>(A, xs, ys, k, init, s);
      done += 2;
    } else {
      st = xt_cols ? launch_split_spmm_k<V, 1>(A, xs, ys, k, init, s)
                   : launch_spmm_k<V, PRE, 1>(A, xs, ys, k, init, s);
      done += 1;
    }
    if (st != MPH_OK) return st;
  }
  if (xt_cols) {
    // the x-tiled rows: groups of 4 / 2 columns through the k-wide pass (a
    // plan with the value copy; MPH_XTILE_K=0 keeps one pass per column),
    // the rest one column at a time
    int j = 0;
    if (kw > 1) {
      for (; kw >= 4 && j + 4 <= k; j += 4)
        MPH_CHECK_STATUS(xtile_multiply_k(A, A->rb.xt, 4, static_cast<const Vec*>(x) + j,
                                          static_cast<Vec*>(y) + j, init, k, s));
      for (; j + 2 <= k; j += 2)
        MPH_CHECK_STATUS(xtile_multiply_k(A, A->rb.xt, 2, static_cast<const Vec*>(x) + j,
                                          static_cast<Vec*>(y) + j, init, k, s));
    }
    for (; j < k; j++)
      MPH_CHECK_STATUS(xtile_multiply(A, A->rb.xt, static_cast<const Vec*>(x) + j,
                                      static_cast<Vec*>(y) + j, init, k, s));
  }
  return MPH_OK;
}

template <typename V>
int dispatch_pre(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s) {
  // (a HYB over its merged copy, like COO and CSR, has no prefix part)
  int pre = A->format == MPH_HDC                    ? PRE_DIA
            : A->format == MPH_HYB && !A->rb.m_ro ? PRE_ELL
                                                    : PRE_NONE;
  // an HDC without diagonals / a HYB without ELL slots is its CSR-shaped part
  if ((pre == PRE_DIA && A->ndiag == 0) || (pre == PRE_ELL && A->ell_loop == 0)) pre = PRE_NONE;
  if (k == 1) {
    if (pre == PRE_DIA) return launch_spmv<V, PRE_DIA>(A, x, y, init, s);
    if (pre == PRE_ELL) return launch_spmv<V, PRE_ELL>(A, x, y, init, s);
    return launch_spmv<V, PRE_NONE>(A, x, y, init, s);
  }
  if (pre != PRE_NONE && A->rb.prefix_apart) {
    // the ELL / DIA part in its own pass, the tail / CSR part added to it
    MPH_CHECK_STATUS(pre == PRE_ELL ? ell_multiply(A, k, x, y, init, s)
                                    : dia_multiply(A, k, x, y, init, s));
    return launch_spmm<V, PRE_NONE>(A, k, x, y, 0, s);
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

int coo_row_offsets(const int32_t* ri, int64_t nnz, int64_t nrows, int32_t* ro, hipStream_t s) {
  hipLaunchKernelGGL(tail_offsets_kernel, dim3((unsigned)ceil_div(nnz + 1, 256)), dim3(256), 0, s,
                     ri, nnz, nrows, ro);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

namespace {

template <typename V>
int xtile_multiply_t(mph_matrix_impl* A, const XTilePlan& xp, const void* x, void* y, int init,
                     int64_t ld, hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  constexpr bool comp = std::is_same<Acc, float>::value;
  RBArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.ld = (int32_t)ld;
  a.init = init;
  XTArgs t = xt_args(xp);
  const int64_t tiles8 = ceil_div((int64_t)xp.ntiles, (int64_t)kNumXcd) * kNumXcd;
  const bool direct = xp.rpos != nullptr;
  const size_t xlds = xtile_lds_bytes(sizeof(typename vtraits<V>::vec), sizeof(Acc), xp.tile, direct);
  if (direct)
    hipLaunchKernelGGL((xtile_direct_kernel<V, comp>), dim3((unsigned)(tiles8 * xp.groups)),
                       dim3(kXTileThreads), xlds, s, a, t);
  else
    hipLaunchKernelGGL((xtile_kernel<V, comp>), dim3((unsigned)(tiles8 * xp.groups)),
                       dim3(kXTileThreads), xlds, s, a, t);
  hipLaunchKernelGGL((xtile_combine_kernel<V, PRE_NONE, comp>),
                     dim3((unsigned)ceil_div(xp.rows, kXCombRows)), dim3(256), 0, s, a, t);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

// The k-wide x-tiled pass over columns [0, KR) of X / Y (leading dimension
// ld): its partials buffer and LDS attribute are set up by the first call.
template <typename V, int KR>
int xtile_multiply_k_t(mph_matrix_impl* A, XTilePlan& xp, const void* x, void* y, int init,
                       int64_t ld, hipStream_t s) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  constexpr bool comp = std::is_same<Acc, float>::value;
  const size_t xlds = xtile_k_lds_bytes(sizeof(Vec), xp.tile, KR);
  if (xp.partial_k_cols < KR) {
    void* d = nullptr;
    const size_t bytes = sizeof(Acc) * (comp ? 2 : 1) * (size_t)xp.ntiles * (size_t)xp.rows * KR;
    MPH_HIP_CHECK(hipMalloc(&d, bytes));
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&xtile_kernel_k<V, comp, KR, kXTileKThreads>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)xlds);
    if (e != hipSuccess) {
      dfree(d);
      set_error(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
      return MPH_ERR_HIP;
    }
    dfree(xp.partial_k);
    xp.partial_k = d;
    xp.partial_k_cols = KR;
  } else {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&xtile_kernel_k<V, comp, KR, kXTileKThreads>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)xlds);
    if (e != hipSuccess) {
      set_error(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
      return MPH_ERR_HIP;
    }
  }
  RBArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.x = static_cast<const Vec*>(x);
  a.y = static_cast<Vec*>(y);
  a.ld = (int32_t)ld;
  a.init = init;
  XTArgs t = xt_args(xp);
  const int64_t tiles8 = ceil_div((int64_t)xp.ntiles, (int64_t)kNumXcd) * kNumXcd;
  hipLaunchKernelGGL((xtile_kernel_k<V, comp, KR, kXTileKThreads>), dim3((unsigned)(tiles8 * xp.groups)),
                     dim3(kXTileKThreads), xlds, s, a, t, xp.partial_k);
  hipLaunchKernelGGL((xtile_combine_k_kernel<V, comp, KR>),
                     dim3((unsigned)ceil_div(xp.rows, kXCombRows)), dim3(256), 0, s, a, t,
                     (const void*)xp.partial_k);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

}  // namespace

// The widest k-wide pass (4, 2, or 1: none) whose LDS tile fits the device's
// per-workgroup limit.
int xtile_k_cols(const mph_matrix_impl* A, const XTilePlan& xp) {
  int dev = 0, lim = 0, optin = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    lim = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess)
    optin = 0;
  const size_t cap = (size_t)std::max(lim, optin);
  const size_t vec = vector_size(A->vt);
  if (xtile_k_lds_bytes(vec, xp.tile, 4) <= cap) return 4;
  if (xtile_k_lds_bytes(vec, xp.tile, 2) <= cap) return 2;
  return 1;
}

// KR = 2 or 4 columns of a copy plan's x-tiled pass (the caller checks both).
int xtile_multiply_k(mph_matrix_impl* A, XTilePlan& xp, int kr, const void* x, void* y, int init,
                     int64_t ld, hipStream_t s) {
  if (xp.rows == 0) return MPH_OK;
  if (xp.rpos || (kr != 2 && kr != 4)) return runtime_error("k-wide x-tiled pass: unsupported plan");
  switch (A->vt) {
    case MPH_F64:
      return kr == 4 ? xtile_multiply_k_t<double, 4>(A, xp, x, y, init, ld, s)
                     : xtile_multiply_k_t<double, 2>(A, xp, x, y, init, ld, s);
    case MPH_F32:
      return kr == 4 ? xtile_multiply_k_t<float, 4>(A, xp, x, y, init, ld, s)
                     : xtile_multiply_k_t<float, 2>(A, xp, x, y, init, ld, s);
    case MPH_BF16:
      return kr == 4 ? xtile_multiply_k_t<bf16, 4>(A, xp, x, y, init, ld, s)
                     : xtile_multiply_k_t<bf16, 2>(A, xp, x, y, init, ld, s);
  }
  return invalid("unknown value type");
}

int xtile_multiply(mph_matrix_impl* A, const XTilePlan& xp, const void* x, void* y, int init,
                   int64_t ld, hipStream_t s) {
  if (xp.rows == 0) return MPH_OK;
  switch (A->vt) {
    case MPH_F64: return xtile_multiply_t<double>(A, xp, x, y, init, ld, s);
    case MPH_F32: return xtile_multiply_t<float>(A, xp, x, y, init, ld, s);
    case MPH_BF16: return xtile_multiply_t<bf16>(A, xp, x, y, init, ld, s);
  }
  return invalid("unknown value type");
}

}  // namespace mph
