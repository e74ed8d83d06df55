// batch.hip -- batched SpMM: many matrices of uneven size and structure,
// each with k right-hand sides, in ONE persistent launch that walks a work
// queue of (matrix, row block) tiles; replayable through a hipGraph.
//
// There is no counterpart in the reference (its multiply takes one matrix and
// one vector, core/src/Morpheus_Multiply.hpp:57-78); this is the GPU form of
// running Morpheus::multiply over a set of DynamicMatrix objects.
//
// Design (gfx950):
//  * X / Y are row-major with k columns.  A row is served by KL lanes, each
//    owning Q consecutive columns (16 bytes: 2 FP64 / 4 FP32 values; fewer
//    when k is small), so every x / y access of a wave is whole 16-byte
//    pieces of contiguous X / Y rows: one instruction per entry instead of
//    KL, and rows of adjacent lanes are adjacent in memory.
//  * Every tile loads a batch of 8 entries (diagonals, slots or items) per
//    lane -- values, column indices and X pieces all issued before the first
//    multiply-add -- so a lane keeps 8 gathers in flight and a tile costs one
//    or a few memory round trips.
//  * Tiles: a slab of 256 / KL rows of a DIA or ELL matrix (one row per lane
//    group); a block of 256 / KL consecutive rows of a CSR-shaped matrix
//    (CSR, row-sorted COO, HYB's COO tail with its ELL part, HDC's CSR part
//    with its DIA part), where a lane group takes one row of at most
//    kGroupRowMax entries and leaves longer rows to their own tiles; up to
//    four rows of kGroupRowMax + 1 .. kWaveRowMax entries, one wave each; a
//    row of more entries, the whole workgroup.  FP32 / BF16 rows of more
//    than kCompensateLen entries are summed as Dot2 pairs in a launch of
//    their own (batch_comp_rows_kernel).
//  * Each matrix goes whole to one of eight sub-queues (one per XCD:
//    workgroup w runs on XCD w % 8 and drains sub-queue w % 8 first), chosen
//    by longest-processing-time packing; a workgroup takes `grab` tiles per
//    counter update and, once its sub-queue is drained, the others'.
//  * The counters are cleared by a small kernel before the main kernel, so
//    runs stay capturable (no memset nodes); the direct run and the two
//    captured graphs (init / accumulate) each have counters of their own.
#include <algorithm>
#include <cstdlib>
#include <queue>
#include <vector>

#include "mph_internal.hpp"

namespace mph {

enum { TILE_ROWS = 0, TILE_WAVE = 1, TILE_WHOLE = 2, TILE_DIA = 3, TILE_ELL = 4 };
enum { BPRE_NONE = 0, BPRE_DIA = 1, BPRE_ELL = 2 };

struct BatchMat {
  int64_t nrows, ncols;
  const int32_t* ro;  // CSR-shaped part (nullptr: none)
  const int32_t* ci;
  const void* val;
  int32_t pre;        // BPRE_*: a DIA / ELL part added to every row
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

// TILE_WAVE: r0 is the first entry of the tile's row ids in the batch's row
// list and r1 their count (<= 4); other kinds: rows [r0, r1) (TILE_WHOLE and
// the Dot2 rows: one row).
struct BatchTile {
  int32_t mat;
  int32_t kind;
  int32_t r0;
  int32_t r1;
};

constexpr int kBT = 256;
constexpr int kQueues = 8;         // one sub-queue per XCD
constexpr int kB = 8;              // entries per lane per batch
constexpr int kGroupRowMax = 32;   // rows a lane group takes in a row block
constexpr int kWaveRowMax = 512;   // rows one wave takes

// Lane layout for k = KR columns: Q values per lane, KL lanes per row.
template <typename V, int KR>
struct Lay {
  using Vec = typename vtraits<V>::vec;
  using Acc = typename vtraits<V>::acc;
  static constexpr int Q = (KR * (int)sizeof(Vec) >= 16) ? 16 / (int)sizeof(Vec) : KR;
  static constexpr int KL = KR / Q;
};

// the Q accumulators of a lane
template <typename V, int KR>
using AccQ = typename Lay<V, KR>::Acc[Lay<V, KR>::Q];

// Global-memory views of the pointers a matrix record holds.  The records
// are read from memory, so the compiler cannot prove where their pointers
// point and would emit flat loads and stores, whose waits also cover the
// scalar loads (every wait became vmcnt(0) lgkmcnt(0)).
template <typename T>
using Gp = const __attribute__((address_space(1))) T*;
template <typename T>
using Gm = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ Gp<T> gp(const void* p) {
  return (Gp<T>)p;
}
template <typename T>
__device__ __forceinline__ Gm<T> gm(void* p) {
  return (Gm<T>)p;
}
template <typename T>
__device__ __forceinline__ T gld(Gp<T> p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ bf16 gld<bf16>(Gp<bf16> p) {
  bf16 r;
  r.bits = p->bits;
  return r;
}

// Q consecutive values of a row of X (aligned to Q * sizeof(Vec)), widened
// to the accumulation type.
template <typename Vec, int Q, typename Acc>
__device__ __forceinline__ void load_piece(Gp<Vec> p, Acc (&out)[Q]) {
  if constexpr (Q * sizeof(Vec) == 16) {
    typedef Vec v16 __attribute__((ext_vector_type(16 / sizeof(Vec))));
    const v16 w = *(Gp<v16>)p;
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (Acc)w[q];
  } else if constexpr (Q == 2 && sizeof(Vec) == 4) {
    typedef Vec v8 __attribute__((ext_vector_type(2)));
    const v8 w = *(Gp<v8>)p;
    out[0] = (Acc)w[0];
    out[1] = (Acc)w[1];
  } else {
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (Acc)p[q];
  }
}

template <typename Vec, int Q, typename Acc>
__device__ __forceinline__ void store_piece(Gm<Vec> p, const Acc (&t)[Q], int init) {
  Acc o[Q];
  if (!init) load_piece<Vec, Q, Acc>((Gp<Vec>)p, o);
  if constexpr (Q * sizeof(Vec) == 16) {
    typedef Vec v16 __attribute__((ext_vector_type(16 / sizeof(Vec))));
    v16 w;
#pragma unroll
    for (int q = 0; q < Q; q++) w[q] = (Vec)(init ? t[q] : o[q] + t[q]);
    *(Gm<v16>)p = w;
  } else {
#pragma unroll
    for (int q = 0; q < Q; q++) p[q] = (Vec)(init ? t[q] : o[q] + t[q]);
  }
}

// t += sum over the batch of v[w] * X[c[w]] (the lane's Q columns), for the
// entries with bit w of ok set; every X piece is requested before the first
// product, and c[w] is a valid row of X even where the bit is clear (its
// product is dropped).
template <typename V, int KR>
__device__ __forceinline__ void fma_batch(AccQ<V, KR>& t,
                                          V (&v)[kB], const int32_t (&c)[kB], uint32_t ok,
                                          Gp<typename Lay<V, KR>::Vec> x, int l) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  Acc xp[kB][L::Q];
#pragma unroll
  for (int w = 0; w < kB; w++)
    load_piece<typename L::Vec, L::Q, Acc>(x + (int64_t)c[w] * KR + l * L::Q, xp[w]);
  hold_raw(v);  // BF16: widened below the gathers
#pragma unroll
  for (int w = 0; w < kB; w++) {
    const Acc a = vtraits<V>::widen(v[w]);
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
      const Acc f = __builtin_fma(a, xp[w][q], t[q]);
      t[q] = ((ok >> w) & 1u) ? f : t[q];
    }
  }
}

// DIA part of row r over diagonals d = d0, d0 + step, ... (live: the row
// counts; r is a valid row either way)
template <typename V, int KR>
__device__ __forceinline__ void dia_part(const BatchMat& m, int64_t r, bool live, int d0, int step,
                                         AccQ<V, KR>& t, int l) {
  using L = Lay<V, KR>;
  const Gp<V> dv = gp<V>(m.dval);
  const Gp<int32_t> doff = gp<int32_t>(m.doff);
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  for (int db = d0; db < m.ndiag; db += kB * step) {
    V v[kB];
    int32_t c[kB];
    uint32_t ok = 0;
#pragma unroll
    for (int w = 0; w < kB; w++) {
      const int d = db + w * step;
      const int dc = min(d, m.ndiag - 1);
      const int64_t col = r + doff[dc];
      const bool in = live && d < m.ndiag && col >= 0 && col < m.ncols;
      ok |= (uint32_t)in << w;
      c[w] = in ? (int32_t)col : 0;
      // (d < ndiag is the same for the whole wave when step == 1)
      v[w] = (step > 1 || d < m.ndiag) ? gld<V>(dv + r + m.dpitch * dc) : V{};
    }
    fma_batch<V, KR>(t, v, c, ok, x, l);
  }
}

// ELL part of row r over slots n = n0, n0 + step, ...
template <typename V, int KR>
__device__ __forceinline__ void ell_part(const BatchMat& m, int64_t r, bool live, int n0, int step,
                                         AccQ<V, KR>& t, int l) {
  using L = Lay<V, KR>;
  const Gp<V> ev = gp<V>(m.evalp);
  const Gp<int32_t> eci = gp<int32_t>(m.eci);
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  for (int nb = n0; nb < m.eloop; nb += kB * step) {
    V v[kB];
    int32_t ci[kB];
#pragma unroll
    for (int w = 0; w < kB; w++) {
      const int64_t slot = r + m.epitch * min(nb + w * step, m.eloop - 1);
      // (nb + w < eloop is the same for the whole wave when step == 1)
      const bool any = step > 1 || nb + w < m.eloop;
      ci[w] = any ? gld<int32_t>(eci + slot) : -1;
      v[w] = any ? gld<V>(ev + slot) : V{};
    }
    int32_t c[kB];
    uint32_t ok = 0;
#pragma unroll
    for (int w = 0; w < kB; w++) {
      const bool in = live && nb + w * step < m.eloop && ci[w] >= 0;
      ok |= (uint32_t)in << w;
      c[w] = in ? ci[w] : 0;
    }
    fma_batch<V, KR>(t, v, c, ok, x, l);
  }
}

// CSR-shaped items st + j0, st + j0 + step, ... of a row ending at en
template <typename V, int KR>
__device__ __forceinline__ void csr_part(const BatchMat& m, int64_t st, int64_t en, int j0,
                                         int step, AccQ<V, KR>& t,
                                         int l) {
  using L = Lay<V, KR>;
  const Gp<V> val = gp<V>(m.val);
  const Gp<int32_t> cip = gp<int32_t>(m.ci);
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  for (int64_t jb = st + j0; jb < en; jb += (int64_t)kB * step) {
    V v[kB];
    int32_t ci[kB];
#pragma unroll
    for (int w = 0; w < kB; w++) {
      const int64_t j = jb + (int64_t)w * step;
      const int64_t jj = j < en ? j : jb;  // jb < en: a valid item
      ci[w] = gld<int32_t>(cip + jj);
      v[w] = gld<V>(val + jj);
    }
    int32_t c[kB];
    uint32_t ok = 0;
#pragma unroll
    for (int w = 0; w < kB; w++) {
      const bool in = jb + (int64_t)w * step < en;
      ok |= (uint32_t)in << w;
      c[w] = in ? ci[w] : 0;
    }
    fma_batch<V, KR>(t, v, c, ok, x, l);
  }
}

// the DIA / ELL slots a row adds to its CSR-shaped part
__device__ __forceinline__ int64_t prefix_len(const BatchMat& m) {
  return m.pre == BPRE_DIA ? m.ndiag : m.pre == BPRE_ELL ? m.eloop : 0;
}

// a slab of a DIA / ELL matrix, or a block of rows of a CSR-shaped one: lane
// group g takes row r0 + g
template <typename V, int KR, int KIND>
__device__ void tile_rows(const BatchMat& m, const BatchTile& t, int init) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  const int g = threadIdx.x / L::KL, l = threadIdx.x % L::KL;
  const int64_t r = t.r0 + g;
  const bool valid = r < t.r1;
  const int64_t rc = valid ? r : t.r0;  // a row of the tile, for in-bounds loads
  Acc acc[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; q++) acc[q] = 0;
  bool mine = valid;
  if constexpr (KIND == TILE_DIA) {
    dia_part<V, KR>(m, rc, valid, 0, 1, acc, l);
  } else if constexpr (KIND == TILE_ELL) {
    ell_part<V, KR>(m, rc, valid, 0, 1, acc, l);
  } else {
    const Gp<int32_t> ro = gp<int32_t>(m.ro);
    const int64_t st = ro[rc], en = ro[rc + 1];
    mine = valid && (en - st) + prefix_len(m) <= kGroupRowMax;
    if (m.pre == BPRE_DIA) dia_part<V, KR>(m, rc, mine, 0, 1, acc, l);
    else if (m.pre == BPRE_ELL) ell_part<V, KR>(m, rc, mine, 0, 1, acc, l);
    if (mine) csr_part<V, KR>(m, st, en, 0, 1, acc, l);
  }
  if (mine)
    store_piece<typename L::Vec, L::Q, Acc>(gm<typename L::Vec>(m.y) + r * KR + l * L::Q,
                                            acc, init);
}

// one row per wave: the wave's G = 64 / KL item groups stride over the row
template <typename V, int KR>
__device__ void tile_wave(const BatchMat& m, const BatchTile& t, const int32_t* __restrict__ rows,
                          int init) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  constexpr int G = kWave / L::KL;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (wave >= t.r1) return;  // whole waves: no barrier below
  const int64_t r = rows[t.r0 + wave];
  const int g = lane / L::KL, l = lane % L::KL;
  Acc acc[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; q++) acc[q] = 0;
  if (m.pre == BPRE_DIA) dia_part<V, KR>(m, r, true, g, G, acc, l);
  else if (m.pre == BPRE_ELL) ell_part<V, KR>(m, r, true, g, G, acc, l);
  if (m.ro) csr_part<V, KR>(m, gp<int32_t>(m.ro)[r], gp<int32_t>(m.ro)[r + 1], g, G, acc, l);
#pragma unroll
  for (int q = 0; q < L::Q; q++)
    for (int s = kWave >> 1; s >= L::KL; s >>= 1) acc[q] += shfl_xor(acc[q], s);
  if (g == 0)
    store_piece<typename L::Vec, L::Q, Acc>(gm<typename L::Vec>(m.y) + r * KR + l * L::Q,
                                            acc, init);
}

// one row per workgroup: its G = 256 / KL item groups stride over the row
template <typename V, int KR>
__device__ void tile_whole(const BatchMat& m, const BatchTile& t, int init,
                           typename Lay<V, KR>::Acc* s_red) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  constexpr int G = kBT / L::KL;
  const int g = threadIdx.x / L::KL, l = threadIdx.x % L::KL;
  const int64_t r = t.r0;
  Acc acc[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; q++) acc[q] = 0;
  if (m.pre == BPRE_DIA) dia_part<V, KR>(m, r, true, g, G, acc, l);
  else if (m.pre == BPRE_ELL) ell_part<V, KR>(m, r, true, g, G, acc, l);
  if (m.ro) csr_part<V, KR>(m, gp<int32_t>(m.ro)[r], gp<int32_t>(m.ro)[r + 1], g, G, acc, l);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int q = 0; q < L::Q; q++) {
    for (int s = kWave >> 1; s >= L::KL; s >>= 1) acc[q] += shfl_xor(acc[q], s);
    if (lane < L::KL) s_red[(wave * L::KL + lane) * L::Q + q] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < L::KL) {
    Acc tot[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
      tot[q] = 0;
      for (int w = 0; w < kBT / kWave; w++) tot[q] += s_red[(w * L::KL + threadIdx.x) * L::Q + q];
    }
    store_piece<typename L::Vec, L::Q, Acc>(
        gm<typename L::Vec>(m.y) + r * KR + threadIdx.x * L::Q, tot, init);
  }
}

// One row reduced by the whole workgroup as Dot2 pairs per column (FP32 /
// BF16 rows of more than kCompensateLen items, kept out of the queue):
// lane-strided pairs, then a fixed tree over lanes and waves; the old y
// joins through TwoSum when accumulating.
template <typename V, int KR>
__device__ void block_row_comp(const BatchMat& m, int64_t r, int init, float* s_red) {
  using Vec = typename vtraits<V>::vec;
  csum<float> cs[KR];
#pragma unroll
  for (int q = 0; q < KR; q++) cs[q] = csum<float>{0, 0};
  const Gp<Vec> x = gp<Vec>(m.x);
  if (m.pre == BPRE_DIA) {
    const Gp<V> dv = gp<V>(m.dval);
    const Gp<int32_t> doff = gp<int32_t>(m.doff);
    for (int d = threadIdx.x; d < m.ndiag; d += kBT) {
      const int64_t col = r + doff[d];
      if (col >= 0 && col < m.ncols) {
        const float v = vtraits<V>::widen(gld<V>(dv + r + m.dpitch * d));
#pragma unroll
        for (int q = 0; q < KR; q++) cs[q].add_prod(v, (float)x[col * KR + q]);
      }
    }
  } else if (m.pre == BPRE_ELL) {
    const Gp<V> ev = gp<V>(m.evalp);
    const Gp<int32_t> eci = gp<int32_t>(m.eci);
    for (int n = threadIdx.x; n < m.eloop; n += kBT) {
      const int32_t c = gld<int32_t>(eci + r + m.epitch * n);
      if (c >= 0) {
        const float v = vtraits<V>::widen(gld<V>(ev + r + m.epitch * n));
#pragma unroll
        for (int q = 0; q < KR; q++) cs[q].add_prod(v, (float)x[(int64_t)c * KR + q]);
      }
    }
  }
  if (m.ro) {
    const Gp<V> val = gp<V>(m.val);
    const Gp<int32_t> ro = gp<int32_t>(m.ro), cip = gp<int32_t>(m.ci);
    for (int64_t j = ro[r] + threadIdx.x; j < ro[r + 1]; j += kBT) {
      const float v = vtraits<V>::widen(gld<V>(val + j));
      const int64_t c = gld<int32_t>(cip + j);
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
    const Gm<Vec> y = gm<Vec>(m.y) + r * KR + q;
    if (!init) t.combine((float)*y, 0.0f);
    *y = (Vec)t.value();
  }
}

__global__ void batch_reset_kernel(int32_t* counters) {
  if (threadIdx.x < kQueues) counters[threadIdx.x] = 0;
}

template <typename V, int KR>
__device__ __forceinline__ void run_tile(const BatchMat& m, const BatchTile& t,
                                         const int32_t* __restrict__ rows, int init,
                                         typename vtraits<V>::acc* s_red) {
  switch (t.kind) {
    case TILE_ROWS: tile_rows<V, KR, TILE_ROWS>(m, t, init); break;
    case TILE_DIA: tile_rows<V, KR, TILE_DIA>(m, t, init); break;
    case TILE_ELL: tile_rows<V, KR, TILE_ELL>(m, t, init); break;
    case TILE_WAVE: tile_wave<V, KR>(m, t, rows, init); break;
    default: tile_whole<V, KR>(m, t, init, s_red); break;
  }
}

// Static schedule: workgroup w walks its own slice woff[w] .. woff[w+1] of
// the tile list, filled on the host by longest-processing-time bin packing
// (over the workgroups of the matrix's XCD).  No counters, no atomics.
template <typename V, int KR>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(6))) void batch_spmm_static_kernel(
    const BatchMat* __restrict__ mats, const BatchTile* __restrict__ tiles,
    const int32_t* __restrict__ woff, const int32_t* __restrict__ rows, int init) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[kBT / kWave * Lay<V, KR>::KL * Lay<V, KR>::Q];
  const int32_t b = woff[blockIdx.x], e = woff[blockIdx.x + 1];
  int32_t cur = -1;
  BatchMat m{};
  for (int32_t i = b; i < e; i++) {
    const BatchTile t = tiles[i];
    if (t.mat != cur) {
      m = mats[t.mat];
      cur = t.mat;
    }
    run_tile<V, KR>(m, t, rows, init, s_red);
    if (t.kind == TILE_WHOLE) __syncthreads();  // s_red is reused
  }
}

// qoff[q] .. qoff[q+1]: the tiles of sub-queue q.  A workgroup drains its
// home sub-queue, then the others; a relaxed load skips a drained sub-queue
// without a read-modify-write on its counter.
template <typename V, int KR>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(6))) void batch_spmm_kernel(const BatchMat* __restrict__ mats,
                                                         const BatchTile* __restrict__ tiles,
                                                         const int32_t* __restrict__ qoff,
                                                         const int32_t* __restrict__ rows,
                                                         int32_t* counters, int32_t grab,
                                                         int init) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[kBT / kWave * Lay<V, KR>::KL * Lay<V, KR>::Q];
  __shared__ int32_t s_next;
  const int home = blockIdx.x % kQueues;
  // the matrix record of the last tile stays in registers: a sub-queue holds
  // each matrix's tiles together, so most tiles skip that round trip
  int32_t cur = -1;
  BatchMat m{};
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
      BatchTile t = tiles[qb + first];
      for (int32_t i = first; i < last; i++) {
        // the next record is requested before this tile's work
        const BatchTile tn = tiles[qb + min(i + 1, last - 1)];
        if (t.mat != cur) {
          m = mats[t.mat];
          cur = t.mat;
        }
        run_tile<V, KR>(m, t, rows, init, s_red);
        if (t.kind == TILE_WHOLE) __syncthreads();  // s_red is reused
        t = tn;
      }
    }
  }
}

// One workgroup per tile (MPH_BATCH_SCHEDULE=2): the hardware dispatcher
// hands out the tiles, largest first.
template <typename V, int KR>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(6))) void batch_spmm_grid_kernel(const BatchMat* __restrict__ mats,
                                                              const BatchTile* __restrict__ tiles,
                                                              const int32_t* __restrict__ rows,
                                                              int init) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[kBT / kWave * Lay<V, KR>::KL * Lay<V, KR>::Q];
  // each XCD takes a contiguous range of the tiles (in matrix and row order),
  // so neighbouring row blocks of a matrix share that XCD's L2
  const BatchTile t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  run_tile<V, KR>(mats[t.mat], t, rows, init, s_red);
}

// FP32 / BF16 rows of more than kCompensateLen items: one Dot2 row per
// workgroup turn, in a launch of its own so that the pairs do not raise the
// register count of the queue kernel.
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
  mph::BatchTile* d_tiles_all = nullptr;  // every tile, largest first (grid)
  int32_t* d_woff = nullptr;              // nwg + 1 list offsets
  int32_t* d_qoff = nullptr;
  int32_t* d_rows = nullptr;              // row ids of the one-wave tiles
  int sched = 2;                          // 0 queues, 1 static lists, 2 grid (default)
  int32_t grab = 4;                       // tiles per counter update
  // queue counters: kQueues for the direct run, then kQueues per captured
  // graph (init, accumulate), so a replay never shares counters with another
  // form of the run
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

// estimated fixed cost of one tile, in bytes moved
constexpr int64_t kTileOverhead = 4096;

template <typename V, int KR>
int launch_batch(mph_batch_s* b, int init, int32_t* counters, hipStream_t s) {
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
                       s, b->d_mats, b->d_tiles_lpt, b->d_woff, b->d_rows, init);
  } else if (b->sched == 2) {
    hipLaunchKernelGGL((batch_spmm_grid_kernel<V, KR>), dim3((unsigned)b->ntiles), dim3(kBT), 0,
                       s, b->d_mats, b->d_tiles_all, b->d_rows, init);
  } else {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL((batch_spmm_kernel<V, KR>), dim3((unsigned)b->nwg), dim3(kBT), 0, s,
                       b->d_mats, b->d_tiles, b->d_qoff, b->d_rows, counters, b->grab, init);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V>
int launch_k(mph_batch_s* b, int init, int32_t* counters, hipStream_t s) {
  switch (b->k) {
    case 1: return launch_batch<V, 1>(b, init, counters, s);
    case 2: return launch_batch<V, 2>(b, init, counters, s);
    case 4: return launch_batch<V, 4>(b, init, counters, s);
    case 8: return launch_batch<V, 8>(b, init, counters, s);
    case 16: return launch_batch<V, 16>(b, init, counters, s);
  }
  return invalid("batched k must be 1, 2, 4, 8 or 16");
}

int batch_launch(mph_batch_s* b, int init, int32_t* counters, hipStream_t s) {
  if (b->ntiles == 0 && b->ncomp == 0) return MPH_OK;
  switch (b->vt) {
    case MPH_F64: return launch_k<double>(b, init, counters, s);
    case MPH_F32: return launch_k<float>(b, init, counters, s);
    case MPH_BF16: return launch_k<bf16>(b, init, counters, s);
  }
  return invalid("unknown value type");
}

// Workgroups of a persistent batch kernel that fit on the device at once:
// more would leave their whole tile lists (static schedule) to run after
// others finish.
template <typename V, int KR>
int resident_wg(int sched) {
  int per_cu = 0;
  const hipError_t e =
      sched == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                       &per_cu, reinterpret_cast<const void*>(&batch_spmm_static_kernel<V, KR>), kBT, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                       &per_cu, reinterpret_cast<const void*>(&batch_spmm_kernel<V, KR>), kBT, 0);
  if (e != hipSuccess || per_cu <= 0) per_cu = 4;
  return per_cu * device_cu_count();
}

template <typename V>
int resident_wg_k(int k, int sched) {
  switch (k) {
    case 1: return resident_wg<V, 1>(sched);
    case 2: return resident_wg<V, 2>(sched);
    case 4: return resident_wg<V, 4>(sched);
    case 8: return resident_wg<V, 8>(sched);
    default: return resident_wg<V, 16>(sched);
  }
}

int resident_workgroups(int vt, int k, int sched) {
  return vt == MPH_F64 ? resident_wg_k<double>(k, sched)
         : vt == MPH_F32 ? resident_wg_k<float>(k, sched)
                          : resident_wg_k<bf16>(k, sched);
}

// lanes per row of the k-column layout (16 bytes of a row of X per lane)
int lanes_per_row(int vt, int k) {
  const int vec = (int)vector_size(vt);
  const int q = k * vec >= 16 ? 16 / vec : k;
  return k / q;
}

// The tiles of one member: DIA / ELL slabs, or row blocks of a CSR-shaped
// part (host row offsets `ro`, `pre` slots per row in front) with one-wave,
// whole-workgroup and Dot2 tiles for its longer rows.  cost: bytes moved.
struct TileSet {
  std::vector<BatchTile> tiles;
  std::vector<int64_t> cost;
  std::vector<BatchTile> comp;
  std::vector<int32_t> rows;  // row ids of one-wave tiles (r0 indexes this list)
};

void tile_member(TileSet& ts, int32_t i, int64_t nrows, int kind_slab, int64_t slab_len,
                 const std::vector<int32_t>* ro, int64_t pre, int vt, int k) {
  const int kl = lanes_per_row(vt, k);
  const int64_t rpb = kBT / kl;
  const int64_t vsz = (int64_t)value_size(vt), xsz = (int64_t)vector_size(vt) * k;
  const bool comp_type = vt != MPH_F64;
  if (kind_slab >= 0) {
    const bool long_rows = comp_type && slab_len > kCompensateLen;
    for (int64_t r = 0; r < nrows; r += long_rows ? 1 : rpb) {
      if (long_rows) {
        ts.comp.push_back(BatchTile{i, TILE_WHOLE, (int32_t)r, (int32_t)r + 1});
        continue;
      }
      const int64_t r1 = std::min(nrows, r + rpb);
      ts.tiles.push_back(BatchTile{i, kind_slab, (int32_t)r, (int32_t)r1});
      ts.cost.push_back((r1 - r) * (slab_len * (4 + vsz + xsz) + xsz) + kTileOverhead);
    }
    return;
  }
  std::vector<int32_t> wave;
  int64_t wave_items = 0;
  auto flush_wave = [&]() {
    if (wave.empty()) return;
    ts.tiles.push_back(BatchTile{i, TILE_WAVE, (int32_t)ts.rows.size(), (int32_t)wave.size()});
    ts.cost.push_back(wave_items * (4 + vsz + xsz) + (int64_t)wave.size() * xsz + kTileOverhead);
    ts.rows.insert(ts.rows.end(), wave.begin(), wave.end());
    wave.clear();
    wave_items = 0;
  };
  for (int64_t r0 = 0; r0 < nrows; r0 += rpb) {
    const int64_t r1 = std::min(nrows, r0 + rpb);
    int64_t items = 0;
    for (int64_t r = r0; r < r1; r++) {
      const int64_t len = (int64_t)(*ro)[r + 1] - (*ro)[r] + pre;
      if (len <= kGroupRowMax) {
        items += len;
      } else if (comp_type && len > kCompensateLen) {
        ts.comp.push_back(BatchTile{i, TILE_WHOLE, (int32_t)r, (int32_t)r + 1});
      } else if (len > kWaveRowMax) {
        ts.tiles.push_back(BatchTile{i, TILE_WHOLE, (int32_t)r, (int32_t)r + 1});
        ts.cost.push_back(len * (4 + vsz + xsz) + xsz + kTileOverhead);
      } else {
        wave.push_back((int32_t)r);
        wave_items += len;
        if ((int)wave.size() == kBT / kWave) flush_wave();
      }
    }
    // the block's tile covers every row of [r0, r1); its lane groups leave
    // the longer rows to their own tiles
    ts.tiles.push_back(BatchTile{i, TILE_ROWS, (int32_t)r0, (int32_t)r1});
    ts.cost.push_back(items * (4 + vsz + xsz) + (r1 - r0) * (xsz + 8) + kTileOverhead);
  }
  flush_wave();
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
  TileSet ts;
  int vt = nmat > 0 ? mats[0]->vt : MPH_F32;
  const size_t piece = std::min<size_t>(16, (size_t)k * vector_size(vt));
  // COO members are batched as the CSR view of their row-sorted entries: row
  // offsets derived on the device, owned by the batch
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
    if (A->nrows > 0 && (!X[i] || !Y[i])) return invalid("null X / Y in batch");
    if ((reinterpret_cast<uintptr_t>(X[i]) % piece) != 0 ||
        (reinterpret_cast<uintptr_t>(Y[i]) % piece) != 0)
      return invalid("batched X / Y rows must be aligned to min(16, k * element) bytes");
    BatchMat m{};
    m.nrows = A->nrows;
    m.ncols = A->ncols;
    m.x = X[i];
    m.y = Y[i];
    m.pre = BPRE_NONE;
    std::vector<int32_t> ro;
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
      m.ro = d_ro;
      m.ci = A->coo_ci;
      m.val = A->coo_v;
    } else {
      if (!A->analysed) MPH_CHECK_STATUS(mph_matrix_analyse(mats[i], stream));
      if (A->format == MPH_HYB && A->rb.tail_ro == nullptr && A->coo_nnz > 0)
        return invalid("HYB with an unsorted COO part cannot be batched");
      m.ndiag = (int32_t)A->ndiag;
      m.dpitch = A->dia_pitch;
      m.doff = A->dia_off;
      m.dval = A->dia_v;
      m.eloop = (int32_t)A->ell_loop;
      m.epitch = A->ell_pitch;
      m.eci = A->ell_ci;
      m.evalp = A->ell_v;
      // an HDC without CSR entries is its DIA part, a HYB without COO
      // entries its ELL part: slabs like DIA / ELL matrices
      const bool as_dia = A->format == MPH_DIA || (A->format == MPH_HDC && A->csr_nnz == 0);
      const bool as_ell = A->format == MPH_ELL || (A->format == MPH_HYB && A->coo_nnz == 0);
      if (as_dia || as_ell) {
        m.ro = nullptr;
        m.pre = as_dia ? BPRE_DIA : BPRE_ELL;
        tile_member(ts, i, A->nrows, as_dia ? TILE_DIA : TILE_ELL,
                    as_dia ? A->ndiag : A->ell_loop, nullptr, 0, vt, k);
        hm[i] = m;
        continue;
      }
      m.ro = A->format == MPH_HYB ? A->rb.tail_ro : A->csr_ro;
      m.ci = A->format == MPH_HYB ? A->coo_ci : A->csr_ci;
      m.val = A->format == MPH_HYB ? A->coo_v : A->csr_v;
      m.pre = A->format == MPH_HDC ? BPRE_DIA : A->format == MPH_HYB ? BPRE_ELL : BPRE_NONE;
      if ((m.pre == BPRE_DIA && m.ndiag == 0) || (m.pre == BPRE_ELL && m.eloop == 0))
        m.pre = BPRE_NONE;
    }
    if (A->nrows > 0) {
      ro.resize(A->nrows + 1);
      MPH_HIP_CHECK(hipMemcpyAsync(ro.data(), m.ro, sizeof(int32_t) * (A->nrows + 1),
                                   hipMemcpyDeviceToHost, s));
      MPH_HIP_CHECK(hipStreamSynchronize(s));
    }
    const int64_t pre = m.pre == BPRE_DIA ? m.ndiag : m.pre == BPRE_ELL ? m.eloop : 0;
    tile_member(ts, i, A->nrows, -1, 0, &ro, pre, vt, k);
    hm[i] = m;
  }
  std::vector<BatchTile>& tiles = ts.tiles;
  std::vector<int64_t>& cost = ts.cost;
  // largest tiles first: the dynamic queue then drains evenly
  std::vector<size_t> order(tiles.size());
  for (size_t q = 0; q < order.size(); q++) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return cost[a] > cost[b]; });
  // Each matrix goes whole to one sub-queue (one XCD: workgroup w runs on
  // XCD w % 8 and drains sub-queue w % 8 first), chosen by longest-
  // processing-time packing of the matrices' costs, so that a matrix's X is
  // brought into one XCD's L2 instead of all eight.  (MPH_BATCH_AFFINITY=0
  // deals the tiles round-robin instead.)
  // (measured on the 512-matrix batch: no gain from the affinity, and with
  // one large matrix it leaves a single counter to every workgroup)
  bool affinity = false;
  if (const char* env = std::getenv("MPH_BATCH_AFFINITY")) affinity = std::atoi(env) != 0;
  std::vector<int> mat_queue(std::max<int32_t>(nmat, 1), 0);
  if (affinity) {
    std::vector<int64_t> mcost(std::max<int32_t>(nmat, 1), 0);
    for (size_t q = 0; q < tiles.size(); q++) mcost[tiles[q].mat] += cost[q];
    std::vector<int32_t> morder(nmat);
    for (int32_t i = 0; i < nmat; i++) morder[i] = i;
    std::stable_sort(morder.begin(), morder.end(),
                     [&](int32_t a, int32_t b) { return mcost[a] > mcost[b]; });
    int64_t load[kQueues] = {};
    for (int32_t i : morder) {
      const int q = (int)(std::min_element(load, load + kQueues) - load);
      mat_queue[i] = q;
      load[q] += mcost[i];
    }
  }
  // a sub-queue lists its matrices largest first, each matrix's tiles
  // together (largest first): consecutive tiles then mostly share the matrix
  // record the kernel keeps in registers
  std::vector<int64_t> mtot(std::max<int32_t>(nmat, 1), 0);
  for (size_t q = 0; q < tiles.size(); q++) mtot[tiles[q].mat] += cost[q];
  std::vector<size_t> qorder = order;
  std::stable_sort(qorder.begin(), qorder.end(), [&](size_t a, size_t b) {
    const int32_t ma = tiles[a].mat, mb = tiles[b].mat;
    if (ma != mb) return mtot[ma] != mtot[mb] ? mtot[ma] > mtot[mb] : ma < mb;
    return cost[a] > cost[b];
  });
  std::vector<BatchTile> sorted;
  sorted.reserve(tiles.size());
  std::vector<int32_t> qoff(kQueues + 1, 0);
  for (int q = 0; q < kQueues; q++) {
    qoff[q] = (int32_t)sorted.size();
    if (affinity) {
      for (size_t j : qorder)
        if (mat_queue[tiles[j].mat] == q) sorted.push_back(tiles[j]);
    } else {
      for (size_t j = q; j < order.size(); j += kQueues) sorted.push_back(tiles[order[j]]);
    }
  }
  qoff[kQueues] = (int32_t)sorted.size();

  mph_batch_s* b = new mph_batch_s();
  b->owned = owned;
  guard.keep = true;
  b->vt = vt;
  b->k = k;
  b->ntiles = (int32_t)sorted.size();
  if (const char* env = std::getenv("MPH_BATCH_SCHEDULE")) b->sched = std::atoi(env);
  if (b->sched < 0 || b->sched > 2) b->sched = 2;
  b->nwg = (int32_t)std::max<int64_t>(
      1, std::min<int64_t>(resident_workgroups(vt, k, b->sched), (int64_t)b->ntiles));
  hipError_t e = hipMalloc(&b->d_mats, sizeof(BatchMat) * std::max<int32_t>(nmat, 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_tiles, sizeof(BatchTile) * std::max<int32_t>(b->ntiles, 1));
  if (e == hipSuccess)
    e = hipMalloc(&b->d_rows, sizeof(int32_t) * std::max<size_t>(ts.rows.size(), 1));
  // longest-processing-time packing into per-workgroup lists (with
  // affinity: a matrix's tiles to the workgroups of its XCD only)
  std::vector<std::vector<int32_t>> lists(b->nwg);
  {
    using Item = std::pair<int64_t, int32_t>;
    const int nheap = affinity ? kQueues : 1;
    std::vector<std::priority_queue<Item, std::vector<Item>, std::greater<Item>>> heaps(nheap);
    for (int32_t w = 0; w < b->nwg; w++) heaps[affinity ? w % kQueues : 0].push({0, w});
    for (size_t q = 0; q < order.size(); q++) {
      int h = affinity ? mat_queue[tiles[order[q]].mat] : 0;
      if (heaps[h].empty()) h = 0;  // fewer workgroups than XCDs
      auto top = heaps[h].top();
      heaps[h].pop();
      lists[top.second].push_back((int32_t)order[q]);
      heaps[h].push({top.first + cost[order[q]], top.second});
    }
  }
  std::vector<BatchTile> packed;
  packed.reserve(order.size());
  std::vector<int32_t> woff(b->nwg + 1, 0);
  for (int32_t w = 0; w < b->nwg; w++) {
    woff[w] = (int32_t)packed.size();
    std::stable_sort(lists[w].begin(), lists[w].end(),
                     [&](int32_t a, int32_t c) { return tiles[a].mat < tiles[c].mat; });
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
  if (e == hipSuccess && !ts.rows.empty())
    e = hipMemcpyAsync(b->d_rows, ts.rows.data(), sizeof(int32_t) * ts.rows.size(),
                       hipMemcpyHostToDevice, s);
  // every tile in matrix and row order (one workgroup per tile)
  std::vector<BatchTile> all = tiles;
  if (e == hipSuccess) e = hipMalloc(&b->d_tiles_all, sizeof(BatchTile) * std::max<int32_t>(b->ntiles, 1));
  if (e == hipSuccess && b->ntiles > 0)
    e = hipMemcpyAsync(b->d_tiles_all, all.data(), sizeof(BatchTile) * b->ntiles,
                       hipMemcpyHostToDevice, s);
  // about four grabs per workgroup per sub-queue, at least 8 tiles each:
  // one counter update per tile serialises the workgroups on the counters
  b->grab = (int32_t)std::max<int64_t>(
      8, std::min<int64_t>(32, (int64_t)b->ntiles / std::max<int64_t>(1, (int64_t)b->nwg * 4)));
  if (e == hipSuccess) e = hipMalloc(&b->d_counter, sizeof(int32_t) * kQueues * 3);
  if (e == hipSuccess) e = hipMalloc(&b->d_qoff, sizeof(int32_t) * (kQueues + 1));
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_qoff, qoff.data(), sizeof(int32_t) * (kQueues + 1),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && nmat > 0)
    e = hipMemcpyAsync(b->d_mats, hm.data(), sizeof(BatchMat) * nmat, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && b->ntiles > 0)
    e = hipMemcpyAsync(b->d_tiles, sorted.data(), sizeof(BatchTile) * b->ntiles,
                       hipMemcpyHostToDevice, s);
  b->ncomp = (int32_t)ts.comp.size();
  if (e == hipSuccess && b->ncomp > 0) e = hipMalloc(&b->d_comp, sizeof(BatchTile) * b->ncomp);
  if (e == hipSuccess && b->ncomp > 0)
    e = hipMemcpyAsync(b->d_comp, ts.comp.data(), sizeof(BatchTile) * b->ncomp,
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
  dfree(b->d_tiles_all);
  dfree(b->d_woff);
  dfree(b->d_qoff);
  dfree(b->d_rows);
  dfree(b->d_counter);
  dfree(b->d_comp);
  for (void* p : b->owned) dfree(p);
  delete b;
  return MPH_OK;
}

int mph_batch_run(mph_batch_t b, int init, mph_stream_t stream) {
  if (!b) return invalid("null batch");
  return batch_launch(b, init ? 1 : 0, b->d_counter, reinterpret_cast<hipStream_t>(stream));
}

int mph_batch_run_graph(mph_batch_t b, int init, mph_stream_t stream) {
  if (!b) return invalid("null batch");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int g = init ? 1 : 0;
  if (!b->exec[g]) {
    hipStream_t cap = nullptr;
    MPH_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cap);
      MPH_HIP_CHECK(e);
    }
    const int st = batch_launch(b, g, b->d_counter + kQueues * (1 + g), cap);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cap, &graph);
    (void)hipStreamDestroy(cap);
    if (st != MPH_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (st != MPH_OK) return st;
      MPH_HIP_CHECK(e);
    }
    e = hipGraphInstantiate(&b->exec[g], graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      b->exec[g] = nullptr;
      MPH_HIP_CHECK(e);
    }
    b->graph[g] = graph;
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
