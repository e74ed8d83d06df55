// batch.hip -- batched SpMM: many matrices of uneven size and structure,
// each with k right-hand sides, in ONE launch over a list of (matrix, row
// block) tiles; replayable through a hipGraph.
//
// There is no counterpart in the reference (its multiply takes one matrix and
// one vector, core/src/Morpheus_Multiply.hpp:57-78); this is the GPU form of
// running Morpheus::multiply over a set of DynamicMatrix objects.
//
// Design (gfx950):
//  * X / Y are row-major with k columns.  A row is served by KL lanes, each
//    owning Q consecutive columns (16 bytes: 2 FP64 / 4 FP32 values; fewer
//    when k is small), so every x / y access of a wave is whole 16-byte
//    pieces of contiguous X / Y rows, and rows of adjacent lanes are adjacent
//    in memory.
//  * One workgroup per tile; each XCD takes a contiguous range of the tiles
//    (matrix and row order), so neighbouring row blocks share its L2.  A tile
//    record carries its matrix's record: the workgroup's first round trip
//    brings both.  DIA / ELL members and CSR-shaped members (CSR, COO, HYB,
//    HDC) have kernels of their own, each compiled for the registers its
//    tiles need, the first on a second stream of the batch's, so that the
//    two run side by side.
//  * Row tiles (DIA and ELL matrices, and the rows of at most kGroupRowMax
//    entries of CSR-shaped ones): a lane group takes R rows, 256 / KL rows
//    apart, and every batch loads 8 entries across them (R rows x 8 / R
//    entries) -- values, column indices and X pieces all issued before the
//    first multiply-add -- so a lane keeps 8 gathers in flight and the fixed
//    costs of a tile (its record, the store of Y) are spread over R times as
//    many rows.  Each row still accumulates its entries in order into one
//    sum per column, whatever R.
//  * The short rows of a CSR-shaped member without a DIA / ELL part go to row
//    tiles over a row list sorted by item count, so the rows of a wave have
//    about the same length.  CSR-shaped rows of kShortRowMax + 1 ..
//    kGroupRowMax entries: list tiles, one row per lane group; of
//    kGroupRowMax + 1 .. kSubRowMax entries (members without a DIA / ELL
//    part): kSubRows rows per wave; up to kWaveRowMax entries: one wave per
//    row; longer rows: the whole workgroup.
//  * DIA members whose offsets span at most kStageSpan stage the X rows of a
//    tile in LDS (each is read by every diagonal).  FP32 / BF16
//    rows of more than kCompensateLen entries are summed as Dot2 pairs by a
//    kernel of their own (one wave per row), launched on a second stream
//    forked from the caller's, so it runs beside the tile kernel.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "mph_internal.hpp"

namespace mph {

enum {
  TILE_ROWS = 0,
  TILE_WAVE = 1,
  TILE_WHOLE = 2,
  TILE_ROWS_STAGED = 3,
  TILE_LIST = 4,
  TILE_SORTED = 5,
  TILE_SUB = 6
};
enum { BPRE_NONE = 0, BPRE_DIA = 1, BPRE_ELL = 2 };

// A member: an optional DIA / ELL part (pre) added to every row, and an
// optional CSR-shaped part (ro != nullptr).
struct BatchMat {
  int64_t nrows, ncols;
  const int32_t* ro;  // CSR-shaped part (nullptr: none)
  const int32_t* ci;
  const void* val;
  int32_t pre;        // BPRE_*
  int32_t npre;       // diagonals (DIA) / slots per row (ELL)
  int64_t ppitch;     // DIA / ELL pitch
  const int32_t* pidx;  // DIA offsets / ELL column indices
  const void* pval;     // DIA / ELL values
  const void* x;
  void* y;
  int32_t smin, smax;  // DIA: smallest and largest offset (staged tiles)
};

// TILE_ROWS: rows [r0, r1); TILE_WAVE: entries r0 .. r0 + r1 - 1 of the
// batch's row list; TILE_WHOLE and the Dot2 rows: row r0.
struct BatchTile {
  BatchMat m;
  int32_t kind;
  int32_t r0;
  int32_t r1;
  int32_t pad;
};

constexpr int kBT = 256;
constexpr int kB = 8;              // entries per lane per batch
constexpr int kGroupRowMax = 32;   // rows a lane group takes in a row tile
// with more than one row per lane group, row tiles of CSR-shaped members take
// the rows of at most kShortRowMax entries (a batch covers a wave's rows
// until its longest is done); the rows of kShortRowMax + 1 .. kGroupRowMax
// entries go to list tiles, one row per lane group
constexpr int kShortRowMax = 8;
constexpr int kWaveRowMax = 512;   // rows one wave takes
constexpr int kWaveRows = 1;       // rows per wave of a one-wave tile
constexpr int kSubRows = 4;        // rows per wave of a sub-wave tile
constexpr int kSubRowMax = 128;    // rows of kGroupRowMax + 1 .. this: sub-wave tiles
// DIA row tiles whose offsets span at most kStageSpan (banded matrices)
// stage the X rows they read in LDS: each X row is then fetched once per
// tile instead of once per diagonal.  (Staging only the X rows of a group of
// diagonals -- the three central ones of a 2D stencil -- put a branch per
// entry into the batch and more registers: measured, no faster.)
constexpr int kStageSpan = 64;
// Rows per lane group of a DIA / ELL row tile (r) and of a CSR-shaped one
// (rc).
struct BatchShape {
  int r = 4, rc = 4;
};

// Lane layout for k = KR columns: Q values per lane, KL lanes per row.
template <typename V, int KR>
struct Lay {
  using Vec = typename vtraits<V>::vec;
  using Acc = typename vtraits<V>::acc;
  static constexpr int Q = (KR * (int)sizeof(Vec) >= 16) ? 16 / (int)sizeof(Vec) : KR;
  static constexpr int KL = KR / Q;
  static constexpr int RS = kBT / KL;  // rows of a workgroup slab
};

// the Q accumulators of a lane for each of R rows
template <typename V, int KR, int R>
using AccRQ = typename Lay<V, KR>::Acc[R][Lay<V, KR>::Q];

// Global-memory views of the pointers a tile record holds.  The records are
// read from memory, so the compiler cannot prove where their pointers point
// and would emit flat loads and stores, whose waits also cover the scalar
// loads (every wait became vmcnt(0) lgkmcnt(0)).
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

// t[w % R] += v[w] * X[c[w]] (the lane's Q columns) over the batch, for the
// entries with bit w of ok set; every X piece is requested before the first
// product, and c[w] is a valid row of X even where the bit is clear (its
// product is dropped).  R = 1: one row's sum.
template <typename V, int KR, int R>
__device__ __forceinline__ void fma_batch(AccRQ<V, KR, R>& t, V (&v)[kB], const int32_t (&c)[kB], uint32_t ok,
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
      Acc& s = t[w % R][q];
      const Acc f = __builtin_fma(a, xp[w][q], s);
      s = ((ok >> w) & 1u) ? f : s;
    }
  }
}

// fma_batch with the X rows staged in LDS (c[w]: the row within the stage)
template <typename V, int KR, int R>
__device__ __forceinline__ void fma_batch_lds(AccRQ<V, KR, R>& t, V (&v)[kB],
                                              const int32_t (&c)[kB], uint32_t ok,
                                              const typename Lay<V, KR>::Vec* xs, int l) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  using Vec = typename L::Vec;
  Acc xp[kB][L::Q];
#pragma unroll
  for (int w = 0; w < kB; w++) {
    const Vec* p = xs + c[w] * KR + l * L::Q;
    if constexpr (L::Q * sizeof(Vec) == 16) {
      typedef Vec v16 __attribute__((ext_vector_type(16 / sizeof(Vec))));
      const v16 u = *(const v16*)p;
#pragma unroll
      for (int q = 0; q < L::Q; q++) xp[w][q] = (Acc)u[q];
    } else {
#pragma unroll
      for (int q = 0; q < L::Q; q++) xp[w][q] = (Acc)p[q];
    }
  }
  hold_raw(v);
#pragma unroll
  for (int w = 0; w < kB; w++) {
    const Acc a = vtraits<V>::widen(v[w]);
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
      Acc& s = t[w % R][q];
      const Acc f = __builtin_fma(a, xp[w][q], s);
      s = ((ok >> w) & 1u) ? f : s;
    }
  }
}

// wave-wide maximum (a tile's loop bound, so that its batches stay uniform)
__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
  for (int s = kWave >> 1; s > 0; s >>= 1) v = max(v, shfl_xor(v, s));
  return v;
}

// A row tile: lane group g takes rows r0 + g + RS * i (i < R), and batch b
// covers entries b * E .. b * E + E - 1 (E = 8 / R) of each of them: first
// the PRE part's P slots (DIA diagonals / ELL slots), then, with CSR, the
// rows' items.  An entry's slot or item number depends on the batch and
// slot only (uniform across the wave); out-of-range entries load a valid
// one and are masked, so no load sits behind a branch.
template <typename V, int KR, int R, int PRE, bool CSR, bool STAGE = false, bool LIST = false>
__device__ void tile_rows(const BatchTile& t, int init, const int32_t* __restrict__ rows = nullptr) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  constexpr int E = kB / R;
  const BatchMat& m = t.m;
  const int g = threadIdx.x / L::KL, l = threadIdx.x % L::KL;
  int32_t row[R], len[R], st[R];
  bool mine[R];
  int32_t lmax = 0;
#pragma unroll
  for (int i = 0; i < R; i++) {
    int32_t r;
    bool valid;
    if constexpr (LIST) {  // lane group g: entries r0 + g + RS i of the row list (r1 entries)
      const int32_t e = g + L::RS * i;
      valid = e < t.r1;
      r = rows[t.r0 + (valid ? e : 0)];
    } else {
      r = t.r0 + g + L::RS * i;
      valid = r < t.r1;
    }
    row[i] = valid ? r : (LIST ? r : t.r0);
    len[i] = 0;
    st[i] = 0;
    mine[i] = valid;
    if constexpr (CSR) {
      const Gp<int32_t> ro = gp<int32_t>(m.ro);
      st[i] = ro[row[i]];
      len[i] = ro[row[i] + 1] - st[i];
      // longer rows have tiles of their own
      constexpr int kMax = R > 1 ? kShortRowMax : kGroupRowMax;
      mine[i] = valid && (LIST || len[i] + (PRE != BPRE_NONE ? m.npre : 0) <= kMax);
      if (mine[i]) lmax = max(lmax, len[i]);
    }
  }
  Acc acc[R][L::Q];
#pragma unroll
  for (int i = 0; i < R; i++)
#pragma unroll
    for (int q = 0; q < L::Q; q++) acc[i][q] = 0;
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  // STAGE (DIA rows, offsets smin .. smax within kStageSpan): X rows base ..
  // base + n - 1 in LDS, zero where outside X
  constexpr int kStageRows = R * L::RS + kStageSpan;
  __shared__ typename L::Vec xs[STAGE ? kStageRows * KR : 1];
  int32_t base = 0;
  if constexpr (STAGE) {
    base = t.r0 + m.smin;
    const int32_t n = (t.r1 - t.r0) + (m.smax - m.smin);
    for (int p = threadIdx.x; p < n * L::KL; p += kBT) {
      const int32_t xr = base + p / L::KL;
      typename L::Acc u[L::Q];
      if (xr >= 0 && xr < m.ncols) {
        load_piece<typename L::Vec, L::Q, typename L::Acc>(
            x + (int64_t)xr * KR + (p % L::KL) * L::Q, u);
      } else {
#pragma unroll
        for (int q = 0; q < L::Q; q++) u[q] = 0;
      }
#pragma unroll
      for (int q = 0; q < L::Q; q++) xs[p * L::Q + q] = (typename L::Vec)u[q];
    }
    __syncthreads();
  }
  if constexpr (PRE != BPRE_NONE) {
    const int32_t P = m.npre;
    for (int32_t eb = 0; eb < P; eb += E) {
      V v[kB];
      int32_t c[kB];
      uint32_t ok = 0;
#pragma unroll
      for (int w = 0; w < kB; w++) {
        const int i = w % R;
        const int32_t e = eb + w / R;
        const int32_t ec = min(e, P - 1);
        const int64_t slot = m.ppitch * ec + row[i];
        bool in;
        if constexpr (PRE == BPRE_DIA) {
          const int32_t col = row[i] + gp<int32_t>(m.pidx)[ec];
          in = mine[i] && e < P && col >= 0 && col < m.ncols;
          c[w] = in ? (STAGE ? col - base : col) : 0;
        } else {
          const int32_t ci = gld<int32_t>(gp<int32_t>(m.pidx) + slot);
          in = mine[i] && e < P && ci >= 0;
          c[w] = in ? ci : 0;
        }
        v[w] = gld<V>(gp<V>(m.pval) + slot);
        ok |= (uint32_t)in << w;
      }
      if constexpr (STAGE)
        fma_batch_lds<V, KR, R>(acc, v, c, ok, xs, l);
      else
        fma_batch<V, KR, R>(acc, v, c, ok, x, l);
    }
  }
  if constexpr (CSR) {
    lmax = wave_max(lmax);
    for (int32_t eb = 0; eb < lmax; eb += E) {
      V v[kB];
      int32_t c[kB];
      uint32_t ok = 0;
#pragma unroll
      for (int w = 0; w < kB; w++) {
        const int i = w % R;
        const int32_t e = eb + w / R;
        const bool in = mine[i] && e < len[i];
        // (lmax > 0: the matrix has items, and item 0 is a valid one)
        const uint32_t j = in ? (uint32_t)(st[i] + e) : 0u;
        c[w] = gld<int32_t>(gp<int32_t>(m.ci) + j);
        v[w] = gld<V>(gp<V>(m.val) + j);
        ok |= (uint32_t)in << w;
      }
      fma_batch<V, KR, R>(acc, v, c, ok, x, l);
    }
  }
#pragma unroll
  for (int i = 0; i < R; i++)
    if (mine[i])
      store_piece<typename L::Vec, L::Q, Acc>(
          gm<typename L::Vec>(m.y) + (int64_t)row[i] * KR + l * L::Q, acc[i], init);
}

// The entries of row r that a lane takes when G groups stride over it:
// prefix slots e0, e0 + step, ... then CSR items st + e0, st + e0 + step, ...
// 8 per batch, each batch handed to sink(v, c, ok).
template <typename V, int PRE, typename Sink>
__device__ __forceinline__ void row_entries(const BatchMat& m, int32_t r, int64_t st, int64_t len,
                                            int e0, int step, Sink&& sink) {
  if constexpr (PRE != BPRE_NONE) {
    const int32_t P = m.npre;
    const Gp<V> pv = gp<V>(m.pval) + r;
    const Gp<int32_t> pc = gp<int32_t>(m.pidx);
    for (int32_t eb = e0; eb < P; eb += kB * step) {
      V v[kB];
      int32_t c[kB];
      uint32_t ok = 0;
#pragma unroll
      for (int w = 0; w < kB; w++) {
        const int32_t e = eb + w * step;
        const int32_t ec = min(e, P - 1);  // a valid slot for the loads
        const int64_t off = m.ppitch * ec;
        bool in;
        if constexpr (PRE == BPRE_DIA) {
          const int32_t col = r + pc[ec];
          in = e < P && col >= 0 && col < m.ncols;
          c[w] = in ? col : 0;
        } else {
          const int32_t ci = gld<int32_t>(pc + r + off);
          in = e < P && ci >= 0;
          c[w] = in ? ci : 0;
        }
        v[w] = gld<V>(pv + off);
        ok |= (uint32_t)in << w;
      }
      sink(v, c, ok);
    }
  }
  if (len > 0) {
    const Gp<V> cv = gp<V>(m.val) + st;
    const Gp<int32_t> cc = gp<int32_t>(m.ci) + st;
    for (int64_t jb = e0; jb < len; jb += (int64_t)kB * step) {
      V v[kB];
      int32_t c[kB];
      uint32_t ok = 0;
#pragma unroll
      for (int w = 0; w < kB; w++) {
        const int64_t j = jb + (int64_t)w * step;
        const int64_t jj = j < len ? j : jb;  // jb < len: a valid item
        c[w] = gld<int32_t>(cc + jj);
        v[w] = gld<V>(cv + jj);
        ok |= (uint32_t)(j < len) << w;
      }
      sink(v, c, ok);
    }
  }
}

// [st, st + len): the CSR items of row r (none without a CSR part)
__device__ __forceinline__ void csr_extent(const BatchMat& m, int32_t r, int64_t& st,
                                           int64_t& len) {
  st = 0;
  len = 0;
  if (m.ro) {
    const Gp<int32_t> ro = gp<int32_t>(m.ro);
    st = ro[r];
    len = ro[r + 1] - st;
  }
}

// one row per wave (up to kWaveRows rows each): the wave's G = 64 / KL entry
// groups stride over the row
template <typename V, int KR, int PRE>
__device__ void tile_wave(const BatchTile& t, const int32_t* __restrict__ rows, int init) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  constexpr int G = kWave / L::KL;
  const BatchMat& m = t.m;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int g = lane / L::KL, l = lane % L::KL;
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  for (int i = wave; i < t.r1; i += kBT / kWave) {
    const int32_t r = rows[t.r0 + i];
    int64_t st, len;
    csr_extent(m, r, st, len);
    Acc acc[1][L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; q++) acc[0][q] = 0;
    row_entries<V, PRE>(m, r, st, len, g, G, [&](V(&v)[kB], const int32_t(&c)[kB], uint32_t ok) {
      fma_batch<V, KR, 1>(acc, v, c, ok, x, l);
    });
#pragma unroll
    for (int q = 0; q < L::Q; q++)
      for (int s = kWave >> 1; s >= L::KL; s >>= 1) acc[0][q] += shfl_xor(acc[0][q], s);
    if (g == 0)
      store_piece<typename L::Vec, L::Q, Acc>(
          gm<typename L::Vec>(m.y) + (int64_t)r * KR + l * L::Q, acc[0], init);
  }
}

// kSubRows rows per wave (a row list): each row's 64 / kSubRows lanes, G =
// 64 / (kSubRows KL) entry groups, stride over it.  For rows of a few dozen
// items a whole wave per row leaves most of its lanes' batch slots empty and
// takes four times the tiles.
template <typename V, int KR, int PRE>
__device__ void tile_sub(const BatchTile& t, const int32_t* __restrict__ rows, int init) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  constexpr int S = (kWave / L::KL < kSubRows) ? kWave / L::KL : kSubRows;
  constexpr int LW = kWave / S;  // lanes per row
  constexpr int G = LW / L::KL;  // entry groups per row
  const BatchMat& m = t.m;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int sub = lane / LW, sl = lane % LW;
  const int g = sl / L::KL, l = sl % L::KL;
  const int i = wave * S + sub;
  const bool valid = i < t.r1;
  const int32_t r = rows[t.r0 + (valid ? i : 0)];
  int64_t st, len;
  csr_extent(m, r, st, len);
  if (!valid) len = 0;
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  Acc acc[1][L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; q++) acc[0][q] = 0;
  row_entries<V, PRE>(m, r, st, len, g, G, [&](V(&v)[kB], const int32_t(&c)[kB], uint32_t ok) {
    fma_batch<V, KR, 1>(acc, v, c, ok, x, l);
  });
  // (xor distances below LW stay inside the row's lanes)
#pragma unroll
  for (int q = 0; q < L::Q; q++)
    for (int s = LW >> 1; s >= L::KL; s >>= 1) acc[0][q] += shfl_xor(acc[0][q], s);
  if (valid && g == 0)
    store_piece<typename L::Vec, L::Q, Acc>(
        gm<typename L::Vec>(m.y) + (int64_t)r * KR + l * L::Q, acc[0], init);
}

// one row per workgroup: its G = 256 / KL entry groups stride over the row
template <typename V, int KR, int PRE>
__device__ void tile_whole(const BatchTile& t, int init, typename Lay<V, KR>::Acc* s_red) {
  using L = Lay<V, KR>;
  using Acc = typename L::Acc;
  constexpr int G = kBT / L::KL;
  const BatchMat& m = t.m;
  const int g = threadIdx.x / L::KL, l = threadIdx.x % L::KL;
  const int32_t r = t.r0;
  const Gp<typename L::Vec> x = gp<typename L::Vec>(m.x);
  int64_t st, len;
  csr_extent(m, r, st, len);
  Acc acc[1][L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; q++) acc[0][q] = 0;
  row_entries<V, PRE>(m, r, st, len, g, G, [&](V(&v)[kB], const int32_t(&c)[kB], uint32_t ok) {
    fma_batch<V, KR, 1>(acc, v, c, ok, x, l);
  });
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int q = 0; q < L::Q; q++) {
    for (int s = kWave >> 1; s >= L::KL; s >>= 1) acc[0][q] += shfl_xor(acc[0][q], s);
    if (lane < L::KL) s_red[(wave * L::KL + lane) * L::Q + q] = acc[0][q];
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
        gm<typename L::Vec>(m.y) + (int64_t)r * KR + threadIdx.x * L::Q, tot, init);
  }
}

// Tile kinds of the kernel's switch: the tile's shape and its member's parts.
constexpr int tile_code(int kind, int pre, bool csr) { return kind * 8 + pre * 2 + (csr ? 1 : 0); }

// One workgroup per tile; each XCD takes a contiguous range of the tiles (in
// matrix and row order), so neighbouring row blocks of a matrix share that
// XCD's L2.  Two kernels, each compiled for the registers its tiles need:
// the row tiles of DIA / ELL members (R rows per lane group, 6 waves per
// SIMD), and the tiles of CSR-shaped members (RC rows per lane group, one-
// wave and whole-workgroup rows; 5 waves per SIMD).
template <typename V, int KR, int R>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(6))) void batch_slab_kernel(
    const BatchTile* __restrict__ tiles, int init) {
  const BatchTile t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  switch (t.kind) {
    case tile_code(TILE_ROWS, BPRE_DIA, false): tile_rows<V, KR, R, BPRE_DIA, false>(t, init); break;
    case tile_code(TILE_ROWS_STAGED, BPRE_DIA, false):
      tile_rows<V, KR, R, BPRE_DIA, false, true>(t, init);
      break;
    case tile_code(TILE_ROWS, BPRE_ELL, false): tile_rows<V, KR, R, BPRE_ELL, false>(t, init); break;
    default: tile_rows<V, KR, R, BPRE_NONE, false>(t, init); break;  // no entries: y = 0
  }
}

template <typename V, int KR, int RC>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(5))) void batch_rows_kernel(
    const BatchTile* __restrict__ tiles, const int32_t* __restrict__ rows, int init) {
  using Acc = typename vtraits<V>::acc;
  __shared__ Acc s_red[kBT / kWave * Lay<V, KR>::KL * Lay<V, KR>::Q];
  const BatchTile t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  switch (t.kind) {
    case tile_code(TILE_ROWS, BPRE_NONE, true): tile_rows<V, KR, RC, BPRE_NONE, true>(t, init); break;
    case tile_code(TILE_ROWS, BPRE_DIA, true): tile_rows<V, KR, RC, BPRE_DIA, true>(t, init); break;
    case tile_code(TILE_ROWS, BPRE_ELL, true): tile_rows<V, KR, RC, BPRE_ELL, true>(t, init); break;
    case tile_code(TILE_LIST, BPRE_NONE, true):
      tile_rows<V, KR, 1, BPRE_NONE, true, false, true>(t, init, rows);
      break;
    case tile_code(TILE_SORTED, BPRE_NONE, true):
      tile_rows<V, KR, RC, BPRE_NONE, true, false, true>(t, init, rows);
      break;
    case tile_code(TILE_SUB, BPRE_NONE, true): tile_sub<V, KR, BPRE_NONE>(t, rows, init); break;
    case tile_code(TILE_LIST, BPRE_DIA, true):
      tile_rows<V, KR, 1, BPRE_DIA, true, false, true>(t, init, rows);
      break;
    case tile_code(TILE_LIST, BPRE_ELL, true):
      tile_rows<V, KR, 1, BPRE_ELL, true, false, true>(t, init, rows);
      break;
    case tile_code(TILE_WAVE, BPRE_NONE, true): tile_wave<V, KR, BPRE_NONE>(t, rows, init); break;
    case tile_code(TILE_WAVE, BPRE_DIA, true): tile_wave<V, KR, BPRE_DIA>(t, rows, init); break;
    case tile_code(TILE_WAVE, BPRE_ELL, true): tile_wave<V, KR, BPRE_ELL>(t, rows, init); break;
    case tile_code(TILE_WHOLE, BPRE_NONE, true): tile_whole<V, KR, BPRE_NONE>(t, init, s_red); break;
    case tile_code(TILE_WHOLE, BPRE_DIA, true): tile_whole<V, KR, BPRE_DIA>(t, init, s_red); break;
    case tile_code(TILE_WHOLE, BPRE_ELL, true): tile_whole<V, KR, BPRE_ELL>(t, init, s_red); break;
    default: break;
  }
}

// FP32 / BF16 rows of more than kCompensateLen entries, one per wave, summed
// as Dot2 pairs per column: the wave's G = 64 / KL entry groups stride over
// the row (8 entries per lane per batch, X pieces as in the tile kernel),
// then a fixed tree over the groups; the old y joins through TwoSum when
// accumulating.
template <typename V, int KR, int PRE>
__device__ void comp_row(const BatchTile& t, int init) {
  using L = Lay<V, KR>;
  using Vec = typename L::Vec;
  constexpr int G = kWave / L::KL;
  const BatchMat& m = t.m;
  const int lane = threadIdx.x % kWave, g = lane / L::KL, l = lane % L::KL;
  const int32_t r = t.r0;
  int64_t st, len;
  csr_extent(m, r, st, len);
  const Gp<Vec> x = gp<Vec>(m.x);
  csum<float> cs[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; q++) cs[q] = csum<float>{0, 0};
  row_entries<V, PRE>(m, r, st, len, g, G, [&](V(&v)[kB], const int32_t(&c)[kB], uint32_t ok) {
    float xp[kB][L::Q];
#pragma unroll
    for (int w = 0; w < kB; w++)
      load_piece<Vec, L::Q, float>(x + (int64_t)c[w] * KR + l * L::Q, xp[w]);
    hold_raw(v);
#pragma unroll
    for (int w = 0; w < kB; w++) {
      if ((ok >> w) & 1u) {
        const float a = vtraits<V>::widen(v[w]);
#pragma unroll
        for (int q = 0; q < L::Q; q++) cs[q].add_prod(a, xp[w][q]);
      }
    }
  });
#pragma unroll
  for (int q = 0; q < L::Q; q++)
    for (int s = kWave >> 1; s >= L::KL; s >>= 1) {
      const float s2 = shfl_xor(cs[q].s, s), c2 = shfl_xor(cs[q].c, s);
      cs[q].combine(s2, c2);
    }
  if (g == 0) {
    const Gm<Vec> y = gm<Vec>(m.y) + (int64_t)r * KR + l * L::Q;
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
      csum<float> tq = cs[q];
      if (!init) tq.combine((float)y[q], 0.0f);
      y[q] = (Vec)tq.value();
    }
  }
}

template <typename V, int KR>
__global__ __launch_bounds__(kBT) void batch_comp_rows_kernel(const BatchTile* __restrict__ tiles,
                                                              int32_t ntiles, int init) {
  const int32_t i = __builtin_amdgcn_readfirstlane(
      (int32_t)(blockIdx.x * (kBT / kWave) + threadIdx.x / kWave));
  if (i >= ntiles) return;
  const BatchTile t = tiles[i];
  switch (t.m.pre) {
    case BPRE_DIA: comp_row<V, KR, BPRE_DIA>(t, init); break;
    case BPRE_ELL: comp_row<V, KR, BPRE_ELL>(t, init); break;
    default: comp_row<V, KR, BPRE_NONE>(t, init); break;
  }
}

}  // namespace mph

// ------------------------------------------------------------------ host
struct mph_batch_s {
  int vt = MPH_F32;
  int k = 1;
  mph::BatchShape shape;
  int32_t nslab = 0;  // the first nslab tiles: DIA / ELL row tiles (slab kernel)
  int32_t ntiles = 0;
  mph::BatchTile* d_tiles = nullptr;  // every tile, matrix and row order
  int32_t* d_rows = nullptr;          // row ids of the one-wave tiles
  // FP32 / BF16 rows of more than kCompensateLen entries (Dot2 kernel),
  // launched on `side`, forked from and joined to the caller's stream
  mph::BatchTile* d_comp = nullptr;
  int32_t ncomp = 0;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // the slab kernel's stream when both kernels run
  hipStream_t side2 = nullptr;
  hipEvent_t fork2 = nullptr, join2 = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<void*> owned;  // row offsets derived for COO members
};

namespace mph {
namespace {

template <typename V, int KR>
int launch_batch(mph_batch_s* b, int init, hipStream_t s) {
  const bool comp = !std::is_same<V, double>::value && b->ncomp > 0;
  if (comp) {
    // the Dot2 rows run beside the tile kernel (different rows of Y)
    MPH_HIP_CHECK(hipEventRecord(b->fork, s));
    MPH_HIP_CHECK(hipStreamWaitEvent(b->side, b->fork, 0));
    const unsigned blocks = (unsigned)((b->ncomp + kBT / kWave - 1) / (kBT / kWave));
    hipLaunchKernelGGL((batch_comp_rows_kernel<V, KR>), dim3(blocks), dim3(kBT), 0, b->side,
                       b->d_comp, b->ncomp, init);
    MPH_HIP_CHECK(hipGetLastError());
    MPH_HIP_CHECK(hipEventRecord(b->join, b->side));
  }
  const BatchShape& sh = b->shape;
  const bool slabs = b->nslab > 0, csr = b->ntiles > b->nslab;
  if (slabs && csr) {
    // the DIA / ELL members' kernel runs beside the CSR-shaped members' one
    MPH_HIP_CHECK(hipEventRecord(b->fork2, s));
    MPH_HIP_CHECK(hipStreamWaitEvent(b->side2, b->fork2, 0));
  }
  if (slabs) {
    hipStream_t ss = csr ? b->side2 : s;
    const dim3 grid((unsigned)b->nslab), block(kBT);
    if (sh.r == 1)
      hipLaunchKernelGGL((batch_slab_kernel<V, KR, 1>), grid, block, 0, ss, b->d_tiles, init);
    else
      hipLaunchKernelGGL((batch_slab_kernel<V, KR, 4>), grid, block, 0, ss, b->d_tiles, init);
    MPH_HIP_CHECK(hipGetLastError());
    if (csr) MPH_HIP_CHECK(hipEventRecord(b->join2, b->side2));
  }
  if (csr) {
    const dim3 grid((unsigned)(b->ntiles - b->nslab)), block(kBT);
    const BatchTile* tiles = b->d_tiles + b->nslab;
    if (sh.rc == 4)
      hipLaunchKernelGGL((batch_rows_kernel<V, KR, 4>), grid, block, 0, s, tiles, b->d_rows, init);
    else
      hipLaunchKernelGGL((batch_rows_kernel<V, KR, 1>), grid, block, 0, s, tiles, b->d_rows, init);
    MPH_HIP_CHECK(hipGetLastError());
  }
  if (slabs && csr) MPH_HIP_CHECK(hipStreamWaitEvent(s, b->join2, 0));
  if (comp) MPH_HIP_CHECK(hipStreamWaitEvent(s, b->join, 0));
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

// lanes per row of the k-column layout (16 bytes of a row of X per lane)
int lanes_per_row(int vt, int k) {
  const int vec = (int)vector_size(vt);
  const int q = k * vec >= 16 ? 16 / vec : k;
  return k / q;
}

// The tiles of one member (record m): row tiles of R * 256 / KL rows, and
// for a CSR-shaped part (host row offsets `ro`; P prefix slots per row) the
// one-wave, whole-workgroup and Dot2 tiles of its longer rows.
struct TileSet {
  std::vector<BatchTile> slabs;  // DIA / ELL row tiles
  std::vector<BatchTile> tiles;  // tiles of CSR-shaped members
  std::vector<BatchTile> comp;
  std::vector<int32_t> rows;  // row ids of one-wave tiles (r0 indexes this list)
};

// MPH_BATCH_SORT_ROWS=0: short rows of CSR-shaped members in row order
bool getenv_sorted() {
  const char* env = std::getenv("MPH_BATCH_SORT_ROWS");
  return !env || std::atoi(env) != 0;
}
// MPH_BATCH_SUB_MAX: the longest row of the sub-wave tiles (0: none)
int64_t getenv_sub_max() {
  const char* env = std::getenv("MPH_BATCH_SUB_MAX");
  return env ? std::max(0, std::atoi(env)) : kSubRowMax;
}

void tile_member(TileSet& ts, const BatchMat& m, const std::vector<int32_t>* ro, int vt, int k,
                 const BatchShape& sh, int wave_rows, bool stage_ok) {
  const int R = ro ? sh.rc : sh.r;
  const int64_t rs = kBT / lanes_per_row(vt, k);
  const int64_t P = m.pre != BPRE_NONE ? m.npre : 0;
  const bool comp_type = vt != MPH_F64;
  auto tile = [&](int kind, int64_t r0, int64_t r1) {
    BatchTile t{};
    t.m = m;
    t.kind = tile_code(kind, m.pre, m.ro != nullptr);
    t.r0 = (int32_t)r0;
    t.r1 = (int32_t)r1;
    return t;
  };
  if (!ro) {  // a DIA / ELL matrix: every row has the same P slots
    if (comp_type && P > kCompensateLen) {
      for (int64_t r = 0; r < m.nrows; r++) ts.comp.push_back(tile(TILE_WHOLE, r, r + 1));
      return;
    }
    const bool stage = m.pre == BPRE_DIA && stage_ok && m.smax - m.smin <= kStageSpan;
    for (int64_t r = 0; r < m.nrows; r += R * rs)
      ts.slabs.push_back(tile(stage ? TILE_ROWS_STAGED : TILE_ROWS, r, std::min(m.nrows, r + R * rs)));
    return;
  }
  std::vector<int32_t> wave;
  auto flush_wave = [&]() {
    if (wave.empty()) return;
    ts.tiles.push_back(tile(TILE_WAVE, (int64_t)ts.rows.size(), (int64_t)wave.size()));
    ts.rows.insert(ts.rows.end(), wave.begin(), wave.end());
    wave.clear();
  };
  std::vector<int32_t> subl;
  auto flush_sub = [&]() {
    if (subl.empty()) return;
    ts.tiles.push_back(tile(TILE_SUB, (int64_t)ts.rows.size(), (int64_t)subl.size()));
    ts.rows.insert(ts.rows.end(), subl.begin(), subl.end());
    subl.clear();
  };
  const int64_t sub_max = P == 0 ? getenv_sub_max() : 0;
  std::vector<int32_t> list;
  auto flush_list = [&]() {
    if (list.empty()) return;
    ts.tiles.push_back(tile(TILE_LIST, (int64_t)ts.rows.size(), (int64_t)list.size()));
    ts.rows.insert(ts.rows.end(), list.begin(), list.end());
    list.clear();
  };
  const int64_t row_max = R > 1 ? kShortRowMax : kGroupRowMax;
  // a member without a DIA / ELL part, R rows per lane group: its short rows
  // (at most row_max items) sorted by item count into row-list tiles, so a
  // wave's rows have about the same length and its batches end with them
  // (in row order a wave of power-law rows mostly runs to its one longest)
  const bool sorted = R > 1 && P == 0 && getenv_sorted();
  if (sorted) {
    std::vector<std::vector<int32_t>> by_len((size_t)row_max + 1);
    for (int64_t r = 0; r < m.nrows; r++) {
      const int64_t items = (int64_t)(*ro)[r + 1] - (*ro)[r];
      if (items <= row_max) by_len[(size_t)items].push_back((int32_t)r);
    }
    std::vector<int32_t> order;
    order.reserve((size_t)m.nrows);
    for (const auto& b : by_len) order.insert(order.end(), b.begin(), b.end());
    const int64_t per = R * rs;
    for (size_t i = 0; i < order.size(); i += (size_t)per) {
      const int64_t cnt = std::min<int64_t>(per, (int64_t)(order.size() - i));
      ts.tiles.push_back(tile(TILE_SORTED, (int64_t)ts.rows.size(), cnt));
      ts.rows.insert(ts.rows.end(), order.begin() + (int64_t)i, order.begin() + (int64_t)i + cnt);
    }
  }
  for (int64_t r0 = 0; r0 < m.nrows; r0 += R * rs) {
    const int64_t r1 = std::min(m.nrows, r0 + R * rs);
    for (int64_t r = r0; r < r1; r++) {
      const int64_t len = (int64_t)(*ro)[r + 1] - (*ro)[r] + P;
      if (len <= row_max) continue;
      if (len <= kGroupRowMax) {
        list.push_back((int32_t)r);
        if ((int64_t)list.size() == rs) flush_list();
        continue;
      }
      if (comp_type && len > kCompensateLen) {
        ts.comp.push_back(tile(TILE_WHOLE, r, r + 1));
      } else if (len > kWaveRowMax) {
        ts.tiles.push_back(tile(TILE_WHOLE, r, r + 1));
      } else if (len <= sub_max) {
        subl.push_back((int32_t)r);
        if ((int)subl.size() == kSubRows * kBT / kWave) flush_sub();
      } else {
        wave.push_back((int32_t)r);
        if ((int)wave.size() == wave_rows * kBT / kWave) flush_wave();
      }
    }
    // the block's tile covers every row of [r0, r1); its lane groups leave
    // the longer rows to their own tiles (sorted: the short rows have theirs)
    if (!sorted) ts.tiles.push_back(tile(TILE_ROWS, r0, r1));
  }
  flush_wave();
  flush_sub();
  flush_list();
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
  BatchShape sh;
  int wave_rows = kWaveRows;
  if (const char* env = std::getenv("MPH_BATCH_ROWS_PER_GROUP")) sh.r = std::atoi(env) == 1 ? 1 : 4;
  if (const char* env = std::getenv("MPH_BATCH_CSR_ROWS_PER_GROUP")) sh.rc = std::atoi(env) == 1 ? 1 : 4;
  if (const char* env = std::getenv("MPH_BATCH_WAVE_ROWS")) wave_rows = std::max(1, std::min(8, std::atoi(env)));
  bool stage_ok = true;
  if (const char* env = std::getenv("MPH_BATCH_STAGE")) stage_ok = std::atoi(env) != 0;
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
      const bool dia = A->format == MPH_DIA || A->format == MPH_HDC;
      const bool ell = A->format == MPH_ELL || A->format == MPH_HYB;
      if (dia && A->ndiag > 0) {
        m.pre = BPRE_DIA;
        m.npre = (int32_t)A->ndiag;
        m.ppitch = A->dia_pitch;
        m.pidx = A->dia_off;
        m.pval = A->dia_v;
        std::vector<int32_t> off((size_t)A->ndiag);
        MPH_HIP_CHECK(hipMemcpyAsync(off.data(), A->dia_off, sizeof(int32_t) * A->ndiag,
                                     hipMemcpyDeviceToHost, s));
        MPH_HIP_CHECK(hipStreamSynchronize(s));
        m.smin = *std::min_element(off.begin(), off.end());
        m.smax = *std::max_element(off.begin(), off.end());
      } else if (ell && A->ell_loop > 0) {
        m.pre = BPRE_ELL;
        m.npre = (int32_t)A->ell_loop;
        m.ppitch = A->ell_pitch;
        m.pidx = A->ell_ci;
        m.pval = A->ell_v;
      }
      // an HDC without CSR entries is its DIA part, a HYB without COO
      // entries its ELL part
      if (A->format == MPH_CSR || (A->format == MPH_HDC && A->csr_nnz > 0)) {
        m.ro = A->csr_ro;
        m.ci = A->csr_ci;
        m.val = A->csr_v;
      } else if (A->format == MPH_HYB && A->coo_nnz > 0) {
        m.ro = A->rb.tail_ro;
        m.ci = A->coo_ci;
        m.val = A->coo_v;
      }
    }
    if (m.ncols == 0) {
      // no entry has a column: y = 0 (init) or unchanged, and X is empty
      m.pre = BPRE_NONE;
      m.ro = nullptr;
    }
    if (!m.ro) {
      tile_member(ts, m, nullptr, vt, k, sh, wave_rows, stage_ok);
      continue;
    }
    std::vector<int32_t> ro;
    if (A->nrows > 0) {
      ro.resize(A->nrows + 1);
      MPH_HIP_CHECK(hipMemcpyAsync(ro.data(), m.ro, sizeof(int32_t) * (A->nrows + 1),
                                   hipMemcpyDeviceToHost, s));
      MPH_HIP_CHECK(hipStreamSynchronize(s));
    }
    tile_member(ts, m, &ro, vt, k, sh, wave_rows, stage_ok);
  }

  mph_batch_s* b = new mph_batch_s();
  b->owned = owned;
  guard.keep = true;
  b->vt = vt;
  b->k = k;
  b->shape = sh;
  b->nslab = (int32_t)ts.slabs.size();
  ts.slabs.insert(ts.slabs.end(), ts.tiles.begin(), ts.tiles.end());
  ts.tiles.swap(ts.slabs);
  b->ntiles = (int32_t)ts.tiles.size();
  b->ncomp = (int32_t)ts.comp.size();
  hipError_t e = hipSuccess;
  if (b->ntiles > 0) e = hipMalloc(&b->d_tiles, sizeof(BatchTile) * b->ntiles);
  if (e == hipSuccess && b->ntiles > 0)
    e = hipMemcpyAsync(b->d_tiles, ts.tiles.data(), sizeof(BatchTile) * b->ntiles,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !ts.rows.empty()) e = hipMalloc(&b->d_rows, sizeof(int32_t) * ts.rows.size());
  if (e == hipSuccess && !ts.rows.empty())
    e = hipMemcpyAsync(b->d_rows, ts.rows.data(), sizeof(int32_t) * ts.rows.size(),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && b->ncomp > 0) e = hipMalloc(&b->d_comp, sizeof(BatchTile) * b->ncomp);
  if (e == hipSuccess && b->ncomp > 0)
    e = hipMemcpyAsync(b->d_comp, ts.comp.data(), sizeof(BatchTile) * b->ncomp,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && b->ncomp > 0) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
  if (e == hipSuccess && b->ncomp > 0) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess && b->ncomp > 0) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
  const bool two = b->nslab > 0 && b->ntiles > b->nslab;
  if (e == hipSuccess && two) e = hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking);
  if (e == hipSuccess && two) e = hipEventCreateWithFlags(&b->fork2, hipEventDisableTiming);
  if (e == hipSuccess && two) e = hipEventCreateWithFlags(&b->join2, hipEventDisableTiming);
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
  if (b->side) (void)hipStreamSynchronize(b->side);
  if (b->side2) (void)hipStreamSynchronize(b->side2);
  dfree(b->d_tiles);
  dfree(b->d_rows);
  dfree(b->d_comp);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->fork2) (void)hipEventDestroy(b->fork2);
  if (b->join2) (void)hipEventDestroy(b->join2);
  if (b->side2) (void)hipStreamDestroy(b->side2);
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
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cap);
      MPH_HIP_CHECK(e);
    }
    const int st = batch_launch(b, g, cap);
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
  if (nwg) *nwg = b->ncomp;
  return MPH_OK;
}

}  // extern "C"
