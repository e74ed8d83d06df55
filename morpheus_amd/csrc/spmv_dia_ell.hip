// spmv_dia_ell.hip -- DIA and ELL SpMV / SpMM.
//
// Replaces core/src/impl/Dia/HIP/Morpheus_Multiply_Impl.hpp:41-85 (kernel
// core/src/impl/Dia/Kernels/Morpheus_Multiply_Impl.hpp:37-84) and
// core/src/impl/Ell/HIP/Morpheus_Multiply_Impl.hpp:41-80 (kernel
// core/src/impl/Ell/Kernels/Morpheus_Multiply_Impl.hpp:37-65).
//
// Both are one thread per row over column-major (LayoutLeft) storage, so
// every diagonal / slot is one fully coalesced, non-temporal stream across a
// wavefront.  DIA offsets are wave-uniform and come through scalar loads;
// the x entries of clusters of nearby diagonals are staged in LDS per
// workgroup (dia_staged_kernel), and y is written once (the reference zeroes
// y in a separate launch and then read-modify-writes it per 256-diagonal
// chunk, and skips the zeroing when ndiag == 0).  ELL visits entries_per_row slots,
// not the 32-padded width, when the analysis has confirmed that the slots
// beyond entries_per_row hold only the -1 padding.  FP32 / BF16 SpMV takes
// two adjacent rows per thread (row-pair kernels below); FP32 / BF16 rows of
// more than kCompensateLen diagonals / slots take a Dot2-compensated kernel.
// SpMM gives each row a lane group (16 bytes of its columns per lane); with
// 4 or more lanes per row DIA gives each group 4 consecutive rows and loads
// the X rows of each run of consecutive diagonals once.  A split FP64 HDC
// whose CSR rows are short adds them in the staged DIA pass.
#include <algorithm>
#include <cstdlib>

#include "mph_internal.hpp"

namespace mph {

template <typename V>
struct DiaArgs {
  int64_t nrows, ncols;
  int32_t ndiag;
  int64_t pitch;
  const int32_t* __restrict__ off;
  const V* __restrict__ val;
  const typename vtraits<V>::vec* __restrict__ x;
  typename vtraits<V>::vec* __restrict__ y;
  int32_t ld;
  int init;
  // dia_staged_kernel<V, true>: HDC's CSR part, added per row after the
  // diagonals
  const int32_t* __restrict__ tro;
  const int32_t* __restrict__ tci;
  const V* __restrict__ tval;
};

// t[u][q] += v[u][w] * x[c[u][w]][q] over a batch of 8 diagonals / slots,
// skipping c < 0.  The x loads are unconditional (c < 0 reads x[0]) and the
// sums are selected rather than branched, so the batch's gathers are all in
// flight before the first product: under a branch each gather was waited on
// before the next was issued.  Same fused multiply-adds in the same order
// as t += v * x.  With many right-hand sides (KR * RPT > 2) the gathers of
// one diagonal / slot are issued together instead, to bound registers.
template <typename V, int KR, int RPT>
__device__ __forceinline__ void batch_fma(typename vtraits<V>::acc (&t)[RPT][KR],
                                          const V (&v)[RPT][8], const int64_t (&c)[RPT][8],
                                          const typename vtraits<V>::vec* __restrict__ x,
                                          int32_t ld) {
  using Acc = typename vtraits<V>::acc;
  if constexpr (KR * RPT <= 2) {
    Acc xq[RPT][8][KR];
#pragma unroll
    for (int w = 0; w < 8; w++)
#pragma unroll
      for (int u = 0; u < RPT; u++)
#pragma unroll
        for (int q = 0; q < KR; q++) xq[u][w][q] = (Acc)x[(c[u][w] >= 0 ? c[u][w] : 0) * ld + q];
#pragma unroll
    for (int w = 0; w < 8; w++)
#pragma unroll
      for (int u = 0; u < RPT; u++)
#pragma unroll
        for (int q = 0; q < KR; q++) {
          const Acc f = __builtin_fma(vtraits<V>::widen(v[u][w]), xq[u][w][q], t[u][q]);
          t[u][q] = c[u][w] >= 0 ? f : t[u][q];
        }
  } else {
#pragma unroll
    for (int w = 0; w < 8; w++)
#pragma unroll
      for (int u = 0; u < RPT; u++) {
        const int64_t xc = c[u][w] >= 0 ? c[u][w] : 0;
#pragma unroll
        for (int q = 0; q < KR; q++) {
          const Acc xq = (Acc)x[xc * ld + q];
          if (c[u][w] >= 0) t[u][q] += vtraits<V>::widen(v[u][w]) * xq;
        }
      }
  }
}

template <typename V, int KR, int RPT>
__global__ __launch_bounds__(256) void dia_kernel(DiaArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row0 = ((int64_t)b * 256 * RPT) + threadIdx.x;
  Acc t[RPT][KR];
#pragma unroll
  for (int u = 0; u < RPT; u++)
#pragma unroll
    for (int q = 0; q < KR; q++) t[u][q] = 0;
  // batches of 8 diagonals: offsets, values and x entries of the batch are
  // all requested before any of them is consumed
  // (The branch-free loads of ell_kernel made this kernel slower: S7
  // 31.6 -> 34.0 us, S27 210 -> 223 us in FP64; its value loads already
  // stream at the read rate of the device.)
  for (int d0 = 0; d0 < a.ndiag; d0 += 8) {
    V v[RPT][8];
    int64_t c[RPT][8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const int64_t o = d < a.ndiag ? (int64_t)a.off[d] : 0;
#pragma unroll
      for (int u = 0; u < RPT; u++) {
        const int64_t row = row0 + u * 256;
        const int64_t col = row + o;
        const bool ok = d < a.ndiag && row < a.nrows && col >= 0 && col < a.ncols;
        c[u][w] = ok ? col : -1;
        v[u][w] = ok ? ld_stream(a.val + a.pitch * d + row) : V{};
      }
    }
#pragma unroll
    for (int u = 0; u < RPT; u++) hold_raw(v[u]);
#pragma unroll
    for (int w = 0; w < 8; w++)
#pragma unroll
      for (int u = 0; u < RPT; u++)
        if (c[u][w] >= 0)
#pragma unroll
          for (int q = 0; q < KR; q++)
            t[u][q] += vtraits<V>::widen(v[u][w]) * (Acc)a.x[c[u][w] * a.ld + q];
  }
#pragma unroll
  for (int u = 0; u < RPT; u++) {
    const int64_t row = row0 + u * 256;
    if (row < a.nrows) {
#pragma unroll
      for (int q = 0; q < KR; q++) {
        Vec* yp = a.y + row * a.ld + q;
        *yp = a.init ? (Vec)t[u][q] : (Vec)((Acc)*yp + t[u][q]);
      }
    }
  }
}

// DIA SpMV with x staged in LDS per cluster of nearby diagonals (the
// clusters come from dia_analyse).  For each cluster of two or more
// diagonals the workgroup's 256 rows read x[r0 + omin .. r0 + 255 + omax]
// once, coalesced, into LDS; every diagonal of the cluster then takes its x
// entries from there instead of a global load (on the 27-point stencil the
// x loads per row drop from 27 to about 7).  Diagonals are summed in storage
// order, as dia_kernel sums them: the two kernels give identical results.
template <typename V, bool TAIL = false>
__global__ __launch_bounds__(256) void dia_staged_kernel(DiaArgs<V> a, DiaStage cl) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  __shared__ Vec xs[kDiaStageMax];
  const int64_t r0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256;
  const int64_t row = r0 + threadIdx.x;
  for (int c = 0; c < cl.n; c++)
    for (int j = threadIdx.x; j < cl.len[c]; j += 256) {
      const int64_t col = r0 + cl.omin[c] + j;
      xs[cl.base[c] + j] = (col >= 0 && col < a.ncols) ? a.x[col] : (Vec)0;
    }
  __syncthreads();
  Acc t = 0;
  for (int d0 = 0; d0 < a.ndiag; d0 += 8) {
    V v[8];
    int64_t col[8];
    bool ok[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const int64_t o = d < a.ndiag ? (int64_t)a.off[d] : 0;
      col[w] = row + o;
      ok[w] = d < a.ndiag && row < a.nrows && col[w] >= 0 && col[w] < a.ncols;
      v[w] = ok[w] ? ld_stream(a.val + a.pitch * d + row) : V{};
    }
    hold_raw(v);
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const int32_t lb = d < a.ndiag ? cl.lb[d] : kDiaDirect;  // uniform
      if (lb != kDiaDirect) {
        // (inside the cluster's window whenever ok)
        if (ok[w]) t += vtraits<V>::widen(v[w]) * (Acc)xs[lb + (int32_t)(col[w] - r0)];
      } else {
        if (ok[w]) t += vtraits<V>::widen(v[w]) * (Acc)a.x[col[w]];
      }
    }
  }
  if constexpr (TAIL) {
    // HDC: the row's CSR entries (a few at most), in order after the
    // diagonals -- the reference's DIA product followed by its CSR product
    // accumulating onto it
    if (row < a.nrows) {
      const int32_t s0 = a.tro[row], s1 = a.tro[row + 1];
      for (int32_t j = s0; j < s1; j++) t += vtraits<V>::widen(a.tval[j]) * (Acc)a.x[a.tci[j]];
    }
  }
  if (row < a.nrows) a.y[row] = a.init ? (Vec)t : (Vec)((Acc)a.y[row] + t);
}

template <typename V>
struct EllArgs {
  int64_t nrows;
  int64_t ncols;  // set for ell_kernel (its x loads read x[0] for empty slots)
  int32_t loop;  // slots visited per row
  int64_t pitch;
  const int32_t* __restrict__ ci;
  const V* __restrict__ val;
  const typename vtraits<V>::vec* __restrict__ x;
  typename vtraits<V>::vec* __restrict__ y;
  int32_t ld;
  int init;
  // ell_kernel only: per row, 1 + its last slot holding a column (nullptr:
  // every row visits `loop` slots)
  const uint8_t* __restrict__ cnt;
};

template <typename V, int KR, int RPT>
__global__ __launch_bounds__(256) void ell_kernel(EllArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row0 = ((int64_t)b * 256 * RPT) + threadIdx.x;
  Acc t[RPT][KR];
#pragma unroll
  for (int u = 0; u < RPT; u++)
#pragma unroll
    for (int q = 0; q < KR; q++) t[u][q] = 0;
  // slots each row visits, and the wave's loop bound (uniform)
  int32_t lim[RPT];
  int32_t wl = a.loop;
  if (a.cnt) {
    int32_t m = 0;
#pragma unroll
    for (int u = 0; u < RPT; u++) {
      const int64_t row = row0 + u * 256;
      lim[u] = row < a.nrows ? (int32_t)a.cnt[row] : 0;
      m = max(m, lim[u]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    wl = __builtin_amdgcn_readfirstlane(m);
  } else {
#pragma unroll
    for (int u = 0; u < RPT; u++) lim[u] = a.loop;
  }
  // batches of 8 slots: indices first, then values and x entries
  for (int n0 = 0; n0 < wl; n0 += 8) {
    // index and value loads are unconditional (slots outside the matrix, or
    // beyond the row's count, read slot 0 and are discarded): a load under a
    // branch is waited on where the branch joins, which left one index load
    // in flight at a time
    int64_t c[RPT][8];
    int64_t slot[RPT][8];
#pragma unroll
    for (int w = 0; w < 8; w++)
#pragma unroll
      for (int u = 0; u < RPT; u++) {
        const int64_t row = row0 + u * 256;
        const int n = n0 + w;
        const bool in = n < lim[u] && row < a.nrows;
        slot[u][w] = in ? row + a.pitch * n : 0;
        const int32_t ci = ld_stream(a.ci + slot[u][w]);
        c[u][w] = in ? (int64_t)ci : -1;
      }
    V v[RPT][8];
#pragma unroll
    for (int w = 0; w < 8; w++)
#pragma unroll
      for (int u = 0; u < RPT; u++) v[u][w] = ld_stream(a.val + slot[u][w]);
#pragma unroll
    for (int u = 0; u < RPT; u++) hold_raw(v[u]);
    // x loads unconditional as well (padding slots read x[0], unused)
    if (a.ncols > 0) batch_fma<V, KR, RPT>(t, v, c, a.x, a.ld);
  }
#pragma unroll
  for (int u = 0; u < RPT; u++) {
    const int64_t row = row0 + u * 256;
    if (row < a.nrows) {
#pragma unroll
      for (int q = 0; q < KR; q++) {
        Vec* yp = a.y + row * a.ld + q;
        *yp = a.init ? (Vec)t[u][q] : (Vec)((Acc)*yp + t[u][q]);
      }
    }
  }
}

// ----------------------------------------------------------- SpMM lane groups
// k right-hand sides (row-major X[ncols][ld], Y[nrows][ld]): KL lanes per
// row, each owning 16 bytes of the row's columns (2 FP64 / 4 FP32), so a
// row of x or y moves as KL adjacent 16-byte pieces, and a wave moves 64 / KL
// rows.  DIA's x rows of adjacent rows are adjacent, so a wave's x load
// reads one contiguous 1 KiB span; the y stores are contiguous the same way.
// (One thread per row with k scalar loads instead puts 64 rows 8k bytes
// apart into every load instruction: 64 cache lines per instruction.)  The
// KL lanes of a row load the same value / index (one broadcast line).  Sums
// run over diagonals / slots in the same order as the one-thread kernels.
template <typename Vec>
struct Piece;
template <>
struct Piece<double> {
  using T = double __attribute__((ext_vector_type(2)));
  static constexpr int Q = 2;
};
template <>
struct Piece<float> {
  using T = float __attribute__((ext_vector_type(4)));
  static constexpr int Q = 4;
};

template <typename V, int KL>
__global__ __launch_bounds__(256) void dia_spmm_kernel(DiaArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  using P = Piece<Vec>;
  constexpr int Q = P::Q;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = (int64_t)b * (256 / KL) + threadIdx.x / KL;
  const int l = threadIdx.x % KL;
  // rows past the end stay in the wave (their lanes may hold the batch's
  // offsets read back below) and do no work
  const bool live = row < a.nrows;
  Acc t[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) t[q] = 0;
  for (int d0 = 0; d0 < a.ndiag; d0 += 8) {
    // the batch's eight offsets in one round trip: lane w loads offset
    // d0 + w and the wave reads them back lane by lane (left to the
    // compiler, the BF16 kernel loaded them one at a time, each waited on
    // before the next)
    const int lane = threadIdx.x % kWave;
    const int32_t mo = lane < 8 && d0 + lane < a.ndiag ? a.off[d0 + lane] : 0;
    int32_t o8[8];
#pragma unroll
    for (int w = 0; w < 8; w++) o8[w] = __builtin_amdgcn_readlane(mo, w);
    V v[8];
    int64_t c[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const int64_t o = o8[w];
      const int64_t col = row + o;
      const bool ok = live && d < a.ndiag && col >= 0 && col < a.ncols;
      c[w] = ok ? col : -1;
      v[w] = ok ? ld_stream(a.val + a.pitch * d + row) : V{};
    }
    typename P::T xv[8];
#pragma unroll
    for (int w = 0; w < 8; w++)
      xv[w] = c[w] >= 0 ? *reinterpret_cast<const typename P::T*>(a.x + c[w] * a.ld + l * Q)
                        : (typename P::T)0;
    hold_raw(v);
#pragma unroll
    for (int w = 0; w < 8; w++)
      if (c[w] >= 0) {
        const Acc vw = vtraits<V>::widen(v[w]);
#pragma unroll
        for (int q = 0; q < Q; q++) t[q] += vw * (Acc)xv[w][q];
      }
  }
  if (!live) return;
  typename P::T* yp = reinterpret_cast<typename P::T*>(a.y + row * a.ld + l * Q);
  typename P::T out;
  if (a.init) {
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (Vec)t[q];
  } else {
    const typename P::T old = *yp;
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (Vec)((Acc)old[q] + t[q]);
  }
  *yp = out;
}

// DIA SpMM with R consecutive rows per lane group (R = 4): for each run of
// diagonals whose offsets step by one (o, o + 1, ..., at most kDiaRunMax),
// the group loads the R + m - 1 X pieces the run covers once and slides
// over them, instead of one X piece per row and diagonal (the 27-point
// stencil: 6 X loads for 4 rows x 3 diagonals instead of 12).  Each row
// still sums its diagonals in storage order, as dia_spmm_kernel does, with
// the same skips, so the two give identical results.  Loads are
// unconditional (columns outside x read column 0, a row past the matrix its
// last row) and the products selected; the workgroup loads the run's values
// for its rows into LDS together, so no lane group leaves before the end.
constexpr int kDiaRunMax = 3;

template <typename V, int KL, int R>
__global__ __launch_bounds__(256) void dia_spmm_runs_kernel(DiaArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  using P = Piece<Vec>;
  constexpr int Q = P::Q;
  constexpr int W = R + kDiaRunMax - 1;
  constexpr int kRows = 256 / KL * R;  // rows of the workgroup
  // the run's values for the workgroup's rows, loaded by all its threads in
  // one coalesced pass per diagonal (each lane then reads its R rows'
  // values from here instead of R loads of its own)
  __shared__ V s_val[kDiaRunMax][kRows];
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wr0 = (int64_t)b * kRows;
  const int gl = threadIdx.x / KL;
  const int64_t row0 = wr0 + (int64_t)gl * R;
  const int l = threadIdx.x % KL;
  const int64_t last = a.nrows - 1;
  Acc t[R][Q];
#pragma unroll
  for (int i = 0; i < R; i++)
#pragma unroll
    for (int q = 0; q < Q; q++) t[i][q] = 0;
  for (int d = 0; d < a.ndiag;) {
    const int64_t o = a.off[d];
    int m = 1;
    while (m < kDiaRunMax && d + m < a.ndiag && (int64_t)a.off[d + m] == o + m) m++;
    // the X rows first: the value stores below wait for every older load
    typename P::T xw[W];
#pragma unroll
    for (int i = 0; i < W; i++) {
      const int64_t col = row0 + o + i;
      const bool in = i < R + m - 1 && col >= 0 && col < a.ncols;
      xw[i] = *reinterpret_cast<const typename P::T*>(a.x + (in ? col : 0) * a.ld + l * Q);
    }
    __syncthreads();  // the previous run's values are read
    for (int j = 0; j < m; j++)
      for (int e = threadIdx.x; e < kRows; e += 256)
        s_val[j][e] = ld_stream(a.val + a.pitch * (d + j) + min(wr0 + e, last));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kDiaRunMax; j++) {
      if (j < m) {
#pragma unroll
        for (int i = 0; i < R; i++) {
          const int64_t col = row0 + i + o + j;
          const bool ok = row0 + i < a.nrows && col >= 0 && col < a.ncols;
          const Acc vw = vtraits<V>::widen(s_val[j][gl * R + i]);
#pragma unroll
          for (int q = 0; q < Q; q++) {
            const Acc f = t[i][q] + vw * (Acc)xw[i + j][q];
            t[i][q] = ok ? f : t[i][q];
          }
        }
      }
    }
    d += m;
  }
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int64_t row = row0 + i;
    if (row < a.nrows) {
      typename P::T* yp = reinterpret_cast<typename P::T*>(a.y + row * a.ld + l * Q);
      typename P::T out;
      if (a.init) {
#pragma unroll
        for (int q = 0; q < Q; q++) out[q] = (Vec)t[i][q];
      } else {
        const typename P::T old = *yp;
#pragma unroll
        for (int q = 0; q < Q; q++) out[q] = (Vec)((Acc)old[q] + t[i][q]);
      }
      *yp = out;
    }
  }
}

template <typename V, int KL>
__global__ __launch_bounds__(256) void ell_spmm_kernel(EllArgs<V> a) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  using P = Piece<Vec>;
  constexpr int Q = P::Q;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = (int64_t)b * (256 / KL) + threadIdx.x / KL;
  const int l = threadIdx.x % KL;
  if (row >= a.nrows) return;
  Acc t[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) t[q] = 0;
  for (int n0 = 0; n0 < a.loop; n0 += 8) {
    int64_t c[8];
#pragma unroll
    for (int w = 0; w < 8; w++)
      c[w] = n0 + w < a.loop ? (int64_t)ld_stream(a.ci + row + a.pitch * (n0 + w)) : -1;
    V v[8];
#pragma unroll
    for (int w = 0; w < 8; w++)
      v[w] = c[w] >= 0 ? ld_stream(a.val + row + a.pitch * (n0 + w)) : V{};
    typename P::T xv[8];
#pragma unroll
    for (int w = 0; w < 8; w++)
      xv[w] = c[w] >= 0 ? *reinterpret_cast<const typename P::T*>(a.x + c[w] * a.ld + l * Q)
                        : (typename P::T)0;
    hold_raw(v);
#pragma unroll
    for (int w = 0; w < 8; w++)
      if (c[w] >= 0) {
        const Acc vw = vtraits<V>::widen(v[w]);
#pragma unroll
        for (int q = 0; q < Q; q++) t[q] += vw * (Acc)xv[w][q];
      }
  }
  typename P::T* yp = reinterpret_cast<typename P::T*>(a.y + row * a.ld + l * Q);
  typename P::T out;
  if (a.init) {
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (Vec)t[q];
  } else {
    const typename P::T old = *yp;
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (Vec)((Acc)old[q] + t[q]);
  }
  *yp = out;
}

// ----------------------------------------------------------- row pairs
// FP32 / BF16 SpMV: a thread takes the adjacent rows 2t and 2t + 1, so that
// one load moves the two rows' values of a diagonal / slot (8 B in FP32, 4 B
// in BF16) and, in ELL, their two column indices (8 B).  With one row per
// thread a 16-bit value stream moved 128 B per wave load, and BF16 ran
// slower than FP32.  Used when the pitch is even (a pair never straddles
// two diagonals / slots; an odd row count then leaves the pair's second
// element inside the pitch) and the arrays are aligned to the pair.
template <typename V>
__global__ __launch_bounds__(256) void dia_pair_kernel(DiaArgs<V> a) {
  using P = Pair<V>;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = ((int64_t)b * 256 + threadIdx.x) * 2;
  const typename P::T* pv = reinterpret_cast<const typename P::T*>(a.val);
  float t0 = 0, t1 = 0;
  for (int d0 = 0; d0 < a.ndiag; d0 += 8) {
    float v0[8], v1[8];
    int64_t c0[8], c1[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const bool in = d < a.ndiag && row < a.nrows;
      const int64_t o = d < a.ndiag ? (int64_t)a.off[d] : 0;
      const typename P::T pw = in ? ld_stream(pv + (a.pitch * d + row) / 2) : (typename P::T)0;
      P::split(pw, v0[w], v1[w]);
      const int64_t col0 = row + o, col1 = row + 1 + o;
      c0[w] = in && col0 >= 0 && col0 < a.ncols ? col0 : -1;
      c1[w] = in && row + 1 < a.nrows && col1 >= 0 && col1 < a.ncols ? col1 : -1;
    }
#pragma unroll
    for (int w = 0; w < 8; w++) {
      if (c0[w] >= 0) t0 += v0[w] * a.x[c0[w]];
      if (c1[w] >= 0) t1 += v1[w] * a.x[c1[w]];
    }
  }
  if (row < a.nrows) a.y[row] = a.init ? t0 : a.y[row] + t0;
  if (row + 1 < a.nrows) a.y[row + 1] = a.init ? t1 : a.y[row + 1] + t1;
}

// dia_pair_kernel with x staged in LDS per cluster of nearby diagonals (as
// dia_staged_kernel, over the workgroup's 512 rows); same sums in the same
// order as dia_pair_kernel.
template <typename V>
__global__ __launch_bounds__(256) void dia_pair_staged_kernel(DiaArgs<V> a, DiaStage cl) {
  using P = Pair<V>;
  __shared__ float xs[2 * kDiaStageMax];
  const int64_t r0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 512;
  const int64_t row = r0 + 2 * threadIdx.x;
  for (int c = 0; c < cl.n; c++) {
    const int32_t base = cl.base[c] + 256 * c, len = cl.len[c] + 256;
    for (int j = threadIdx.x; j < len; j += 256) {
      const int64_t col = r0 + cl.omin[c] + j;
      xs[base + j] = (col >= 0 && col < a.ncols) ? a.x[col] : 0.0f;
    }
  }
  __syncthreads();
  const typename P::T* pv = reinterpret_cast<const typename P::T*>(a.val);
  float t0 = 0, t1 = 0;
  for (int d0 = 0; d0 < a.ndiag; d0 += 8) {
    float v0[8], v1[8];
    int64_t c0[8], c1[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const bool in = d < a.ndiag && row < a.nrows;
      const int64_t o = d < a.ndiag ? (int64_t)a.off[d] : 0;
      const typename P::T pw = in ? ld_stream(pv + (a.pitch * d + row) / 2) : (typename P::T)0;
      P::split(pw, v0[w], v1[w]);
      const int64_t col0 = row + o, col1 = row + 1 + o;
      c0[w] = in && col0 >= 0 && col0 < a.ncols ? col0 : -1;
      c1[w] = in && row + 1 < a.nrows && col1 >= 0 && col1 < a.ncols ? col1 : -1;
    }
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int d = d0 + w;
      const int32_t lb = d < a.ndiag ? cl.lb2[d] : kDiaDirect;  // uniform
      if (lb != kDiaDirect) {
        // (inside the cluster's window whenever the column is)
        if (c0[w] >= 0) t0 += v0[w] * xs[lb + (int32_t)(c0[w] - r0)];
        if (c1[w] >= 0) t1 += v1[w] * xs[lb + (int32_t)(c1[w] - r0)];
      } else {
        if (c0[w] >= 0) t0 += v0[w] * a.x[c0[w]];
        if (c1[w] >= 0) t1 += v1[w] * a.x[c1[w]];
      }
    }
  }
  if (row < a.nrows) a.y[row] = a.init ? t0 : a.y[row] + t0;
  if (row + 1 < a.nrows) a.y[row + 1] = a.init ? t1 : a.y[row + 1] + t1;
}

template <typename V>
__global__ __launch_bounds__(256) void ell_pair_kernel(EllArgs<V> a) {
  using P = Pair<V>;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = ((int64_t)b * 256 + threadIdx.x) * 2;
  const typename P::T* pv = reinterpret_cast<const typename P::T*>(a.val);
  const uint64_t* pc = reinterpret_cast<const uint64_t*>(a.ci);
  float t0 = 0, t1 = 0;
  const bool in = row < a.nrows;
  for (int n0 = 0; n0 < a.loop; n0 += 8) {
    int32_t c0[8], c1[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int n = n0 + w;
      const uint64_t cw = in && n < a.loop ? ld_stream(pc + (a.pitch * n + row) / 2) : ~0ull;
      c0[w] = (int32_t)(uint32_t)cw;
      // the pair's second row may lie past nrows, inside the pitch
      c1[w] = row + 1 < a.nrows ? (int32_t)(uint32_t)(cw >> 32) : -1;
    }
    float v0[8], v1[8];
#pragma unroll
    for (int w = 0; w < 8; w++) {
      const int n = n0 + w;
      const typename P::T pw =
          (c0[w] >= 0 || c1[w] >= 0) ? ld_stream(pv + (a.pitch * n + row) / 2) : (typename P::T)0;
      P::split(pw, v0[w], v1[w]);
    }
#pragma unroll
    for (int w = 0; w < 8; w++) {
      if (c0[w] >= 0) t0 += v0[w] * a.x[c0[w]];
      if (c1[w] >= 0) t1 += v1[w] * a.x[c1[w]];
    }
  }
  if (row < a.nrows) a.y[row] = a.init ? t0 : a.y[row] + t0;
  if (row + 1 < a.nrows) a.y[row + 1] = a.init ? t1 : a.y[row + 1] + t1;
}

// FP32 / BF16 rows of more than kCompensateLen diagonals / slots: thread
// (row, column q) sums its row as Dot2 pairs (TwoSum on the additions, the
// FMA TwoProduct on the products) and adds the old y through TwoSum when
// accumulating.  Diagonals / slots in storage order, as the other kernels.
template <typename V, bool DIA>
__global__ __launch_bounds__(256) void dia_ell_comp_kernel(DiaArgs<V> d, EllArgs<V> e, int32_t k) {
  using Acc = typename vtraits<V>::acc;
  using Vec = typename vtraits<V>::vec;
  const int64_t nrows = DIA ? d.nrows : e.nrows;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows * k) return;
  const int64_t row = t / k, q = t % k;
  csum<Acc> cs{0, 0};
  if constexpr (DIA) {
    for (int32_t n = 0; n < d.ndiag; n++) {
      const int64_t col = row + d.off[n];
      if (col >= 0 && col < d.ncols)
        cs.add_prod(vtraits<V>::widen(ld_stream(d.val + d.pitch * n + row)),
                    (Acc)d.x[col * d.ld + q]);
    }
  } else {
    for (int32_t n = 0; n < e.loop; n++) {
      const int64_t slot = row + e.pitch * n;
      const int32_t c = ld_stream(e.ci + slot);
      if (c >= 0) cs.add_prod(vtraits<V>::widen(ld_stream(e.val + slot)), (Acc)e.x[(int64_t)c * e.ld + q]);
    }
  }
  Vec* yp = (DIA ? d.y : e.y) + row * (DIA ? d.ld : e.ld) + q;
  if (!(DIA ? d.init : e.init)) cs.combine((Acc)*yp, (Acc)0);
  *yp = (Vec)cs.value();
}

namespace {

template <typename V, bool DIA>
int launch_comp(mph_matrix_impl* A, int k, const void* x, void* y, int init, hipStream_t s) {
  DiaArgs<V> d{};
  EllArgs<V> e{};
  if constexpr (DIA) {
    d.nrows = A->nrows;
    d.ncols = A->ncols;
    d.ndiag = (int32_t)A->ndiag;
    d.pitch = A->dia_pitch;
    d.off = A->dia_off;
    d.val = static_cast<const V*>(A->dia_v);
    d.x = static_cast<const typename vtraits<V>::vec*>(x);
    d.y = static_cast<typename vtraits<V>::vec*>(y);
    d.ld = k;
    d.init = init;
  } else {
    e.nrows = A->nrows;
    e.ncols = A->ncols;
    e.loop = (int32_t)A->ell_loop;
    e.pitch = A->ell_pitch;
    e.ci = A->ell_ci;
    e.val = static_cast<const V*>(A->ell_v);
    e.x = static_cast<const typename vtraits<V>::vec*>(x);
    e.y = static_cast<typename vtraits<V>::vec*>(y);
    e.ld = k;
    e.init = init;
  }
  const int64_t n = A->nrows * (int64_t)k;
  if (n == 0) return MPH_OK;
  hipLaunchKernelGGL((dia_ell_comp_kernel<V, DIA>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                     s, d, e, (int32_t)k);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, int KR, int RPT>
int launch_dia_k(mph_matrix_impl* A, const void* x, void* y, int ld, int init,
                 hipStream_t s) {
  DiaArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.ndiag = (int32_t)A->ndiag;
  a.pitch = A->dia_pitch;
  a.off = A->dia_off;
  a.val = static_cast<const V*>(A->dia_v);
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.ld = ld;
  a.init = init;
  const int64_t nb = ceil_div(A->nrows, 256 * RPT);
  if (nb == 0) return MPH_OK;
  hipLaunchKernelGGL((dia_kernel<V, KR, RPT>), dim3((unsigned)nb), dim3(256), 0,
                     s, a);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V>
int launch_dia_staged(mph_matrix_impl* A, const void* x, void* y, int init, hipStream_t s) {
  DiaArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.ndiag = (int32_t)A->ndiag;
  a.pitch = A->dia_pitch;
  a.off = A->dia_off;
  a.val = static_cast<const V*>(A->dia_v);
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.ld = 1;
  a.init = init;
  const int64_t nb = ceil_div(A->nrows, 256);
  if (nb == 0) return MPH_OK;
  hipLaunchKernelGGL(dia_staged_kernel<V>, dim3((unsigned)nb), dim3(256), 0, s, a, A->dia_stage);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, int KR, int RPT>
int launch_ell_k(mph_matrix_impl* A, const void* x, void* y, int ld, int init,
                 hipStream_t s) {
  EllArgs<V> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.loop = (int32_t)A->ell_loop;
  a.pitch = A->ell_pitch;
  a.ci = A->ell_ci;
  a.val = static_cast<const V*>(A->ell_v);
  a.x = static_cast<const typename vtraits<V>::vec*>(x);
  a.y = static_cast<typename vtraits<V>::vec*>(y);
  a.ld = ld;
  a.init = init;
  a.cnt = A->rb.ell_cnt;
  const int64_t nb = ceil_div(A->nrows, 256 * RPT);
  if (nb == 0) return MPH_OK;
  hipLaunchKernelGGL((ell_kernel<V, KR, RPT>), dim3((unsigned)nb), dim3(256), 0,
                     s, a);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, bool DIA, int KL>
int launch_spmm_lanes(mph_matrix_impl* A, const void* x, void* y, int ld, int init,
                      hipStream_t s) {
  const int64_t nb = ceil_div(A->nrows, 256 / KL);
  if (nb == 0) return MPH_OK;
  if constexpr (DIA) {
    DiaArgs<V> a{};
    a.nrows = A->nrows;
    a.ncols = A->ncols;
    a.ndiag = (int32_t)A->ndiag;
    a.pitch = A->dia_pitch;
    a.off = A->dia_off;
    a.val = static_cast<const V*>(A->dia_v);
    a.x = static_cast<const typename vtraits<V>::vec*>(x);
    a.y = static_cast<typename vtraits<V>::vec*>(y);
    a.ld = ld;
    a.init = init;
    // the run form (4 rows per lane group) where it was measured faster: 4
    // or more lanes per row (S27 FP64 k = 8 587 -> 423 us, k = 16 1108 ->
    // 630 us; FP32 k = 16 415 -> 374 us; BF16 k = 16 383 -> 347 us); not with
    // fewer lanes (FP64 k = 2 213 -> 255 us, k = 4 even; FP32 k = 4 124 ->
    // 188 us, k = 8 209 -> 258 us; BF16 k = 8 209 -> 254 us);
    // MPH_DIA_SPMM_RUNS=0 / 1 forces it off / on
    const char* env = std::getenv("MPH_DIA_SPMM_RUNS");
    const int mode = env ? std::atoi(env) : -1;
    constexpr bool wide = KL >= 4;
    const bool runs = A->ncols > 0 && (mode > 0 || (mode < 0 && wide));
    if (runs) {
      constexpr int R = 4;
      const int64_t nbr = ceil_div(A->nrows, (int64_t)(256 / KL) * R);
      hipLaunchKernelGGL((dia_spmm_runs_kernel<V, KL, R>), dim3((unsigned)nbr), dim3(256), 0, s, a);
    } else {
      hipLaunchKernelGGL((dia_spmm_kernel<V, KL>), dim3((unsigned)nb), dim3(256), 0, s, a);
    }
  } else {
    EllArgs<V> a{};
    a.nrows = A->nrows;
    a.loop = (int32_t)A->ell_loop;
    a.pitch = A->ell_pitch;
    a.ci = A->ell_ci;
    a.val = static_cast<const V*>(A->ell_v);
    a.x = static_cast<const typename vtraits<V>::vec*>(x);
    a.y = static_cast<typename vtraits<V>::vec*>(y);
    a.ld = ld;
    a.init = init;
    hipLaunchKernelGGL((ell_spmm_kernel<V, KL>), dim3((unsigned)nb), dim3(256), 0, s, a);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <typename V, bool DIA, int RPT>
int launch_fmt(mph_matrix_impl* A, int k, const void* x, void* y, int init,
               hipStream_t s) {
  using Vec = typename vtraits<V>::vec;
  constexpr int Q = Piece<Vec>::Q;
  // lane groups when every row of X / Y starts on 16 bytes
  const bool pieces = k > 1 && (k * sizeof(Vec)) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(y) % 16 == 0;
  int done = 0;
  while (done < k) {
    const int rem = k - done;
    const Vec* xs = static_cast<const Vec*>(x) + done;
    Vec* ys = static_cast<Vec*>(y) + done;
    if (pieces && rem >= Q) {
      const int kl = rem >= 16 * Q ? 16 : rem >= 8 * Q ? 8 : rem >= 4 * Q ? 4 : rem >= 2 * Q ? 2 : 1;
      int st;
      switch (kl) {
        case 16: st = launch_spmm_lanes<V, DIA, 16>(A, xs, ys, k, init, s); break;
        case 8: st = launch_spmm_lanes<V, DIA, 8>(A, xs, ys, k, init, s); break;
        case 4: st = launch_spmm_lanes<V, DIA, 4>(A, xs, ys, k, init, s); break;
        case 2: st = launch_spmm_lanes<V, DIA, 2>(A, xs, ys, k, init, s); break;
        default: st = launch_spmm_lanes<V, DIA, 1>(A, xs, ys, k, init, s); break;
      }
      if (st != MPH_OK) return st;
      done += kl * Q;
      continue;
    }
    int kr = rem >= 16 ? 16 : rem >= 8 ? 8 : rem >= 4 ? 4 : rem >= 2 ? 2 : 1;
    if (k == 1) kr = 1;
    int st;
#define MPH_LAUNCH(KRV)                                             \
  st = DIA ? launch_dia_k<V, KRV, RPT>(A, xs, ys, k, init, s)       \
           : launch_ell_k<V, KRV, RPT>(A, xs, ys, k, init, s)
    switch (kr) {
      case 16: MPH_LAUNCH(16); break;
      case 8: MPH_LAUNCH(8); break;
      case 4: MPH_LAUNCH(4); break;
      case 2: MPH_LAUNCH(2); break;
      default: MPH_LAUNCH(1); break;
    }
#undef MPH_LAUNCH
    if (st != MPH_OK) return st;
    done += kr;
  }
  return MPH_OK;
}

// the row-pair kernels: FP32 / BF16 SpMV with the default launch shape, an
// even pitch and arrays aligned to the pair loads
template <typename V, bool DIA>
int launch_pair(mph_matrix_impl* A, const void* x, void* y, int init, hipStream_t s) {
  const int64_t nb = ceil_div(A->nrows, 512);
  if (nb == 0) return MPH_OK;
  if constexpr (DIA) {
    DiaArgs<V> a{};
    a.nrows = A->nrows;
    a.ncols = A->ncols;
    a.ndiag = (int32_t)A->ndiag;
    a.pitch = A->dia_pitch;
    a.off = A->dia_off;
    a.val = static_cast<const V*>(A->dia_v);
    a.x = static_cast<const float*>(x);
    a.y = static_cast<float*>(y);
    a.ld = 1;
    a.init = init;
    if (A->dia_stage.n > 0 && A->dia_stage.pair)
      hipLaunchKernelGGL(dia_pair_staged_kernel<V>, dim3((unsigned)nb), dim3(256), 0, s, a,
                         A->dia_stage);
    else
      hipLaunchKernelGGL(dia_pair_kernel<V>, dim3((unsigned)nb), dim3(256), 0, s, a);
  } else {
    EllArgs<V> a{};
    a.nrows = A->nrows;
    a.loop = (int32_t)A->ell_loop;
    a.pitch = A->ell_pitch;
    a.ci = A->ell_ci;
    a.val = static_cast<const V*>(A->ell_v);
    a.x = static_cast<const float*>(x);
    a.y = static_cast<float*>(y);
    a.ld = 1;
    a.init = init;
    hipLaunchKernelGGL(ell_pair_kernel<V>, dim3((unsigned)nb), dim3(256), 0, s, a);
  }
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

template <bool DIA>
bool pair_ok(const mph_matrix_impl* A, int k) {
  if (k != 1 || A->tuning.rows_per_thread != 0) return false;
  if (A->vt != MPH_F32 && A->vt != MPH_BF16) return false;
  if (!DIA && A->rb.ell_cnt) return false;  // the counted one-row kernel
  const int64_t pitch = DIA ? A->dia_pitch : A->ell_pitch;
  const uintptr_t v = reinterpret_cast<uintptr_t>(DIA ? A->dia_v : A->ell_v);
  const size_t pair_bytes = 2 * value_size(A->vt);
  if (pitch % 2 != 0 || v % pair_bytes != 0) return false;
  return DIA || reinterpret_cast<uintptr_t>(A->ell_ci) % 8 == 0;
}

template <bool DIA>
int dispatch(mph_matrix_impl* A, int k, const void* x, void* y, int init,
             hipStream_t s) {
  if (A->vt != MPH_F64 && (DIA ? A->ndiag : A->ell_loop) > kCompensateLen)
    return A->vt == MPH_F32 ? launch_comp<float, DIA>(A, k, x, y, init, s)
                            : launch_comp<bf16, DIA>(A, k, x, y, init, s);
  if (pair_ok<DIA>(A, k))
    return A->vt == MPH_F32 ? launch_pair<float, DIA>(A, x, y, init, s)
                            : launch_pair<bf16, DIA>(A, x, y, init, s);
  if (DIA && k == 1 && A->dia_stage.n > 0 && A->tuning.rows_per_thread == 0)
    return A->vt == MPH_F64   ? launch_dia_staged<double>(A, x, y, init, s)
           : A->vt == MPH_F32 ? launch_dia_staged<float>(A, x, y, init, s)
                              : launch_dia_staged<bf16>(A, x, y, init, s);
  const bool two = A->tuning.rows_per_thread == 2;
  switch (A->vt) {
    case MPH_F64:
      return two ? launch_fmt<double, DIA, 2>(A, k, x, y, init, s)
                 : launch_fmt<double, DIA, 1>(A, k, x, y, init, s);
    case MPH_F32:
      return two ? launch_fmt<float, DIA, 2>(A, k, x, y, init, s)
                 : launch_fmt<float, DIA, 1>(A, k, x, y, init, s);
    case MPH_BF16:
      return two ? launch_fmt<bf16, DIA, 2>(A, k, x, y, init, s)
                 : launch_fmt<bf16, DIA, 1>(A, k, x, y, init, s);
  }
  return invalid("unknown value type");
}

}  // namespace

// Clusters of consecutive diagonals (storage order) whose offsets span at
// most kDiaClusterSpan; a cluster of two or more diagonals is staged.  More
// than kDiaStageMaxDiags diagonals, more than kDiaMaxClusters staged
// clusters, windows beyond kDiaStageMax entries or nothing worth staging:
// the plain kernel (n = 0).
int dia_analyse(mph_matrix_impl* A, hipStream_t s) {
  A->dia_stage = DiaStage{};
  const int64_t nd = A->ndiag;
  if (nd < 2 || nd > kDiaStageMaxDiags || A->nrows == 0 || A->ncols == 0) return MPH_OK;
  int mode = 2;  // MPH_DIA_STAGE: 0 off, 1 the one-row kernels only, 2 also the row pairs
  if (const char* env = std::getenv("MPH_DIA_STAGE")) mode = std::atoi(env);
  if (mode <= 0) return MPH_OK;
  std::vector<int32_t> off((size_t)nd);
  MPH_HIP_CHECK(hipMemcpyAsync(off.data(), A->dia_off, sizeof(int32_t) * nd,
                               hipMemcpyDeviceToHost, s));
  MPH_HIP_CHECK(hipStreamSynchronize(s));
  DiaStage st{};
  int32_t used = 0;
  int64_t staged_diags = 0;
  for (int64_t i = 0; i < nd;) {
    int64_t lo = off[i], hi = off[i], j = i + 1;
    while (j < nd && std::max<int64_t>(hi, off[j]) - std::min<int64_t>(lo, off[j]) <= kDiaClusterSpan) {
      lo = std::min<int64_t>(lo, off[j]);
      hi = std::max<int64_t>(hi, off[j]);
      j++;
    }
    int32_t lb = kDiaDirect, lb2 = kDiaDirect;
    if (j - i >= 2) {
      const int32_t len = (int32_t)(256 + hi - lo);
      if (st.n == kDiaMaxClusters || used + len > kDiaStageMax ||
          lo < -(int64_t)(INT32_MAX - 4 * kDiaStageMax))
        return MPH_OK;
      const int c = st.n++;
      st.omin[c] = (int32_t)lo;
      st.base[c] = used;
      st.len[c] = len;
      lb = used - (int32_t)lo;
      lb2 = used + 256 * c - (int32_t)lo;
      used += len;
      staged_diags += j - i;
    }
    for (int64_t d = i; d < j; d++) {
      st.lb[d] = lb;
      st.lb2[d] = lb2;
    }
    i = j;
  }
  if (st.n == 0) return MPH_OK;
  // the row-pair layout: windows 256 entries longer
  // the row pairs move half the bytes per x entry and finish a workgroup's
  // rows sooner: staging pays there only when the staged clusters average
  // kDiaPairMinDiags diagonals or more (27-point stencil: 9 per cluster, FP32
  // 125 -> 108 us; 7-point: one cluster of 5, BF16 21.5 -> 23.7 us)
  st.pair = mode >= 2 && used + 256 * st.n <= 2 * kDiaStageMax &&
            staged_diags >= (int64_t)kDiaPairMinDiags * st.n;
  A->dia_stage = st;
  return MPH_OK;
}

int dia_multiply(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s) {
  return dispatch<true>(A, k, x, y, init, s);
}

int hdc_tail_multiply(mph_matrix_impl* A, const void* x, void* y, int init, hipStream_t s) {
  if (A->format != MPH_HDC || A->vt != MPH_F64 || A->dia_stage.n == 0)
    return runtime_error("HDC tail pass: unsupported plan");
  DiaArgs<double> a{};
  a.nrows = A->nrows;
  a.ncols = A->ncols;
  a.ndiag = (int32_t)A->ndiag;
  a.pitch = A->dia_pitch;
  a.off = A->dia_off;
  a.val = static_cast<const double*>(A->dia_v);
  a.x = static_cast<const double*>(x);
  a.y = static_cast<double*>(y);
  a.ld = 1;
  a.init = init;
  a.tro = A->csr_ro;
  a.tci = A->csr_ci;
  a.tval = static_cast<const double*>(A->csr_v);
  const int64_t nb = ceil_div(A->nrows, 256);
  if (nb == 0) return MPH_OK;
  hipLaunchKernelGGL((dia_staged_kernel<double, true>), dim3((unsigned)nb), dim3(256), 0, s, a,
                     A->dia_stage);
  MPH_HIP_CHECK(hipGetLastError());
  return MPH_OK;
}

int ell_multiply(mph_matrix_impl* A, int k, const void* x, void* y, int init,
                 hipStream_t s) {
  return dispatch<false>(A, k, x, y, init, s);
}

}  // namespace mph
