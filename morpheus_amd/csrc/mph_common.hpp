// mph_common.hpp -- shared device/host helpers for the MI355X (gfx950) SpMV
// hot path: value-type traits, wave64 primitives, XCD-aware block mapping,
// streaming loads and error plumbing.
#pragma once

#include <type_traits>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/morpheus_amd.h"

namespace mph {

// ---------------------------------------------------------------- dtypes
// BF16 storage: the upper 16 bits of an IEEE binary32.  Arithmetic is done in
// FP32 after widening, so only the bit pattern is needed here.
struct bf16 {
  uint16_t bits;
};

template <typename V>
struct vtraits;

template <>
struct vtraits<double> {
  using acc = double;  // accumulation type
  using vec = double;  // x / y element type
  static constexpr int code = MPH_F64;
  __device__ __forceinline__ static double widen(double v) { return v; }
};

template <>
struct vtraits<float> {
  using acc = float;
  using vec = float;
  static constexpr int code = MPH_F32;
  __device__ __forceinline__ static float widen(float v) { return v; }
};

template <>
struct vtraits<bf16> {
  using acc = float;
  using vec = float;
  static constexpr int code = MPH_BF16;
  __device__ __forceinline__ static float widen(bf16 v) {
    return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
  }
};

inline size_t value_size(int vt) {
  return vt == MPH_F64 ? 8 : (vt == MPH_F32 || vt == MPH_I32) ? 4 : 2;
}
inline size_t vector_size(int vt) { return vt == MPH_F64 ? 8 : 4; }

// ---------------------------------------------------------------- loads
// Matrix streams are read exactly once per SpMV: load them non-temporally so
// they do not evict x (which is gathered and benefits from L2 / MALL reuse).
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
// 16-bit values take a plain load: measured, the non-temporal form of the
// sub-dword load ran 3-5% slower in the row-block SpMV (B9, Z15 in BF16)
template <>
__device__ __forceinline__ bf16 ld_stream<bf16>(const bf16* p) {
  bf16 r;
  r.bits = p->bits;
  return r;
}

// BF16 values loaded in a batch are kept raw until their products: widened
// right after its load, each value made the compiler wait for that load
// before issuing the next one (two loads in flight instead of the batch).
// Called after the batch's remaining loads (the x gathers) are issued, the
// empty asm statement, which takes each raw word as an in-out operand, keeps
// the widening below it.  No-op for FP64 / FP32.
template <typename V, int N>
__device__ __forceinline__ void hold_raw(V (&v)[N]) {
  if constexpr (std::is_same<V, bf16>::value) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      uint32_t w = v[k].bits;
      asm volatile("" : "+v"(w));
      v[k].bits = (uint16_t)w;
    }
  }
}

// Two adjacent FP32 / BF16 values loaded as one word (8 B / 4 B) and split.
template <typename V>
struct Pair;
template <>
struct Pair<float> {
  using T = uint64_t;
  __device__ __forceinline__ static void split(T w, float& lo, float& hi) {
    lo = __uint_as_float((uint32_t)w);
    hi = __uint_as_float((uint32_t)(w >> 32));
  }
};
template <>
struct Pair<bf16> {
  using T = uint32_t;
  __device__ __forceinline__ static void split(T w, float& lo, float& hi) {
    lo = __uint_as_float(w << 16);
    hi = __uint_as_float(w & 0xffff0000u);
  }
};

// ---------------------------------------------------------------- wave64
constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m) {
  return __shfl_xor(v, m, kWave);
}
template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) {
  return __shfl_up(v, d, kWave);
}
template <typename T>
__device__ __forceinline__ T shfl_down(T v, int d) {
  return __shfl_down(v, d, kWave);
}
template <typename T>
__device__ __forceinline__ T shfl(T v, int src) {
  return __shfl(v, src, kWave);
}

// DPP lane moves of the GFX9 wavefront (no LDS, no address registers).
// Controls: row_shr:n = 0x110 + n (within rows of 16 lanes), row_bcast:15 =
// 0x142 (lane 15 of a row to the next row; with row mask 0xA rows 1 and 3
// are written), row_bcast:31 = 0x143 (lane 31 to rows 2 and 3, row mask
// 0xC), wave_shr:1 = 0x138 (lane i gets lane i - 1), wave_shl:1 = 0x130
// (lane i gets lane i + 1).  Lanes without a source, or in rows the mask
// leaves out, read 0.
enum : int {
  kDppRowShr1 = 0x111,
  kDppRowShr2 = 0x112,
  kDppRowShr4 = 0x114,
  kDppRowShr8 = 0x118,
  kDppRowBcast15 = 0x142,
  kDppRowBcast31 = 0x143,
  kDppWaveShr1 = 0x138,
  kDppWaveShl1 = 0x130,
};

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int32_t dpp_i32(int32_t v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_val(float v) {
  return __builtin_bit_cast(float, dpp_i32<CTRL, ROWS>(__builtin_bit_cast(int32_t, v)));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_val(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)dpp_i32<CTRL, ROWS>((int32_t)(uint32_t)u);
  const uint32_t hi = (uint32_t)dpp_i32<CTRL, ROWS>((int32_t)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over aligned groups of L lanes (L a power of two <= 64).
template <typename A>
__device__ __forceinline__ A group_sum(A v, int L) {
  for (int m = L >> 1; m > 0; m >>= 1) v += shfl_xor(v, m);
  return v;
}

// Knuth's TwoSum: s + e == a + b exactly.  Contraction is off in the error
// free transformations: hipcc contracts a * b + c into an FMA by default,
// which would fold a product into the sum and leave the error term wrong.
template <typename A>
__device__ __forceinline__ void two_sum(A a, A b, A& s, A& e) {
#pragma clang fp contract(off)
  s = a + b;
  const A bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// TwoProduct through the FMA: p + e == a * b exactly (no underflow).
template <typename A>
__device__ __forceinline__ void two_prod(A a, A b, A& p, A& e) {
#pragma clang fp contract(off)
  p = a * b;
  e = __builtin_fma(a, b, -p);
}

// Compensated (sum, error) pair.  add_prod() is one step of Dot2 (Ogita,
// Rump, Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6),
// 2005, Algorithm 5.3): the rounding errors of the product (TwoProduct) and
// of the addition (TwoSum) both go into the correction c, so a row sum is as
// accurate as if computed in twice the working precision and then rounded:
// |result - exact| <= u |exact| + O(n^2 u^2) sum |a x| (in practice about
// n u^2 sum |a x|).  combine() joins two pairs (cross-lane reductions).
template <typename A>
struct csum {
  A s, c;
  __device__ __forceinline__ void add(A v) {
#pragma clang fp contract(off)
    A t, e;
    two_sum(s, v, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void add_prod(A a, A b) {
#pragma clang fp contract(off)
    A p, ep, t, e;
    two_prod(a, b, p, ep);
    two_sum(s, p, t, e);
    s = t;
    c += e + ep;
  }
  __device__ __forceinline__ void combine(A s2, A c2) {
#pragma clang fp contract(off)
    A t, e;
    two_sum(s, s2, t, e);
    s = t;
    c += e + c2;
  }
  __device__ __forceinline__ A value() const {
#pragma clang fp contract(off)
    return s + c;
  }
};

// ---------------------------------------------------------------- XCDs
// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remapping blockIdx so that each XCD walks one
// contiguous range of logical blocks keeps neighbouring rows (and the x
// entries a stencil row shares with them) inside one L2.
constexpr int kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
  const uint32_t xcd = bid % kNumXcd;
  const uint32_t idx = bid / kNumXcd;
  const uint32_t q = nblocks / kNumXcd;
  const uint32_t r = nblocks % kNumXcd;
  return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);

#define MPH_HIP_CHECK(expr)                                                 \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      ::mph::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));  \
      return (_e == hipErrorOutOfMemory) ? MPH_ERR_OUT_OF_MEMORY            \
                                         : MPH_ERR_HIP;                     \
    }                                                                       \
  } while (0)

#define MPH_CHECK_STATUS(expr) \
  do {                         \
    int _s = (expr);           \
    if (_s != MPH_OK) return _s; \
  } while (0)

inline int invalid(const std::string& msg) {
  set_error(msg);
  return MPH_ERR_INVALID_INPUT;
}

inline int runtime_error(const std::string& msg) {
  set_error(msg);
  return MPH_ERR_RUNTIME;
}

// device free whose status is deliberately ignored (release paths)
inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t pad_size(int64_t len, int64_t align) {
  return align * ((len + align - 1) / align);
}

}  // namespace mph
