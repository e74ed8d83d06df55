// Morpheus_AMD_Algorithms.hpp -- Morpheus::copy, create_mirror[_container],
// multiply and convert of the MI355X drop-in, over the C-ABI of
// include/morpheus_amd.h.
//
// Reference front-ends:
//   copy                 core/src/Morpheus_Copy.hpp:51-80
//   create_mirror[...]   core/src/Morpheus_MirrorContainers.hpp:62-192
//   multiply             core/src/Morpheus_Multiply.hpp:57-78
//   convert              core/src/Morpheus_Convert.hpp:54-57, hub
//                        core/src/impl/Morpheus_Convert_Impl.hpp:83-106
//   convert (in place)   core/src/Morpheus_Convert.hpp:88-141
//   dynamic dispatch     core/src/impl/Dynamic/Morpheus_{Multiply,Convert,Copy}_Impl.hpp
//
// Execution spaces.  Morpheus::HIP and Morpheus::Generic::HIP run the
// MI355X kernels on HIPSpace containers.  Morpheus::Serial and
// Morpheus::OpenMP on HostSpace containers run the same kernels: the matrix
// structure is staged into HBM once (kept with the container's analysis
// until it is modified), its values and x / y travel per call.  Conversion on the device succeeds for
// every format pair (the reference HIP converters throw) and is bit-exact
// with the reference Serial converters.
//
// All device work is ordered on the default stream, as in the reference.
#pragma once

#include <cmath>

#include "Morpheus_AMD_Containers.hpp"

namespace Morpheus {

namespace Impl {

template <class T>
inline const void* vp(const T* p) {
  return static_cast<const void*>(p);
}

// ------------------------------------------------------------ copy kernels
template <class A, class B>
inline void copy_array(const A& src, B& dst, const char* what) {
  using SM = typename A::memory_space;
  using DM = typename B::memory_space;
  static_assert(sizeof(typename A::value_type) == sizeof(typename B::value_type),
                "Morpheus::copy between different value types");
  if (src.size() != dst.size())
    throw RuntimeException(std::string("Morpheus::copy: ") + what + " sizes differ (" +
                           std::to_string(src.size()) + " != " + std::to_string(dst.size()) +
                           ")");
  A::copy_bytes(dst.data(), src.data(), src.size() * sizeof(typename A::value_type),
                DM::is_device, SM::is_device);
}

// DenseMatrix deep copy; transposes through host memory when the storage
// layouts differ.
template <class S, class D>
inline void copy_dense_matrix(const S& src, D& dst) {
  if (src.nrows() != dst.nrows() || src.ncols() != dst.ncols())
    throw RuntimeException("Morpheus::copy: DenseMatrix shapes differ");
  using SM = typename S::memory_space;
  using DM = typename D::memory_space;
  using T = typename S::non_const_value_type;
  if (S::is_left == D::is_left || src.nrows() <= 1 || src.ncols() <= 1) {
    copy_array(src.const_view(), dst.view(), "DenseMatrix");
    return;
  }
  const size_t m = src.nrows(), n = src.ncols();
  std::vector<T> a(m * n), b(m * n);
  S::value_array_type::copy_bytes(a.data(), src.data(), a.size() * sizeof(T), false, SM::is_device);
  for (size_t i = 0; i < m; i++)
    for (size_t j = 0; j < n; j++)
      b[D::is_left ? i + j * m : i * n + j] = a[S::is_left ? i + j * m : i * n + j];
  S::value_array_type::copy_bytes(dst.data(), b.data(), b.size() * sizeof(T), DM::is_device, false);
}

template <class S, class D>
inline void check_shape(const S& src, const D& dst) {
  if (dst.nrows() != src.nrows() || dst.ncols() != src.ncols())
    throw RuntimeException("Morpheus::copy: destination matrix must have equal shape to the source");
  if (dst.nnnz() != src.nnnz())
    throw RuntimeException(
        "Morpheus::copy: destination matrix must have equal number of non-zeros to the source");
}

template <class S, class D>
inline void copy_sparse(const S& src, D& dst) {
  static_assert(S::format_id == D::format_id,
                "Morpheus::copy is only available between containers of the same format");
  check_shape(src, dst);
  if constexpr (S::format_id == COO_FORMAT) {
    copy_array(src.crow_indices().const_view(), dst.row_indices().view(), "row_indices");
    copy_array(src.ccolumn_indices().const_view(), dst.column_indices().view(), "column_indices");
    copy_array(src.cvalues().const_view(), dst.values().view(), "values");
  } else if constexpr (S::format_id == CSR_FORMAT) {
    copy_array(src.crow_offsets().const_view(), dst.row_offsets().view(), "row_offsets");
    copy_array(src.ccolumn_indices().const_view(), dst.column_indices().view(), "column_indices");
    copy_array(src.cvalues().const_view(), dst.values().view(), "values");
  } else if constexpr (S::format_id == DIA_FORMAT) {
    copy_array(src.cdiagonal_offsets().const_view(), dst.diagonal_offsets().view(),
               "diagonal_offsets");
    copy_dense_matrix(src.cvalues(), dst.values());
  } else if constexpr (S::format_id == ELL_FORMAT) {
    copy_dense_matrix(src.ccolumn_indices(), dst.column_indices());
    copy_dense_matrix(src.cvalues(), dst.values());
  } else if constexpr (S::format_id == HYB_FORMAT) {
    copy_sparse(src.cell(), dst.ell());
    copy_sparse(src.ccoo(), dst.coo());
  } else {
    copy_sparse(src.cdia(), dst.dia());
    copy_sparse(src.ccsr(), dst.csr());
  }
  dst.touch();
}

// ------------------------------------------------------------ int64 / int staging
// The kernels take int32 indices and FP64 / FP32 / BF16 values.  A
// long long-indexed container reaches them through int32 copies of its index
// arrays (every index and size below 2^31, checked), an int-valued one
// through FP64 copies of its values, recast before every use (exact: int32
// values are FP64 numbers, and so are the products and sums while they stay
// below 2^53 in magnitude).

// device scratch, released with its owner
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) {
    if (bytes) hip_check(hipMalloc(&p, bytes), "Morpheus: device scratch");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

template <class I>
inline constexpr bool wide_index_v = std::is_same_v<std::remove_const_t<I>, long long>;

// The buffers a handle is built over, and the value casts around its use.
struct Stage {
  std::vector<std::shared_ptr<DevBuf>> bufs;
  std::vector<std::function<void()>> refresh, write_back;
};

// int32 view of n device indices of type I (the array itself for int)
template <class I>
inline const int* stage_index(Stage& st, const I* p, size_t n) {
  if constexpr (wide_index_v<I>) {
    auto b = std::make_shared<DevBuf>(sizeof(int32_t) * std::max<size_t>(n, 1));
    check(mph_index_narrow(reinterpret_cast<const int64_t*>(p), (int64_t)n,
                           static_cast<int32_t*>(b->p), nullptr),
          "Morpheus: indices");
    st.bufs.push_back(b);
    return static_cast<const int*>(b->p);
  } else {
    return p;
  }
}

// The values the kernels read for n device values of type V (FP64 copies
// for int, recast by refresh and cast back by write_back)
template <class V>
inline const void* stage_values(Stage& st, const V* p, size_t n) {
  if constexpr (std::is_same_v<std::remove_const_t<V>, int>) {
    auto b = std::make_shared<DevBuf>(sizeof(double) * std::max<size_t>(n, 1));
    st.bufs.push_back(b);
    double* d = static_cast<double*>(b->p);
    st.refresh.push_back([p, d, n]() {
      check(mph_values_cast(p, MPH_I32, d, MPH_F64, (int64_t)n, nullptr), "Morpheus: values");
    });
    st.write_back.push_back([p, d, n]() {
      check(mph_values_cast(d, MPH_F64, const_cast<V*>(p), MPH_I32, (int64_t)n, nullptr),
            "Morpheus: values");
    });
    return d;
  } else {
    return p;
  }
}

// the value type a handle computes in
template <class V>
inline constexpr int handle_code_v = std::is_same_v<V, int> ? MPH_F64 : value_code<V>::value;

// An int32 device array that the kernels write n entries of, widened into
// dst by done() (dst itself for int).  in_place: dst's entries are read too.
template <class I>
struct IdxOut {
  int* p = nullptr;
  I* dst;
  size_t n;
  std::shared_ptr<DevBuf> buf;
  IdxOut(I* d, size_t count, bool in_place = false) : dst(d), n(count) {
    if constexpr (wide_index_v<I>) {
      buf = std::make_shared<DevBuf>(sizeof(int32_t) * std::max<size_t>(n, 1));
      p = static_cast<int*>(buf->p);
      if (in_place)
        check(mph_index_narrow(reinterpret_cast<const int64_t*>(dst), (int64_t)n, p, nullptr),
              "Morpheus: indices");
    } else {
      p = d;
    }
  }
  void done() {
    if constexpr (wide_index_v<I>)
      check(mph_index_widen(p, (int64_t)n, reinterpret_cast<int64_t*>(dst), nullptr),
            "Morpheus: indices");
  }
};

template <class I>
inline IdxOut<I> idx_out(I* d, size_t n, bool in_place = false) {
  return IdxOut<I>(d, n, in_place);
}

// n entries of a device vector of V as FP64 (int) or as itself: fn(pointer)
// runs on them; `in`: the entries are read, `out`: written back
template <class V, class Fn>
inline void as_compute_vector(V* p, size_t n, bool in, bool out, Fn&& fn) {
  if constexpr (std::is_same_v<V, int>) {
    DevBuf d(sizeof(double) * std::max<size_t>(n, 1));
    if (in) check(mph_values_cast(p, MPH_I32, d.p, MPH_F64, (int64_t)n, nullptr), "Morpheus: vector");
    fn(d.p);
    if (out) check(mph_values_cast(d.p, MPH_F64, p, MPH_I32, (int64_t)n, nullptr), "Morpheus: vector");
  } else {
    fn(static_cast<void*>(p));
  }
}

// ------------------------------------------------------------ handles
template <class M>
inline std::vector<const void*> signature(const M& A) {
  std::vector<const void*> s{reinterpret_cast<const void*>(A.nrows()),
                             reinterpret_cast<const void*>(A.ncols()),
                             reinterpret_cast<const void*>(A.nnnz())};
  if constexpr (M::format_id == COO_FORMAT) {
    s.insert(s.end(), {vp(A.crow_indices().data()), vp(A.ccolumn_indices().data()),
                       vp(A.cvalues().data())});
  } else if constexpr (M::format_id == CSR_FORMAT) {
    s.insert(s.end(), {vp(A.crow_offsets().data()), vp(A.ccolumn_indices().data()),
                       vp(A.cvalues().data())});
  } else if constexpr (M::format_id == DIA_FORMAT) {
    s.insert(s.end(), {vp(A.cdiagonal_offsets().data()), vp(A.cvalues().data()),
                       reinterpret_cast<const void*>(A.cvalues().nrows()),
                       reinterpret_cast<const void*>(A.cdiagonal_offsets().size())});
  } else if constexpr (M::format_id == ELL_FORMAT) {
    s.insert(s.end(), {vp(A.ccolumn_indices().data()), vp(A.cvalues().data()),
                       reinterpret_cast<const void*>(A.cvalues().ncols()),
                       reinterpret_cast<const void*>(A.entries_per_row())});
  } else if constexpr (M::format_id == HYB_FORMAT) {
    auto a = signature(A.cell()), b = signature(A.ccoo());
    s.insert(s.end(), a.begin(), a.end());
    s.insert(s.end(), b.begin(), b.end());
  } else {
    auto a = signature(A.cdia()), b = signature(A.ccsr());
    s.insert(s.end(), a.begin(), a.end());
    s.insert(s.end(), b.begin(), b.end());
  }
  return s;
}

// C-ABI handle over the arrays of a HIPSpace container (through the int32
// / FP64 copies of `st` for long long indices / int values).
template <class M>
inline mph_matrix_t create_handle(const M& A, Stage& st) {
  static_assert(M::memory_space::is_device, "handle over host arrays");
  using V = typename M::non_const_value_type;
  constexpr int vt = handle_code_v<V>;
  const int64_t nr = (int64_t)A.nrows(), nc = (int64_t)A.ncols();
  auto dm_size = [](const auto& d) { return (size_t)d.nrows() * d.ncols(); };
  mph_matrix_t h = nullptr;
  int status = MPH_OK;
  if constexpr (M::format_id == COO_FORMAT) {
    status = mph_matrix_create_coo(&h, nr, nc, (int64_t)A.nnnz(),
                                   stage_index(st, A.crow_indices().data(), A.nnnz()),
                                   stage_index(st, A.ccolumn_indices().data(), A.nnnz()),
                                   stage_values(st, A.cvalues().data(), A.nnnz()), vt);
  } else if constexpr (M::format_id == CSR_FORMAT) {
    status = mph_matrix_create_csr(&h, nr, nc, (int64_t)A.nnnz(),
                                   stage_index(st, A.crow_offsets().data(), A.nrows() + 1),
                                   stage_index(st, A.ccolumn_indices().data(), A.nnnz()),
                                   stage_values(st, A.cvalues().data(), A.nnnz()), vt);
  } else if constexpr (M::format_id == DIA_FORMAT) {
    status = mph_matrix_create_dia(
        &h, nr, nc, (int64_t)A.nnnz(), (int64_t)A.cdiagonal_offsets().size(),
        (int64_t)A.cvalues().nrows(),
        stage_index(st, A.cdiagonal_offsets().data(), A.cdiagonal_offsets().size()),
        stage_values(st, A.cvalues().data(), dm_size(A.cvalues())), vt);
  } else if constexpr (M::format_id == ELL_FORMAT) {
    status = mph_matrix_create_ell(
        &h, nr, nc, (int64_t)A.nnnz(), (int64_t)A.entries_per_row(), (int64_t)A.cvalues().ncols(),
        (int64_t)A.cvalues().nrows(),
        stage_index(st, A.ccolumn_indices().data(), dm_size(A.ccolumn_indices())),
        stage_values(st, A.cvalues().data(), dm_size(A.cvalues())), vt);
  } else if constexpr (M::format_id == HYB_FORMAT) {
    const auto& e = A.cell();
    const auto& c = A.ccoo();
    status = mph_matrix_create_hyb(
        &h, nr, nc, (int64_t)e.nnnz(), (int64_t)e.entries_per_row(), (int64_t)e.cvalues().ncols(),
        (int64_t)e.cvalues().nrows(),
        stage_index(st, e.ccolumn_indices().data(), dm_size(e.ccolumn_indices())),
        stage_values(st, e.cvalues().data(), dm_size(e.cvalues())), (int64_t)c.nnnz(),
        stage_index(st, c.crow_indices().data(), c.nnnz()),
        stage_index(st, c.ccolumn_indices().data(), c.nnnz()),
        stage_values(st, c.cvalues().data(), c.nnnz()), vt);
  } else {
    const auto& d = A.cdia();
    const auto& c = A.ccsr();
    status = mph_matrix_create_hdc(
        &h, nr, nc, (int64_t)d.nnnz(), (int64_t)d.cdiagonal_offsets().size(),
        (int64_t)d.cvalues().nrows(),
        stage_index(st, d.cdiagonal_offsets().data(), d.cdiagonal_offsets().size()),
        stage_values(st, d.cvalues().data(), dm_size(d.cvalues())), (int64_t)c.nnnz(),
        stage_index(st, c.crow_offsets().data(), c.nrows() + 1),
        stage_index(st, c.ccolumn_indices().data(), c.nnnz()),
        stage_values(st, c.cvalues().data(), c.nnnz()), vt);
  }
  check(status, "Morpheus::multiply");
  // the analysis keeps no copy of any value: products always read the
  // container's value arrays, so values written in place through a View
  // kept from an earlier accessor call are seen by the next multiply (as
  // with the reference's containers, which hold no derived state)
  status = mph_matrix_set_policy(h, MPH_POLICY_NO_VALUE_COPIES);
  if (status != MPH_OK) mph_matrix_destroy(h);
  check(status, "Morpheus::multiply");
  for (auto& f : st.refresh) f();
  return h;
}

// The value arrays of a host container copied into its staged device copy
// (same format and shape).
template <class S, class D>
inline void copy_values(const S& src, D& dst) {
  if constexpr (S::format_id == COO_FORMAT || S::format_id == CSR_FORMAT) {
    copy_array(src.cvalues().const_view(), dst.values().view(), "values");
  } else if constexpr (S::format_id == DIA_FORMAT || S::format_id == ELL_FORMAT) {
    copy_dense_matrix(src.cvalues(), dst.values());
  } else if constexpr (S::format_id == HYB_FORMAT) {
    copy_values(src.cell(), dst.ell());
    copy_values(src.ccoo(), dst.coo());
  } else {
    copy_values(src.cdia(), dst.dia());
    copy_values(src.ccsr(), dst.csr());
  }
}

template <class M>
using device_type_of = typename M::template rebind<typename M::non_const_value_type,
                                                   typename M::non_const_index_type,
                                                   typename M::array_layout, HIPSpace>;

// Handle of a concrete sparse container, built (or rebuilt after a change)
// on demand.  Host containers are staged into HBM first.
template <class M>
inline mph_matrix_t matrix_handle(const M& A) {
  PlanCache& pc = A.plan_cache();
  auto sig = signature(A);
  if (pc.handle && !pc.stale && pc.signature == sig) {
    if constexpr (!M::memory_space::is_device) {
      // the host values may have been written through a kept reference: the
      // staged copy takes them on every call (the structure stays cached)
      using DM = device_type_of<M>;
      copy_values(A, *std::static_pointer_cast<DM>(pc.device_copy));
    }
    if (pc.refresh) pc.refresh();
    return pc.handle;
  }
  pc.reset();
  auto st = std::make_shared<Stage>();
  if constexpr (M::memory_space::is_device) {
    pc.handle = create_handle(A, *st);
  } else {
    using DM = device_type_of<M>;
    auto dev = std::make_shared<DM>();
    dev->allocate(A);
    copy_sparse(A, *dev);
    pc.handle = create_handle(*dev, *st);
    pc.device_copy = dev;
  }
  if (!st->refresh.empty()) {
    pc.refresh = [st]() {
      for (auto& f : st->refresh) f();
    };
    pc.write_back = [st]() {
      for (auto& f : st->write_back) f();
    };
  }
  pc.stage = st;
  pc.signature = std::move(sig);
  pc.stale = false;
  return pc.handle;
}

// ------------------------------------------------------------ multiply
template <class M, class VX, class VY>
inline void multiply_concrete(const M& A, const VX& x, VY& y, bool init) {
  using V = typename M::non_const_value_type;
  static_assert(std::is_same_v<typename VX::non_const_value_type, V> &&
                    std::is_same_v<typename VY::non_const_value_type, V>,
                "Morpheus::multiply: matrix and vectors must hold the same value type");
  if (x.size() != A.ncols() || y.size() != A.nrows())
    throw InvalidInputException("Morpheus::multiply: x must have ncols entries and y nrows entries (" +
                                std::to_string(A.nrows()) + "x" + std::to_string(A.ncols()) +
                                " matrix, x " + std::to_string(x.size()) + ", y " +
                                std::to_string(y.size()) + ")");
  mph_matrix_t h = matrix_handle(A);
  // int vectors travel as FP64 (the handle of an int container computes in FP64)
  auto run = [&](const V* xd, V* yd) {
    as_compute_vector(const_cast<V*>(xd), x.size(), true, false, [&](void* xc) {
      as_compute_vector(yd, y.size(), !init, true, [&](void* yc) {
        check(mph_multiply(h, xc, yc, init ? 1 : 0, nullptr), "Morpheus::multiply");
      });
    });
  };
  if constexpr (M::memory_space::is_device) {
    run(x.data(), y.data());
  } else {
    DenseVector<V, HIPSpace> xd(x.size()), yd(y.size());
    copy_array(x.const_view(), xd.view(), "x");
    if (!init) copy_array(y.const_view(), yd.view(), "y");
    run(xd.data(), yd.data());
    copy_array(yd.const_view(), y.view(), "y");
  }
}

template <class Exec, class M>
constexpr void check_access() {
  static_assert(Exec::is_device == M::memory_space::is_device,
                "the execution space cannot access the container's memory space");
}

// ------------------------------------------------------------ convert
inline int64_t i64(size_t v) { return (int64_t)v; }

template <class S, class D>
inline void convert_device(const S& src, D& dst);

// COO -> every format (core/src/impl/<Fmt>/Serial/Morpheus_Convert_Impl.hpp).
// Indices reach the kernels as int32 (long long containers: narrowed copies
// in, widened copies out); values move as raw words of their own type.
template <class S, class D>
inline void convert_from_coo(const S& src, D& dst) {
  constexpr int vt = value_code<typename S::non_const_value_type>::value;
  const int64_t n = i64(src.nrows()), m = i64(src.ncols()), nnz = i64(src.nnnz());
  Stage st;
  const int* ri = stage_index(st, src.crow_indices().data(), src.nnnz());
  const int* ci = stage_index(st, src.ccolumn_indices().data(), src.nnnz());
  const void* v = src.cvalues().data();
  auto dm_size = [](const auto& d) { return (size_t)d.nrows() * d.ncols(); };
  if constexpr (D::format_id == CSR_FORMAT) {
    dst.resize(n, m, nnz);
    auto ro = idx_out(dst.row_offsets().data(), (size_t)n + 1);
    auto co = idx_out(dst.column_indices().data(), (size_t)nnz);
    check(mph_convert_coo_to_csr(n, nnz, ri, ci, v, vt, ro.p, co.p, dst.values().data(), nullptr),
          "Morpheus::convert");
    ro.done();
    co.done();
  } else if constexpr (D::format_id == DIA_FORMAT) {
    int64_t nd = 0;
    check(mph_convert_coo_to_dia_size(n, m, nnz, ri, ci, &nd, nullptr), "Morpheus::convert");
    dst.resize(n, m, nnz, nd, 32);
    auto off = idx_out(dst.diagonal_offsets().data(), (size_t)nd);
    check(mph_convert_coo_to_dia(n, m, nnz, ri, ci, v, vt, nd, i64(dst.values().nrows()), off.p,
                                 dst.values().data(), nullptr),
          "Morpheus::convert");
    off.done();
  } else if constexpr (D::format_id == ELL_FORMAT) {
    int64_t K = 0;
    check(mph_convert_coo_to_ell_size(n, nnz, ri, &K, nullptr), "Morpheus::convert");
    dst.resize(n, m, nnz, K, 32);
    auto eci = idx_out(dst.column_indices().data(), dm_size(dst.column_indices()));
    check(mph_convert_coo_to_ell(n, nnz, ri, ci, v, vt, K, i64(dst.values().ncols()),
                                 i64(dst.values().nrows()), eci.p, dst.values().data(), nullptr),
          "Morpheus::convert");
    eci.done();
  } else if constexpr (D::format_id == HYB_FORMAT) {
    int64_t K = 0, enz = 0, cnz = 0;
    check(mph_convert_coo_to_hyb_size(n, nnz, ri, &K, &enz, &cnz, nullptr), "Morpheus::convert");
    dst.resize(n, m, enz, cnz, K, 32);
    auto& e = dst.ell();
    auto& c = dst.coo();
    auto eci = idx_out(e.column_indices().data(), dm_size(e.column_indices()));
    auto cri = idx_out(c.row_indices().data(), (size_t)cnz);
    auto cci = idx_out(c.column_indices().data(), (size_t)cnz);
    check(mph_convert_coo_to_hyb(n, nnz, ri, ci, v, vt, K, i64(e.values().ncols()),
                                 i64(e.values().nrows()), eci.p, e.values().data(), cri.p, cci.p,
                                 c.values().data(), nullptr),
          "Morpheus::convert");
    eci.done();
    cri.done();
    cci.done();
  } else if constexpr (D::format_id == HDC_FORMAT) {
    int64_t nd = 0, dnz = 0, cnz = 0;
    check(mph_convert_coo_to_hdc_size(n, m, nnz, ri, ci, &nd, &dnz, &cnz, nullptr),
          "Morpheus::convert");
    dst.resize(n, m, dnz, cnz, nd, 32);
    auto& d = dst.dia();
    auto& c = dst.csr();
    auto off = idx_out(d.diagonal_offsets().data(), (size_t)nd);
    auto cro = idx_out(c.row_offsets().data(), (size_t)n + 1);
    auto cci = idx_out(c.column_indices().data(), (size_t)cnz);
    check(mph_convert_coo_to_hdc(n, m, nnz, ri, ci, v, vt, nd, i64(d.values().nrows()), off.p,
                                 d.values().data(), cro.p, cci.p, c.values().data(), nullptr),
          "Morpheus::convert");
    off.done();
    cro.done();
    cci.done();
  }
  dst.touch();
}

// every format -> COO
template <class S, class D>
inline void convert_to_coo(const S& src, D& dst) {
  constexpr int vt = value_code<typename S::non_const_value_type>::value;
  const int64_t n = i64(src.nrows()), m = i64(src.ncols()), nnz = i64(src.nnnz());
  dst.resize(n, m, nnz);
  Stage st;
  auto dm_size = [](const auto& d) { return (size_t)d.nrows() * d.ncols(); };
  auto ro = idx_out(dst.row_indices().data(), (size_t)nnz);
  auto co = idx_out(dst.column_indices().data(), (size_t)nnz);
  int* ri = ro.p;
  int* ci = co.p;
  void* v = dst.values().data();
  int status = MPH_OK;
  if constexpr (S::format_id == CSR_FORMAT) {
    status = mph_convert_csr_to_coo(n, nnz, stage_index(st, src.crow_offsets().data(), src.nrows() + 1),
                                    stage_index(st, src.ccolumn_indices().data(), src.nnnz()),
                                    src.cvalues().data(), vt, ri, ci, v, nullptr);
  } else if constexpr (S::format_id == DIA_FORMAT) {
    status = mph_convert_dia_to_coo(
        n, m, nnz, i64(src.cdiagonal_offsets().size()), i64(src.cvalues().nrows()),
        stage_index(st, src.cdiagonal_offsets().data(), src.cdiagonal_offsets().size()),
        src.cvalues().data(), vt, ri, ci, v, nullptr);
  } else if constexpr (S::format_id == ELL_FORMAT) {
    status = mph_convert_ell_to_coo(
        n, nnz, i64(src.cvalues().ncols()), i64(src.cvalues().nrows()),
        stage_index(st, src.ccolumn_indices().data(), dm_size(src.ccolumn_indices())),
        src.cvalues().data(), vt, ri, ci, v, nullptr);
  } else if constexpr (S::format_id == HYB_FORMAT) {
    const auto& e = src.cell();
    const auto& c = src.ccoo();
    status = mph_convert_hyb_to_coo(
        n, nnz, i64(e.cvalues().ncols()), i64(e.cvalues().nrows()),
        stage_index(st, e.ccolumn_indices().data(), dm_size(e.ccolumn_indices())),
        e.cvalues().data(), i64(c.nnnz()), stage_index(st, c.crow_indices().data(), c.nnnz()),
        stage_index(st, c.ccolumn_indices().data(), c.nnnz()), c.cvalues().data(), vt, ri, ci, v,
        nullptr);
  } else if constexpr (S::format_id == HDC_FORMAT) {
    const auto& d = src.cdia();
    const auto& c = src.ccsr();
    status = mph_convert_hdc_to_coo(
        n, m, nnz, i64(d.cdiagonal_offsets().size()), i64(d.cvalues().nrows()),
        stage_index(st, d.cdiagonal_offsets().data(), d.cdiagonal_offsets().size()),
        d.cvalues().data(), i64(c.nnnz()), stage_index(st, c.crow_offsets().data(), c.nrows() + 1),
        stage_index(st, c.ccolumn_indices().data(), c.nnnz()), c.cvalues().data(), vt, ri, ci, v,
        nullptr);
  }
  check(status, "Morpheus::convert");
  ro.done();
  co.done();
  dst.touch();
}

// Same-format convert is a deep copy into freshly sized storage
// (core/src/impl/Csr/Serial/Morpheus_Convert_Impl.hpp:38-64).
template <class S, class D>
inline void convert_device(const S& src, D& dst) {
  static_assert(std::is_same_v<typename S::non_const_value_type, typename D::non_const_value_type>,
                "Morpheus::convert between value types is not supported");
  if constexpr (S::format_id == D::format_id) {
    dst.allocate(src);
    copy_sparse(src, dst);
  } else if constexpr (S::format_id == COO_FORMAT) {
    convert_from_coo(src, dst);
  } else if constexpr (D::format_id == COO_FORMAT) {
    convert_to_coo(src, dst);
  } else {
    // core/src/impl/Morpheus_Convert_Impl.hpp:83-106: through a temporary COO
    CooMatrix<typename S::non_const_value_type, typename S::non_const_index_type,
              typename S::array_layout, HIPSpace>
        tmp;
    convert_to_coo(src, tmp);
    convert_from_coo(tmp, dst);
  }
}

template <class S, class D>
inline void convert_concrete(const S& src, D& dst) {
  if constexpr (S::memory_space::is_device) {
    convert_device(src, dst);
  } else {
    // host containers: the conversion runs on the device and the result
    // comes back into dst
    device_type_of<S> sd;
    sd.allocate(src);
    copy_sparse(src, sd);
    device_type_of<D> dd;
    convert_device(sd, dd);
    dst.allocate(dd);
    copy_sparse(dd, dst);
  }
}

}  // namespace Impl

// ============================================================ copy
template <class V1, class... P1, class V2, class... P2>
void copy(const DenseVector<V1, P1...>& src, DenseVector<V2, P2...>& dst) {
  Impl::copy_array(src.const_view(), dst.view(), "DenseVector");
}
template <class V1, class... P1, class V2, class... P2>
void copy(const DenseMatrix<V1, P1...>& src, DenseMatrix<V2, P2...>& dst) {
  Impl::copy_dense_matrix(src, dst);
}
// sliced DenseVector copy (core/src/Morpheus_Copy.hpp:70-80)
template <class V1, class... P1, class V2, class... P2>
void copy(const DenseVector<V1, P1...>& src, DenseVector<V2, P2...>& dst, const int src_begin,
          const int src_end, const int dst_begin, const int dst_end) {
  using SM = typename DenseVector<V1, P1...>::memory_space;
  using DM = typename DenseVector<V2, P2...>::memory_space;
  if (src_end - src_begin != dst_end - dst_begin || src_begin < 0 || dst_begin < 0 ||
      src_end < src_begin || (size_t)src_end > src.size() || (size_t)dst_end > dst.size())
    throw RuntimeException("Morpheus::copy: invalid slice ranges");
  Impl::Array<V1, SM>::copy_bytes(dst.data() + dst_begin, src.data() + src_begin,
                                  (size_t)(src_end - src_begin) * sizeof(V1), DM::is_device,
                                  SM::is_device);
}

template <class S, class D>
std::enable_if_t<is_sparse_matrix_container_v<S> && is_sparse_matrix_container_v<D>> copy(
    const S& src, D& dst) {
  Impl::copy_sparse(src, dst);
}

// DynamicMatrix copies: the active formats must agree
// (core/src/impl/Dynamic/Morpheus_Copy_Impl.hpp:30-120).
template <class S, class D>
std::enable_if_t<(is_dynamic_matrix_format_container_v<S> ||
                  is_dynamic_matrix_format_container_v<D>) &&
                 !is_dense_vector_format_container_v<S>>
copy(const S& src, D& dst) {
  if (src.format_index() != dst.format_index())
    throw FormatConversionException(
        "Morpheus::copy() is only available between the same container types (" +
        std::to_string(src.format_index()) + " != " + std::to_string(dst.format_index()) +
        "). Please use Morpheus::convert() instead to perform conversions between different "
        "types.");
  auto one = [&](const auto& s) {
    using SC = std::decay_t<decltype(s)>;
    if constexpr (is_dynamic_matrix_format_container_v<D>) {
      std::visit(
          [&](auto& d) {
            using DC = std::decay_t<decltype(d)>;
            if constexpr (SC::format_id == DC::format_id) Impl::copy_sparse(s, d);
          },
          dst.formats());
    } else {
      if constexpr (SC::format_id == D::format_id) Impl::copy_sparse(s, dst);
    }
  };
  if constexpr (is_dynamic_matrix_format_container_v<S>)
    std::visit(one, src.const_formats());
  else
    one(src);
}

// ============================================================ mirrors
template <template <class, class...> class C, class T, class... P>
typename C<T, P...>::HostMirror create_mirror(const C<T, P...>& src) {
  typename C<T, P...>::HostMirror m;
  m.allocate(src);
  return m;
}

template <class Space, template <class, class...> class C, class T, class... P>
auto create_mirror(const C<T, P...>& src) {
  using S = C<T, P...>;
  using M = C<typename S::non_const_value_type, typename S::non_const_index_type,
              typename S::array_layout, typename memory_space_of<Space>::type>;
  M m;
  m.allocate(src);
  return m;
}

template <template <class, class...> class C, class T, class... P>
typename C<T, P...>::HostMirror create_mirror_container(const C<T, P...>& src) {
  using S = C<T, P...>;
  if constexpr (!S::memory_space::is_device &&
                std::is_same_v<typename S::HostMirror::array_layout, typename S::array_layout>) {
    return typename S::HostMirror(src);
  } else {
    return create_mirror(src);
  }
}

template <class Space, template <class, class...> class C, class T, class... P>
auto create_mirror_container(const C<T, P...>& src) {
  using S = C<T, P...>;
  using Mem = typename memory_space_of<Space>::type;
  using M = C<typename S::non_const_value_type, typename S::non_const_index_type,
              typename S::array_layout, Mem>;
  if constexpr (std::is_same_v<Mem, typename S::memory_space>) {
    return M(src);
  } else {
    M m;
    m.allocate(src);
    return m;
  }
}

// ============================================================ multiply
// y = A x (init) or y += A x (core/src/Morpheus_Multiply.hpp:57-78)
template <typename ExecSpace, typename Matrix, typename VectorIn, typename VectorOut>
void multiply(const Matrix& A, const VectorIn& x, VectorOut& y, const bool init) {
  static_assert(is_dense_vector_format_container_v<VectorIn> &&
                    is_dense_vector_format_container_v<VectorOut>,
                "Morpheus::multiply: x and y must be DenseVector containers");
  if constexpr (is_dynamic_matrix_format_container_v<Matrix>) {
    std::visit(
        [&](const auto& a) {
          using M = std::decay_t<decltype(a)>;
          Impl::check_access<ExecSpace, M>();
          Impl::multiply_concrete(a, x, y, init);
        },
        A.const_formats());
  } else {
    Impl::check_access<ExecSpace, Matrix>();
    Impl::multiply_concrete(A, x, y, init);
  }
}
template <typename ExecSpace, typename Matrix, typename VectorIn, typename VectorOut>
void multiply(const Matrix& A, const VectorIn& x, VectorOut& y) {
  multiply<ExecSpace>(A, x, y, true);
}

// ============================================================ convert
template <typename ExecSpace, typename SourceType, typename DestinationType>
std::enable_if_t<!std::is_same_v<DestinationType, formats_e> && !std::is_same_v<DestinationType, int>>
convert(const SourceType& src, DestinationType& dst) {
  constexpr bool sdyn = is_dynamic_matrix_format_container_v<SourceType>;
  constexpr bool ddyn = is_dynamic_matrix_format_container_v<DestinationType>;
  if constexpr (sdyn && ddyn) {
    // core/src/impl/Dynamic/Morpheus_Convert_Impl.hpp:71-83
    dst.set_nrows(src.nrows());
    dst.set_ncols(src.ncols());
    dst.set_nnnz(src.nnnz());
    std::visit(
        [&](const auto& s, auto& d) {
          Impl::check_access<ExecSpace, std::decay_t<decltype(s)>>();
          Impl::convert_concrete(s, d);
        },
        src.const_formats(), dst.formats());
  } else if constexpr (sdyn) {
    std::visit(
        [&](const auto& s) {
          Impl::check_access<ExecSpace, std::decay_t<decltype(s)>>();
          Impl::convert_concrete(s, dst);
        },
        src.const_formats());
  } else if constexpr (ddyn) {
    dst.set_nrows(src.nrows());
    dst.set_ncols(src.ncols());
    dst.set_nnnz(src.nnnz());
    Impl::check_access<ExecSpace, SourceType>();
    std::visit([&](auto& d) { Impl::convert_concrete(src, d); }, dst.formats());
  } else {
    Impl::check_access<ExecSpace, SourceType>();
    Impl::check_access<ExecSpace, DestinationType>();
    Impl::convert_concrete(src, dst);
  }
}

// In-place switch of a DynamicMatrix (core/src/Morpheus_Convert.hpp:88-141):
// through a temporary COO; failures are reported, not thrown.
template <typename ExecSpace, typename SourceType>
conversion_error_e convert(SourceType& src, const formats_e index) {
  static_assert(is_dynamic_matrix_format_container_v<SourceType>,
                "Container must be a DynamicMatrix.");
  using Coo = std::variant_alternative_t<0, typename SourceType::variant_type>;
  if (src.active_index() == static_cast<int>(index)) return CONV_SUCCESS;
  Coo temp;
  if (src.active_index() != COO_FORMAT) {
    try {
      Morpheus::convert<ExecSpace>(src, temp);
    } catch (...) {
      return DYNAMIC_TO_PROXY;
    }
  } else {
    temp = std::get<0>(src.const_formats());
  }
  SourceType dynamic_temp;
  dynamic_temp.activate(index);
  try {
    Morpheus::convert<ExecSpace>(temp, dynamic_temp);
  } catch (...) {
    return PROXY_TO_DYNAMIC;
  }
  src = dynamic_temp;
  return CONV_SUCCESS;
}
template <typename ExecSpace, typename SourceType>
conversion_error_e convert(SourceType& src, const int index) {
  return Morpheus::convert<ExecSpace>(src, static_cast<formats_e>(index));
}

// ============================================================ sort / analytics
namespace Impl {
// Runs f on a HIPSpace view of a COO container (the container itself, or a
// staged device copy whose contents are copied back when write_back).
template <class M, class F>
inline void on_device_coo(M& A, bool write_back, F&& f) {
  if constexpr (M::memory_space::is_device) {
    f(A);
  } else {
    device_type_of<M> d;
    d.allocate(A);
    copy_sparse(A, d);
    f(d);
    if (write_back) copy_sparse(d, A);
  }
}
}  // namespace Impl

// Stable sort of COO entries by (row, column)
// (core/src/Morpheus_Sort.hpp:32-40, core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:111-166).
template <typename ExecSpace, typename Matrix>
void sort_by_row_and_column(Matrix& mat) {
  static_assert(Matrix::format_id == COO_FORMAT, "sort_by_row_and_column takes a CooMatrix");
  Impl::check_access<ExecSpace, Matrix>();
  Impl::on_device_coo(mat, true, [](auto& d) {
    constexpr int vt = Impl::value_code<typename Matrix::non_const_value_type>::value;
    auto ri = Impl::idx_out(d.row_indices().data(), d.nnnz(), true);
    auto ci = Impl::idx_out(d.column_indices().data(), d.nnnz(), true);
    Impl::check(mph_coo_sort_by_row_and_column((int64_t)d.nnnz(), ri.p, ci.p, d.values().data(),
                                               vt, nullptr),
                "Morpheus::sort_by_row_and_column");
    ri.done();
    ci.done();
  });
  mat.touch();
}

// core/src/Morpheus_Sort.hpp:42-46, core/src/impl/Coo/Serial/Morpheus_Sort_Impl.hpp:168-189
template <typename ExecSpace, typename Matrix>
bool is_sorted(Matrix& mat) {
  static_assert(Matrix::format_id == COO_FORMAT, "is_sorted takes a CooMatrix");
  Impl::check_access<ExecSpace, Matrix>();
  int sorted = 0;
  Impl::on_device_coo(mat, false, [&](auto& d) {
    Impl::Stage st;
    Impl::check(mph_coo_is_sorted((int64_t)d.nnnz(), Impl::stage_index(st, d.crow_indices().data(), d.nnnz()),
                                  Impl::stage_index(st, d.ccolumn_indices().data(), d.nnnz()),
                                  &sorted, nullptr),
                "Morpheus::is_sorted");
  });
  return sorted != 0;
}

// ============================================================ dense vectors
namespace Impl {
template <class Vec>
constexpr int vector_code() {
  using V = typename Vec::non_const_value_type;
  static_assert(std::is_same_v<V, double> || std::is_same_v<V, float>,
                "the vector algorithms take double or float vectors");
  return std::is_same_v<V, double> ? MPH_F64 : MPH_F32;
}
// fn(device pointer) for a vector in either space (host vectors are staged;
// written back when `out`)
template <class Vec, class Fn>
inline void with_device_vector(Vec& v, bool in, bool out, Fn&& fn) {
  if constexpr (Vec::memory_space::is_device) {
    fn(v.data());
  } else {
    DenseVector<typename Vec::non_const_value_type, HIPSpace> d(v.size());
    if (in) copy_array(v.const_view(), d.view(), "vector");
    fn(d.data());
    if (out) copy_array(d.const_view(), v.view(), "vector");
  }
}
}  // namespace Impl

// core/src/Morpheus_Dot.hpp:45-57: the sum of x_i y_i over the first n
// entries (formed in FP64 and added in a fixed order)
template <typename ExecSpace, typename Vector1, typename Vector2>
typename Vector2::value_type dot(typename Vector1::size_type n, const Vector1& x,
                                 const Vector2& y) {
  static_assert(is_dense_vector_format_container_v<Vector1> &&
                    is_dense_vector_format_container_v<Vector2>,
                "x and y must be DenseVector containers");
  Impl::check_access<ExecSpace, Vector1>();
  if (x.size() < n || y.size() < n) throw InvalidInputException("Morpheus::dot: vectors too short");
  double r = 0;
  auto& xm = const_cast<Vector1&>(x);
  auto& ym = const_cast<Vector2&>(y);
  Impl::with_device_vector(xm, true, false, [&](auto* px) {
    Impl::with_device_vector(ym, true, false, [&](auto* py) {
      Impl::check(mph_dot((int64_t)n, px, py, Impl::vector_code<Vector1>(), &r, nullptr),
                  "Morpheus::dot");
    });
  });
  return (typename Vector2::value_type)r;
}

// core/src/Morpheus_WAXPBY.hpp:48-65: w = alpha x + beta y on the first n
// entries, each product and sum rounded on its own
template <typename ExecSpace, typename Vector1, typename Vector2, typename Vector3>
void waxpby(const size_t n, const typename Vector1::value_type alpha, const Vector1& x,
            const typename Vector2::value_type beta, const Vector2& y, Vector3& w) {
  static_assert(is_dense_vector_format_container_v<Vector1> &&
                    is_dense_vector_format_container_v<Vector2> &&
                    is_dense_vector_format_container_v<Vector3>,
                "x, y and w must be DenseVector containers");
  Impl::check_access<ExecSpace, Vector1>();
  if (x.size() < n || y.size() < n || w.size() < n)
    throw InvalidInputException("Morpheus::waxpby: vectors too short");
  auto& xm = const_cast<Vector1&>(x);
  auto& ym = const_cast<Vector2&>(y);
  Impl::with_device_vector(xm, true, false, [&](auto* px) {
    Impl::with_device_vector(ym, true, false, [&](auto* py) {
      Impl::with_device_vector(w, false, true, [&](auto* pw) {
        Impl::check(mph_waxpby((int64_t)n, (double)alpha, px, (double)beta, py, pw,
                               Impl::vector_code<Vector1>(), nullptr),
                    "Morpheus::waxpby");
      });
    });
  });
}

// core/src/Morpheus_Reduction.hpp:44-52: the sum of the first `size` entries
template <typename ExecSpace, typename Vector>
typename Vector::value_type reduce(const Vector& in, typename Vector::size_type size) {
  static_assert(is_dense_vector_format_container_v<Vector>, "in must be a DenseVector container");
  Impl::check_access<ExecSpace, Vector>();
  if (in.size() < size) throw InvalidInputException("Morpheus::reduce: vector too short");
  double r = 0;
  auto& vm = const_cast<Vector&>(in);
  Impl::with_device_vector(vm, true, false, [&](auto* p) {
    Impl::check(mph_reduce((int64_t)size, p, Impl::vector_code<Vector>(), &r, nullptr),
                "Morpheus::reduce");
  });
  return (typename Vector::value_type)r;
}

// ============================================================ analytics
// core/src/Morpheus_MatrixAnalytics.hpp:49-104
template <typename Matrix>
typename Matrix::size_type number_of_rows(const Matrix& A) {
  return A.nrows();
}
template <typename Matrix>
typename Matrix::size_type number_of_columns(const Matrix& A) {
  return A.ncols();
}
template <typename Matrix>
typename Matrix::size_type number_of_nnz(const Matrix& A) {
  return A.nnnz();
}
template <typename Matrix>
double average_nnnz(const Matrix& A) {
  return A.nnnz() / (double)A.nrows();
}
template <typename Matrix>
double density(const Matrix& A) {
  return A.nnnz() / (double)(A.nrows() * A.ncols());
}

namespace Impl {
// a concrete container of A (the active one of a DynamicMatrix)
template <class M, class Fn>
inline void with_concrete(const M& A, Fn&& fn) {
  if constexpr (is_dynamic_matrix_format_container_v<M>)
    std::visit([&](const auto& a) { fn(a); }, A.const_formats());
  else
    fn(A);
}

// fn(int* device counters) over an int DenseVector in either space: host
// vectors travel through a device copy (in when `in`, always back out)
template <class Vec, class Fn>
inline void with_device_ints(Vec& v, bool in, Fn&& fn) {
  static_assert(std::is_same_v<typename Vec::non_const_value_type, int>,
                "the analytics count into a DenseVector<int>");
  if constexpr (Vec::memory_space::is_device) {
    fn(v.data());
  } else {
    DenseVector<int, HIPSpace> d(v.size());
    if (in) copy_array(v.const_view(), d.view(), "counters");
    fn(d.data());
    copy_array(d.const_view(), v.view(), "counters");
  }
}

template <class Vec>
inline mph_i32_stats_t int_stats(const Vec& v, size_t n, double mean, int threshold) {
  static_assert(std::is_same_v<typename Vec::non_const_value_type, int>,
                "the vector analytics take a DenseVector<int>");
  mph_i32_stats_t st{};
  if constexpr (Vec::memory_space::is_device) {
    check(mph_i32_stats(v.data(), (int64_t)n, mean, threshold, &st, nullptr), "Morpheus::max/min/std");
  } else {
    DenseVector<int, HIPSpace> d(v.size());
    copy_array(v.const_view(), d.view(), "vector");
    check(mph_i32_stats(d.data(), (int64_t)n, mean, threshold, &st, nullptr), "Morpheus::max/min/std");
  }
  return st;
}

template <class Exec, class M>
inline DenseVector<int, HIPSpace> rows_counted(const M& A) {
  DenseVector<int, HIPSpace> c(A.nrows());
  with_concrete(A, [&](const auto& a) {
    check_access<Exec, std::decay_t<decltype(a)>>();
    check(mph_matrix_count_nnz_per_row(matrix_handle(a), c.data(), 1, nullptr),
          "Morpheus::count_nnz_per_row");
  });
  return c;
}

template <class Exec, class M>
inline DenseVector<int, HIPSpace> diagonals_counted(const M& A) {
  DenseVector<int, HIPSpace> c(A.nrows() + A.ncols() - 1);
  with_concrete(A, [&](const auto& a) {
    check_access<Exec, std::decay_t<decltype(a)>>();
    check(mph_matrix_count_nnz_per_diagonal(matrix_handle(a), c.data(), 1, nullptr),
          "Morpheus::count_nnz_per_diagonal");
  });
  return c;
}
}  // namespace Impl

// Non-zeros per row into a DenseVector<int> of nrows counters, per format as
// core/src/impl/<Fmt>/Serial/Morpheus_NonZeros_Per_Row_Impl.hpp counts them
// (core/src/Morpheus_MatrixAnalytics.hpp:112-116).  init = false adds.
template <typename ExecSpace, typename Matrix, typename Vector>
void count_nnz_per_row(const Matrix& A, Vector& nnz_per_row, const bool init = true) {
  if (nnz_per_row.size() != A.nrows())
    throw InvalidInputException(
        "count_nnz_per_row: destination vector must have equal size to the source matrix "
        "number of rows");
  Impl::with_concrete(A, [&](const auto& a) {
    Impl::check_access<ExecSpace, std::decay_t<decltype(a)>>();
    Impl::with_device_ints(nnz_per_row, !init, [&](int* out) {
      Impl::check(mph_matrix_count_nnz_per_row(Impl::matrix_handle(a), out, init ? 1 : 0, nullptr),
                  "Morpheus::count_nnz_per_row");
    });
  });
}

// Non-zeros per diagonal into a DenseVector<int> of nrows + ncols - 1
// counters, diagonal d = col - row + nrows - 1
// (core/src/Morpheus_MatrixAnalytics.hpp:206-211; per format as
// core/src/impl/<Fmt>/Kokkos/Morpheus_NonZeros_Per_Diagonal_Impl.hpp).
template <typename ExecSpace, typename Matrix, typename Vector>
void count_nnz_per_diagonal(const Matrix& A, Vector& nnz_per_diagonal, const bool init = true) {
  if (nnz_per_diagonal.size() != A.nrows() + A.ncols() - 1)
    throw InvalidInputException(
        "count_nnz_per_diagonal: destination vector must have equal size to the source matrix "
        "number of diagonals (i.e NROWS + NCOLS - 1)");
  Impl::with_concrete(A, [&](const auto& a) {
    Impl::check_access<ExecSpace, std::decay_t<decltype(a)>>();
    Impl::with_device_ints(nnz_per_diagonal, !init, [&](int* out) {
      Impl::check(mph_matrix_count_nnz_per_diagonal(Impl::matrix_handle(a), out, init ? 1 : 0,
                                                    nullptr),
                  "Morpheus::count_nnz_per_diagonal");
    });
  });
}

// core/src/Morpheus_MatrixAnalytics.hpp:126-140
template <typename ExecSpace, typename Matrix>
typename Matrix::size_type max_nnnz(const Matrix& A) {
  auto c = Impl::rows_counted<ExecSpace>(A);
  return (typename Matrix::size_type)Impl::int_stats(c, c.size(), 0.0, 0).max;
}
// core/src/Morpheus_MatrixAnalytics.hpp:151-165
template <typename ExecSpace, typename Matrix>
typename Matrix::size_type min_nnnz(const Matrix& A) {
  auto c = Impl::rows_counted<ExecSpace>(A);
  return (typename Matrix::size_type)Impl::int_stats(c, c.size(), 0.0, 0).min;
}
// core/src/Morpheus_MatrixAnalytics.hpp:177-191: the deviation around the
// integer mean nnnz / nrows, returned in the size type (truncated)
template <typename ExecSpace, typename Matrix>
typename Matrix::size_type std_nnnz(const Matrix& A) {
  auto c = Impl::rows_counted<ExecSpace>(A);
  const double mean = (double)(A.nnnz() / A.nrows());
  const auto st = Impl::int_stats(c, c.size(), mean, 0);
  return (typename Matrix::size_type)std::sqrt(st.sq_dev / (double)c.size());
}
// core/src/Morpheus_MatrixAnalytics.hpp:220-234
template <typename ExecSpace, typename Matrix>
typename Matrix::size_type count_diagonals(const Matrix& A) {
  auto c = Impl::diagonals_counted<ExecSpace>(A);
  return (typename Matrix::size_type)Impl::int_stats(c, c.size(), 0.0, 0).count;
}
// core/src/Morpheus_MatrixAnalytics.hpp:245-259
template <typename ExecSpace, typename Matrix>
typename Matrix::size_type count_true_diagonals(const Matrix& A,
                                                typename Matrix::index_type threshold) {
  auto c = Impl::diagonals_counted<ExecSpace>(A);
  return (typename Matrix::size_type)Impl::int_stats(c, c.size(), 0.0, (int)threshold).count;
}

// Vector analytics over a DenseVector<int> (core/src/Morpheus_VectorAnalytics.hpp:48-130)
template <typename ExecSpace, typename Vector>
typename Vector::value_type max(const Vector& vec, typename Vector::size_type size) {
  return (typename Vector::value_type)Impl::int_stats(vec, size, 0.0, 0).max;
}
template <typename ExecSpace, typename Vector>
typename Vector::value_type min(const Vector& vec, typename Vector::size_type size) {
  return (typename Vector::value_type)Impl::int_stats(vec, size, 0.0, 0).min;
}
template <typename ExecSpace, typename Vector>
double std(const Vector& vec, typename Vector::size_type size, double mean) {
  return std::sqrt(Impl::int_stats(vec, size, mean, 0).sq_dev / (double)size);
}
template <typename ExecSpace, typename Vector>
typename Vector::size_type count_nnz(const Vector& vec, typename Vector::value_type threshold = 0) {
  return (typename Vector::size_type)Impl::int_stats(vec, vec.size(), 0.0, (int)threshold).count;
}

// ============================================================ matrix operations
// core/src/Morpheus_MatrixOperations.hpp:54-58, with the Serial semantics of
// core/src/impl/<Fmt>/Serial/Morpheus_MatrixOperations_Impl.hpp (CSR / ELL
// the first entry of row i in column i, COO every diagonal entry, DIA the
// main diagonal's slots with stored zeros kept, HYB = ELL + COO, HDC = DIA +
// CSR).  A's values change in place; its cached analysis is rebuilt on the
// next product.
template <typename ExecSpace, typename Matrix, typename Vector>
void update_diagonal(Matrix& A, const Vector& diagonal) {
  static_assert(is_dense_vector_format_container_v<Vector>, "diagonal must be a DenseVector");
  if (diagonal.size() < std::min(A.nrows(), A.ncols()))
    throw InvalidInputException("Morpheus::update_diagonal: diagonal too short");
  if constexpr (is_dynamic_matrix_format_container_v<Matrix>) {
    std::visit([&](auto& a) { update_diagonal<ExecSpace>(a, diagonal); }, A.formats());
  } else {
    Impl::check_access<ExecSpace, Matrix>();
    auto& dm = const_cast<Vector&>(diagonal);
    const size_t nd = diagonal.size();
    if constexpr (Matrix::memory_space::is_device) {
      Impl::with_device_vector(dm, true, false, [&](auto* pd) {
        Impl::as_compute_vector(pd, nd, true, false, [&](void* pc) {
          Impl::check(mph_matrix_update_diagonal(Impl::matrix_handle(A), pc, nullptr),
                      "Morpheus::update_diagonal");
        });
      });
      // int values: the FP64 copy the handle updated, cast back
      if (A.plan_cache().write_back) A.plan_cache().write_back();
    } else {
      Impl::device_type_of<Matrix> dev;
      dev.allocate(A);
      Impl::copy_sparse(A, dev);
      Impl::Stage stage;
      mph_matrix_t h = Impl::create_handle(dev, stage);
      int st = MPH_OK;
      Impl::with_device_vector(dm, true, false, [&](auto* pd) {
        Impl::as_compute_vector(pd, nd, true, false, [&](void* pc) {
          st = mph_matrix_update_diagonal(h, pc, nullptr);
        });
      });
      mph_matrix_destroy(h);
      Impl::check(st, "Morpheus::update_diagonal");
      for (auto& f : stage.write_back) f();
      Impl::copy_sparse(dev, A);
      A.touch();  // the staged copy of A held the old values
    }
  }
}

// The diagonal the product sees: the sum of every stored (i, i) entry, into
// min(nrows, ncols) entries.  The reference's get_diagonal
// (core/src/Morpheus_MatrixOperations.hpp:69-73) throws NotImplementedException.
template <typename ExecSpace, typename Matrix, typename Vector>
void get_diagonal(const Matrix& A, Vector& diagonal) {
  static_assert(is_dense_vector_format_container_v<Vector>, "diagonal must be a DenseVector");
  const size_t n = std::min(A.nrows(), A.ncols());
  if (diagonal.size() < n) throw InvalidInputException("Morpheus::get_diagonal: diagonal too short");
  Impl::with_concrete(A, [&](const auto& a) {
    Impl::check_access<ExecSpace, std::decay_t<decltype(a)>>();
    Impl::with_device_vector(diagonal, false, true, [&](auto* pd) {
      Impl::as_compute_vector(pd, diagonal.size(), false, true, [&](void* pc) {
        Impl::check(mph_matrix_get_diagonal(Impl::matrix_handle(a), pc, nullptr),
                    "Morpheus::get_diagonal");
      });
    });
  });
}

// ============================================================ extensions
// Launch-shape autotuning of one HIPSpace matrix (no reference counterpart):
// times the admissible launch shapes of A on x / y and keeps the fastest.
// Returns the best time per product in milliseconds.
template <typename Matrix, typename VectorIn, typename VectorOut>
float tune_multiply(const Matrix& A, const VectorIn& x, VectorOut& y, int reps = 10) {
  static_assert(Matrix::memory_space::is_device, "tune_multiply takes HIPSpace containers");
  if (x.size() != A.ncols() || y.size() != A.nrows())
    throw InvalidInputException("Morpheus::tune_multiply: vector sizes do not match A");
  float ms = 0.f;
  Impl::check(mph_matrix_autotune(Impl::matrix_handle(A), 1, x.data(), y.data(), reps, nullptr, &ms),
              "Morpheus::tune_multiply");
  return ms;
}

// Chooses the active format of a HIPSpace DynamicMatrix by timing every
// admissible format on the device (the measured counterpart of the
// --morpheus-format flag, core/src/impl/Morpheus_Core.cpp:92-140) and
// switches A to it.  Returns the chosen format.
template <typename ExecSpace, typename Dynamic>
formats_e select_format(Dynamic& A, int reps = 10) {
  static_assert(is_dynamic_matrix_format_container_v<Dynamic>, "select_format takes a DynamicMatrix");
  static_assert(ExecSpace::is_device && Dynamic::memory_space::is_device,
                "select_format runs on HIPSpace containers");
  using Csr = CsrMatrix<typename Dynamic::non_const_value_type,
                        typename Dynamic::non_const_index_type, typename Dynamic::array_layout,
                        HIPSpace>;
  Csr csr;
  convert<ExecSpace>(A, csr);
  int best = CSR_FORMAT;
  constexpr int vt = Impl::value_code<typename Dynamic::non_const_value_type>::value;
  Impl::check(mph_autotune_format((int64_t)csr.nrows(), (int64_t)csr.ncols(), (int64_t)csr.nnnz(),
                                  csr.crow_offsets().data(), csr.ccolumn_indices().data(),
                                  csr.cvalues().data(), vt, 1, reps, nullptr, &best, nullptr),
              "Morpheus::select_format");
  if (convert<ExecSpace>(A, best) != CONV_SUCCESS)
    throw FormatConversionException("Morpheus::select_format: switch to the chosen format failed");
  return static_cast<formats_e>(best);
}

}  // namespace Morpheus

#ifndef MORPHEUS_AMD_NO_KOKKOS_SHIM
namespace Kokkos {
template <class Space>
using Random_XorShift64_Pool = Morpheus::RandomPool<Space>;
}  // namespace Kokkos
#endif
