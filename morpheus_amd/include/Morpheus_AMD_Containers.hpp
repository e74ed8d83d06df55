// Morpheus_AMD_Containers.hpp -- the Morpheus container types of the SpMV
// path (DenseVector, DenseMatrix, Coo/Csr/Dia/Ell/Hyb/HdcMatrix and
// DynamicMatrix) over host memory or MI355X HBM.
//
// Semantics follow the reference containers:
//   core/src/Morpheus_DenseVector.hpp:81-367, Morpheus_DenseMatrix.hpp,
//   Morpheus_CooMatrix.hpp:90-566, Morpheus_CsrMatrix.hpp:90-537,
//   Morpheus_DiaMatrix.hpp:96-570, Morpheus_EllMatrix.hpp:90-564,
//   Morpheus_HybMatrix.hpp:101-477, Morpheus_HdcMatrix.hpp:107-468,
//   Morpheus_DynamicMatrix.hpp:102-462, Morpheus_MatrixBase.hpp:68-188,
//   defaults core/src/Morpheus_ContainerTraits.hpp:59-120.
// Copies and assignments are shallow (shared, reference-counted storage);
// new storage is zero-initialised; resize keeps the leading contents.
//
// Storage on the device.  DIA values and ELL column indices / values are
// kept column-major on HIPSpace whatever the declared array_layout (the
// layout the MI355X kernels stream: diagonal-major for DIA, slot-major for
// ELL).  Host containers keep the declared layout; Morpheus::copy transposes
// between the two.
//
// Each sparse container carries a shared analysis cache: the C-ABI handle
// (row blocks, carry scratch, tuned launch shape) of its device arrays, or,
// for host containers, a device copy of the matrix plus that handle.  Calls
// of the non-const index and value accessors mark the cache stale (the
// analysis of a device container holds copies of values: the x-tiled long
// rows, a HYB's merged copy); resize detaches it.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

#include "Morpheus_AMD_Spaces.hpp"

namespace Morpheus {

// core/src/Morpheus_MatrixOptions.hpp:29-41
typedef enum {
  MATSTR_NONE = 0,
  MATSTR_SYMMETRIC = 1,
  MATSTR_STRUCTURALLY_SYMMETRIC = 2,
  MATSTR_HERMITIAN = 3,
  MATSTR_SPD = 4
} MatrixStructure;
typedef enum { MATOPT_NONE = 0, MATOPT_SHORT_ROWS = 1 } MatrixOptions;

namespace Impl {

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw RuntimeException(std::string(what) + ": " + hipGetErrorString(e));
}

// core/src/impl/Dia/Morpheus_Utils_Impl.hpp:33-36
template <typename T>
inline T get_pad_size(T len, T alignment) {
  return alignment * ((len + alignment - 1) / alignment);
}

// ------------------------------------------------------------ properties
template <class T>
struct is_layout : std::false_type {};
template <>
struct is_layout<LayoutLeft> : std::true_type {};
template <>
struct is_layout<LayoutRight> : std::true_type {};

template <class T, class = void>
struct is_space : std::false_type {};
template <class T>
struct is_space<T, std::void_t<decltype(T::is_device)>> : std::true_type {};

template <class Index, class Layout, class Space, class... P>
struct parse_properties {
  using index_type = Index;
  using array_layout = Layout;
  using space = Space;
};
template <class Index, class Layout, class Space, class H, class... P>
struct parse_properties<Index, Layout, Space, H, P...>
    : parse_properties<std::conditional_t<std::is_integral_v<H>, H, Index>,
                       std::conditional_t<is_layout<H>::value, H, Layout>,
                       std::conditional_t<is_space<H>::value, H, Space>, P...> {
  static_assert(std::is_integral_v<H> || is_layout<H>::value || is_space<H>::value,
                "container property must be an index type, a layout or a space");
};

template <class M>
struct default_exec;
template <>
struct default_exec<HostSpace> {
  using type = Serial;
};
template <>
struct default_exec<HIPSpace> {
  using type = HIP;
};

// core/src/Morpheus_ContainerTraits.hpp:59-82: index int, space
// DefaultExecutionSpace (HIP here), layout the execution space's default.
template <class V, class... P>
struct ContainerTraits {
  using parsed = parse_properties<void, void, void, P...>;
  using value_type = V;
  using non_const_value_type = std::remove_const_t<V>;
  using index_type = std::conditional_t<std::is_void_v<typename parsed::index_type>, int,
                                        typename parsed::index_type>;
  using non_const_index_type = std::remove_const_t<index_type>;
  using size_type = size_t;
  using space_given = std::conditional_t<std::is_void_v<typename parsed::space>,
                                         DefaultExecutionSpace, typename parsed::space>;
  using memory_space = typename memory_space_of<space_given>::type;
  using execution_space =
      std::conditional_t<std::is_same_v<space_given, memory_space>,
                         typename default_exec<memory_space>::type, space_given>;
  using array_layout =
      std::conditional_t<std::is_void_v<typename parsed::array_layout>,
                         typename memory_space::default_layout, typename parsed::array_layout>;
  static constexpr bool is_device = memory_space::is_device;

  // (int64 indices reach the kernels narrowed to int32 -- every index and
  // size below 2^31 -- and come back widened)
  static_assert(std::is_same_v<non_const_index_type, int> ||
                    std::is_same_v<non_const_index_type, long long>,
                "IndexType must be int or long long");
  static_assert(std::is_arithmetic_v<non_const_value_type>,
                "ValueType must be an arithmetic type such as int or double");
};

// Value types of the C-ABI: double and float containers compute in their
// own type; int containers are converted as raw words (MPH_I32) and
// multiplied exactly through FP64 (staged values and vectors).
template <class V>
struct value_code {
  static_assert(sizeof(V) == 0,
                "Morpheus multiply / convert on the MI355X take double, float or int values");
};
template <>
struct value_code<int> {
  static constexpr int value = MPH_I32;
};
template <>
struct value_code<double> {
  static constexpr int value = MPH_F64;
};
template <>
struct value_code<float> {
  static constexpr int value = MPH_F32;
};

// Storage layout of a container's 2-D arrays in a memory space.
template <class Mem, class Layout>
using storage_layout_t = std::conditional_t<Mem::is_device, LayoutLeft, Layout>;

// Two containers can share storage (shallow copy) when value type, index
// type, memory space and storage layout agree (on HIPSpace every declared
// layout stores column-major, so device containers differing only in the
// layout tag share storage).
template <class A, class B>
inline constexpr bool is_compatible_v =
    std::is_same_v<typename A::non_const_value_type, typename B::non_const_value_type> &&
    std::is_same_v<typename A::non_const_index_type, typename B::non_const_index_type> &&
    std::is_same_v<typename A::memory_space, typename B::memory_space> &&
    std::is_same_v<storage_layout_t<typename A::memory_space, typename A::array_layout>,
                   storage_layout_t<typename B::memory_space, typename B::array_layout>>;

// ------------------------------------------------------------ storage
// Reference-counted 1-D array in host memory or HBM (the role of a
// Kokkos::View in the reference).  Zero-initialised; resize keeps contents.
template <class T, class Mem>
class Array {
 public:
  using value_type = T;
  using memory_space = Mem;

  Array() = default;
  explicit Array(size_t n) { allocate_(n); }
  // unmanaged: wraps caller memory, never freed here
  Array(T* user, size_t n) : ptr_(user, [](T*) {}), n_(n) {}

  size_t size() const { return n_; }
  size_t extent(int) const { return n_; }
  T* data() const { return ptr_.get(); }
  long use_count() const { return ptr_.use_count(); }

  T& operator()(size_t i) const {
    static_assert(!Mem::is_device, "element access of a HIPSpace array from host code");
    return ptr_.get()[i];
  }
  T& operator[](size_t i) const { return (*this)(i); }

  void resize(size_t n) {
    if (n == n_ && ptr_) return;
    Array next(n);
    const size_t keep = std::min(n, n_);
    if (keep) copy_bytes(next.data(), data(), keep * sizeof(T), Mem::is_device, Mem::is_device);
    *this = next;
  }

  static void copy_bytes(void* dst, const void* src, size_t bytes, bool dst_dev, bool src_dev) {
    if (!bytes) return;
    if (!dst_dev && !src_dev) {
      std::memcpy(dst, src, bytes);
      return;
    }
    const hipMemcpyKind kind = dst_dev ? (src_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)
                                       : hipMemcpyDeviceToHost;
    hip_check(hipMemcpy(dst, src, bytes, kind), "Morpheus::copy");
  }

 private:
  void allocate_(size_t n) {
    n_ = n;
    if (n == 0) {
      ptr_.reset();
      return;
    }
    if constexpr (Mem::is_device) {
      void* p = nullptr;
      hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
      hip_check(hipMemset(p, 0, n * sizeof(T)), "hipMemset");
      ptr_ = std::shared_ptr<T>(static_cast<T*>(p), [](T* q) { (void)hipFree(q); });
    } else {
      ptr_ = std::shared_ptr<T>(new T[n](), std::default_delete<T[]>());
    }
  }

  std::shared_ptr<T> ptr_;
  size_t n_ = 0;
};

// Analysis cache shared by shallow copies of one sparse container.
struct PlanCache {
  mph_matrix_t handle = nullptr;
  std::vector<const void*> signature;
  bool stale = true;
  std::shared_ptr<void> device_copy;  // host containers: the staged device matrix
  // int64-indexed / int-valued containers: the int32 indices and FP64 values
  // the handle is built over; refresh() recasts the values before every use
  // of the handle, write_back() casts them into the container after an
  // operation that writes values
  std::shared_ptr<void> stage;
  std::function<void()> refresh, write_back;
  ~PlanCache() {
    if (handle) mph_matrix_destroy(handle);
  }
  void reset() {
    if (handle) mph_matrix_destroy(handle);
    handle = nullptr;
    device_copy.reset();
    stage.reset();
    refresh = nullptr;
    write_back = nullptr;
    stale = true;
  }
};

// core/src/Morpheus_MatrixBase.hpp:68-188
class MatrixBase {
 public:
  using size_type = size_t;
  MatrixBase() = default;
  MatrixBase(size_type rows, size_type cols, size_type entries = 0)
      : _m(rows), _n(cols), _nnz(entries) {}

  void resize(size_type rows, size_type cols, size_type entries) {
    _m = rows;
    _n = cols;
    _nnz = entries;
  }
  size_type nrows() const { return _m; }
  size_type ncols() const { return _n; }
  size_type nnnz() const { return _nnz; }
  void set_nrows(size_type rows) { _m = rows; }
  void set_ncols(size_type cols) { _n = cols; }
  void set_nnnz(size_type nnz) { _nnz = nnz; }
  MatrixStructure structure() const { return _structure; }
  MatrixOptions options() const { return _options; }
  void set_structure(MatrixStructure op) { _structure = op; }
  void set_options(MatrixOptions op) { _options = op; }

  // analysis cache (shared by shallow copies; detached by resize)
  PlanCache& plan_cache() const { return *_plan; }
  void touch() const { _plan->stale = true; }
  void detach_plan() { _plan = std::make_shared<PlanCache>(); }

 private:
  size_type _m = 0, _n = 0, _nnz = 0;
  MatrixStructure _structure = MATSTR_NONE;
  MatrixOptions _options = MATOPT_NONE;
  std::shared_ptr<PlanCache> _plan = std::make_shared<PlanCache>();
};

}  // namespace Impl

// ------------------------------------------------------------ random pool
// Seeded generator for DenseVector(n, pool, lo, hi) and assign(n, pool, lo,
// hi), the role of Kokkos::Random_XorShift64_Pool in the reference
// (core/src/Morpheus_DenseVector.hpp:191-196,271-277).  xorshift64* (Vigna,
// "An experimental exploration of Marsaglia's xorshift generators, scrambled",
// 2016); the stream differs from Kokkos' and nothing depends on it.
template <class Space = HostSpace>
class RandomPool {
 public:
  explicit RandomPool(uint64_t seed) : s_(seed ? seed : 0x9E3779B97F4A7C15ull) {}
  double next() {
    s_ ^= s_ >> 12;
    s_ ^= s_ << 25;
    s_ ^= s_ >> 27;
    const uint64_t r = s_ * 0x2545F4914F6CDD1Dull;
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
  }

 private:
  uint64_t s_;
};

// ============================================================ DenseVector
template <class ValueType, class... Properties>
class DenseVector {
  using traits = Impl::ContainerTraits<ValueType, Properties...>;

 public:
  using value_type = typename traits::value_type;
  using non_const_value_type = typename traits::non_const_value_type;
  using index_type = typename traits::index_type;
  using non_const_index_type = typename traits::non_const_index_type;
  using size_type = typename traits::size_type;
  using array_layout = typename traits::array_layout;
  using memory_space = typename traits::memory_space;
  using execution_space = typename traits::execution_space;
  using value_array_type = Impl::Array<non_const_value_type, memory_space>;
  using value_array_pointer = non_const_value_type*;
  using value_array_reference = non_const_value_type&;
  using type = DenseVector;
  using HostMirror = DenseVector<non_const_value_type, non_const_index_type, array_layout, HostSpace>;
  static constexpr formats_e format_id = NFORMATS;  // not a sparse format

  DenseVector() = default;
  DenseVector(const DenseVector&) = default;
  DenseVector(DenseVector&&) = default;
  DenseVector& operator=(const DenseVector&) = default;
  DenseVector& operator=(DenseVector&&) = default;

  // core/src/Morpheus_DenseVector.hpp:158-161
  DenseVector(const size_type n, value_type val = 0) : _size(n), _values(n) { assign(n, val); }

  // unmanaged wrap of user memory (core/src/Morpheus_DenseVector.hpp:171-179)
  template <typename ValuePtr>
  DenseVector(const size_type n, ValuePtr ptr,
              std::enable_if_t<std::is_pointer_v<ValuePtr>>* = nullptr)
      : _size(n), _values(ptr, n) {
    static_assert(std::is_same_v<value_array_pointer, ValuePtr>,
                  "Constructing DenseVector to wrap user memory must supply matching pointer type");
  }

  template <class Space>
  DenseVector(const size_type n, RandomPool<Space> pool, const value_type lo, const value_type hi)
      : _size(n), _values(n) {
    assign(n, pool, lo, hi);
  }

  // shallow copy from a compatible vector (same value/index/layout/space)
  template <class VR, class... PR,
            class = std::enable_if_t<
                std::is_same_v<typename DenseVector<VR, PR...>::memory_space, memory_space> &&
                std::is_same_v<typename DenseVector<VR, PR...>::non_const_value_type,
                               non_const_value_type> &&
                !std::is_same_v<DenseVector<VR, PR...>, DenseVector>>>
  DenseVector(const DenseVector<VR, PR...>& src) : _size(src.size()), _values(src.const_view()) {}

  template <class VR, class... PR>
  DenseVector& allocate(const DenseVector<VR, PR...>& src) {
    resize(src.size());
    return *this;
  }

  // core/src/Morpheus_DenseVector.hpp:253-259
  void assign(const size_type n, const value_type val) {
    if (n > _size) resize(n);
    if constexpr (memory_space::is_device) {
      std::vector<non_const_value_type> h(n, val);
      value_array_type::copy_bytes(_values.data(), h.data(), n * sizeof(value_type), true, false);
    } else {
      std::fill(_values.data(), _values.data() + n, val);
    }
  }

  template <class Space>
  void assign(const size_type n, RandomPool<Space>& pool, const value_type lo,
              const value_type hi) {
    if (n > _size) resize(n);
    std::vector<non_const_value_type> h(n);
    for (auto& v : h) v = (value_type)(lo + (hi - lo) * pool.next());
    value_array_type::copy_bytes(_values.data(), h.data(), n * sizeof(value_type),
                                 memory_space::is_device, false);
  }
  template <class Space>
  void assign(const size_type n, RandomPool<Space>&& pool, const value_type lo,
              const value_type hi) {
    assign(n, pool, lo, hi);
  }

  value_array_reference operator()(const size_type i) const { return _values(i); }
  value_array_reference operator[](const size_type i) const { return _values(i); }
  size_type size() const { return _size; }
  value_array_pointer data() const { return _values.data(); }
  value_array_type& view() { return _values; }
  const value_array_type& const_view() const { return _values; }

  void resize(const size_type n) {
    _values.resize(n);
    _size = n;
  }
  void resize(const size_type n, const value_type val) {
    resize(n);
    assign(n, val);
  }
  template <class VR, class... PR>
  void resize(const DenseVector<VR, PR...>& src) {
    resize(src.size());
  }

 private:
  size_type _size = 0;
  value_array_type _values;
};

// ============================================================ DenseMatrix
template <class ValueType, class... Properties>
class DenseMatrix {
  using traits = Impl::ContainerTraits<ValueType, Properties...>;

 public:
  using value_type = typename traits::value_type;
  using non_const_value_type = typename traits::non_const_value_type;
  using index_type = typename traits::index_type;
  using non_const_index_type = typename traits::non_const_index_type;
  using size_type = typename traits::size_type;
  using array_layout = typename traits::array_layout;
  using memory_space = typename traits::memory_space;
  using execution_space = typename traits::execution_space;
  using value_array_type = Impl::Array<non_const_value_type, memory_space>;
  using value_array_pointer = non_const_value_type*;
  using value_array_reference = non_const_value_type&;
  using HostMirror = DenseMatrix<non_const_value_type, non_const_index_type, array_layout, HostSpace>;
  static constexpr bool is_left = std::is_same_v<array_layout, LayoutLeft>;

  DenseMatrix() = default;
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<DenseMatrix<VR, PR...>, DenseMatrix> &&
                                     Impl::is_compatible_v<DenseMatrix, DenseMatrix<VR, PR...>>>>
  DenseMatrix(const DenseMatrix<VR, PR...>& src)
      : _m(src.nrows()), _n(src.ncols()), _values(src.const_view()) {}
  DenseMatrix(const size_type rows, const size_type cols, const value_type val = 0)
      : _m(rows), _n(cols), _values(rows * cols) {
    assign(rows, cols, val);
  }
  template <typename ValuePtr>
  DenseMatrix(const size_type rows, const size_type cols, ValuePtr ptr,
              std::enable_if_t<std::is_pointer_v<ValuePtr>>* = nullptr)
      : _m(rows), _n(cols), _values(ptr, rows * cols) {}

  template <class VR, class... PR>
  DenseMatrix& allocate(const DenseMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols());
    return *this;
  }

  void assign(const size_type rows, const size_type cols, const value_type val) {
    if (rows != _m || cols != _n) resize(rows, cols);
    const size_t n = rows * cols;
    if constexpr (memory_space::is_device) {
      std::vector<non_const_value_type> h(n, val);
      value_array_type::copy_bytes(_values.data(), h.data(), n * sizeof(value_type), true, false);
    } else {
      std::fill(_values.data(), _values.data() + n, val);
    }
  }

  // element (i, j); LayoutLeft: i + j*nrows, LayoutRight: i*ncols + j
  size_t offset(const size_type i, const size_type j) const {
    return is_left ? i + j * _m : i * _n + j;
  }
  value_array_reference operator()(const size_type i, const size_type j) const {
    return _values(offset(i, j));
  }
  size_type nrows() const { return _m; }
  size_type ncols() const { return _n; }
  size_type size() const { return _m * _n; }
  value_array_pointer data() const { return _values.data(); }
  value_array_type& view() { return _values; }
  const value_array_type& const_view() const { return _values; }

  // Layout-preserving resize: entries (i, j) inside both shapes keep their values.
  void resize(const size_type rows, const size_type cols) {
    if (rows == _m && cols == _n) return;
    value_array_type next(rows * cols);
    const size_type mr = std::min(rows, _m), mc = std::min(cols, _n);
    if (mr && mc) {
      std::vector<non_const_value_type> a(size()), b(rows * cols, value_type(0));
      value_array_type::copy_bytes(a.data(), data(), size() * sizeof(value_type), false,
                                   memory_space::is_device);
      for (size_type i = 0; i < mr; i++)
        for (size_type j = 0; j < mc; j++)
          b[is_left ? i + j * rows : i * cols + j] = a[offset(i, j)];
      value_array_type::copy_bytes(next.data(), b.data(), b.size() * sizeof(value_type),
                                   memory_space::is_device, false);
    }
    _values = next;
    _m = rows;
    _n = cols;
  }
  template <class VR, class... PR>
  void resize(const DenseMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols());
  }

 private:
  size_type _m = 0, _n = 0;
  value_array_type _values;
};

// Common type block of the sparse containers.
#define MORPHEUS_AMD_SPARSE_TYPES(NAME)                                                        \
  using traits = Impl::ContainerTraits<ValueType, Properties...>;                             \
                                                                                               \
 public:                                                                                       \
  using value_type = typename traits::value_type;                                             \
  using non_const_value_type = typename traits::non_const_value_type;                         \
  using index_type = typename traits::index_type;                                             \
  using non_const_index_type = typename traits::non_const_index_type;                         \
  using size_type = typename traits::size_type;                                               \
  using array_layout = typename traits::array_layout;                                         \
  using memory_space = typename traits::memory_space;                                         \
  using execution_space = typename traits::execution_space;                                   \
  using type = NAME;                                                                           \
  using HostMirror = NAME<non_const_value_type, non_const_index_type, array_layout, HostSpace>; \
  using index_array_type = DenseVector<non_const_index_type, non_const_index_type, array_layout, \
                                       memory_space>;                                          \
  using value_array_1d = DenseVector<non_const_value_type, non_const_index_type, array_layout, \
                                     memory_space>;                                            \
  using index_array_reference = non_const_index_type&;                                        \
  using value_array_reference = non_const_value_type&;                                        \
  template <class V2, class... P2>                                                             \
  using rebind = NAME<V2, P2...>;

// ============================================================ CooMatrix
template <class ValueType, class... Properties>
class CooMatrix : public Impl::MatrixBase {
  MORPHEUS_AMD_SPARSE_TYPES(CooMatrix)
  using value_array_type = value_array_1d;
  static constexpr formats_e format_id = COO_FORMAT;

  CooMatrix() = default;
  // shallow copy from a compatible container with other property spelling
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<CooMatrix<VR, PR...>, CooMatrix> &&
                                     Impl::is_compatible_v<CooMatrix, CooMatrix<VR, PR...>>>>
  CooMatrix(const CooMatrix<VR, PR...>& src)
      : MatrixBase(src.nrows(), src.ncols(), src.nnnz()),
        _row_indices(src.crow_indices()),
        _column_indices(src.ccolumn_indices()),
        _values(src.cvalues()) {}
  // core/src/Morpheus_CooMatrix.hpp:178-185
  CooMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries)
      : MatrixBase(num_rows, num_cols, num_entries),
        _row_indices(num_entries),
        _column_indices(num_entries),
        _values(num_entries) {}
  // unmanaged wrap of user pointers
  CooMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries,
            non_const_index_type* rind, non_const_index_type* cind, non_const_value_type* vals)
      : MatrixBase(num_rows, num_cols, num_entries),
        _row_indices(num_entries, rind),
        _column_indices(num_entries, cind),
        _values(num_entries, vals) {}
  CooMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries,
            const index_array_type& rind, const index_array_type& cind,
            const value_array_type& vals)
      : MatrixBase(num_rows, num_cols, num_entries),
        _row_indices(rind),
        _column_indices(cind),
        _values(vals) {}

  void resize(const size_type num_rows, const size_type num_cols, const size_type num_entries) {
    MatrixBase::resize(num_rows, num_cols, num_entries);
    _row_indices.resize(num_entries);
    _column_indices.resize(num_entries);
    _values.resize(num_entries);
    detach_plan();
  }
  template <class VR, class... PR>
  void resize(const CooMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz());
  }
  template <class VR, class... PR>
  CooMatrix& allocate(const CooMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz());
    return *this;
  }

  formats_e format_enum() const { return COO_FORMAT; }
  int format_index() const { return static_cast<int>(COO_FORMAT); }

  index_array_reference row_indices(size_type n) {
    touch();
    return _row_indices(n);
  }
  index_array_reference column_indices(size_type n) {
    touch();
    return _column_indices(n);
  }
  value_array_reference values(size_type n) {
    touch();
    return _values(n);
  }
  const non_const_index_type& crow_indices(size_type n) const { return _row_indices(n); }
  const non_const_index_type& ccolumn_indices(size_type n) const { return _column_indices(n); }
  const non_const_value_type& cvalues(size_type n) const { return _values(n); }
  index_array_type& row_indices() {
    touch();
    return _row_indices;
  }
  index_array_type& column_indices() {
    touch();
    return _column_indices;
  }
  value_array_type& values() {
    touch();
    return _values;
  }
  const index_array_type& crow_indices() const { return _row_indices; }
  const index_array_type& ccolumn_indices() const { return _column_indices; }
  const value_array_type& cvalues() const { return _values; }

 private:
  index_array_type _row_indices, _column_indices;
  value_array_type _values;
};

// ============================================================ CsrMatrix
template <class ValueType, class... Properties>
class CsrMatrix : public Impl::MatrixBase {
  MORPHEUS_AMD_SPARSE_TYPES(CsrMatrix)
  using value_array_type = value_array_1d;
  static constexpr formats_e format_id = CSR_FORMAT;

  CsrMatrix() = default;
  // shallow copy from a compatible container with other property spelling
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<CsrMatrix<VR, PR...>, CsrMatrix> &&
                                     Impl::is_compatible_v<CsrMatrix, CsrMatrix<VR, PR...>>>>
  CsrMatrix(const CsrMatrix<VR, PR...>& src)
      : MatrixBase(src.nrows(), src.ncols(), src.nnnz()),
        _row_offsets(src.crow_offsets()),
        _column_indices(src.ccolumn_indices()),
        _values(src.cvalues()) {}
  // core/src/Morpheus_CsrMatrix.hpp:178-185
  CsrMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries)
      : MatrixBase(num_rows, num_cols, num_entries),
        _row_offsets(num_rows + 1),
        _column_indices(num_entries),
        _values(num_entries) {}
  // unmanaged wrap of user pointers (core/src/Morpheus_CsrMatrix.hpp:186-201)
  CsrMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries,
            non_const_index_type* roff, non_const_index_type* cind, non_const_value_type* vals)
      : MatrixBase(num_rows, num_cols, num_entries),
        _row_offsets(num_rows + 1, roff),
        _column_indices(num_entries, cind),
        _values(num_entries, vals) {}
  CsrMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries,
            const index_array_type& roff, const index_array_type& cind,
            const value_array_type& vals)
      : MatrixBase(num_rows, num_cols, num_entries),
        _row_offsets(roff),
        _column_indices(cind),
        _values(vals) {}

  // core/src/Morpheus_CsrMatrix.hpp:363-371
  void resize(const size_type num_rows, const size_type num_cols, const size_type num_entries) {
    MatrixBase::resize(num_rows, num_cols, num_entries);
    _row_offsets.resize(num_rows + 1);
    _column_indices.resize(num_entries);
    _values.resize(num_entries);
    detach_plan();
  }
  template <class VR, class... PR>
  void resize(const CsrMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz());
  }
  template <class VR, class... PR>
  CsrMatrix& allocate(const CsrMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz());
    return *this;
  }

  formats_e format_enum() const { return CSR_FORMAT; }
  int format_index() const { return static_cast<int>(CSR_FORMAT); }

  index_array_reference row_offsets(size_type n) {
    touch();
    return _row_offsets(n);
  }
  index_array_reference column_indices(size_type n) {
    touch();
    return _column_indices(n);
  }
  value_array_reference values(size_type n) {
    touch();
    return _values(n);
  }
  const non_const_index_type& crow_offsets(size_type n) const { return _row_offsets(n); }
  const non_const_index_type& ccolumn_indices(size_type n) const { return _column_indices(n); }
  const non_const_value_type& cvalues(size_type n) const { return _values(n); }
  index_array_type& row_offsets() {
    touch();
    return _row_offsets;
  }
  index_array_type& column_indices() {
    touch();
    return _column_indices;
  }
  value_array_type& values() {
    touch();
    return _values;
  }
  const index_array_type& crow_offsets() const { return _row_offsets; }
  const index_array_type& ccolumn_indices() const { return _column_indices; }
  const value_array_type& cvalues() const { return _values; }

 private:
  index_array_type _row_offsets, _column_indices;
  value_array_type _values;
};

// ============================================================ DiaMatrix
template <class ValueType, class... Properties>
class DiaMatrix : public Impl::MatrixBase {
  MORPHEUS_AMD_SPARSE_TYPES(DiaMatrix)
  using value_array_type =
      DenseMatrix<non_const_value_type, non_const_index_type,
                  Impl::storage_layout_t<memory_space, array_layout>, memory_space>;
  static constexpr formats_e format_id = DIA_FORMAT;

  DiaMatrix() = default;
  // shallow copy from a compatible container with other property spelling
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<DiaMatrix<VR, PR...>, DiaMatrix> &&
                                     Impl::is_compatible_v<DiaMatrix, DiaMatrix<VR, PR...>>>>
  DiaMatrix(const DiaMatrix<VR, PR...>& src)
      : MatrixBase(src.nrows(), src.ncols(), src.nnnz()),
        _ndiags(src.ndiags()),
        _alignment(src.alignment()),
        _diagonal_offsets(src.cdiagonal_offsets()),
        _values(src.cvalues()) {}
  // core/src/Morpheus_DiaMatrix.hpp:192-201
  DiaMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries,
            const size_type num_diagonals, const size_type alignment = 32)
      : MatrixBase(num_rows, num_cols, num_entries),
        _ndiags(num_diagonals),
        _alignment(alignment),
        _diagonal_offsets(num_diagonals) {
    _values.resize(Impl::get_pad_size<size_type>(num_rows, alignment), num_diagonals);
  }

  // core/src/Morpheus_DiaMatrix.hpp:388-403 (throws on the fill tolerance)
  void resize(const size_type num_rows, const size_type num_cols, const size_type num_entries,
              const size_type num_diagonals, const size_type alignment = 32) {
    MatrixBase::resize(num_rows, num_cols, num_entries);
    _ndiags = num_diagonals;
    _alignment = alignment;
    if (mph_exceeds_tolerance((int64_t)num_rows, (int64_t)num_entries, (int64_t)_ndiags))
      throw FormatConversionException("DiaMatrix fill-in would exceed maximum tolerance");
    _diagonal_offsets.resize(num_diagonals);
    _values.resize(Impl::get_pad_size<size_type>(num_rows, alignment), num_diagonals);
    detach_plan();
  }
  template <class VR, class... PR>
  void resize(const DiaMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz(), src.ndiags(), src.alignment());
  }
  template <class VR, class... PR>
  DiaMatrix& allocate(const DiaMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz(), src.ndiags(), src.alignment());
    return *this;
  }

  formats_e format_enum() const { return DIA_FORMAT; }
  int format_index() const { return static_cast<int>(DIA_FORMAT); }

  index_array_reference diagonal_offsets(size_type n) {
    touch();
    return _diagonal_offsets(n);
  }
  // values(i, d) = A(i, i + offset[d])
  value_array_reference values(size_type i, size_type j) {
    touch();
    return _values(i, j);
  }
  const non_const_index_type& cdiagonal_offsets(size_type n) const { return _diagonal_offsets(n); }
  const non_const_value_type& cvalues(size_type i, size_type j) const { return _values(i, j); }
  index_array_type& diagonal_offsets() {
    touch();
    return _diagonal_offsets;
  }
  value_array_type& values() {
    touch();
    return _values;
  }
  const index_array_type& cdiagonal_offsets() const { return _diagonal_offsets; }
  const value_array_type& cvalues() const { return _values; }

  size_type ndiags() const { return _ndiags; }
  size_type alignment() const { return _alignment; }
  void set_ndiags(const size_type num_diagonals) { _ndiags = num_diagonals; }
  void set_alignment(const size_type alignment) { _alignment = alignment; }

 private:
  size_type _ndiags = 0, _alignment = 32;
  index_array_type _diagonal_offsets;
  value_array_type _values;
};

// ============================================================ EllMatrix
template <class ValueType, class... Properties>
class EllMatrix : public Impl::MatrixBase {
  MORPHEUS_AMD_SPARSE_TYPES(EllMatrix)
  using value_array_type =
      DenseMatrix<non_const_value_type, non_const_index_type,
                  Impl::storage_layout_t<memory_space, array_layout>, memory_space>;
  using index_matrix_type =
      DenseMatrix<non_const_index_type, non_const_index_type,
                  Impl::storage_layout_t<memory_space, array_layout>, memory_space>;
  static constexpr formats_e format_id = ELL_FORMAT;

  EllMatrix() = default;
  // shallow copy from a compatible container with other property spelling
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<EllMatrix<VR, PR...>, EllMatrix> &&
                                     Impl::is_compatible_v<EllMatrix, EllMatrix<VR, PR...>>>>
  EllMatrix(const EllMatrix<VR, PR...>& src)
      : MatrixBase(src.nrows(), src.ncols(), src.nnnz()),
        _entries_per_row(src.entries_per_row()),
        _alignment(src.alignment()),
        _column_indices(src.ccolumn_indices()),
        _values(src.cvalues()) {}
  // core/src/Morpheus_EllMatrix.hpp:189-202
  EllMatrix(const size_type num_rows, const size_type num_cols, const size_type num_entries,
            const size_type num_entries_per_row, const size_type alignment = 32)
      : MatrixBase(num_rows, num_cols, num_entries),
        _entries_per_row(num_entries_per_row),
        _alignment(alignment) {
    const size_type w = Impl::get_pad_size<size_type>(num_entries_per_row, alignment);
    _column_indices.resize(num_rows, w);
    _values.resize(num_rows, w);
  }

  // core/src/Morpheus_EllMatrix.hpp:372-385
  void resize(const size_type num_rows, const size_type num_cols, const size_type num_entries,
              const size_type num_entries_per_row, const size_type alignment = 32) {
    MatrixBase::resize(num_rows, num_cols, num_entries);
    _entries_per_row = num_entries_per_row;
    _alignment = alignment;
    const size_type w = Impl::get_pad_size<size_type>(num_entries_per_row, alignment);
    _column_indices.resize(num_rows, w);
    _values.resize(num_rows, w);
    detach_plan();
  }
  template <class VR, class... PR>
  void resize(const EllMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz(), src.entries_per_row(), src.alignment());
  }
  template <class VR, class... PR>
  EllMatrix& allocate(const EllMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.nnnz(), src.entries_per_row(), src.alignment());
    return *this;
  }

  formats_e format_enum() const { return ELL_FORMAT; }
  int format_index() const { return static_cast<int>(ELL_FORMAT); }

  index_array_reference column_indices(size_type i, size_type j) {
    touch();
    return _column_indices(i, j);
  }
  value_array_reference values(size_type i, size_type j) {
    touch();
    return _values(i, j);
  }
  const non_const_index_type& ccolumn_indices(size_type i, size_type j) const {
    return _column_indices(i, j);
  }
  const non_const_value_type& cvalues(size_type i, size_type j) const { return _values(i, j); }
  index_matrix_type& column_indices() {
    touch();
    return _column_indices;
  }
  value_array_type& values() {
    touch();
    return _values;
  }
  const index_matrix_type& ccolumn_indices() const { return _column_indices; }
  const value_array_type& cvalues() const { return _values; }

  size_type entries_per_row() const { return _entries_per_row; }
  size_type alignment() const { return _alignment; }
  index_type invalid_index() const { return -1; }
  void set_entries_per_row(const size_type n) { _entries_per_row = n; }
  void set_alignment(const size_type alignment) { _alignment = alignment; }

 private:
  size_type _entries_per_row = 0, _alignment = 32;
  index_matrix_type _column_indices;
  value_array_type _values;
};

// ============================================================ HybMatrix
template <class ValueType, class... Properties>
class HybMatrix : public Impl::MatrixBase {
  MORPHEUS_AMD_SPARSE_TYPES(HybMatrix)
  using ell_matrix_type = EllMatrix<ValueType, Properties...>;
  using coo_matrix_type = CooMatrix<ValueType, Properties...>;
  static constexpr formats_e format_id = HYB_FORMAT;

  HybMatrix() = default;
  // shallow copy from a compatible container with other property spelling
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<HybMatrix<VR, PR...>, HybMatrix> &&
                                     Impl::is_compatible_v<HybMatrix, HybMatrix<VR, PR...>>>>
  HybMatrix(const HybMatrix<VR, PR...>& src)
      : MatrixBase(src.nrows(), src.ncols(), src.nnnz()),
        _ell(src.cell()),
        _coo(src.ccoo()),
        _alignment(src.alignment()) {}
  // core/src/Morpheus_HybMatrix.hpp:182-198 (K defaults to nnz / nrows)
  HybMatrix(const size_type num_rows, const size_type num_cols, const size_type num_ell_entries,
            const size_type num_coo_entries, const size_type num_entries_per_row = 0,
            const size_type alignment = 32)
      : MatrixBase(num_rows, num_cols, num_ell_entries + num_coo_entries),
        _coo(num_rows, num_cols, num_coo_entries),
        _alignment(alignment) {
    size_type k = num_entries_per_row;
    if (k == 0) k = num_rows ? (num_ell_entries + num_coo_entries) / num_rows : 0;
    _ell.resize(num_rows, num_cols, num_ell_entries, k, alignment);
  }
  HybMatrix(const ell_matrix_type& ell, const coo_matrix_type& coo)
      : MatrixBase(ell.nrows(), ell.ncols(), ell.nnnz() + coo.nnnz()),
        _ell(ell),
        _coo(coo),
        _alignment(ell.alignment()) {}

  // core/src/Morpheus_HybMatrix.hpp:359-376
  void resize(const size_type num_rows, const size_type num_cols, const size_type num_ell_entries,
              const size_type num_coo_entries, const size_type num_entries_per_row = 0,
              const size_type alignment = 32) {
    MatrixBase::resize(num_rows, num_cols, num_ell_entries + num_coo_entries);
    size_type k = num_entries_per_row;
    if (k == 0) k = num_rows ? (num_ell_entries + num_coo_entries) / num_rows : 0;
    _ell.resize(num_rows, num_cols, num_ell_entries, k, alignment);
    _coo.resize(num_rows, num_cols, num_coo_entries);
    _alignment = alignment;
    detach_plan();
  }
  template <class VR, class... PR>
  void resize(const HybMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.cell().nnnz(), src.ccoo().nnnz(),
           src.cell().entries_per_row(), src.alignment());
  }
  template <class VR, class... PR>
  HybMatrix& allocate(const HybMatrix<VR, PR...>& src) {
    resize(src);
    return *this;
  }

  formats_e format_enum() const { return HYB_FORMAT; }
  int format_index() const { return static_cast<int>(HYB_FORMAT); }

  ell_matrix_type& ell() {
    touch();
    return _ell;
  }
  coo_matrix_type& coo() {
    touch();
    return _coo;
  }
  const ell_matrix_type& cell() const { return _ell; }
  const coo_matrix_type& ccoo() const { return _coo; }
  size_type alignment() const { return _alignment; }
  void set_alignment(const size_type alignment) { _alignment = alignment; }

 private:
  ell_matrix_type _ell;
  coo_matrix_type _coo;
  size_type _alignment = 32;
};

// ============================================================ HdcMatrix
template <class ValueType, class... Properties>
class HdcMatrix : public Impl::MatrixBase {
  MORPHEUS_AMD_SPARSE_TYPES(HdcMatrix)
  using dia_matrix_type = DiaMatrix<ValueType, Properties...>;
  using csr_matrix_type = CsrMatrix<ValueType, Properties...>;
  static constexpr formats_e format_id = HDC_FORMAT;

  HdcMatrix() = default;
  // shallow copy from a compatible container with other property spelling
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<HdcMatrix<VR, PR...>, HdcMatrix> &&
                                     Impl::is_compatible_v<HdcMatrix, HdcMatrix<VR, PR...>>>>
  HdcMatrix(const HdcMatrix<VR, PR...>& src)
      : MatrixBase(src.nrows(), src.ncols(), src.nnnz()),
        _dia(src.cdia()),
        _csr(src.ccsr()),
        _alignment(src.alignment()) {}
  // core/src/Morpheus_HdcMatrix.hpp:189-198
  HdcMatrix(const size_type num_rows, const size_type num_cols, const size_type num_dia_entries,
            const size_type num_csr_entries, const size_type num_diagonals,
            const size_type alignment = 32)
      : MatrixBase(num_rows, num_cols, num_dia_entries + num_csr_entries),
        _dia(num_rows, num_cols, num_dia_entries, num_diagonals, alignment),
        _csr(num_rows, num_cols, num_csr_entries),
        _alignment(alignment) {}
  HdcMatrix(const dia_matrix_type& dia, const csr_matrix_type& csr)
      : MatrixBase(dia.nrows(), dia.ncols(), dia.nnnz() + csr.nnnz()),
        _dia(dia),
        _csr(csr),
        _alignment(dia.alignment()) {}

  // core/src/Morpheus_HdcMatrix.hpp:356-367
  void resize(const size_type num_rows, const size_type num_cols, const size_type num_dia_entries,
              const size_type num_csr_entries, const size_type num_diagonals,
              const size_type alignment = 32) {
    MatrixBase::resize(num_rows, num_cols, num_dia_entries + num_csr_entries);
    _dia.resize(num_rows, num_cols, num_dia_entries, num_diagonals, alignment);
    _csr.resize(num_rows, num_cols, num_csr_entries);
    _alignment = alignment;
    detach_plan();
  }
  template <class VR, class... PR>
  void resize(const HdcMatrix<VR, PR...>& src) {
    resize(src.nrows(), src.ncols(), src.cdia().nnnz(), src.ccsr().nnnz(), src.cdia().ndiags(),
           src.alignment());
  }
  template <class VR, class... PR>
  HdcMatrix& allocate(const HdcMatrix<VR, PR...>& src) {
    resize(src);
    return *this;
  }

  formats_e format_enum() const { return HDC_FORMAT; }
  int format_index() const { return static_cast<int>(HDC_FORMAT); }

  dia_matrix_type& dia() {
    touch();
    return _dia;
  }
  csr_matrix_type& csr() {
    touch();
    return _csr;
  }
  const dia_matrix_type& cdia() const { return _dia; }
  const csr_matrix_type& ccsr() const { return _csr; }
  size_type alignment() const { return _alignment; }
  void set_alignment(const size_type alignment) { _alignment = alignment; }

 private:
  dia_matrix_type _dia;
  csr_matrix_type _csr;
  size_type _alignment = 32;
};

#undef MORPHEUS_AMD_SPARSE_TYPES

// ============================================================ DynamicMatrix
enum conversion_error_e { DYNAMIC_TO_PROXY = 0, PROXY_TO_DYNAMIC, CONV_SUCCESS };

template <class ValueType, class... Properties>
class DynamicMatrix : public Impl::MatrixBase {
  using traits = Impl::ContainerTraits<ValueType, Properties...>;

 public:
  using value_type = typename traits::value_type;
  using non_const_value_type = typename traits::non_const_value_type;
  using index_type = typename traits::index_type;
  using non_const_index_type = typename traits::non_const_index_type;
  using size_type = typename traits::size_type;
  using array_layout = typename traits::array_layout;
  using memory_space = typename traits::memory_space;
  using execution_space = typename traits::execution_space;
  using HostMirror =
      DynamicMatrix<non_const_value_type, non_const_index_type, array_layout, HostSpace>;
  // core/src/Morpheus_FormatsRegistry.hpp:40-51 (order of formats_e)
  using variant_type =
      std::variant<CooMatrix<ValueType, Properties...>, CsrMatrix<ValueType, Properties...>,
                   DiaMatrix<ValueType, Properties...>, EllMatrix<ValueType, Properties...>,
                   HybMatrix<ValueType, Properties...>, HdcMatrix<ValueType, Properties...>>;
  using variant = variant_type;

  DynamicMatrix() = default;

  // shallow copy from a compatible DynamicMatrix with other property spelling
  // (core/src/Morpheus_DynamicMatrix.hpp:218-246)
  template <class VR, class... PR,
            class = std::enable_if_t<!std::is_same_v<DynamicMatrix<VR, PR...>, DynamicMatrix> &&
                                     Impl::is_compatible_v<DynamicMatrix, DynamicMatrix<VR, PR...>>>>
  DynamicMatrix(const DynamicMatrix<VR, PR...>& src) {
    set_nrows(src.nrows());
    set_ncols(src.ncols());
    set_nnnz(src.nnnz());
    std::visit(
        [&](const auto& s) {
          using S = std::decay_t<decltype(s)>;
          using D = typename S::template rebind<ValueType, Properties...>;
          _formats = D(s);
        },
        src.const_formats());
  }

  // shallow assignment of a concrete container
  // (core/src/Morpheus_DynamicMatrix.hpp:180-216)
  template <class Matrix, class = std::enable_if_t<
                              std::is_constructible_v<variant_type, const Matrix&> &&
                              !std::is_base_of_v<DynamicMatrix, Matrix>>>
  DynamicMatrix(const Matrix& src) {
    *this = src;
  }
  template <class Matrix>
  std::enable_if_t<std::is_constructible_v<variant_type, const Matrix&> &&
                       !std::is_base_of_v<DynamicMatrix, Matrix>,
                   DynamicMatrix&>
  operator=(const Matrix& src) {
    set_nrows(src.nrows());
    set_ncols(src.ncols());
    set_nnnz(src.nnnz());
    _formats = src;
    return *this;
  }

  template <class VR, class... PR>
  DynamicMatrix& allocate(const DynamicMatrix<VR, PR...>& src) {
    activate(src.active_index());
    set_nrows(src.nrows());
    set_ncols(src.ncols());
    set_nnnz(src.nnnz());
    std::visit(
        [&](const auto& s) {
          using S = std::decay_t<decltype(s)>;
          using D = typename S::template rebind<ValueType, Properties...>;
          std::get<D>(_formats).allocate(s);
        },
        src.const_formats());
    return *this;
  }
  template <class VR, class... PR>
  void resize(const DynamicMatrix<VR, PR...>& src) {
    allocate(src);
  }
  template <class Matrix>
  std::enable_if_t<std::is_constructible_v<variant_type, Matrix>> resize(const Matrix& src) {
    activate(src.format_enum());
    set_nrows(src.nrows());
    set_ncols(src.ncols());
    set_nnnz(src.nnnz());
    std::visit(
        [&](auto& d) {
          using D = std::decay_t<decltype(d)>;
          if constexpr (D::format_id == Matrix::format_id) d.resize(src);
        },
        _formats);
  }

  int active_index() const { return static_cast<int>(_formats.index()); }
  int format_index() const { return active_index(); }
  formats_e active_enum() const { return static_cast<formats_e>(_formats.index()); }
  formats_e format_enum() const { return active_enum(); }

  // core/src/Morpheus_DynamicMatrix.hpp:409-442 and
  // core/src/impl/Dynamic/Morpheus_DynamicMatrix_Impl.hpp:190-212: the same
  // index is a no-op; an index above the number of formats warns and keeps
  // the active one; any other index outside [0, NFORMATS) (NFORMATS itself
  // or a negative one) selects COO.  Switching resets the shape to 0 x 0.
  void activate(const formats_e index) {
    const int size = static_cast<int>(NFORMATS);
    const int idx = static_cast<int>(index);
    if (idx == active_index()) return;
    if (idx > size) {
      std::cout << "Warning: There are " << size << " available formats to switch to. "
                << "Selecting to switch to format with index " << idx
                << " will default to no change." << std::endl;
      return;
    }
    set_nrows(0);
    set_ncols(0);
    set_nnnz(0);
    switch (idx) {
      case CSR_FORMAT: _formats = std::variant_alternative_t<1, variant_type>(); break;
      case DIA_FORMAT: _formats = std::variant_alternative_t<2, variant_type>(); break;
      case ELL_FORMAT: _formats = std::variant_alternative_t<3, variant_type>(); break;
      case HYB_FORMAT: _formats = std::variant_alternative_t<4, variant_type>(); break;
      case HDC_FORMAT: _formats = std::variant_alternative_t<5, variant_type>(); break;
      default: _formats = std::variant_alternative_t<0, variant_type>(); break;
    }
  }
  void activate(const int index) { activate(static_cast<formats_e>(index)); }

  const variant_type& const_formats() const { return _formats; }
  variant_type& formats() { return _formats; }

 private:
  variant_type _formats;
};

// ------------------------------------------------------------ type traits
template <class T>
struct is_dynamic_matrix_format_container : std::false_type {};
template <class V, class... P>
struct is_dynamic_matrix_format_container<DynamicMatrix<V, P...>> : std::true_type {};
template <class T>
struct is_dense_vector_format_container : std::false_type {};
template <class V, class... P>
struct is_dense_vector_format_container<DenseVector<V, P...>> : std::true_type {};
template <class T>
struct is_dense_matrix_format_container : std::false_type {};
template <class V, class... P>
struct is_dense_matrix_format_container<DenseMatrix<V, P...>> : std::true_type {};
template <class T, class = void>
struct is_sparse_matrix_container : std::false_type {};
template <class T>
struct is_sparse_matrix_container<T, std::void_t<decltype(T::format_id)>>
    : std::bool_constant<!is_dense_vector_format_container<T>::value &&
                         (T::format_id != NFORMATS)> {};
template <class T>
inline constexpr bool is_dynamic_matrix_format_container_v =
    is_dynamic_matrix_format_container<T>::value;
template <class T>
inline constexpr bool is_dense_vector_format_container_v =
    is_dense_vector_format_container<T>::value;
template <class T>
inline constexpr bool is_sparse_matrix_container_v = is_sparse_matrix_container<T>::value;
template <class T>
inline constexpr bool is_dense_matrix_format_container_v =
    is_dense_matrix_format_container<T>::value;
// a concrete sparse container of one format (core/src/Morpheus_FormatTraits.hpp)
template <class T, int F, class = void>
struct has_format_id : std::false_type {};
template <class T, int F>
struct has_format_id<T, F, std::void_t<decltype(T::format_id)>>
    : std::bool_constant<is_sparse_matrix_container<T>::value && T::format_id == F> {};
template <class T>
inline constexpr bool is_coo_matrix_format_container_v = has_format_id<T, COO_FORMAT>::value;
template <class T>
inline constexpr bool is_csr_matrix_format_container_v = has_format_id<T, CSR_FORMAT>::value;
template <class T>
inline constexpr bool is_dia_matrix_format_container_v = has_format_id<T, DIA_FORMAT>::value;
template <class T>
inline constexpr bool is_ell_matrix_format_container_v = has_format_id<T, ELL_FORMAT>::value;
template <class T>
inline constexpr bool is_hyb_matrix_format_container_v = has_format_id<T, HYB_FORMAT>::value;
template <class T>
inline constexpr bool is_hdc_matrix_format_container_v = has_format_id<T, HDC_FORMAT>::value;

}  // namespace Morpheus
