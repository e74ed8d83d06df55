// C++ drop-in tests: the Morpheus API (containers, copy / mirrors,
// multiply, convert, DynamicMatrix) on the MI355X, written the way the
// reference tests are (core/tests/Test_Multiply.hpp,
// core/tests/Test_Multiply_Dynamic.hpp, core/tests/Serial/Test_Convert.hpp,
// core/tests/Serial/Test_Convert_Dynamic.hpp, core/tests/Test_Copy.hpp,
// core/tests/Test_MirrorContainers.hpp, core/tests/Test_DynamicMatrix.hpp).
//
// Expected containers come from the reference's 10x10 fixture
// (core/tests/utils/Macros_*Matrix.hpp reset_small_container), passed as a
// text file (argv[1]); larger cases are checked against the CPU oracle
// (oracle/mph_oracle.h).  Usage: test_dropin <fixture.txt> [filter]
#include <Morpheus_AMD.hpp>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <map>
#include <random>
#include <sstream>

extern "C" {
#include "mph_oracle.h"
}

// ------------------------------------------------------------ harness
static int g_failures = 0;
static int g_checks = 0;
#define EXPECT_TRUE(c)                                                         \
  do {                                                                         \
    g_checks++;                                                                \
    if (!(c)) {                                                                \
      g_failures++;                                                            \
      std::fprintf(stderr, "%s:%d: EXPECT_TRUE(%s) failed\n", __FILE__, __LINE__, #c); \
    }                                                                          \
  } while (0)
#define EXPECT_EQ(a, b)                                                         \
  do {                                                                          \
    g_checks++;                                                                 \
    if (!((a) == (b))) {                                                        \
      g_failures++;                                                             \
      std::ostringstream os_;                                                   \
      os_ << (a) << " != " << (b);                                              \
      std::fprintf(stderr, "%s:%d: EXPECT_EQ(%s, %s) failed: %s\n", __FILE__, __LINE__, #a, #b, \
                   os_.str().c_str());                                          \
    }                                                                           \
  } while (0)
#define EXPECT_THROW(stmt, E)                                                    \
  do {                                                                           \
    g_checks++;                                                                  \
    bool thrown_ = false;                                                        \
    try {                                                                        \
      stmt;                                                                      \
    } catch (const E&) {                                                         \
      thrown_ = true;                                                            \
    } catch (...) {                                                              \
    }                                                                            \
    if (!thrown_) {                                                              \
      g_failures++;                                                              \
      std::fprintf(stderr, "%s:%d: %s did not throw %s\n", __FILE__, __LINE__, #stmt, #E); \
    }                                                                            \
  } while (0)

struct TestCase {
  const char* name;
  std::function<void()> fn;
};
static std::vector<TestCase>& registry() {
  static std::vector<TestCase> r;
  return r;
}
struct Register {
  Register(const char* n, std::function<void()> f) { registry().push_back({n, std::move(f)}); }
};
#define TEST(name)                                  \
  static void name();                               \
  static Register reg_##name(#name, name);          \
  static void name()

// ------------------------------------------------------------ types
using value_type = double;
using Host = Morpheus::HostSpace;
using Dev = Morpheus::HIPSpace;
using HIPExec = Morpheus::HIP;
using Serial = Morpheus::Serial;

template <class S>
using Coo = Morpheus::CooMatrix<value_type, S>;
template <class S>
using Csr = Morpheus::CsrMatrix<value_type, S>;
template <class S>
using Dia = Morpheus::DiaMatrix<value_type, S>;
template <class S>
using Ell = Morpheus::EllMatrix<value_type, S>;
template <class S>
using Hyb = Morpheus::HybMatrix<value_type, S>;
template <class S>
using Hdc = Morpheus::HdcMatrix<value_type, S>;
template <class S>
using Vec = Morpheus::DenseVector<value_type, S>;
template <class S>
using Dyn = Morpheus::DynamicMatrix<value_type, S>;

// ------------------------------------------------------------ fixture
// Lines: "C <NAME> <value>" for constants, "<Fmt> <array> <i> [<j>] <value>".
struct Fixture {
  std::map<std::string, long> c;
  std::map<std::string, std::vector<std::pair<std::vector<long>, double>>> a;
};
static Fixture g_fx;

static void load_fixture(const char* path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(std::string("cannot open fixture ") + path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream is(line);
    std::string f, k;
    is >> f >> k;
    if (f.empty()) continue;
    if (f == "C") {
      long v;
      is >> v;
      g_fx.c[k] = v;
      continue;
    }
    std::vector<double> nums;
    double d;
    while (is >> d) nums.push_back(d);
    std::vector<long> idx;
    for (size_t i = 0; i + 1 < nums.size(); i++) idx.push_back((long)nums[i]);
    g_fx.a[f + "." + k].push_back({idx, nums.back()});
  }
}
static long C(const char* k) { return g_fx.c.at(k); }

template <class Vector>
static void fill_vec(Vector& v, const std::string& key) {
  for (auto& e : g_fx.a.at(key)) v(e.first[0]) = (typename Vector::value_type)e.second;
}
template <class Mat>
static void fill_mat(Mat& m, const std::string& key) {
  for (auto& e : g_fx.a.at(key)) m(e.first[0], e.first[1]) = (typename Mat::value_type)e.second;
}

static Coo<Host> small_coo() {
  Coo<Host> A(C("SMALL_MATRIX_NROWS"), C("SMALL_MATRIX_NCOLS"), C("SMALL_MATRIX_NNZ"));
  fill_vec(A.row_indices(), "Coo.row_indices");
  fill_vec(A.column_indices(), "Coo.column_indices");
  fill_vec(A.values(), "Coo.values");
  return A;
}
static Csr<Host> small_csr() {
  Csr<Host> A(C("SMALL_MATRIX_NROWS"), C("SMALL_MATRIX_NCOLS"), C("SMALL_MATRIX_NNZ"));
  fill_vec(A.row_offsets(), "Csr.row_offsets");
  fill_vec(A.column_indices(), "Csr.column_indices");
  fill_vec(A.values(), "Csr.values");
  return A;
}
static void fill_dia(Dia<Host>& A, const std::string& p) {
  fill_vec(A.diagonal_offsets(), p + ".diagonal_offsets");
  fill_mat(A.values(), p + ".values");
}
static Dia<Host> small_dia() {
  Dia<Host> A(C("SMALL_MATRIX_NROWS"), C("SMALL_MATRIX_NCOLS"), C("SMALL_MATRIX_NNZ"),
              C("SMALL_DIA_MATRIX_NDIAGS"), C("SMALL_MATRIX_ALIGNMENT"));
  fill_dia(A, "Dia");
  return A;
}
static void fill_ell(Ell<Host>& A, const std::string& p) {
  auto& ci = A.column_indices();
  for (size_t i = 0; i < ci.nrows(); i++)
    for (size_t j = 0; j < ci.ncols(); j++) ci(i, j) = A.invalid_index();
  fill_mat(ci, p + ".column_indices");
  fill_mat(A.values(), p + ".values");
}
static Ell<Host> small_ell() {
  Ell<Host> A(C("SMALL_MATRIX_NROWS"), C("SMALL_MATRIX_NCOLS"), C("SMALL_MATRIX_NNZ"),
              C("SMALL_ELL_ENTRIES_PER_ROW"), C("SMALL_MATRIX_ALIGNMENT"));
  fill_ell(A, "Ell");
  return A;
}
static Hyb<Host> small_hyb() {
  Hyb<Host> A(C("SMALL_MATRIX_NROWS"), C("SMALL_MATRIX_NCOLS"), C("SMALL_HYB_ELL_NNZ"),
              C("SMALL_HYB_COO_NNZ"), C("SMALL_HYB_ENTRIES_PER_ROW"), C("SMALL_MATRIX_ALIGNMENT"));
  fill_ell(A.ell(), "Hyb.ell");
  fill_vec(A.coo().row_indices(), "Hyb.coo.row_indices");
  fill_vec(A.coo().column_indices(), "Hyb.coo.column_indices");
  fill_vec(A.coo().values(), "Hyb.coo.values");
  return A;
}
static Hdc<Host> small_hdc() {
  Hdc<Host> A(C("SMALL_MATRIX_NROWS"), C("SMALL_MATRIX_NCOLS"), C("SMALL_HDC_DIA_NNZ"),
              C("SMALL_HDC_CSR_NNZ"), C("SMALL_HDC_DIA_NDIAG"), C("SMALL_MATRIX_ALIGNMENT"));
  fill_dia(A.dia(), "Hdc.dia");
  fill_vec(A.csr().row_offsets(), "Hdc.csr.row_offsets");
  fill_vec(A.csr().column_indices(), "Hdc.csr.column_indices");
  fill_vec(A.csr().values(), "Hdc.csr.values");
  return A;
}

// ------------------------------------------------------------ comparisons
template <class A, class B>
static bool same_array(const A& a, const B& b) {
  if (a.size() != b.size()) return false;
  return a.size() == 0 ||
         std::memcmp(a.data(), b.data(), a.size() * sizeof(typename A::value_type)) == 0;
}
template <class A, class B>
static bool same_dense(const A& a, const B& b) {
  if (a.nrows() != b.nrows() || a.ncols() != b.ncols()) return false;
  for (size_t i = 0; i < a.nrows(); i++)
    for (size_t j = 0; j < a.ncols(); j++)
      if (std::memcmp(&a(i, j), &b(i, j), sizeof(a(i, j))) != 0) return false;
  return true;
}

// exact equality of two host containers of one format
template <class A, class B>
static bool same_container(const A& a, const B& b) {
  if (a.nrows() != b.nrows() || a.ncols() != b.ncols() || a.nnnz() != b.nnnz()) return false;
  if constexpr (A::format_id == Morpheus::COO_FORMAT) {
    return same_array(a.crow_indices(), b.crow_indices()) &&
           same_array(a.ccolumn_indices(), b.ccolumn_indices()) &&
           same_array(a.cvalues(), b.cvalues());
  } else if constexpr (A::format_id == Morpheus::CSR_FORMAT) {
    return same_array(a.crow_offsets(), b.crow_offsets()) &&
           same_array(a.ccolumn_indices(), b.ccolumn_indices()) &&
           same_array(a.cvalues(), b.cvalues());
  } else if constexpr (A::format_id == Morpheus::DIA_FORMAT) {
    return a.ndiags() == b.ndiags() && same_array(a.cdiagonal_offsets(), b.cdiagonal_offsets()) &&
           same_dense(a.cvalues(), b.cvalues());
  } else if constexpr (A::format_id == Morpheus::ELL_FORMAT) {
    return a.entries_per_row() == b.entries_per_row() &&
           same_dense(a.ccolumn_indices(), b.ccolumn_indices()) &&
           same_dense(a.cvalues(), b.cvalues());
  } else if constexpr (A::format_id == Morpheus::HYB_FORMAT) {
    return same_container(a.cell(), b.cell()) && same_container(a.ccoo(), b.ccoo());
  } else {
    return same_container(a.cdia(), b.cdia()) && same_container(a.ccsr(), b.ccsr());
  }
}

// y_ref = A x in COO entry order (the Serial COO loop), and the per-row
// bound 4 * nnz_row * u * sum |a x| on |y - y_ref|.
struct Ref {
  std::vector<double> y, bound;
};
static Ref dense_ref(const Coo<Host>& A, const Vec<Host>& x, const std::vector<double>* y0) {
  Ref r;
  r.y.assign(A.nrows(), 0.0);
  r.bound.assign(A.nrows(), 0.0);
  std::vector<double> absrow(A.nrows(), 0.0);
  std::vector<int> cnt(A.nrows(), 0);
  if (y0) r.y = *y0;
  for (size_t n = 0; n < A.nnnz(); n++) {
    const int i = A.crow_indices(n), j = A.ccolumn_indices(n);
    r.y[i] += A.cvalues(n) * x(j);
    absrow[i] += std::fabs(A.cvalues(n) * x(j));
    cnt[i]++;
  }
  for (size_t i = 0; i < A.nrows(); i++) {
    const double base = y0 ? std::fabs((*y0)[i]) : 0.0;
    r.bound[i] = 4.0 * (cnt[i] + 1) * 0x1p-53 * (absrow[i] + base) + 1e-300;
  }
  return r;
}
static bool within(const Vec<Host>& y, const Ref& r) {
  for (size_t i = 0; i < y.size(); i++)
    if (!(std::fabs(y(i) - r.y[i]) <= r.bound[i])) {
      std::fprintf(stderr, "  row %zu: %.17g vs %.17g (bound %.3g)\n", i, y(i), r.y[i], r.bound[i]);
      return false;
    }
  return true;
}

static Vec<Host> random_x(size_t n, uint64_t seed) {
  Vec<Host> x(n, Kokkos::Random_XorShift64_Pool<Host>(seed), -1.0, 1.0);
  return x;
}

// every fixture container, by format
template <class F>
static void for_each_small(F&& f) {
  f(small_coo());
  f(small_csr());
  f(small_dia());
  f(small_ell());
  f(small_hyb());
  f(small_hdc());
}
template <int I>
static auto small_of() {
  if constexpr (I == 0) return small_coo();
  else if constexpr (I == 1) return small_csr();
  else if constexpr (I == 2) return small_dia();
  else if constexpr (I == 3) return small_ell();
  else if constexpr (I == 4) return small_hyb();
  else return small_hdc();
}

// ============================================================ containers
TEST(DenseVector_Construct) {
  Vec<Host> v(10, 3.5);
  EXPECT_EQ(v.size(), 10u);
  for (size_t i = 0; i < 10; i++) EXPECT_EQ(v(i), 3.5);
  Vec<Host> w = v;  // shallow
  w(3) = -1.0;
  EXPECT_EQ(v(3), -1.0);
  v.resize(20, 2.0);
  EXPECT_EQ(v.size(), 20u);
  EXPECT_EQ(v[19], 2.0);
  double buf[4] = {1, 2, 3, 4};
  Vec<Host> u(4, buf);  // unmanaged
  u(0) = 9;
  EXPECT_EQ(buf[0], 9.0);
  Vec<Dev> d(1000, 1.25);
  Vec<Host> h = Morpheus::create_mirror(d);
  EXPECT_EQ(h.size(), 1000u);
  Morpheus::copy(d, h);
  EXPECT_EQ(h(999), 1.25);
}

TEST(Sparse_ShallowCopyAndResize) {
  auto A = small_csr();
  Csr<Host> B = A;
  B.values(0) = 42.0;
  EXPECT_EQ(A.cvalues(0), 42.0);
  EXPECT_EQ(B.crow_offsets().data(), A.crow_offsets().data());
  Morpheus::CsrMatrix<double, int, Morpheus::LayoutRight, Morpheus::HostSpace> C2 = A;
  EXPECT_EQ(C2.cvalues().data(), A.cvalues().data());
  B.resize(20, 20, 64);
  EXPECT_EQ(B.nrows(), 20u);
  EXPECT_EQ(B.nnnz(), 64u);
  EXPECT_EQ(A.nrows(), 10u);
  EXPECT_EQ(A.format_enum(), Morpheus::CSR_FORMAT);
  EXPECT_EQ(A.format_index(), 1);
  auto D = small_dia();
  EXPECT_EQ(D.cvalues().nrows(), 32u);  // pad(10, 32) rows
  EXPECT_EQ(D.ndiags(), 9u);
  auto E = small_ell();
  EXPECT_EQ(E.cvalues().ncols(), 32u);  // pad(K, 32) columns
  EXPECT_EQ(E.invalid_index(), -1);
  // DIA fill-in tolerance (core/src/Morpheus_DiaMatrix.hpp:388-403,
  // core/src/impl/Morpheus_Utils.hpp:91-105): 1000 diagonals x 200000 rows
  // is 2e8 slots (> 1e8) for 100 entries (fill > 10); 50 x 20000 is only 1e6.
  Dia<Host> big;
  EXPECT_THROW(big.resize(200000, 200000, 100, 1000), Morpheus::FormatConversionException);
  big.resize(20000, 20000, 100, 50);
  EXPECT_EQ(big.cvalues().nrows(), 20000u);
}

TEST(DynamicMatrix_Activate) {
  Dyn<Host> A;
  EXPECT_EQ(A.active_index(), 0);
  A.activate(Morpheus::DIA_FORMAT);
  EXPECT_EQ(A.active_enum(), Morpheus::DIA_FORMAT);
  A.activate(3);
  EXPECT_EQ(A.format_index(), 3);
  // index == NFORMATS falls to COO; above it keeps the active format
  // (core/src/Morpheus_DynamicMatrix.hpp:409-442)
  A.activate(Morpheus::NFORMATS);
  EXPECT_EQ(A.active_index(), 0);
  A.activate(Morpheus::HDC_FORMAT);
  A.activate(7);
  EXPECT_EQ(A.active_index(), 5);
  A = small_hyb();
  EXPECT_EQ(A.active_index(), 4);
  EXPECT_EQ(A.nrows(), 10u);
  EXPECT_EQ(A.nnnz(), 32u);
  Dyn<Host> B(small_dia());
  EXPECT_EQ(B.active_enum(), Morpheus::DIA_FORMAT);
}

// ============================================================ copy / mirrors
TEST(Copy_MirrorRoundTrip) {
  for_each_small([](auto A) {
    using H = decltype(A);
    auto Ad = Morpheus::create_mirror_container<Dev>(A);
    static_assert(decltype(Ad)::memory_space::is_device, "device mirror");
    Morpheus::copy(A, Ad);
    auto Bh = Morpheus::create_mirror(Ad);
    static_assert(std::is_same_v<typename decltype(Bh)::memory_space, Host>, "host mirror");
    Morpheus::copy(Ad, Bh);
    EXPECT_TRUE(same_container(A, Bh));
    // a host mirror container of a host container is the container itself
    auto Ch = Morpheus::create_mirror_container(A);
    EXPECT_TRUE(same_container(A, Ch));
    (void)sizeof(H);
  });
  // copy checks shapes
  auto A = small_csr();
  Csr<Dev> wrong(10, 10, 31);
  EXPECT_THROW(Morpheus::copy(A, wrong), Morpheus::RuntimeException);
  // DynamicMatrix copies need equal active formats
  Dyn<Host> dh(small_csr());
  Dyn<Dev> dd;
  dd.activate(Morpheus::COO_FORMAT);
  EXPECT_THROW(Morpheus::copy(dh, dd), Morpheus::FormatConversionException);
  Dyn<Dev> dd2 = Morpheus::create_mirror_container<Dev>(dh);
  Morpheus::copy(dh, dd2);
  auto back = Morpheus::create_mirror(dd2);
  Morpheus::copy(dd2, back);
  EXPECT_TRUE(same_container(std::get<1>(back.const_formats()), small_csr()));
}

// ============================================================ multiply
template <class ExecSpace, class Space>
static void multiply_small_all() {
  const auto coo = small_coo();
  const auto x = random_x(coo.ncols(), 5374857);
  const Ref ref = dense_ref(coo, x, nullptr);
  for_each_small([&](auto Ah) {
    Vec<Host> yh(coo.nrows(), 7.0);
    if constexpr (Space::is_device) {
      auto Ad = Morpheus::create_mirror_container<Dev>(Ah);
      Morpheus::copy(Ah, Ad);
      Vec<Dev> xd(x.size()), yd(yh.size(), 7.0);
      Morpheus::copy(x, xd);
      Morpheus::multiply<ExecSpace>(Ad, xd, yd);
      Morpheus::copy(yd, yh);
    } else {
      Morpheus::multiply<ExecSpace>(Ah, x, yh);
    }
    if (!within(yh, ref)) {
      std::fprintf(stderr, "  format %d\n", Ah.format_index());
      EXPECT_TRUE(false);
    }
  });
}
TEST(Multiply_Small_HIP) { multiply_small_all<HIPExec, Dev>(); }
TEST(Multiply_Small_GenericHIP) { multiply_small_all<Morpheus::Generic::HIP, Dev>(); }
TEST(Multiply_Small_Serial) { multiply_small_all<Serial, Host>(); }
TEST(Multiply_Small_OpenMP) { multiply_small_all<Morpheus::OpenMP, Host>(); }

TEST(Multiply_Accumulate_And_Errors) {
  const auto coo = small_coo();
  const auto x = random_x(coo.ncols(), 11);
  std::vector<double> y0(coo.nrows());
  for (size_t i = 0; i < y0.size(); i++) y0[i] = 0.5 * (double)i - 2.0;
  const Ref ref = dense_ref(coo, x, &y0);
  for_each_small([&](auto Ah) {
    Vec<Host> yh(coo.nrows());
    for (size_t i = 0; i < y0.size(); i++) yh(i) = y0[i];
    Morpheus::multiply<Serial>(Ah, x, yh, false);
    EXPECT_TRUE(within(yh, ref));
  });
  auto A = small_csr();
  Vec<Host> xs(9), y(10);
  EXPECT_THROW(Morpheus::multiply<Serial>(A, xs, y), Morpheus::InvalidInputException);
}

TEST(Multiply_Dynamic) {
  const auto coo = small_coo();
  const auto x = random_x(coo.ncols(), 77);
  const Ref ref = dense_ref(coo, x, nullptr);
  for_each_small([&](auto Ah) {
    Dyn<Host> Dh(Ah);
    Dyn<Dev> Dd = Morpheus::create_mirror_container<Dev>(Dh);
    Morpheus::copy(Dh, Dd);
    Vec<Dev> xd(x.size()), yd(coo.nrows());
    Morpheus::copy(x, xd);
    Morpheus::multiply<HIPExec>(Dd, xd, yd);
    Vec<Host> yh(coo.nrows());
    Morpheus::copy(yd, yh);
    EXPECT_TRUE(within(yh, ref));
    Vec<Host> y2(coo.nrows());
    Morpheus::multiply<Serial>(Dh, x, y2);
    EXPECT_TRUE(within(y2, ref));
  });
}

TEST(Multiply_PlanInvalidation) {
  auto A = small_csr();
  const auto x = random_x(A.ncols(), 3);
  Vec<Host> y(A.nrows());
  Morpheus::multiply<Serial>(A, x, y);
  // change a value and the structure through the host accessors
  A.values(0) = 100.0;
  A.column_indices(0) = 2;
  Coo<Host> c;
  Morpheus::convert<Serial>(A, c);
  const Ref ref = dense_ref(c, x, nullptr);
  Morpheus::multiply<Serial>(A, x, y);
  EXPECT_TRUE(within(y, ref));
}

// A device container whose values are rewritten in place through values():
// the cached analysis (here the x-tiled copy of a 2000-entry row, x of
// 2.4 MB) is rebuilt, so the second product sees the new values.
TEST(Multiply_PlanInvalidation_DeviceValues) {
  const size_t nr = 4, nc = 300000, nlong = 2000;
  const size_t nnz = nlong + 3 * 5;
  Csr<Host> A(nr, nc, nnz);
  size_t k = 0;
  A.row_offsets(0) = 0;
  for (size_t r = 0; r < nr; r++) {
    const size_t len = r == 0 ? nlong : 5;
    for (size_t j = 0; j < len; j++, k++) {
      A.column_indices(k) = (int)(j * (nc / len) + r);
      A.values(k) = 1.0 + 0.001 * (double)k;
    }
    A.row_offsets(r + 1) = (int)k;
  }
  auto Ad = Morpheus::create_mirror_container<Dev>(A);
  Morpheus::copy(A, Ad);
  const auto x = random_x(nc, 9);
  Vec<Dev> xd(nc), yd(nr, 0.0);
  Morpheus::copy(x, xd);
  Morpheus::multiply<HIPExec>(Ad, xd, yd);
  for (size_t q = 0; q < nnz; q++) A.values(q) *= -2.0;
  // rewrite the device values in place
  const bool copied = hipMemcpy(Ad.values().data(), A.cvalues().data(), sizeof(double) * nnz,
                                hipMemcpyHostToDevice) == hipSuccess;
  EXPECT_TRUE(copied);
  Morpheus::multiply<HIPExec>(Ad, xd, yd);
  Vec<Host> y(nr, 0.0);
  Morpheus::copy(yd, y);
  Coo<Host> c;
  Morpheus::convert<Serial>(A, c);
  EXPECT_TRUE(within(y, dense_ref(c, x, nullptr)));
}

// Values written through a pointer kept from an accessor call made BEFORE
// the first product (no accessor call afterwards, so the container is never
// told): the analysis keeps no copy of any value, so the second product sees
// them.  CSR with a 2000-entry row and x of 2.4 MB (x-tiled by default
// through the C-ABI), and a 70,000-row HYB whose ELL slots are mostly
// padding (merged into a CSR copy by default).
TEST(Multiply_KeptPointer_DeviceValues) {
  const size_t nr = 4, nc = 300000, nlong = 2000;
  const size_t nnz = nlong + 3 * 5;
  Csr<Host> A(nr, nc, nnz);
  size_t k = 0;
  A.row_offsets(0) = 0;
  for (size_t r = 0; r < nr; r++) {
    const size_t len = r == 0 ? nlong : 5;
    for (size_t j = 0; j < len; j++, k++) {
      A.column_indices(k) = (int)(j * (nc / len) + r);
      A.values(k) = 1.0 + 0.001 * (double)k;
    }
    A.row_offsets(r + 1) = (int)k;
  }
  auto Ad = Morpheus::create_mirror_container<Dev>(A);
  Morpheus::copy(A, Ad);
  double* kept = Ad.values().data();  // kept across the products
  const auto x = random_x(nc, 9);
  Vec<Dev> xd(nc), yd(nr, 0.0);
  Morpheus::copy(x, xd);
  Morpheus::multiply<HIPExec>(Ad, xd, yd);
  for (size_t q = 0; q < nnz; q++) A.values(q) *= -2.0;
  const bool copied =
      hipMemcpy(kept, A.cvalues().data(), sizeof(double) * nnz, hipMemcpyHostToDevice) == hipSuccess;
  EXPECT_TRUE(copied);
  Morpheus::multiply<HIPExec>(Ad, xd, yd);
  Vec<Host> y(nr, 0.0);
  Morpheus::copy(yd, y);
  Coo<Host> c;
  Morpheus::convert<Serial>(A, c);
  EXPECT_TRUE(within(y, dense_ref(c, x, nullptr)));
}

TEST(Multiply_KeptPointer_HybValues) {
  // 70,000 rows: row r has 1 + (r % 97 == 0 ? 60 : 0) entries, so the ELL
  // part (entries_per_row from the reference's HYB split) is mostly padding
  const size_t nr = 70000, nc = 70000;
  std::vector<int> ri, ci;
  std::vector<double> v;
  for (size_t r = 0; r < nr; r++) {
    const size_t len = 1 + (r % 97 == 0 ? 60 : 0);
    for (size_t j = 0; j < len; j++) {
      ri.push_back((int)r);
      ci.push_back((int)((r + j * 1013) % nc));
      v.push_back(1.0 + 0.0001 * (double)(ri.size() % 1000));
    }
  }
  // rows sorted, columns sorted within each row
  Coo<Host> A(nr, nc, ri.size());
  {
    std::vector<size_t> order(ri.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      return ri[a] != ri[b] ? ri[a] < ri[b] : ci[a] < ci[b];
    });
    for (size_t i = 0; i < order.size(); i++) {
      A.row_indices(i) = ri[order[i]];
      A.column_indices(i) = ci[order[i]];
      A.values(i) = v[order[i]];
    }
  }
  Hyb<Host> H;
  Morpheus::convert<Serial>(A, H);
  auto Hd = Morpheus::create_mirror_container<Dev>(H);
  Morpheus::copy(H, Hd);
  double* ell_kept = Hd.ell().values().data();
  double* coo_kept = Hd.coo().values().data();
  const auto x = random_x(nc, 5);
  Vec<Dev> xd(nc), yd(nr, 0.0);
  Morpheus::copy(x, xd);
  Morpheus::multiply<HIPExec>(Hd, xd, yd);
  // scale every value by 3 on the host copy and write it through the kept pointers
  const size_t ell_n = H.cell().cvalues().nrows() * H.cell().cvalues().ncols();
  for (size_t i = 0; i < H.cell().cvalues().nrows(); i++)
    for (size_t j = 0; j < H.cell().cvalues().ncols(); j++) H.ell().values(i, j) *= 3.0;
  for (size_t q = 0; q < H.ccoo().nnnz(); q++) H.coo().values(q) *= 3.0;
  // (through a device mirror of the scaled matrix: the device layout of the
  // ELL values is the one the kept pointer addresses)
  auto Hs = Morpheus::create_mirror_container<Dev>(H);
  Morpheus::copy(H, Hs);
  bool ok = hipMemcpy(ell_kept, Hs.cell().cvalues().data(), sizeof(double) * ell_n,
                      hipMemcpyDeviceToDevice) == hipSuccess;
  ok = ok && (H.ccoo().nnnz() == 0 ||
              hipMemcpy(coo_kept, Hs.ccoo().cvalues().data(), sizeof(double) * H.ccoo().nnnz(),
                        hipMemcpyDeviceToDevice) == hipSuccess);
  EXPECT_TRUE(ok);
  Morpheus::multiply<HIPExec>(Hd, xd, yd);
  Vec<Host> y(nr, 0.0);
  Morpheus::copy(yd, y);
  Coo<Host> c;
  Morpheus::convert<Serial>(H, c);
  EXPECT_TRUE(within(y, dense_ref(c, x, nullptr)));
}

// A host container written through a reference kept from before the first
// product: the staged device copy takes the values on every call.
TEST(Multiply_KeptReference_HostValues) {
  auto A = small_csr();
  double& kept = A.values(0);
  const auto x = random_x(A.ncols(), 3);
  Vec<Host> y(A.nrows());
  Morpheus::multiply<Serial>(A, x, y);
  kept = 100.0;
  Coo<Host> c;
  Morpheus::convert<Serial>(A, c);
  const Ref ref = dense_ref(c, x, nullptr);
  Morpheus::multiply<Serial>(A, x, y);
  EXPECT_TRUE(within(y, ref));
}

// ============================================================ convert
template <int S, int D, class ExecSpace>
static void convert_pair() {
  auto src = small_of<S>();
  auto expect = small_of<D>();
  decltype(expect) out;
  if constexpr (ExecSpace::is_device) {
    auto sd = Morpheus::create_mirror_container<Dev>(src);
    Morpheus::copy(src, sd);
    decltype(Morpheus::create_mirror_container<Dev>(expect)) dd;
    Morpheus::convert<ExecSpace>(sd, dd);
    out.allocate(dd);
    Morpheus::copy(dd, out);
  } else {
    Morpheus::convert<ExecSpace>(src, out);
  }
  if (!same_container(out, expect)) {
    std::fprintf(stderr, "  convert %d -> %d differs from the fixture\n", S, D);
    EXPECT_TRUE(false);
  } else {
    EXPECT_TRUE(true);
  }
}
template <int S, class ExecSpace>
static void convert_row() {
  convert_pair<S, 0, ExecSpace>();
  convert_pair<S, 1, ExecSpace>();
  convert_pair<S, 2, ExecSpace>();
  convert_pair<S, 3, ExecSpace>();
  convert_pair<S, 4, ExecSpace>();
  convert_pair<S, 5, ExecSpace>();
}
template <class ExecSpace>
static void convert_all() {
  convert_row<0, ExecSpace>();
  convert_row<1, ExecSpace>();
  convert_row<2, ExecSpace>();
  convert_row<3, ExecSpace>();
  convert_row<4, ExecSpace>();
  convert_row<5, ExecSpace>();
}
TEST(Convert_Small_AllPairs_HIP) { convert_all<HIPExec>(); }
TEST(Convert_Small_AllPairs_Serial) { convert_all<Serial>(); }

// ------------------------------------------------------------ type lists
// The reference's convert / container type lists (core/tests/utils/
// Utils.hpp:50-60): values {double, int} x indices {long long, int}.  The
// fixture containers retyped element by element (the fixture's values
// 1.11 .. 32.32 become 1 .. 32 as int).
template <class V, class I, class M>
static auto retype(const M& A) {
  using Morpheus::HostSpace;
  if constexpr (M::format_id == Morpheus::COO_FORMAT) {
    Morpheus::CooMatrix<V, I, HostSpace> B(A.nrows(), A.ncols(), A.nnnz());
    for (size_t k = 0; k < A.nnnz(); k++) {
      B.row_indices(k) = (I)A.crow_indices(k);
      B.column_indices(k) = (I)A.ccolumn_indices(k);
      B.values(k) = (V)A.cvalues(k);
    }
    return B;
  } else if constexpr (M::format_id == Morpheus::CSR_FORMAT) {
    Morpheus::CsrMatrix<V, I, HostSpace> B(A.nrows(), A.ncols(), A.nnnz());
    for (size_t r = 0; r <= A.nrows(); r++) B.row_offsets(r) = (I)A.crow_offsets(r);
    for (size_t k = 0; k < A.nnnz(); k++) {
      B.column_indices(k) = (I)A.ccolumn_indices(k);
      B.values(k) = (V)A.cvalues(k);
    }
    return B;
  } else if constexpr (M::format_id == Morpheus::DIA_FORMAT) {
    Morpheus::DiaMatrix<V, I, HostSpace> B(A.nrows(), A.ncols(), A.nnnz(), A.ndiags(),
                                           A.alignment());
    for (size_t d = 0; d < A.ndiags(); d++) B.diagonal_offsets(d) = (I)A.cdiagonal_offsets(d);
    for (size_t i = 0; i < A.cvalues().nrows(); i++)
      for (size_t j = 0; j < A.cvalues().ncols(); j++) B.values(i, j) = (V)A.cvalues(i, j);
    return B;
  } else if constexpr (M::format_id == Morpheus::ELL_FORMAT) {
    Morpheus::EllMatrix<V, I, HostSpace> B(A.nrows(), A.ncols(), A.nnnz(), A.entries_per_row(),
                                           A.alignment());
    for (size_t i = 0; i < A.cvalues().nrows(); i++)
      for (size_t j = 0; j < A.cvalues().ncols(); j++) {
        B.column_indices(i, j) = (I)A.ccolumn_indices(i, j);
        B.values(i, j) = (V)A.cvalues(i, j);
      }
    return B;
  } else if constexpr (M::format_id == Morpheus::HYB_FORMAT) {
    Morpheus::HybMatrix<V, I, HostSpace> B(A.nrows(), A.ncols(), A.cell().nnnz(), A.ccoo().nnnz(),
                                           A.cell().entries_per_row(), A.alignment());
    B.ell() = retype<V, I>(A.cell());
    B.coo() = retype<V, I>(A.ccoo());
    return B;
  } else {
    Morpheus::HdcMatrix<V, I, HostSpace> B(A.nrows(), A.ncols(), A.cdia().nnnz(), A.ccsr().nnnz(),
                                           A.cdia().ndiags(), A.alignment());
    B.dia() = retype<V, I>(A.cdia());
    B.csr() = retype<V, I>(A.ccsr());
    return B;
  }
}

template <class V, class I, int S, int D, class ExecSpace>
static void convert_pair_typed() {
  auto src = retype<V, I>(small_of<S>());
  auto expect = retype<V, I>(small_of<D>());
  decltype(expect) out;
  if constexpr (ExecSpace::is_device) {
    auto sd = Morpheus::create_mirror_container<Dev>(src);
    Morpheus::copy(src, sd);
    decltype(Morpheus::create_mirror_container<Dev>(expect)) dd;
    Morpheus::convert<ExecSpace>(sd, dd);
    out.allocate(dd);
    Morpheus::copy(dd, out);
  } else {
    Morpheus::convert<ExecSpace>(src, out);
  }
  if (!same_container(out, expect)) {
    std::fprintf(stderr, "  convert %d -> %d (%zu-byte values, %zu-byte indices) differs\n", S,
                 D, sizeof(V), sizeof(I));
    EXPECT_TRUE(false);
  } else {
    EXPECT_TRUE(true);
  }
}
template <class V, class I, int S, class ExecSpace>
static void convert_row_typed() {
  convert_pair_typed<V, I, S, 0, ExecSpace>();
  convert_pair_typed<V, I, S, 1, ExecSpace>();
  convert_pair_typed<V, I, S, 2, ExecSpace>();
  convert_pair_typed<V, I, S, 3, ExecSpace>();
  convert_pair_typed<V, I, S, 4, ExecSpace>();
  convert_pair_typed<V, I, S, 5, ExecSpace>();
}
template <class V, class I, class ExecSpace>
static void convert_all_typed() {
  convert_row_typed<V, I, 0, ExecSpace>();
  convert_row_typed<V, I, 1, ExecSpace>();
  convert_row_typed<V, I, 2, ExecSpace>();
  convert_row_typed<V, I, 3, ExecSpace>();
  convert_row_typed<V, I, 4, ExecSpace>();
  convert_row_typed<V, I, 5, ExecSpace>();
}
TEST(Convert_Types_Double_LongLong_HIP) { convert_all_typed<double, long long, HIPExec>(); }
TEST(Convert_Types_Int_LongLong_HIP) { convert_all_typed<int, long long, HIPExec>(); }
TEST(Convert_Types_Int_Int_HIP) { convert_all_typed<int, int, HIPExec>(); }
TEST(Convert_Types_Int_LongLong_Serial) { convert_all_typed<int, long long, Serial>(); }

// Multiply over the same type list: int values are exact (the products of
// the fixture's 1 .. 32 with an integer x in [-8, 8] are small integers).
template <class V, class I>
static void multiply_typed() {
  const auto coo = small_coo();
  Morpheus::DenseVector<V, I, Host> x(coo.ncols());
  Vec<Host> xd_ref(coo.ncols());
  for (size_t j = 0; j < coo.ncols(); j++) {
    x(j) = (V)((int)(j * 7 % 17) - 8);
    xd_ref(j) = (double)x(j);
  }
  auto coo_t = retype<V, I>(coo);
  // the reference product of the retyped matrix, in double
  Coo<Host> coo_d(coo.nrows(), coo.ncols(), coo.nnnz());
  for (size_t k = 0; k < coo.nnnz(); k++) {
    coo_d.row_indices(k) = coo.crow_indices(k);
    coo_d.column_indices(k) = coo.ccolumn_indices(k);
    coo_d.values(k) = (double)coo_t.cvalues(k);
  }
  const Ref ref = dense_ref(coo_d, xd_ref, nullptr);
  auto check_one = [&](const auto& Ah) {
    auto Ad = Morpheus::create_mirror_container<Dev>(Ah);
    Morpheus::copy(Ah, Ad);
    Morpheus::DenseVector<V, I, Dev> xdev(x.size()), ydev(coo.nrows(), (V)7);
    Morpheus::copy(x, xdev);
    Morpheus::multiply<HIPExec>(Ad, xdev, ydev);
    Morpheus::DenseVector<V, I, Host> y(coo.nrows(), (V)0);
    Morpheus::copy(ydev, y);
    Vec<Host> yd(coo.nrows());
    for (size_t i = 0; i < y.size(); i++) yd(i) = (double)y(i);
    const bool ok = std::is_same_v<V, int> ? [&] {
      for (size_t i = 0; i < y.size(); i++)
        if ((double)y(i) != ref.y[i]) return false;
      return true;
    }()
                                           : within(yd, ref);
    if (!ok) std::fprintf(stderr, "  multiply format %d differs\n", Ah.format_index());
    EXPECT_TRUE(ok);
  };
  check_one(retype<V, I>(small_coo()));
  check_one(retype<V, I>(small_csr()));
  check_one(retype<V, I>(small_dia()));
  check_one(retype<V, I>(small_ell()));
  check_one(retype<V, I>(small_hyb()));
  check_one(retype<V, I>(small_hdc()));
}
TEST(Multiply_Types_Double_LongLong) { multiply_typed<double, long long>(); }
TEST(Multiply_Types_Int_LongLong) { multiply_typed<int, long long>(); }
TEST(Multiply_Types_Int_Int) { multiply_typed<int, int>(); }

// core/tests/Serial/Test_Convert_Dynamic.hpp: in-place switches
TEST(Convert_Dynamic_InPlace) {
  for (int s = 0; s < 6; s++) {
    for (int d = 0; d < 6; d++) {
      Dyn<Host> Ah;
      switch (s) {
        case 0: Ah = small_coo(); break;
        case 1: Ah = small_csr(); break;
        case 2: Ah = small_dia(); break;
        case 3: Ah = small_ell(); break;
        case 4: Ah = small_hyb(); break;
        default: Ah = small_hdc(); break;
      }
      Dyn<Dev> Ad = Morpheus::create_mirror_container<Dev>(Ah);
      Morpheus::copy(Ah, Ad);
      EXPECT_EQ(Morpheus::convert<HIPExec>(Ad, d), Morpheus::CONV_SUCCESS);
      EXPECT_EQ(Ad.active_index(), d);
      EXPECT_EQ(Ad.nrows(), 10u);
      auto back = Morpheus::create_mirror(Ad);
      Morpheus::copy(Ad, back);
      bool ok = false;
      std::visit(
          [&](const auto& m) {
            using M = std::decay_t<decltype(m)>;
            if constexpr (M::format_id == 0) ok = same_container(m, small_coo());
            else if constexpr (M::format_id == 1) ok = same_container(m, small_csr());
            else if constexpr (M::format_id == 2) ok = same_container(m, small_dia());
            else if constexpr (M::format_id == 3) ok = same_container(m, small_ell());
            else if constexpr (M::format_id == 4) ok = same_container(m, small_hyb());
            else ok = same_container(m, small_hdc());
          },
          back.const_formats());
      if (!ok) std::fprintf(stderr, "  in-place %d -> %d\n", s, d);
      EXPECT_TRUE(ok);
      // and the same on host containers through Serial
      EXPECT_EQ(Morpheus::convert<Serial>(Ah, d), Morpheus::CONV_SUCCESS);
      EXPECT_EQ(Ah.active_index(), d);
    }
  }
}

// DIA of a matrix whose fill exceeds the tolerance fails, and the in-place
// switch reports it (core/src/Morpheus_Convert.hpp:113-118).
TEST(Convert_Tolerance) {
  const int n = 20000;
  Coo<Host> A(n, n, n);
  for (int i = 0; i < n; i++) {
    A.row_indices(i) = i;
    A.column_indices(i) = (int)((i * 7919ll) % n);
    A.values(i) = 1.0;
  }
  Morpheus::sort_by_row_and_column<Serial>(A);
  EXPECT_TRUE(Morpheus::is_sorted<Serial>(A));
  Coo<Dev> Ad = Morpheus::create_mirror_container<Dev>(A);
  Morpheus::copy(A, Ad);
  Dia<Dev> D;
  EXPECT_THROW(Morpheus::convert<HIPExec>(Ad, D), Morpheus::FormatConversionException);
  Ell<Dev> E;
  Morpheus::convert<HIPExec>(Ad, E);  // one entry per row: admissible
  EXPECT_EQ(E.entries_per_row(), 1u);
  Dyn<Dev> dyn(Ad);
  EXPECT_EQ(Morpheus::convert<HIPExec>(dyn, Morpheus::DIA_FORMAT), Morpheus::PROXY_TO_DYNAMIC);
  EXPECT_EQ(dyn.active_index(), 0);
}

// ============================================================ analytics
// core/tests/Test_MatrixAnalytics.hpp: non-zeros per row on the diagonal
// test matrix and per diagonal on the anti-diagonal one, counted on the
// dense matrix; init and accumulate; min / max / std; count_diagonals; each
// format, on HIPSpace through HIP and on HostSpace through Serial, and
// through a DynamicMatrix.
static Coo<Host> analytics_matrix(int n, bool anti) {
  std::vector<int> diags{0};
  for (int d = 5; d < n; d += 5) {
    diags.push_back(d);
    diags.push_back(-d);
  }
  std::vector<double> dense((size_t)n * n, 0.0);
  std::mt19937_64 rng(n + (anti ? 1 : 0));
  std::uniform_real_distribution<double> u(0.5, 2.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int d : diags)
        if (std::abs(j - i) == std::abs(d)) dense[(size_t)i * n + (anti ? n - j - 1 : j)] = u(rng);
  size_t nnz = 0;
  for (double v : dense) nnz += v != 0;
  Coo<Host> A(n, n, nnz);
  size_t k = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      if (dense[(size_t)i * n + j] != 0) {
        A.row_indices(k) = i;
        A.column_indices(k) = j;
        A.values(k) = dense[(size_t)i * n + j];
        k++;
      }
  return A;
}

template <int F>
static void analytics_format(int n) {
  using IVecH = Morpheus::DenseVector<int, Host>;
  using IVecD = Morpheus::DenseVector<int, Dev>;
  const Coo<Host> R = analytics_matrix(n, false), Q = analytics_matrix(n, true);
  std::vector<int> rows(n, 0), diag(2 * n - 1, 0);
  for (size_t k = 0; k < R.nnnz(); k++) rows[R.crow_indices(k)]++;
  for (size_t k = 0; k < Q.nnnz(); k++) diag[Q.ccolumn_indices(k) - Q.crow_indices(k) + n - 1]++;
  const int mx = *std::max_element(rows.begin(), rows.end());
  const int mn = *std::min_element(rows.begin(), rows.end());
  double sq = 0;
  const double mean = (double)(R.nnnz() / (size_t)n);
  for (int r : rows) sq += (r - mean) * (r - mean);
  const size_t sd = (size_t)std::sqrt(sq / n);
  size_t ndiag = 0, ntrue = 0;
  for (int d : diag) {
    ndiag += d > 0;
    ntrue += d > 3;
  }

  using M = std::decay_t<decltype(small_of<F>())>;
  M Rh, Qh;
  Morpheus::convert<Serial>(R, Rh);
  Morpheus::convert<Serial>(Q, Qh);
  auto Rd = Morpheus::create_mirror_container<Dev>(Rh);
  auto Qd = Morpheus::create_mirror_container<Dev>(Qh);
  Morpheus::copy(Rh, Rd);
  Morpheus::copy(Qh, Qd);

  IVecD cd(n, 1);
  Morpheus::count_nnz_per_row<HIPExec>(Rd, cd);
  IVecH ch(n, 0);
  Morpheus::copy(cd, ch);
  bool ok = true;
  for (int i = 0; i < n; i++) ok = ok && ch(i) == rows[i];
  EXPECT_TRUE(ok);
  IVecH acc(n, 1);
  Morpheus::count_nnz_per_row<Serial>(Rh, acc, false);
  ok = true;
  for (int i = 0; i < n; i++) ok = ok && acc(i) == rows[i] + 1;
  EXPECT_TRUE(ok);
  EXPECT_EQ(Morpheus::max_nnnz<HIPExec>(Rd), (size_t)mx);
  EXPECT_EQ(Morpheus::min_nnnz<Serial>(Rh), (size_t)mn);
  EXPECT_EQ(Morpheus::std_nnnz<HIPExec>(Rd), sd);
  EXPECT_EQ(Morpheus::max<HIPExec>(cd, cd.size()), mx);

  IVecD dd(2 * n - 1, 0);
  Morpheus::count_nnz_per_diagonal<HIPExec>(Qd, dd);
  IVecH dh(2 * n - 1, 0);
  Morpheus::copy(dd, dh);
  ok = true;
  for (int i = 0; i < 2 * n - 1; i++) ok = ok && dh(i) == diag[i];
  EXPECT_TRUE(ok);
  EXPECT_EQ(Morpheus::count_diagonals<HIPExec>(Qd), ndiag);
  EXPECT_EQ(Morpheus::count_true_diagonals<Serial>(Qh, 3), ntrue);
  EXPECT_EQ(Morpheus::count_nnz<HIPExec>(dd), ndiag);

  Dyn<Dev> dyn(Qd);
  EXPECT_EQ(Morpheus::count_diagonals<HIPExec>(dyn), ndiag);
  EXPECT_EQ(Morpheus::number_of_nnz(Rh), R.nnnz());
  EXPECT_TRUE(std::abs(Morpheus::average_nnnz(Rd) - R.nnnz() / (double)n) < 1e-12);
}

TEST(Analytics_AllFormats) {
  for (int n : {10, 100, 1000}) {
    analytics_format<0>(n);
    analytics_format<1>(n);
    analytics_format<2>(n);
    analytics_format<3>(n);
    analytics_format<4>(n);
    analytics_format<5>(n);
  }
}

// core/tests/Test_Dot.hpp, Test_WAXPBY.hpp, Test_Reduction.hpp: device and
// host vectors against the oracle's Serial loops
TEST(DenseVector_Algorithms) {
  const int n = 100000;
  Vec<Host> x = random_x(n, 5), y = random_x(n, 6), w(n, 0.0);
  Vec<Dev> xd(n), yd(n), wd(n);
  Morpheus::copy(x, xd);
  Morpheus::copy(y, yd);
  const double ref = orc_dot(n, x.data(), y.data());
  double absdot = 0;
  for (int i = 0; i < n; i++) absdot += std::abs(x(i) * y(i));
  EXPECT_TRUE(std::abs(Morpheus::dot<HIPExec>(n, xd, yd) - ref) <= 4.0 * n * 1.2e-16 * absdot);
  EXPECT_TRUE(std::abs(Morpheus::dot<Serial>(n, x, y) - ref) <= 4.0 * n * 1.2e-16 * absdot);
  std::vector<double> wref(n);
  orc_waxpby(n, 2.5, x.data(), -0.5, y.data(), wref.data());
  Morpheus::waxpby<HIPExec>(n, 2.5, xd, -0.5, yd, wd);
  Morpheus::copy(wd, w);
  bool same = true;
  for (int i = 0; i < n; i++) same = same && w(i) == wref[i];
  EXPECT_TRUE(same);
  Vec<Host> wh(n, 0.0);
  Morpheus::waxpby<Serial>(n, 2.5, x, -0.5, y, wh);
  same = true;
  for (int i = 0; i < n; i++) same = same && wh(i) == wref[i];
  EXPECT_TRUE(same);
  double absx = 0;
  for (int i = 0; i < n; i++) absx += std::abs(x(i));
  EXPECT_TRUE(std::abs(Morpheus::reduce<HIPExec>(xd, n) - orc_reduce(n, x.data())) <= 4e-16 * absx);
}

// core/tests/Test_MatrixOperations.hpp: update_diagonal in every format, on
// HIPSpace through HIP and on HostSpace through Serial, then the dense
// diagonal is the new one; a product afterwards sees the new values
TEST(MatrixOperations_UpdateDiagonal) {
  const int n = 200;
  Coo<Host> R = analytics_matrix(n, false);
  Vec<Host> d = random_x(n, 17);
  Vec<Dev> dd(n);
  Morpheus::copy(d, dd);
  auto check_diag = [&](const Coo<Host>& C) {
    bool ok = true;
    for (size_t k = 0; k < C.nnnz(); k++)
      if (C.crow_indices(k) == C.ccolumn_indices(k)) ok = ok && C.cvalues(k) == d(C.crow_indices(k));
    return ok;
  };
  for_each_small([&](const auto& proto) {
    using M = std::decay_t<decltype(proto)>;
    M Ah;
    Morpheus::convert<Serial>(R, Ah);
    auto Ad = Morpheus::create_mirror_container<Dev>(Ah);
    Morpheus::copy(Ah, Ad);
    Morpheus::update_diagonal<HIPExec>(Ad, dd);
    M back;
    back = Morpheus::create_mirror(Ad);
    Morpheus::copy(Ad, back);
    Coo<Host> C;
    Morpheus::convert<Serial>(back, C);
    EXPECT_TRUE(check_diag(C));
    Morpheus::update_diagonal<Serial>(Ah, d);
    Coo<Host> C2;
    Morpheus::convert<Serial>(Ah, C2);
    EXPECT_TRUE(check_diag(C2));
    Vec<Dev> gd(n, 0.0);
    Morpheus::get_diagonal<HIPExec>(Ad, gd);
    Vec<Host> gh(n, 0.0);
    Morpheus::copy(gd, gh);
    bool ok = true;
    for (int i = 0; i < n; i++) ok = ok && gh(i) == d(i);
    EXPECT_TRUE(ok);
    // a product on the host container after the update sees the new values
    Vec<Host> x = random_x(n, 3), y(n, 0.0);
    Morpheus::multiply<Serial>(Ah, x, y);
    std::vector<double> yref(n);
    std::vector<int> ri(C2.nnnz()), ci(C2.nnnz());
    for (size_t k = 0; k < C2.nnnz(); k++) {
      ri[k] = C2.crow_indices(k);
      ci[k] = C2.ccolumn_indices(k);
    }
    orc_coo_spmv(n, C2.nnnz(), ri.data(), ci.data(), C2.values().data(), x.data(), yref.data(), 1);
    ok = true;
    for (int i = 0; i < n; i++) ok = ok && std::abs(y(i) - yref[i]) <= 1e-12 * (1 + std::abs(yref[i]));
    EXPECT_TRUE(ok);
  });
}

// ============================================================ larger cases vs the oracle
// A random matrix with ragged rows (one long row), converted on the device
// from CSR to every format and multiplied: conversions are compared with the
// oracle's Serial converters bit for bit, products with the FP64 bound.
TEST(Random_AllFormats_vs_Oracle) {
  const int n = 3000, m = 2500;
  std::mt19937_64 rng(20240601);
  std::vector<int> ro(n + 1, 0);
  std::vector<int> ci;
  std::vector<double> va;
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < n; i++) {
    int len = (i == 17) ? 2400 : (int)(rng() % 24);
    std::vector<int> cols(m);
    for (int j = 0; j < m; j++) cols[j] = j;
    for (int k = 0; k < len; k++) std::swap(cols[k], cols[k + rng() % (m - k)]);
    std::sort(cols.begin(), cols.begin() + len);
    for (int k = 0; k < len; k++) {
      ci.push_back(cols[k]);
      va.push_back(nd(rng));
    }
    ro[i + 1] = (int)ci.size();
  }
  const int nnz = (int)ci.size();
  Csr<Host> A(n, m, nnz);
  std::copy(ro.begin(), ro.end(), A.row_offsets().data());
  std::copy(ci.begin(), ci.end(), A.column_indices().data());
  std::copy(va.begin(), va.end(), A.values().data());
  Csr<Dev> Ad = Morpheus::create_mirror_container<Dev>(A);
  Morpheus::copy(A, Ad);

  // oracle products
  auto x = random_x(m, 99);
  std::vector<double> yref(n), absr(n);
  orc_csr_spmv(n, ro.data(), ci.data(), va.data(), x.data(), yref.data(), 1);
  orc_csr_absrow(n, ro.data(), ci.data(), va.data(), x.data(), absr.data());
  Vec<Dev> xd(m), yd(n);
  Morpheus::copy(x, xd);

  // oracle COO
  std::vector<int> ri(nnz), cc(nnz);
  std::vector<double> cv(nnz);
  orc_csr_to_coo(n, nnz, ro.data(), ci.data(), va.data(), ri.data(), cc.data(), cv.data());

  auto check_y = [&](const char* what) {
    Vec<Host> yh(n);
    Morpheus::copy(yd, yh);
    bool ok = true;
    for (int i = 0; i < n && ok; i++) {
      const double b = 4.0 * (ro[i + 1] - ro[i] + 1) * 0x1p-53 * absr[i] + 1e-300;
      if (!(std::fabs(yh(i) - yref[i]) <= b)) {
        std::fprintf(stderr, "  %s row %d: %.17g vs %.17g\n", what, i, yh(i), yref[i]);
        ok = false;
      }
    }
    EXPECT_TRUE(ok);
  };

  Morpheus::multiply<HIPExec>(Ad, xd, yd);
  check_y("CSR");

  Coo<Dev> Cd;
  Morpheus::convert<HIPExec>(Ad, Cd);
  {
    auto Ch = Morpheus::create_mirror(Cd);
    Morpheus::copy(Cd, Ch);
    EXPECT_TRUE(std::memcmp(Ch.crow_indices().data(), ri.data(), 4 * nnz) == 0);
    EXPECT_TRUE(std::memcmp(Ch.cvalues().data(), cv.data(), 8 * nnz) == 0);
    Morpheus::multiply<HIPExec>(Cd, xd, yd);
    check_y("COO");
  }
  {  // ELL: K = 2400 rows of padding is admissible at this size
    Ell<Dev> Ed;
    Morpheus::convert<HIPExec>(Ad, Ed);
    auto Eh = Morpheus::create_mirror(Ed);
    Morpheus::copy(Ed, Eh);
    const int64_t K = orc_coo_ell_entries_per_row(n, nnz, ri.data());
    EXPECT_EQ((int64_t)Eh.entries_per_row(), K);
    const int64_t w = orc_pad_size(K, 32);
    std::vector<int> eci(w * n);
    std::vector<double> ev(w * n);
    orc_coo_to_ell(n, nnz, ri.data(), cc.data(), cv.data(), K, w, n, eci.data(), ev.data());
    bool ok = true;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < w; j++)
        ok = ok && Eh.ccolumn_indices(i, j) == eci[i + (size_t)n * j] &&
             std::memcmp(&Eh.cvalues(i, j), &ev[i + (size_t)n * j], 8) == 0;
    EXPECT_TRUE(ok);
    Morpheus::multiply<HIPExec>(Ed, xd, yd);
    check_y("ELL");
  }
  {
    Hyb<Dev> Hd;
    Morpheus::convert<HIPExec>(Ad, Hd);
    int64_t K = 0, enz = 0, cnz = 0;
    orc_coo_hyb_sizes(n, nnz, ri.data(), &K, &enz, &cnz);
    EXPECT_EQ((int64_t)Hd.cell().entries_per_row(), K);
    EXPECT_EQ((int64_t)Hd.ccoo().nnnz(), cnz);
    Morpheus::multiply<HIPExec>(Hd, xd, yd);
    check_y("HYB");
  }
  {
    Hdc<Dev> Qd;
    Morpheus::convert<HIPExec>(Ad, Qd);
    int64_t nd = 0, dnz = 0, cnz = 0;
    orc_coo_hdc_sizes(n, m, nnz, ri.data(), cc.data(), &nd, &dnz, &cnz);
    EXPECT_EQ((int64_t)Qd.cdia().ndiags(), nd);
    EXPECT_EQ((int64_t)Qd.ccsr().nnnz(), cnz);
    Morpheus::multiply<HIPExec>(Qd, xd, yd);
    check_y("HDC");
    // HDC -> CSR round trip reproduces A when DIA holds no zeros
    Csr<Dev> back;
    Morpheus::convert<HIPExec>(Qd, back);
    auto bh = Morpheus::create_mirror(back);
    Morpheus::copy(back, bh);
    EXPECT_EQ(bh.nnnz(), (size_t)nnz);
  }
  {  // the launch-shape autotuner keeps results within the bound
    Morpheus::tune_multiply(Ad, xd, yd, 3);
    Morpheus::multiply<HIPExec>(Ad, xd, yd);
    check_y("CSR tuned");
    Dyn<Dev> dyn(Ad);
    const Morpheus::formats_e f = Morpheus::select_format<HIPExec>(dyn, 3);
    EXPECT_EQ(dyn.active_enum(), f);
    Morpheus::multiply<HIPExec>(dyn, xd, yd);
    check_y("Dynamic selected");
  }
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <fixture.txt> [filter]\n", argv[0]);
    return 2;
  }
  Morpheus::InitArguments args;
  Morpheus::initialize(argc, argv, args);
  load_fixture(argv[1]);
  const char* filter = argc > 2 ? argv[2] : nullptr;
  int failed_tests = 0, ran = 0;
  for (auto& t : registry()) {
    if (filter && !std::strstr(t.name, filter)) continue;
    std::printf("[ RUN      ] %s\n", t.name);
    std::fflush(stdout);
    const int before = g_failures;
    try {
      t.fn();
    } catch (const std::exception& e) {
      std::fprintf(stderr, "  unexpected exception: %s\n", e.what());
      g_failures++;
    }
    ran++;
    const bool ok = g_failures == before;
    if (!ok) failed_tests++;
    std::printf("[ %s ] %s\n", ok ? "      OK" : " FAILED ", t.name);
    std::fflush(stdout);
  }
  Morpheus::finalize();
  std::printf("%d tests, %d checks, %d failed tests\n", ran, g_checks, failed_tests);
  return failed_tests ? 1 : 0;
}
