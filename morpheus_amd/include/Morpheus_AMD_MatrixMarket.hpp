// Morpheus_AMD_MatrixMarket.hpp -- MatrixMarket I/O of host containers, the
// input side of the SpMV path: Morpheus::IO::read_matrix_market_{file,stream}
// and write_matrix_market_{file,stream} for CooMatrix, DynamicMatrix (active
// COO), DenseMatrix and DenseVector in HostSpace
// (core/src/Morpheus_MatrixMarket.hpp:62-209,
// core/src/impl/Morpheus_MatrixMarket_Impl.hpp:42-496).
//
// Parsing is host work by nature, so it runs on the host; what it produces
// is what the reference produces:
//  * coordinate files: 1-based indices checked against the size line, then
//    made 0-based; "symmetric" mirrors every off-diagonal entry right after
//    it; "pattern" values are 1; the entries end stably sorted by (row,
//    column), as sort_by_row_and_column leaves them;
//  * array files fill a DenseMatrix in row-major order (row by row, as the
//    reference reads them) and a DenseVector in file order;
//  * complex data, hermitian and skew-symmetric matrices, pattern arrays and
//    non-general arrays raise NotImplementedException; malformed banners,
//    sizes, indices and early ends raise IOException, with the reference's
//    messages;
//  * writers print `coordinate real general` / `array real general` with
//    1-based indices and the stream's own number formatting.
// One divergence: the symmetry "unsymmetric", which the reference's banner
// check admits but its reader turns into a zero-filled matrix, is read as
// "general" here.
#pragma once

#include <algorithm>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "Morpheus_AMD_Algorithms.hpp"

namespace Morpheus {
namespace IO {
namespace Impl {

inline std::vector<std::string> tokens_of(const std::string& s) {
  std::vector<std::string> t;
  const char* ws = "\n\r\t ";
  std::string::size_type b = s.find_first_not_of(ws), e;
  while (b != std::string::npos) {
    e = s.find_first_of(ws, b);
    t.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
    b = e == std::string::npos ? e : s.find_first_not_of(ws, e);
  }
  return t;
}

struct mm_banner {
  std::string storage, type, symmetry;
};

template <typename Stream>
mm_banner read_banner(Stream& input) {
  std::string line;
  std::getline(input, line);
  const auto t = tokens_of(line);
  if (t.size() != 5 || t[0] != "%%MatrixMarket" || t[1] != "matrix")
    throw IOException("invalid MatrixMarket banner");
  mm_banner b{t[2], t[3], t[4]};
  if (b.storage != "array" && b.storage != "coordinate")
    throw IOException("invalid MatrixMarket storage format [" + b.storage + "]");
  if (b.type != "complex" && b.type != "real" && b.type != "integer" && b.type != "pattern")
    throw IOException("invalid MatrixMarket data type [" + b.type + "]");
  if (b.symmetry != "general" && b.symmetry != "symmetric" && b.symmetry != "unsymmetric" &&
      b.symmetry != "hermitian" && b.symmetry != "skew-symmetric")
    throw IOException("invalid MatrixMarket symmetry [" + b.symmetry + "]");
  return b;
}

// the first line after the banner that is not a comment
template <typename Stream>
std::string size_line(Stream& input) {
  std::string line;
  do {
    if (!std::getline(input, line)) {
      line.clear();
      break;
    }
  } while (!line.empty() && line[0] == '%');
  return line;
}

template <class Coo, typename Stream>
void read_coordinate(Coo& coo, Stream& input, const mm_banner& b) {
  using T = typename Coo::non_const_value_type;
  const auto t = tokens_of(size_line(input));
  if (t.size() != 3) throw IOException("invalid MatrixMarket coordinate format");
  size_t nrows = 0, ncols = 0, nnz = 0;
  std::istringstream(t[0]) >> nrows;
  std::istringstream(t[1]) >> ncols;
  std::istringstream(t[2]) >> nnz;
  std::vector<long long> r(nnz, 0), c(nnz, 0);
  std::vector<T> v(nnz, T(1));
  size_t nread = 0;
  if (b.type == "pattern") {
    while (nread < nnz && !input.eof()) {
      input >> r[nread];
      input >> c[nread];
      nread++;
    }
  } else if (b.type == "real" || b.type == "integer") {
    while (nread < nnz && !input.eof()) {
      T value{};
      input >> r[nread];
      input >> c[nread];
      input >> value;
      v[nread] = value;
      nread++;
    }
  } else if (b.type == "complex") {
    throw NotImplementedException("Morpheus does not currently support complex matrices");
  } else {
    throw IOException("invalid MatrixMarket data type");
  }
  if (nread != nnz) throw IOException("unexpected EOF while reading MatrixMarket entries");
  if (nnz > 0) {
    const long long rmin = *std::min_element(r.begin(), r.end());
    const long long rmax = *std::max_element(r.begin(), r.end());
    const long long cmin = *std::min_element(c.begin(), c.end());
    const long long cmax = *std::max_element(c.begin(), c.end());
    if (rmin < 1) throw IOException("found invalid row index (index < 1)");
    if (cmin < 1) throw IOException("found invalid column index (index < 1)");
    if (rmax > (long long)nrows) throw IOException("found invalid row index (index > num_rows)");
    if (cmax > (long long)ncols)
      throw IOException("found invalid column index (index > num_columns)");
  }
  for (size_t n = 0; n < nnz; n++) {
    r[n] -= 1;
    c[n] -= 1;
  }
  if (b.symmetry == "hermitian")
    throw NotImplementedException("MatrixMarket I/O does not currently support hermitian matrices");
  if (b.symmetry == "skew-symmetric")
    throw NotImplementedException(
        "MatrixMarket I/O does not currently support skew-symmetric matrices");
  if (b.symmetry == "symmetric") {
    std::vector<long long> r2, c2;
    std::vector<T> v2;
    r2.reserve(2 * nnz);
    c2.reserve(2 * nnz);
    v2.reserve(2 * nnz);
    for (size_t n = 0; n < nnz; n++) {
      r2.push_back(r[n]);
      c2.push_back(c[n]);
      v2.push_back(v[n]);
      if (r[n] != c[n]) {
        r2.push_back(c[n]);
        c2.push_back(r[n]);
        v2.push_back(v[n]);
      }
    }
    r.swap(r2);
    c.swap(c2);
    v.swap(v2);
  }
  // stable (row, column) order
  std::vector<size_t> perm(r.size());
  std::iota(perm.begin(), perm.end(), size_t(0));
  std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b2) {
    return r[a] != r[b2] ? r[a] < r[b2] : c[a] < c[b2];
  });
  Coo out(nrows, ncols, r.size());
  for (size_t n = 0; n < perm.size(); n++) {
    out.row_indices(n) = (int)r[perm[n]];
    out.column_indices(n) = (int)c[perm[n]];
    out.values(n) = v[perm[n]];
  }
  coo = out;
}

template <class Dense, typename Stream>
void read_array(Dense& mtx, Stream& input, const mm_banner& b) {
  using T = typename Dense::non_const_value_type;
  const auto t = tokens_of(size_line(input));
  if (t.size() != 2) throw IOException("invalid MatrixMarket array format");
  size_t nrows = 0, ncols = 0;
  std::istringstream(t[0]) >> nrows;
  std::istringstream(t[1]) >> ncols;
  Dense dense(nrows, ncols);
  const size_t total = nrows * ncols;
  size_t nread = 0, i = 0, j = 0;
  if (b.type == "pattern") {
    throw NotImplementedException("pattern array MatrixMarket format is not supported");
  } else if (b.type == "real" || b.type == "integer") {
    while (nread < total && !input.eof()) {
      double value = 0;
      input >> value;
      dense(i, j) = T(value);
      nread++;
      if (++j == ncols) {
        i++;
        j = 0;
      }
    }
  } else if (b.type == "complex") {
    throw NotImplementedException("complex type is not supported");
  } else {
    throw IOException("invalid MatrixMarket data type");
  }
  if (nread != total) throw IOException("unexpected EOF while reading MatrixMarket entries");
  if (b.symmetry != "general")
    throw NotImplementedException("only general array symmetric MatrixMarket format is supported");
  mtx = dense;
}

template <class C>
constexpr void require_host() {
  static_assert(!C::memory_space::is_device, "Container must reside in HostSpace");
}

template <class Coo, typename Stream>
void write_coordinate(const Coo& coo, Stream& output) {
  output << "%%MatrixMarket matrix coordinate real general\n";
  output << coo.nrows() << "\t" << coo.ncols() << "\t" << coo.nnnz() << "\n";
  for (size_t i = 0; i < coo.nnnz(); i++) {
    output << (coo.crow_indices(i) + 1) << " ";
    output << (coo.ccolumn_indices(i) + 1) << " ";
    output << coo.cvalues(i);
    output << "\n";
  }
}

}  // namespace Impl

template <typename Container, typename Stream>
void read_matrix_market_stream(Container& container, Stream& input) {
  Impl::require_host<Container>();
  const Impl::mm_banner b = Impl::read_banner(input);
  if constexpr (is_coo_matrix_format_container_v<Container>) {
    if (b.storage != "coordinate")
      throw IOException(
          "only coordinate storage format is supported for reading a MatrixMarket file into "
          "CooMatrix container");
    Impl::read_coordinate(container, input, b);
  } else if constexpr (is_dynamic_matrix_format_container_v<Container>) {
    if (b.storage != "coordinate")
      throw IOException(
          "only coordinate storage format is supported for reading a MatrixMarket file into "
          "CooMatrix container");
    if (container.active_index() != COO_FORMAT)
      throw RuntimeException(
          "read_matrix_market_stream: The active state of a DynamicMatrix must be set to "
          "Morpheus::COO_FORMAT before reading a MatrixMarket stream!");
    CooMatrix<typename Container::value_type, HostSpace> temp;
    Impl::read_coordinate(temp, input, b);
    container = temp;
  } else if constexpr (is_dense_matrix_format_container_v<Container>) {
    if (b.storage != "array")
      throw IOException(
          "only coordinate storage format is supported for reading a MatrixMarket file into "
          "DenseMatrix container");
    Impl::read_array(container, input, b);
  } else {
    static_assert(is_dense_vector_format_container_v<Container>,
                  "Container must be either a CooMatrix, DynamicMatrix, DenseMatrix or "
                  "DenseVector container");
    if (b.storage != "array")
      throw IOException(
          "only coordinate storage format is supported for reading a MatrixMarket file into "
          "DenseVector container");
    DenseMatrix<typename Container::non_const_value_type, HostSpace> temp;
    Impl::read_array(temp, input, b);
    container.resize(temp.nrows() * temp.ncols());
    for (size_t i = 0; i < temp.nrows(); i++)
      for (size_t j = 0; j < temp.ncols(); j++) container(i * temp.ncols() + j) = temp(i, j);
  }
}

template <typename Container>
void read_matrix_market_file(Container& container, const std::string& filename) {
  Impl::require_host<Container>();
  std::ifstream file(filename.c_str());
  if (!file)
    throw IOException(std::string("unable to open file \"") + filename + "\" for reading");
  read_matrix_market_stream(container, file);
}

template <typename Container, typename Stream>
void write_matrix_market_stream(const Container& container, Stream& output) {
  Impl::require_host<Container>();
  if constexpr (is_coo_matrix_format_container_v<Container>) {
    Impl::write_coordinate(container, output);
  } else if constexpr (is_dynamic_matrix_format_container_v<Container>) {
    if (container.active_index() != COO_FORMAT)
      throw RuntimeException(
          "write_matrix_market_stream: The active state of a DynamicMatrix must be set to "
          "Morpheus::COO_FORMAT before writing a MatrixMarket stream!");
    std::visit(
        [&](const auto& m) {
          if constexpr (std::decay_t<decltype(m)>::format_id == COO_FORMAT)
            Impl::write_coordinate(m, output);
        },
        container.const_formats());
  } else if constexpr (is_dense_matrix_format_container_v<Container>) {
    output << "%%MatrixMarket matrix array real general\n";
    output << container.nrows() << "\t" << container.ncols() << "\n";
    for (size_t i = 0; i < container.nrows(); i++)
      for (size_t j = 0; j < container.ncols(); j++) {
        output << container(i, j);
        output << "\n";
      }
  } else {
    static_assert(is_dense_vector_format_container_v<Container>,
                  "Container must be either a CooMatrix, DynamicMatrix, DenseMatrix or "
                  "DenseVector container");
    output << "%%MatrixMarket matrix array real general\n";
    output << container.size() << "\t1\n";
    for (size_t i = 0; i < container.size(); i++) {
      output << container[i];
      output << "\n";
    }
  }
}

template <typename Container>
void write_matrix_market_file(const Container& container, const std::string& filename) {
  Impl::require_host<Container>();
  std::ofstream file(filename.c_str());
  if (!file)
    throw IOException(std::string("unable to open file \"") + filename + "\" for writing");
  write_matrix_market_stream(container, file);
}

}  // namespace IO
}  // namespace Morpheus
