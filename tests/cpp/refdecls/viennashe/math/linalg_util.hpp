// Test declarations under ViennaSHE's own path and names: the matrix type of
// viennashe/math/linalg_util.hpp:67-196 (square, one std::map per row, the
// non-const operator() inserting a zero entry when absent).  Used only to
// compile the GPU solver overload standalone.
#pragma once

#include <cassert>
#include <cstddef>
#include <map>
#include <vector>

namespace viennashe {
namespace math {

template <typename NumericT>
class sparse_matrix {
 public:
  typedef std::size_t size_type;
  typedef std::map<size_type, NumericT> row_type;
  typedef typename row_type::iterator iterator2;
  typedef typename row_type::const_iterator const_iterator2;

  sparse_matrix() {}
  sparse_matrix(size_type rows, size_type cols) : rows_(rows) {
    assert(rows == cols && "Only square matrices supported.");
    (void)cols;
  }
  size_type size1() const { return rows_.size(); }
  size_type size2() const { return rows_.size(); }
  NumericT &operator()(size_type i, size_type j) { return rows_.at(i)[j]; }
  NumericT operator()(size_type i, size_type j) const {
    const_iterator2 it = rows_.at(i).find(j);
    return it == rows_.at(i).end() ? NumericT(0) : it->second;
  }
  row_type const &row(size_type i) const { return rows_.at(i); }
  row_type &row(size_type i) { return rows_.at(i); }
  size_type nnz() const {
    size_type s = 0;
    for (size_type i = 0; i < rows_.size(); ++i) s += rows_[i].size();
    return s;
  }

 private:
  std::vector<row_type> rows_;
};

}  // namespace math
}  // namespace viennashe
