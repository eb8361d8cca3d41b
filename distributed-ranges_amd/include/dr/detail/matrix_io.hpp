// dr/detail/matrix_io.hpp -- COO -> CSR conversion of sorted (row, col)
// tuples, the interface of include/dr/detail/matrix_io.hpp:26-72
// (dr::__detail::convert_to_csr(tuples, shape, nnz, allocator)).
#pragma once

#include <cstddef>
#include <memory>
#include <type_traits>

#include "dr/views/csr_matrix_view.hpp"

namespace dr::__detail {

// `tuples`: ((row, col), value) entries sorted by (row, col); the arrays are
// allocated with `allocator` (rebound for the index type).
template <typename Tuples, typename Allocator>
auto convert_to_csr(Tuples &&tuples, dr::index<> shape, std::size_t nnz,
                    Allocator &&allocator) {
  using entry = std::remove_cvref_t<decltype(*std::begin(tuples))>;
  using T = std::remove_cvref_t<decltype(std::declval<entry>().second)>;
  using I = std::remove_cvref_t<decltype(std::declval<entry>().first.first)>;
  typename std::allocator_traits<std::remove_cvref_t<Allocator>>::template rebind_alloc<I>
      ialloc(allocator);
  T *values = allocator.allocate(nnz);
  I *rowptr = ialloc.allocate(shape[0] + 1);
  I *colind = ialloc.allocate(nnz);
  std::size_t row = 0, q = 0;
  rowptr[0] = 0;
  for (auto &&[ij, v] : tuples) {
    values[q] = v;
    colind[q] = ij.second;
    for (; row < static_cast<std::size_t>(ij.first); row++) rowptr[row + 1] = static_cast<I>(q);
    q++;
  }
  for (; row < shape[0]; row++) rowptr[row + 1] = static_cast<I>(nnz);
  return dr::views::csr_matrix_view<T, I>(values, rowptr, colind,
                                          dr::index<I>(shape[0], shape[1]), nnz, 0);
}

}  // namespace dr::__detail
